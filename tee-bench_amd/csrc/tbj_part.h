// Partition-function and work-descriptor types shared by host and device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbj {

// One contiguous input range partitioned independently (a relation in the
// first pass, a first-pass partition of a relation in the second pass).
//
// The histogram keeps one row of bin counts per tile (row t = tile t of the
// pass, so every tile writes one contiguous row), and the output offsets
// come from a scan over group sums.  Group g of a pass is tiles
// [g * kGroupTiles, (g + 1) * kGroupTiles); a segment owns the ngroups
// groups its tiles touch, and its group sums sit bin-major from hist_base:
// (bin b, group g) at hist_base + b * ngroups + g - first_tile /
// kGroupTiles.  Scanned in that order, the entry of bin b's first group is
// where partition (segment, b) starts.  (8-byte and 16-byte tuples alike.)
struct SegDesc {
    uint64_t hist_base;
    uint32_t in_off;      // first input tuple
    uint32_t n;           // tuples
    uint32_t ntiles;      // tiles of the segment
    uint32_t first_tile;  // global index of its first tile
    uint32_t ngroups;
    uint32_t aux;         // sample-sized second pass: the segment's first-pass
                          // partition (relation * F1 + bin); 0 otherwise
};

constexpr uint32_t kGroupTiles = 32;

// Groups touched by tiles [first_tile, first_tile + ntiles).
inline __host__ __device__ uint32_t seg_groups(uint32_t first_tile, uint32_t ntiles) {
    return ntiles ? (first_tile + ntiles - 1) / kGroupTiles - first_tile / kGroupTiles + 1 : 0;
}

struct TileDesc {
    uint32_t in_off;      // first input tuple of the tile
    uint32_t len;         // tuples in the tile
    uint32_t seg;         // owning segment
    uint32_t aux;         // sample-sized passes: first pass, the tile's row
                          // (tile_row); second pass, its segment's aux
};

// Sample-sized passes.  A first-pass launch hands its n tiles to the 8 XCDs
// in contiguous eighths (xcd_tile); each eighth is cut into rows of
// kRowTiles tiles, and every (bin, row) of a relation owns its own output
// region, so that a region is written by one XCD.  A row's regions average
// kRowTiles * kTile / F1 tuples (16 tiles at 256 bins), so that the second
// pass, which reads them region by region, sees mostly full tiles (one
// partial tile per region; 2048-tile rows ran C2 0.3 % and C3 1 % slower).  The
// sample is counted in chunks of kSampleTiles tiles of a row (the 16-bit
// LDS counters of one chunk cannot overflow).
constexpr uint32_t kRowTiles = 4096;
constexpr uint32_t kSampleTiles = 224;
constexpr uint32_t kChunksPerRow = (kRowTiles + kSampleTiles - 1) / kSampleTiles;

// The sampled 16-tuple group of a 256-tuple block (one in 16, at a hashed
// position).
inline __host__ __device__ uint32_t sampled_group(uint32_t block) {
    return (block * 2654435761u) >> 28;
}

inline __host__ __device__ uint32_t eighth_start(uint32_t x, uint32_t n) {
    const uint32_t q = n / 8, r = n % 8;
    return x * q + (x < r ? x : r);
}

inline __host__ __device__ uint32_t rows_per_eighth(uint32_t n) {
    const uint32_t longest = n / 8 + (n % 8 ? 1 : 0);
    return (longest + kRowTiles - 1) / kRowTiles;
}

// Row of tile u of a launch of n tiles: eighth x (the first n % 8 eighths hold
// n / 8 + 1 tiles, the others n / 8), then the kRowTiles-tile row inside it.
inline __host__ __device__ uint32_t tile_row(uint32_t u, uint32_t n) {
    const uint32_t q = n / 8, r = n % 8;
    const uint32_t x = u < r * (q + 1) ? u / (q + 1) : r + (u - r * (q + 1)) / q;
    return x * rows_per_eighth(n) + (u - eighth_start(x, n)) / kRowTiles;
}

// One build/probe task: an R range (at most the LDS capacity) against an S range.
struct JoinTask {
    uint32_t r_off, r_len, s_off, s_len;
};

// Where the join kernels find their tasks: side rel (0 = R, 1 = S) of
// partition p is the npieces ranges bounds[(rel * nparts + p) * npieces + x]
// (one range, or the per-XCD pieces of sample-sized regions); its R ranges
// are cut into builds of at most cap tuples and its S ranges into chunks of
// at most chunk tuples, and every (build, chunk) pair is a task.  Partition
// p holds tasks [task_off[p], task_off[p + 1]), *ntasks_dev in all; the first
// task_cap of them are also listed in tasks.
struct TaskSrc {
    JoinTask* tasks;
    const uint32_t* task_off;
    const uint2* bounds;
    const uint32_t* ntasks_dev;
    uint32_t task_cap;  // entries of tasks
    uint32_t nparts;
    uint32_t cap;       // R tuples per build
    uint32_t chunk;     // S tuples per task
    uint32_t npieces;   // ranges per partition side (1 or 8)
};

// Piecewise-linear key-CDF model (learned range partitioner).  Knot i
// (key x[i]) has cumulative value y[i] in units of 2^-frac partitions; the
// knots sit at equal-mass sample quantiles between R's exact minimum kmin
// and maximum kmax (x[nseg] = kmax + 1, y[nseg] = 2^32).  The partition
// function reads the model through an equal-width cell table: cells[c] is
// the model at key kmin + (c << cshift) in units of 1/256 partition, and a
// key interpolates linearly (integer arithmetic) between the values at its
// cell's two ends, so the map is monotone: partitions are key ranges in key
// order.  Keys outside [kmin, kmax] cannot match R and are dropped from S.
//
// A sample whose CDF is near-linear (sample_is_linear) gives a linear model
// instead: no knots and no cells, the key range [kmin, kmax] (the sample's,
// widened by 8 expected gaps per side) mapped straight onto the partitions,
// partition = ((key - kmin) * lmul) >> 32 with lmul = floor(2^(32 + bits) /
// (kmax - kmin + 1)): still key ranges in key order, at the cost of one
// multiply.
constexpr int kMaxKnots = 4096;
constexpr int kCells = 4096;
struct CdfModel {
    uint32_t nseg;
    uint32_t frac;        // fixed-point fraction bits of y
    uint32_t kmin, kmax;  // exact R key range (inclusive)
    uint32_t drop_outside;
    uint32_t cshift;      // cell of a key = (key - kmin) >> cshift
    uint32_t ncells;      // cells over [kmin, kmax]; cells[ncells] = 256 << bits
    uint32_t linear;      // 1: the linear model (lmul), no knots or cells
    uint64_t lmul;
    uint64_t x[kMaxKnots + 1];  // knot keys (x[nseg] may be 2^32)
    uint32_t y[kMaxKnots + 1];  // cumulative partition position
    uint32_t cells[kCells + 1];
};

enum BinKind : uint32_t { kBinRadix = 0, kBinLearned = 1 };

// The partition function of one pass: full = radix ? key & (2^bits - 1)
// : model(key); bin = (full >> pass_shift) & pass_mask.
struct BinSpec {
    uint32_t kind;
    uint32_t bits;        // total partition bits
    uint32_t pass_shift;
    uint32_t pass_mask;
    const CdfModel* model;
    uint32_t drop;        // 1: drop keys outside R's key range
    uint32_t pad;
    // R's exact key range {min, max} in device memory for the drop test
    // (NULL: the model's [kmin, kmax])
    const uint32_t* range;
};

}  // namespace tbj
