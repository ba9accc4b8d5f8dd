// Partition-function and work-descriptor types shared by host and device.
#pragma once
#include <stdint.h>

namespace tbj {

// One contiguous input range partitioned independently (a relation in the
// first pass, a first-pass partition of a relation in the second pass).
struct SegDesc {
    uint64_t hist_base;   // first slot of this segment's bin-major histogram
    uint32_t in_off;      // first input tuple
    uint32_t n;           // tuples
    uint32_t ntiles;      // tiles of the segment
    uint32_t first_tile;  // global index of its first tile
};

struct TileDesc {
    uint32_t in_off;      // first input tuple of the tile
    uint32_t len;         // tuples in the tile
    uint32_t seg;         // owning segment
    uint32_t pad;
};

// One build/probe task: an R range (at most the LDS capacity) against an S range.
struct JoinTask {
    uint32_t r_off, r_len, s_off, s_len;
};

// Piecewise-linear key-CDF model (learned range partitioner).  Knot i covers
// keys [x[i], x[i+1]); its cumulative value y[i] is in units of
// 2^-frac partitions.  Keys below x[0] map to partition 0 and keys at or
// above x[nseg] to the last one (the model is clamped, so it stays
// monotone).  kmin/kmax are R's exact key bounds: keys outside are dropped
// from S because they cannot match.
constexpr int kMaxKnots = 4096;
constexpr int kGuideCells = 2048;  // equal-width cells over [kmin, kmax]
struct CdfModel {
    uint32_t nseg;
    uint32_t frac;        // fixed-point fraction bits of y
    uint32_t kmin, kmax;  // exact R key range (inclusive)
    uint32_t drop_outside;
    uint32_t gshift;      // cell of a key = (key - kmin) >> gshift
    uint32_t ncells;
    uint32_t pad;
    // guide[c] = segment holding the first key of cell c; guide[ncells] =
    // nseg - 1.  A key of cell c lies in a segment of [guide[c], guide[c+1]].
    uint16_t guide[kGuideCells + 1];
    uint16_t pad2[7];
    uint64_t x[kMaxKnots + 1];      // knot keys (x[nseg] may be 2^32)
    uint32_t y[kMaxKnots + 1];      // cumulative partition position
    uint64_t slope[kMaxKnots];      // (y[i+1]-y[i]) * 2^32 / (x[i+1]-x[i])
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
    uint32_t drop;        // 1: drop keys outside [model->kmin, model->kmax]
    uint32_t pad;
};

}  // namespace tbj
