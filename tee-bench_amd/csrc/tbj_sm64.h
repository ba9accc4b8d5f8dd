// 64-bit-key sort-merge join (16-byte tuples): shapes, device types and the
// host-side launchers of its kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbj_part.h"

namespace tbj {

// Partition passes: one tile is kSm64Span tuples (128 KiB), so the per-tile
// histograms stay small next to the data; the scatter's workgroup of
// kSm64ScBlock threads stages the whole tile in LDS.  The other 64-bit
// kernels use workgroups of kSm64Block threads.
constexpr int kSm64Block = 512;
constexpr int kSm64Span = 8192;
constexpr int kSm64ScBlock = 1024;
constexpr int kSm64ScItems = 8;
constexpr int kSm64ScTile = kSm64ScBlock * kSm64ScItems;
static_assert(kSm64Span == kSm64ScTile, "the scatter stages a whole tile");
constexpr int kSm64PassBits = 10;
constexpr int kSm64MaxBins = 1 << kSm64PassBits;
// Join: R chunks of at most kSm64Cap tuples are sorted in LDS; a task is one
// R chunk against at most kSm64SChunk S tuples of the same partition.
constexpr int kSm64Cap = 2048;
constexpr int kSm64SItems = 8;
constexpr int kSm64SChunk = kSm64Block * kSm64SItems;  // 4096
// Learned model
constexpr int kSm64Knots = 256;
constexpr int kSm64Cells = 2048;
constexpr int kSm64FitMax = 16384;
// Merge passes for R partitions larger than one chunk
constexpr int kSm64MergeItems = 8;
constexpr int kSm64MergeTile = kSm64Block * kSm64MergeItems;

// Piecewise-linear CDF of R's keys: segment i covers keys [x[i], x[i+1])
// (the last one up to kmax and beyond) and maps key k to partition
// clamp(base[i] + (k - x[i]) * slope[i], plo[i], phi[i]).  Knot i sits at
// sample rank r_i with value 2^bits * (r_i + 8) / (m + 15); keys below x[0]
// continue segment 0's line down to partition 0 and keys above kmax
// continue the last segment up to the last partition, so the ~n/m keys
// outside the sample's range spread over the end partitions instead of
// piling into the first and the last.  The clamps make the map monotone: partitions are key ranges
// in key order.
//
// The partition kernels evaluate it through a cell table: cellv[c] is the
// model at key kmin + (c << gshift) (c <= ncells), and a key inside
// [kmin, kmax] interpolates linearly (integer arithmetic) between its cell's
// two ends; keys outside take segment 0's or the last segment's line, the
// latter never below ymax, the table's value at kmax.  The map stays
// monotone.
//
// A near-linear sample (sample_is_linear) gives the linear model instead:
// partition = min(umulhi(key - kmin, lmul), 2^bits - 1) over [kmin, kmax]
// (the sample's range widened by 8 expected gaps per side; keys beyond it
// take the end partitions), lmul ~ 2^(64 + bits) / (kmax - kmin + 1): key
// ranges in key order for one multiply, no knots and no cells.  When the
// widened range is all of [0, 2^64) (linear = 2) lmul is 2^bits and the map
// is the key's top `bits` bits, taken with a shift.
struct Model64 {
    uint32_t nseg, gshift, ncells, ymax;
    uint64_t kmin, kmax;  // sample key range (the cells span it)
    uint32_t linear, pad;
    uint64_t lmul;
    uint64_t x[kSm64Knots + 1];
    double base[kSm64Knots];
    double slope[kSm64Knots];
    uint32_t plo[kSm64Knots], phi[kSm64Knots];
    uint32_t cellv[kSm64Cells + 1];
};

struct Bin64 {
    uint32_t kind;        // kBinRadix or kBinLearned
    uint32_t bits;        // total partition bits
    uint32_t pass_shift;
    uint32_t pass_mask;
    const Model64* model;
    // non-null: keys outside [range[0], range[1]] are dropped (S against R's
    // range), unless gate is set and *gate <= gate_min
    const unsigned long long* range;
    const uint32_t* gate;
    uint32_t gate_min;
};

struct SmTask {
    uint32_t r_off, r_len;  // R chunk (at most kSm64Cap tuples)
    uint32_t s_off, s_len;  // S range of the same partition
    uint32_t rchunk;        // index of the R chunk
    uint32_t first;         // 1 for the chunk's first S range (writes the sorted chunk)
    uint32_t multi;         // 1 if more than one task probes this R chunk (more than one S range)
    uint32_t pad;
};

// One kSm64MergeTile slice of the output of a merge pass over an R partition.
struct MergeBlock {
    uint32_t seg_off, seg_len, out_begin, pad;
};

size_t sm64_model_lds(const Bin64& b);
// *lin_out: 1 when the fit gave the linear model.
void sm64_fit(hipStream_t st, const ulonglong2* R, uint64_t nR, uint32_t m, uint32_t nseg,
              uint32_t bits, uint64_t* sample, Model64* model, uint32_t* lin_out);
// rows: each tile's aux = its row (tile_row) for sample-sized first passes.
void sm64_make_tiles(hipStream_t st, TileDesc* tiles, SegDesc* segs, uint32_t nR, uint32_t nS,
                     uint32_t nbins, uint32_t* ntiles_dev, bool rows = false);
// Sample of a sample-sized first pass: cnt1 (zeroed, F1 words per row of R,
// then per row of S) += the sampled tuples of each (row, bin).  bss must not
// drop (R's range is not known yet).  s_out (optional, zeroed; bss learned):
// += the sampled S keys outside the model's key range [kmin, kmax] -- the
// keys that R's exact range may let S drop.
void sm64_sample_rows(hipStream_t st, const ulonglong2* R, uint32_t nR, const ulonglong2* S,
                      uint32_t nS, uint32_t nrowR, uint32_t nrowS, const Bin64& bsr,
                      const Bin64& bss, uint32_t F1, uint32_t* cnt1, uint32_t* s_out = nullptr);
// Second-pass segments (one per first-pass partition, 2 * F1) over the tiles
// of sample-sized regions (first: their scanned tile counts in partition order).
void sm64_part_segments(hipStream_t st, const uint32_t* first, uint32_t F1, uint32_t F2,
                        uint32_t nrowR, uint32_t nrowS, SegDesc* segs2);
// Claims of a sample-sized first pass (see k_scatter64): regions cb0 +
// tile row * cmul + bin; mm (optional) receives the launch's key range,
// unless gate is set and *gate <= gate_min.
struct Sm64Claims {
    const uint32_t* starts;
    uint32_t* cursors;
    uint32_t* overflow;
    uint32_t cb0, cmul;
    unsigned long long* mm;
    const uint32_t* gate = nullptr;
    uint32_t gate_min = 0;
};
// Tile t's bin counts go to row t of hist (nbins u16: a tile holds at most
// kSm64Span tuples); mm (optional):
// receives atomic min / max of the keys of the tiles.
void sm64_hist(hipStream_t st, uint32_t grid, const ulonglong2* in, const TileDesc* tiles,
               const uint32_t* ntiles_dev, uint32_t tile_base, uint16_t* hist, const Bin64& b,
               uint32_t nbins, unsigned long long* mm);
// Tiles [tile_base, tile_base + grid) below *ntiles_dev; offs: the rows of
// per-tile offsets (launch_tile_offsets).
void sm64_scatter(hipStream_t st, uint32_t grid, const ulonglong2* in, ulonglong2* out,
                  const TileDesc* tiles, const uint32_t* ntiles_dev, uint32_t tile_base,
                  const uint32_t* offs, const Bin64& b, uint32_t nbins,
                  const Sm64Claims* claims = nullptr);
void sm64_segments(hipStream_t st, const uint32_t* scanned1, const SegDesc* segs1, uint32_t nbins1,
                   uint32_t nbins2, SegDesc* segs2, TileDesc* tiles2, uint32_t* ntiles2_dev,
                   uint32_t max_tiles2);
void sm64_bounds(hipStream_t st, const uint32_t* scanned, const SegDesc* segs, uint32_t outer,
                 uint32_t inner, uint32_t passes, uint32_t nR, uint32_t nS, uint2* bounds);
void sm64_task_counts(hipStream_t st, const uint2* bounds, uint32_t nparts, uint32_t* rcnt,
                      uint32_t* jcnt, uint32_t* max_r);
// One thread per task (nrtasks R chunks, njtasks join tasks).
void sm64_fill_tasks(hipStream_t st, const uint2* bounds, uint32_t nparts, const uint32_t* roff,
                     const uint32_t* joff, uint32_t nrtasks, uint32_t njtasks, SmTask* rtasks,
                     SmTask* jtasks);
// Sorts every R chunk by key in place (bucket sort in LDS).
void sm64_sort_r(hipStream_t st, uint32_t grid, ulonglong2* Rb, const SmTask* rtasks,
                 uint32_t nrtasks);
// Sorts the R chunks whose tasks are several (SmTask::multi) from src into
// dst (same offsets).
void sm64_sort_shared(hipStream_t st, const ulonglong2* src, ulonglong2* dst,
                      const SmTask* rtasks, uint32_t nrtasks);
void sm64_merge(hipStream_t st, const ulonglong2* src, ulonglong2* dst, const MergeBlock* blocks,
                uint32_t nblocks, uint64_t width);
void sm64_fill_u64(hipStream_t st, unsigned long long* p, uint64_t n, unsigned long long v);
// Count.  presorted: R chunks are already sorted in Rs; otherwise the chunks
// of several tasks are already sorted in Rout (sm64_sort_shared), and each
// task of a single-task chunk sorts it in LDS and, materialising, writes the
// sorted chunk to Rout (same offsets); materialising, the task of a small
// single-task chunk may instead stage its pairs in key order at its S
// positions of Rout and mark runs[r_off] (k_emit64 copies them).
void sm64_count(hipStream_t st, uint32_t grid, bool mat, bool presorted, const ulonglong2* Rs,
                ulonglong2* Rout, const ulonglong2* Sb, const SmTask* tasks,
                const uint32_t* ntasks_dev, unsigned long long* runs,
                unsigned long long* chunk_vol, unsigned long long* counters);
// Zeroes the run counters of R chunks that several join tasks share (the
// count kernel writes the others whole).
void sm64_clear_shared_runs(hipStream_t st, const SmTask* tasks, uint32_t ntasks,
                            unsigned long long* runs);
// Exclusive scan of a[0, n) (u64) in place; *total = sum; part needs
// sm64_scan_partials(n) words.
uint32_t sm64_scan_partials(uint32_t n);
void sm64_scan_u64(hipStream_t st, unsigned long long* a, uint32_t n, unsigned long long* part,
                   unsigned long long* total);
// runs[g] (S count of the key run starting at sorted R position g) becomes
// the run's first output slot.
void sm64_run_offsets(hipStream_t st, uint32_t grid, const ulonglong2* Rsorted,
                      const SmTask* rtasks, uint32_t nrtasks,
                      const unsigned long long* chunk_off, const unsigned long long* total,
                      unsigned long long* runs);
// Emit: tasks of shared chunks claim slots in their runs; a chunk's only
// task lays its runs out from chunk_off[rchunk] (the scanned volumes).  Sb
// holds R's partitions too: unless presorted, a chunk's only task with at
// most kSm64Block S tuples reads its R chunk there unsorted (the count kernel
// did not sort it) and orders its S tuples by key instead.
void sm64_emit(hipStream_t st, uint32_t grid, const ulonglong2* Rsorted, const ulonglong2* Sb,
               const SmTask* tasks, const uint32_t* ntasks_dev, unsigned long long* runs,
               const unsigned long long* chunk_off, ulonglong2* pairs, uint64_t cap,
               bool presorted);

}  // namespace tbj
