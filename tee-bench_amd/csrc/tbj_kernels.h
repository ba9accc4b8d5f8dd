// Kernel launch shapes and the host-side launchers of every device kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbj_part.h"

namespace tbj {

// Partition kernels: one workgroup of kPartBlock threads owns a tile of
// kTile tuples (kPartItems per thread).
constexpr int kPartBlock = 512;
constexpr int kPartItems = 8;
constexpr int kTile = kPartBlock * kPartItems;  // 4096 tuples = 32 KiB
// Other tile shapes (tuples per tile) the batch's partition passes may use:
// kTile / 2, kTile and kMaxTile
constexpr int kMaxTile = 2 * kTile;
inline bool valid_tile(uint32_t t) { return t == kTile / 2 || t == kTile || t == kMaxTile; }
constexpr int kMaxBins = 2048;                  // bins of one pass
constexpr int kMaxPassBits = 11;
// Device-wide scan
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanBlock * kScanItems;
// CDF model fit
constexpr int kLdsKnots = 256;
constexpr int kFitBlock = 1024;
constexpr int kFitMax = 16384;
// Join kernels
constexpr int kJoinBlock = 512;
constexpr int kJoinCap = 4096;   // max R tuples of one build (LDS hash table)
constexpr int kJoinProbeItems = 4;
constexpr int kJoinSItems = 16;  // S tuples per thread of one task
constexpr int kJoinSChunk = kJoinBlock * kJoinSItems;  // 8192
// S tuples probed per sub-chunk of a join task: kJoinSChunk counting, half
// of it materialising (the per-item output slots need the registers)
inline int join_s_chunk(bool materialize) { return materialize ? kJoinSChunk / 2 : kJoinSChunk; }
// S tuples per join task (the R range is built once per task)
constexpr int kJoinSTaskChunks = 4;

enum JoinMode : int { kModeChain = 0, kModeHist = 1, kModeSort = 2 };

// Tiles [tile_base, tile_base + ntiles) (and below *ntiles_dev); tile t's
// bin counts go to hist[t * nbins, (t + 1) * nbins) (u16: a tile holds at
// most kMaxTile tuples).  tile: the tiles' shape (valid_tile).
void launch_hist(hipStream_t st, uint32_t ntiles, const uint2* in, const TileDesc* tiles,
                 const uint32_t* ntiles_dev, uint32_t tile_base, uint16_t* hist,
                 const BinSpec& bs, uint32_t nbins, uint32_t tile = kTile);
// Per-tile counts (u16 rows of nbins) -> per-tile output offsets (u32 rows
// in offs), for the tiles below *ntiles_dev (at most max_tiles).  gsum (glen
// words, the last one zero and glen - 1 >= nbins * the groups of all
// segments) receives the scanned group sums: partition (segment s, bin b)
// starts at gsum[s.hist_base + b * s.ngroups].  partial: scan_partials(glen)
// words.
void launch_tile_offsets(hipStream_t st, uint32_t max_tiles, const uint16_t* counts,
                         uint32_t* offs, const TileDesc* tiles, const SegDesc* segs,
                         const uint32_t* ntiles_dev, uint32_t nbins, uint32_t* gsum,
                         uint64_t glen, uint32_t* partial, uint32_t* grand_total);
// Sample-sized regions of a claiming scatter: tile t's run of bin i goes to
// region r = cb0 + tiles[t].aux * cmul + i * cstride + xmul * (the XCD eighth
// the tile runs in), which owns
// [starts[r], starts[r + 1]) and fills from cursors[r] (zeroed beforehand);
// a run that does not fit sets *overflow.
struct ClaimSpec {
    const uint32_t* starts;
    uint32_t* cursors;
    uint32_t* overflow;
    uint32_t cb0, cmul, cstride;
    uint32_t xmul;  // r += xmul * the XCD eighth the tile runs in
    const uint32_t* agg_flag = nullptr;  // aggregate ranks only while *agg_flag (last pass:
                                         // NULL means always)
    uint2* tile_range = nullptr;  // non-null: each tile's {min, max} key, by tile - tile_base
};
// out[r * F2 / 2 + j] = in[r * F2 + j] + in[r * F2 + j + F2 / 2] for rows r
// of F2 counts: the counts of bins j and j + F2 / 2 (bins of one more key
// bit) merged into bin j.
void launch_fold_bins(hipStream_t st, const uint32_t* in, uint32_t rows, uint32_t F2,
                      uint32_t* out);
// minmax[0..1] = {least minimum, greatest maximum} of n tile ranges.
void launch_range_reduce(hipStream_t st, const uint2* tile_range, uint32_t n, uint32_t* minmax);
// offs: the per-tile offsets from launch_tile_offsets (claims == NULL), or
// sample-sized regions claimed per (tile, bin) (claims).  last_pass: the
// plan's last pass, whose segments are narrow key ranges where one hot key can
// fill most of a wave: lanes sharing a bin then take their ranks from one LDS
// atomic.  A first pass of two, whose tiles mix the whole key range, takes
// the plain per-tuple atomics, or the aggregated ranks while
// *claims->agg_flag (a range partition of skewed keys can put most of a tile
// in one bin).
void launch_scatter(hipStream_t st, uint32_t ntiles, const uint2* in, uint2* out,
                    const TileDesc* tiles, const uint32_t* ntiles_dev, uint32_t tile_base,
                    const uint32_t* offs, const BinSpec& bs, uint32_t nbins, bool last_pass,
                    const ClaimSpec* claims = nullptr, uint32_t tile = kTile);
// Exclusive scan of a[0..n) in place; partial needs scan_partials(n) words.
uint32_t scan_partials(uint64_t n);
void launch_scan(hipStream_t st, uint32_t* a, uint64_t n, uint32_t* partial,
                 uint32_t* grand_total);
// Second-pass segments from the first pass's scanned group sums;
// *ngroups2_dev (optional) receives the groups of all of them.
void launch_make_segments(hipStream_t st, const uint32_t* scanned1, const SegDesc* segs1,
                          uint32_t nseg1, uint32_t nbins1, uint32_t nbins2, SegDesc* segs2,
                          TileDesc* tiles2, uint32_t* ntiles2_dev, uint32_t max_tiles2,
                          uint32_t* ngroups2_dev = nullptr, uint32_t tile = kTile);
// Tiles of `tile` tuples over a segment list (segment first_tile / ntiles set).
void launch_tiles_from_segs(hipStream_t st, const SegDesc* segs, uint32_t nseg,
                            const uint32_t* ntiles_dev, TileDesc* tiles, uint32_t max_tiles,
                            uint32_t tile = kTile);
void launch_sample(hipStream_t st, const uint2* in, uint64_t n, uint32_t m, uint32_t* sample);
// Sorts the m sampled keys in place (one workgroup); nothing when model
// (optional) is already the linear model.
void launch_fit_sort(hipStream_t st, uint32_t* sample, uint32_t m,
                     const CdfModel* model = nullptr);
// The linear model from the (unsorted) sample when it is near-linear; sets
// model->linear either way (launch_fit then leaves a linear model alone).
void launch_fit_linear(hipStream_t st, const uint32_t* sample, uint32_t m, uint32_t bits,
                       CdfModel* model, uint32_t* lin_out);
// The model from the sorted sample and R's exact key range (minmax), or the
// sample's range when minmax is NULL.
void launch_fit(hipStream_t st, const uint32_t* sorted, uint32_t m, uint32_t nseg,
                uint32_t bits, const uint32_t* minmax, CdfModel* model);
void launch_minmax(hipStream_t st, const uint2* in, uint64_t n, uint32_t* minmax);
void launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v);
// Zeroes several u32 arrays in one launch.
constexpr uint32_t kZeroArrays = 6;
struct ZeroList {
    uint32_t* p[kZeroArrays];
    uint64_t n[kZeroArrays];
    uint32_t k;
};
void launch_zero_multi(hipStream_t st, const ZeroList& z);
// launch_zero_multi and launch_make_tiles1 in one launch (z must not hold
// *ntiles_dev).
void launch_start_tiles1(hipStream_t st, const ZeroList& z, TileDesc* tiles, SegDesc* segs,
                         uint32_t nR, uint32_t nS, uint32_t nbins, uint32_t* ntiles_dev);
// Sample-sized passes (no histogram read).  k_sample_rows: per first-pass
// tile row the sampled counts of each first-pass bin (cnt1[pbase(rel) + row *
// F1 + b1], pbase(S) = F1 * nrowR) and per relation the sampled (b1, b2)
// counts (fine, 2 * F1 * F2 words, zeroed; F1 * F2 <= 65536).
// which: 0 = R's rows, 1 = S's rows, 2 = both.  The second-pass bin comes
// from the same evaluation of the partition function as the first-pass bin
// (bs2 contributes its pass shift and mask only).  *hot1 (if non-null) is
// set when a sample chunk finds one first-pass bin holding over 1/8 of its
// sampled positions (dropped keys count as positions, not in any bin).
void launch_sample_rows(hipStream_t st, const uint2* R, uint32_t nR, const uint2* S,
                        uint32_t nS, uint32_t nrowR, uint32_t nrowS, const BinSpec& bsr,
                        const BinSpec& bss, const BinSpec& bs2, uint32_t F1, uint32_t F2,
                        uint32_t* cnt1, uint32_t* fine, int which = 2,
                        uint32_t* hot1 = nullptr);
// a[i] += *base for i < n.
void launch_add_base(hipStream_t st, uint32_t* a, uint32_t n, const uint32_t* base);
// caps[i] from 16 * cnt[i] for i < n, caps[n] = 0 (scan to region starts).
void launch_row_caps(hipStream_t st, const uint32_t* cnt, uint32_t n, uint32_t* caps);
// Exact first-pass partition sizes (2 * F1 words) from the filled regions;
// their sum is added to *total.
void launch_part_sizes(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                       uint32_t F1, uint32_t nrowR, uint32_t nrowS, uint32_t* npart,
                       uint32_t* total);
// Estimated largest first-pass R partition (from the row counts) into
// stats[0] (zeroed).
void launch_est_part_max(hipStream_t st, const uint32_t* cnt1, uint32_t F1, uint32_t nrowR,
                         uint32_t* stats);
// Second-pass region caps, 8 per-XCD sub-regions per (partition, bin):
// 8 * nseg2 * nbins2 + 1 words, from fine, npart and the partitions' tile
// spans (first: scanned tile counts in partition order; ntiles2: the total).
// *hot becomes 1 when a second-pass bin holds over 1/8 of its partition's
// sample (a hot key: the second pass then aggregates its ranks per wave).
void launch_region_caps(hipStream_t st, const uint32_t* fine, const uint32_t* npart,
                        uint32_t nseg2, uint32_t nbins2, const uint32_t* first,
                        const uint32_t* ntiles2, uint32_t F1, uint32_t nrowR, uint32_t nrowS,
                        uint32_t* caps, uint32_t* hot);
// Second-pass tiles over the n filled first-pass regions, visited in
// partition order: ntiles (n + 1 words), scanned into first tiles, then the
// descriptors.
// tile: tuples per second-pass tile.
void launch_region_tiles(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                         uint32_t n, uint32_t F1, uint32_t nrowR, uint32_t nrowS,
                         uint32_t* ntiles, uint32_t tile = kTile);
// seg_is_part: a tile's segment is its first-pass partition (rel * F1 + b1)
// rather than its region.
void launch_region_tile_descs(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                              const uint32_t* first, uint32_t n, uint32_t F1, uint32_t nrowR,
                              uint32_t nrowS, TileDesc* tiles, uint32_t tile = kTile,
                              bool seg_is_part = false);
void launch_bounds_est(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                       uint32_t n, uint2* bounds);
// Join bounds (8 pieces per partition side) of the second pass's per-XCD
// sub-regions, laid out ((rel * F1 + b1) * 8 + x) * F2 + b2.
void launch_bounds_x8(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                      uint32_t F1, uint32_t F2, uint2* bounds);
void launch_make_tiles1(hipStream_t st, TileDesc* tiles, SegDesc* segs, uint32_t nR, uint32_t nS,
                        uint32_t nbins, uint32_t* ntiles_dev);
// Zeroes hist[mult * (*count_dev) .. n); max_tail bounds the zeroed length.
void launch_zero_tail(hipStream_t st, uint32_t* hist, uint64_t n, const uint32_t* count_dev,
                      uint32_t mult, uint64_t max_tail);

// Final partition bounds of R and S: bounds[rel * nparts + p] = {start, end}
// (scanned: the scanned group sums).
void launch_part_bounds(hipStream_t st, const uint32_t* scanned, const SegDesc* segs,
                        uint32_t nbins_outer, uint32_t nbins_inner, uint32_t passes,
                        uint32_t nR, uint32_t nS, uint2* bounds);
// Task offsets (task_off needs nparts + 1 words, partial scan_partials(nparts + 1)),
// their total in *ntasks_dev, partition maxima folded into stats[0..1]
// (stats must start at zero).  npieces: ranges per partition side in bounds.
void launch_count_tasks(hipStream_t st, const uint2* bounds, uint32_t nparts, uint32_t cap,
                        uint32_t s_chunk, uint32_t* task_off, uint32_t* partial,
                        uint32_t* ntasks_dev, uint32_t* stats, uint32_t npieces = 1);
// Lists the first ts.task_cap tasks (the grid covers task_cap; the count
// itself stays on the device).
void launch_fill_tasks(hipStream_t st, const TaskSrc& ts);
// Where a materialising join writes (tbj_join.hip).  chunk == 0: every
// group of S tuples claims its slots from counters[2] (dense output).
// chunk > 0: a workgroup claims whole chunks of `chunk` slots and hands out
// slots from its current one; tails[w] records workgroup w's last chunk
// (grid entries), whose unused tail is a gap until launch_out_fix fills it,
// and slots at or above cap go to side (side_cap = grid * chunk slots, as
// many as the gaps can hold).
struct OutTail {
    unsigned long long base;
    uint32_t used, pad;
};
struct OutBuf {
    uint2* pairs;
    uint64_t cap;
    uint2* side;
    uint64_t side_cap;
    OutTail* tails;
    uint32_t chunk, pad;
};
// The chunk of a join in `mode` over nS S tuples on grid workgroups (0: one
// claim per group of S tuples).
uint32_t out_chunk(int mode, uint64_t nS, uint32_t grid);
// Persistent join over the tasks of ts (grid: resident workgroups; those that
// find no task left exit).  rel: materialise (key, R.payload) tuples instead
// of (R.payload, S.payload) pairs.
void launch_join(hipStream_t st, int mode, bool materialize, bool rel, uint32_t grid,
                 const uint2* Rb, const uint2* Sb, const TaskSrc& ts, uint32_t* task_ctr,
                 uint32_t hshift, const OutBuf& out, unsigned long long* counters);
// After a chunk-mode join: moves the pairs in slots at or above
// W = min(counters[0], cap) (matches; pairs and side buffer) into the gaps
// below W, so that pairs[0, W) holds every written pair.  ntails: the join's
// grid; work: out_fix_bytes(ntails) bytes of device memory.
size_t out_fix_bytes(uint32_t ntails);
void launch_out_fix(hipStream_t st, const OutBuf& out, uint32_t ntails,
                    const unsigned long long* counters, void* work);
int join_occupancy(int mode, bool materialize);

// No-partitioning joins (tbj_nopart.hip)
void launch_fill_u4(hipStream_t st, void* p, uint64_t bytes, uint32_t v);
size_t pht_bucket_bytes();
size_t pht_node_bytes();
void launch_pht_build(hipStream_t st, const uint2* R, uint64_t nR, void* tab, uint32_t mask,
                      void* pool, uint32_t* pool_ctr);
// Bulk build: one block of 2^hbits buckets (hbits <= kPhtPlaceHBits of
// tbj_pht.h) per partition of Rb (partition_relation on the low pbits bits,
// first pass b1 bits; bounds: its partition bounds), overflow nodes claimed
// from *pool_ctr (zeroed).  The table (2^(pbits + hbits) buckets) needs no
// clearing.
void launch_pht_place(hipStream_t st, const uint2* Rb, const uint2* bounds, uint32_t b1,
                      uint32_t pbits, uint32_t hbits, void* tab, void* pool, uint32_t* pool_ctr);
// pbits, hbits: the table's layout (pht_slot); pbits = 0 for a table built by
// launch_pht_build.
void launch_pht_probe(hipStream_t st, bool mat, const uint2* S, uint64_t nS, const void* tab,
                      uint32_t mask, uint32_t pbits, uint32_t hbits, const void* pool,
                      uint2* pairs, uint64_t cap, unsigned long long* counters);
// CHT state on the device: overflow count (zeroed by the caller), phase-A
// spills, whether the dense array is exactly full, and the number of set bits.
struct ChtState {
    uint32_t spill_ctr, spill_a, dense, nset;
};
// CHT build: words (nwords {bitmap word, rank}, zeroed), cnt (nwords + 1),
// taken ((|R| + 2 + 31) / 32 words, zeroed), arr (|R| + 2 tuples), spill
// (|R|), spilled ((|R| + 31) / 32 words, zeroed).
void launch_cht_build(hipStream_t st, const uint2* R, uint64_t nR, uint32_t bmask,
                      uint32_t pmask, uint2* words, uint32_t* cnt, uint32_t* partial,
                      uint32_t* taken, uint2* arr, uint32_t* spill, uint32_t* spilled,
                      ChtState* state);
// CHT bulk build: Rb = R partitioned (partition_relation) on bitmap position
// bits [sbits, sbits + pbits) (sbits <= 12, and no more than log2 of the CHT
// partition size), bounds its partition bounds, b1 its first pass's bits;
// words (zeroed or not), segcnt (2^pbits + 1), partial
// (scan_partials(2^pbits + 1)), taken (zeroed), arr, spill (Rb positions),
// spilled (zeroed), state (zeroed).
void launch_cht_build_bulk(hipStream_t st, const uint2* Rb, const uint2* bounds, uint32_t b1,
                           uint32_t pbits, uint32_t sbits, uint32_t bmask, uint32_t pmask,
                           uint2* words, uint32_t* segcnt, uint32_t* partial, uint32_t* taken,
                           uint2* arr, uint32_t* spill, uint32_t* spilled, ChtState* state);
void launch_cht_overflow(hipStream_t st, const uint2* R, const uint32_t* spill,
                         const ChtState* state, uint32_t nspill, uint32_t ov_mask,
                         uint32_t* ov_used, uint2* ov);
// dense: ChtState::dense as read back after the build.
void launch_cht_probe(hipStream_t st, bool mat, bool dense, const uint2* S, uint64_t nS,
                      uint32_t bmask, uint32_t pmask, uint32_t ov_mask, const uint2* words,
                      const uint32_t* taken, const uint2* arr, uint32_t nslots,
                      const uint32_t* ov_used, const uint2* ov, const ChtState* state,
                      uint2* pairs, uint64_t cap, unsigned long long* counters);

}  // namespace tbj
