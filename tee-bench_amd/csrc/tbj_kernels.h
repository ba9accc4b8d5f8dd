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
constexpr int kMaxBins = 2048;                  // bins of one pass
constexpr int kMaxPassBits = 11;
// Device-wide scan
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanBlock * kScanItems;
// CDF model fit
constexpr int kLdsKnots = 1024;
constexpr int kFitBlock = 1024;
constexpr int kFitMax = 16384;
// Join kernels
constexpr int kJoinBlock = 512;
constexpr int kJoinCap = 4096;   // max R tuples of one build (LDS hash table)
constexpr int kJoinProbeItems = 4;

enum JoinMode : int { kModeChain = 0, kModeHist = 1 };

size_t lds_model_bytes();

void launch_hist(hipStream_t st, uint32_t grid, const uint2* in, const TileDesc* tiles,
                 const SegDesc* segs, const uint32_t* ntiles_dev, uint32_t tile_base,
                 uint32_t* hist, const BinSpec& bs, uint32_t nbins, uint32_t* minmax);
void launch_scatter(hipStream_t st, uint32_t grid, const uint2* in, uint2* out,
                    const TileDesc* tiles, const SegDesc* segs, const uint32_t* ntiles_dev,
                    uint32_t tile_base, const uint32_t* scanned, const BinSpec& bs,
                    uint32_t nbins);
// Exclusive scan of a[0..n) in place; partial needs scan_partials(n) words.
uint32_t scan_partials(uint64_t n);
void launch_scan(hipStream_t st, uint32_t* a, uint64_t n, uint32_t* partial,
                 uint32_t* grand_total);
void launch_make_segments(hipStream_t st, const uint32_t* scanned1, const SegDesc* segs1,
                          uint32_t nseg1, uint32_t nbins1, uint32_t nbins2, SegDesc* segs2,
                          TileDesc* tiles2, uint32_t* ntiles2_dev, uint32_t max_tiles2);
void launch_sample(hipStream_t st, const uint2* in, uint64_t n, uint32_t m, uint32_t* sample);
void launch_fit(hipStream_t st, const uint32_t* sample, uint32_t m, uint32_t nseg,
                uint32_t bits, CdfModel* model);
void launch_set_bounds(hipStream_t st, CdfModel* model, const uint32_t* minmax, uint32_t drop);
void launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v);
void launch_make_tiles1(hipStream_t st, TileDesc* tiles, SegDesc* segs, uint32_t nR, uint32_t nS,
                        uint32_t nbins, uint32_t* ntiles_dev);
// Zeroes hist[mult * (*count_dev) .. n); max_tail bounds the zeroed length.
void launch_zero_tail(hipStream_t st, uint32_t* hist, uint64_t n, const uint32_t* count_dev,
                      uint32_t mult, uint64_t max_tail);

// Final partition bounds of R and S: bounds[rel * nparts + p] = {start, end}.
void launch_part_bounds(hipStream_t st, const uint32_t* scanned, const SegDesc* segs,
                        uint32_t nbins_outer, uint32_t nbins_inner, uint32_t passes,
                        uint32_t nR, uint32_t nS, uint2* bounds);
void launch_count_tasks(hipStream_t st, const uint2* bounds, uint32_t nparts, uint32_t cap,
                        uint32_t s_chunk, uint32_t* task_off, uint32_t* partial,
                        uint32_t* ntasks_dev, uint32_t* stats);
void launch_fill_tasks(hipStream_t st, const uint2* bounds, uint32_t nparts, uint32_t cap,
                       uint32_t s_chunk, const uint32_t* task_off, JoinTask* tasks);
void launch_join(hipStream_t st, int mode, bool materialize, uint32_t grid, const uint2* Rb,
                 const uint2* Sb, const JoinTask* tasks, const uint32_t* ntasks_dev,
                 uint32_t* task_ctr, uint32_t hshift, uint2* pairs, uint64_t cap,
                 unsigned long long* counters);
int join_occupancy(int mode, bool materialize);

}  // namespace tbj
