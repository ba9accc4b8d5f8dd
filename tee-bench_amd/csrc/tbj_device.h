// Device-side helpers shared by the join kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TBJ_WAVE 64
#define TBJ_DROP 0xFFFFFFFFu

namespace tbj {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (TBJ_WAVE - 1); }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x / TBJ_WAVE; }

// LDS counters: increments ctr[idx] for each active lane with idx !=
// TBJ_DROP and returns the value before the lane's own increment (its rank).
// Lanes sharing the idx of the wave's first such lane are served by one
// atomic (same-address atomics of a wave serialise, and skewed keys put most
// of a wave on one counter); the others add one each.  Every active lane of
// the wave must call it.
__device__ __forceinline__ uint32_t agg_inc(uint32_t* ctr, uint32_t idx) {
    const uint64_t any = __ballot(idx != TBJ_DROP);
    if (any == 0) return 0;
    const uint32_t fl = (uint32_t)__ffsll((unsigned long long)any) - 1;
    const uint32_t lead = __shfl(idx, fl, TBJ_WAVE);
    const uint64_t same = __ballot(idx == lead);
    const uint32_t lane = lane_id();
    uint32_t r = 0;
    if (idx == lead) {
        if (lane == fl) r = atomicAdd(&ctr[idx], (uint32_t)__popcll(same));
    } else if (idx != TBJ_DROP) {
        r = atomicAdd(&ctr[idx], 1u);
    }
    const uint32_t b = __shfl(r, fl, TBJ_WAVE);
    return idx == lead ? b + (uint32_t)__popcll(same & ((1ull << lane) - 1)) : r;
}

// agg_inc without the rank.
__device__ __forceinline__ void agg_add(uint32_t* ctr, uint32_t idx) {
    const uint64_t any = __ballot(idx != TBJ_DROP);
    if (any == 0) return;
    const uint32_t fl = (uint32_t)__ffsll((unsigned long long)any) - 1;
    const uint32_t lead = __shfl(idx, fl, TBJ_WAVE);
    const uint64_t same = __ballot(idx == lead);
    if (idx == lead) {
        if (lane_id() == fl) atomicAdd(&ctr[idx], (uint32_t)__popcll(same));
    } else if (idx != TBJ_DROP) {
        atomicAdd(&ctr[idx], 1u);
    }
}

// Inclusive prefix sum across the 64 lanes of a wave, in DPP moves (no LDS
// crossbar round trips): shifts by 1, 2, 4, 8 inside each row of 16 lanes
// (lanes shifted in from outside the row read 0), then row 0's total is
// broadcast into row 1 and row 2's into row 3 (row_bcast:15, rows 1 and 3
// written), then the total of rows 0-1 into rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    constexpr int kRowShr = 0x110, kBcast15 = 0x142, kBcast31 = 0x143;
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr + 1, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr + 2, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr + 4, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr + 8, 0xf, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kBcast15, 0xa, 0xf, false);
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kBcast31, 0xc, 0xf, false);
    return x;
}

// One DPP step of a 64-bit value (both halves moved alike).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, CTRL, ROWS, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), CTRL, ROWS, 0xf, false);
    return ((uint64_t)hi << 32) | lo;
}

// 64-bit inclusive prefix sum across a wave, the same DPP steps as above.
__device__ __forceinline__ uint64_t wave_inclusive_scan_u64(uint64_t x) {
    x += dpp_u64<0x111, 0xf>(x);
    x += dpp_u64<0x112, 0xf>(x);
    x += dpp_u64<0x114, 0xf>(x);
    x += dpp_u64<0x118, 0xf>(x);
    x += dpp_u64<0x142, 0xa>(x);
    x += dpp_u64<0x143, 0xc>(x);
    return x;
}

// One DPP step of a 64-bit value whose invalid source lanes read `fill`.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_u64_fill(uint64_t x, uint64_t fill) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(
        (int)(uint32_t)fill, (int)(uint32_t)x, CTRL, ROWS, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(
        (int)(uint32_t)(fill >> 32), (int)(uint32_t)(x >> 32), CTRL, ROWS, 0xf, false);
    return ((uint64_t)hi << 32) | lo;
}

// Minimum and maximum of 64-bit values across a wave, in the scans' DPP
// steps (no LDS crossbar round trips); every lane receives both.
__device__ __forceinline__ void wave_minmax_u64(uint64_t& mn, uint64_t& mx) {
    mn = min(mn, dpp_u64_fill<0x111, 0xf>(mn, ~0ull));
    mx = max(mx, dpp_u64_fill<0x111, 0xf>(mx, 0));
    mn = min(mn, dpp_u64_fill<0x112, 0xf>(mn, ~0ull));
    mx = max(mx, dpp_u64_fill<0x112, 0xf>(mx, 0));
    mn = min(mn, dpp_u64_fill<0x114, 0xf>(mn, ~0ull));
    mx = max(mx, dpp_u64_fill<0x114, 0xf>(mx, 0));
    mn = min(mn, dpp_u64_fill<0x118, 0xf>(mn, ~0ull));
    mx = max(mx, dpp_u64_fill<0x118, 0xf>(mx, 0));
    mn = min(mn, dpp_u64_fill<0x142, 0xa>(mn, ~0ull));
    mx = max(mx, dpp_u64_fill<0x142, 0xa>(mx, 0));
    mn = min(mn, dpp_u64_fill<0x143, 0xc>(mn, ~0ull));
    mx = max(mx, dpp_u64_fill<0x143, 0xc>(mx, 0));
    // lane 63 holds the wave's extremes
    mn = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(mn >> 32), 63) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)mn, 63);
    mx = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(mx >> 32), 63) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)mx, 63);
}

__device__ __forceinline__ uint64_t wave_reduce_u64(uint64_t x) {
#pragma unroll
    for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, TBJ_WAVE);
    return x;
}

// Exclusive prefix sum across a block of BLOCK threads.  `scratch` must hold
// BLOCK/64 + 1 words of LDS.  Returns the thread's exclusive prefix; *total
// receives the block-wide sum.  Contains two barriers; every thread of the
// block must call it.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* scratch,
                                                         uint32_t* total) {
    constexpr int NW = BLOCK / TBJ_WAVE;
    const uint32_t lane = lane_id(), wid = wave_id();
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == TBJ_WAVE - 1) scratch[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        uint32_t w = lane < NW ? scratch[lane] : 0;
        uint32_t wi = wave_inclusive_scan(w);
        if (lane < NW) scratch[lane] = wi - w;  // exclusive wave offsets
        if (lane == NW - 1) scratch[NW] = wi;
    }
    __syncthreads();
    uint32_t r = scratch[wid] + inc - v;
    *total = scratch[NW];
    return r;
}

// In-place exclusive scan of `n` LDS words (n <= BLOCK * 64) by a block.
// Returns the total.  Ends with a barrier.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_scan_lds(uint32_t* a, uint32_t n, uint32_t* scratch) {
    const uint32_t per = (n + BLOCK - 1) / BLOCK;
    const uint32_t beg = threadIdx.x * per;
    const uint32_t end = beg + per < n ? beg + per : n;
    uint32_t s = 0;
    for (uint32_t i = beg; i < end; i++) s += a[i];
    uint32_t total;
    uint32_t run = block_exclusive_scan<BLOCK>(s, scratch, &total);
    for (uint32_t i = beg; i < end; i++) {
        uint32_t c = a[i];
        a[i] = run;
        run += c;
    }
    __syncthreads();
    return total;
}

// Tile of block b among a launch's n tiles (one tile per workgroup).
// Workgroups go to the 8 XCDs round-robin by blockIdx.x; block b < n takes a
// tile of the (b % 8)-th eighth of the n tiles, in order, so each XCD walks
// one contiguous range of tiles.  Partitions are laid out with the tiles'
// runs in tile order, so each partition then holds one run per XCD and its
// cache lines fill inside one L2 rather than in two XCDs' partial
// write-backs.  n must be the launch's actual tile count: eighths of an
// upper bound could leave XCDs idle.
constexpr uint32_t kNumXcd = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
    const uint32_t x = b % kNumXcd, q = n / kNumXcd, r = n % kNumXcd;
    return x * q + min(x, r) + b / kNumXcd;
}

// Near-linear test of a key sample, for one whole workgroup (every thread
// calls it; it ends with a barrier).  Splits the sample's range [kmin, kmax]
// into kLinBuckets equal-width buckets (the last one may be narrower) and
// answers true when none holds more than mean + mean / 4 + 4 sqrt(mean) of
// the m keys: at that resolution the sample's CDF stays within a quarter
// bucket of the straight line from kmin to kmax, so a linear map of the key
// range spreads R over the partitions like the fitted model would.
// lds: kLinBuckets + 5 words of 8 bytes.
constexpr uint32_t kLinBuckets = 64;
template <typename K>
__device__ bool sample_is_linear(const K* __restrict__ a, uint32_t m, unsigned long long* lds,
                                 K* kmin_out, K* kmax_out) {
    unsigned long long* hist = lds;
    unsigned long long* mn = lds + kLinBuckets;
    unsigned long long* mx = mn + 1;
    unsigned long long* flag = mn + 2;
    if (threadIdx.x < kLinBuckets) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        *mn = ~0ull;
        *mx = 0;
    }
    __syncthreads();
    K lo = (K)~(K)0, hi = 0;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        lo = min(lo, a[i]);
        hi = max(hi, a[i]);
    }
    atomicMin(mn, (unsigned long long)lo);
    atomicMax(mx, (unsigned long long)hi);
    __syncthreads();
    const K kmin = (K)*mn, kmax = (K)*mx;
    const K w = (K)((kmax - kmin) / kLinBuckets + 1);
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x)
        atomicAdd(&hist[(uint32_t)((a[i] - kmin) / w)], 1ull);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double mean = (double)m / kLinBuckets;
        const double lim = mean + 0.25 * mean + 4.0 * sqrt(mean);
        unsigned long long top = 0;
        for (uint32_t b = 0; b < kLinBuckets; b++) top = max(top, hist[b]);
        *flag = (double)top <= lim;
    }
    __syncthreads();
    *kmin_out = kmin;
    *kmax_out = kmax;
    return *flag != 0;
}

__device__ __forceinline__ uint32_t next_pow2_u32(uint32_t v) {
    if (v <= 1) return 1;
    return 1u << (32 - __clz(v - 1));
}

}  // namespace tbj
