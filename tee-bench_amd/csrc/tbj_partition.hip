// Partitioning kernels: tile histograms, device-wide scan, LDS-staged
// scatter, and the learned CDF model (sample + fit).
//
// The first pass restates parallel_radix_partition (radix_join.c:629-703):
// per-chunk histograms, exclusive offsets, scatter.  Here a chunk is a tile
// of TILE tuples owned by one workgroup; the offsets of all (bin, tile)
// pairs come from one bin-major device-wide scan, so every partition is one
// contiguous run and a tile writes each bin as one coalesced run staged in
// LDS (the role of the reference's software write-combining buffers,
// radix_join.c:733-827).  The second pass restates serial_radix_partition
// (radix_join.c:573-619) with every first-pass partition as its own segment.
#include "tbj_device.h"
#include "tbj_kernels.h"

namespace tbj {

// ---------------------------------------------------------------------------
// Partition function

__device__ __forceinline__ uint32_t model_eval(const uint32_t* __restrict__ xs,
                                               const uint32_t* __restrict__ ys,
                                               const uint64_t* __restrict__ sl, uint32_t nseg,
                                               uint32_t frac, uint32_t key) {
    // largest i in [0, nseg) with xs[i] <= key (xs[0] <= key is ensured by clamping)
    uint32_t lo = 0, hi = nseg;  // invariant: xs[lo] <= key, answer < hi
    if (key < xs[0]) return 0;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (xs[mid] <= key) lo = mid;
        else hi = mid;
    }
    uint64_t d = (uint64_t)(key - xs[lo]);
    uint64_t s = sl[lo];
    // (d * s) >> 32 with a 128-bit product; the result stays below y[lo+1]-y[lo]
    uint64_t lo64 = d * s;
    uint64_t hi64 = __umul64hi(d, s);
    uint64_t add = (hi64 << 32) | (lo64 >> 32);
    uint64_t v = (uint64_t)ys[lo] + add;
    uint64_t vmax = (uint64_t)ys[lo + 1] - (ys[lo + 1] > ys[lo] ? 1 : 0);
    if (v > vmax) v = vmax;
    return (uint32_t)(v >> frac);
}

// LDS copy of a CdfModel's knots (at most kLdsKnots segments).
struct LdsModel {
    uint32_t x[kLdsKnots + 1];
    uint32_t y[kLdsKnots + 1];
    uint64_t s[kLdsKnots];
};

__device__ __forceinline__ void load_model(LdsModel* lm, const CdfModel* m) {
    const uint32_t n = m->nseg;
    for (uint32_t i = threadIdx.x; i <= n; i += blockDim.x) {
        lm->x[i] = i < n ? (uint32_t)m->x[i] : 0xFFFFFFFFu;
        lm->y[i] = m->y[i];
        if (i < n) lm->s[i] = m->slope[i];
    }
}

struct BinFn {
    BinSpec bs;
    const LdsModel* lm;
    uint32_t nseg, frac, kmin, kmax;

    __device__ __forceinline__ uint32_t operator()(uint32_t key) const {
        uint32_t full;
        if (bs.kind == kBinRadix) {
            full = bs.bits >= 32 ? key : (key & ((1u << bs.bits) - 1));
        } else {
            if (bs.drop && (key < kmin || key > kmax)) return TBJ_DROP;
            full = model_eval(lm->x, lm->y, lm->s, nseg, frac, key);
        }
        return (full >> bs.pass_shift) & bs.pass_mask;
    }
};

__device__ __forceinline__ BinFn make_binfn(const BinSpec& bs, LdsModel* lm) {
    BinFn f;
    f.bs = bs;
    f.lm = lm;
    f.nseg = f.frac = f.kmin = f.kmax = 0;
    if (bs.kind == kBinLearned) {
        load_model(lm, bs.model);
        f.nseg = bs.model->nseg;
        f.frac = bs.model->frac;
        f.kmin = bs.model->kmin;
        f.kmax = bs.model->kmax;
    }
    return f;
}

// ---------------------------------------------------------------------------
// Histogram: one workgroup per tile, LDS-privatised counts, written
// bin-major per segment.  With minmax != NULL the tile also folds its key range
// into minmax[0..1] (used to pin the learned model's exact R bounds).

__global__ __launch_bounds__(kPartBlock) void k_hist(const uint2* __restrict__ in,
                                                     const TileDesc* __restrict__ tiles,
                                                     const SegDesc* __restrict__ segs,
                                                     const uint32_t* __restrict__ ntiles_dev,
                                                     uint32_t tile_base, uint32_t* __restrict__ hist,
                                                     BinSpec bs, uint32_t nbins,
                                                     uint32_t* __restrict__ minmax) {
    const uint32_t t = blockIdx.x + tile_base;
    if (t >= *ntiles_dev) return;
    __shared__ uint32_t cnt[kMaxBins];
    extern __shared__ uint8_t dyn_lds[];
    LdsModel& lm = *reinterpret_cast<LdsModel*>(dyn_lds);
    __shared__ uint32_t wmin[kPartBlock / TBJ_WAVE], wmax[kPartBlock / TBJ_WAVE];
    for (uint32_t i = threadIdx.x; i < nbins; i += kPartBlock) cnt[i] = 0;
    BinFn bin = make_binfn(bs, &lm);
    __syncthreads();
    const TileDesc td = tiles[t];
    uint32_t kmn = 0xFFFFFFFFu, kmx = 0;
#pragma unroll 4
    for (uint32_t k = 0; k < kPartItems; k++) {
        uint32_t j = k * kPartBlock + threadIdx.x;
        if (j < td.len) {
            uint32_t key = in[td.in_off + j].x;
            kmn = min(kmn, key);
            kmx = max(kmx, key);
            uint32_t b = bin(key);
            if (b != TBJ_DROP) atomicAdd(&cnt[b], 1u);
        }
    }
    if (minmax) {
        for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) {
            kmn = min(kmn, (uint32_t)__shfl_xor(kmn, d, TBJ_WAVE));
            kmx = max(kmx, (uint32_t)__shfl_xor(kmx, d, TBJ_WAVE));
        }
        if (lane_id() == 0) {
            wmin[wave_id()] = kmn;
            wmax[wave_id()] = kmx;
        }
    }
    __syncthreads();
    if (minmax && threadIdx.x == 0) {
        for (int w = 1; w < kPartBlock / TBJ_WAVE; w++) {
            kmn = min(kmn, wmin[w]);
            kmx = max(kmx, wmax[w]);
        }
        kmn = min(kmn, wmin[0]);
        kmx = max(kmx, wmax[0]);
        atomicMin(&minmax[0], kmn);
        atomicMax(&minmax[1], kmx);
    }
    const SegDesc sd = segs[td.seg];
    const uint32_t tis = t - sd.first_tile;
    for (uint32_t i = threadIdx.x; i < nbins; i += kPartBlock)
        hist[sd.hist_base + (uint64_t)i * sd.ntiles + tis] = cnt[i];
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of a u32 array (reduce, scan partials, downsweep).

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ a,
                                                            uint64_t n,
                                                            uint32_t* __restrict__ partial) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint32_t s = 0;
    for (uint32_t k = 0; k < kScanItems; k++) {
        uint64_t i = base + (uint64_t)k * kScanBlock + threadIdx.x;
        if (i < n) s += a[i];
    }
    uint32_t total;
    (void)block_exclusive_scan<kScanBlock>(s, scratch, &total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the partial sums, carried across rounds.
__global__ __launch_bounds__(kScanBlock) void k_scan_partials(uint32_t* __restrict__ partial,
                                                              uint32_t np,
                                                              uint32_t* __restrict__ grand_total) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < np; base += kScanBlock) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < np ? partial[i] : 0;
        uint32_t total;
        uint32_t ex = block_exclusive_scan<kScanBlock>(v, scratch, &total);
        if (i < np) partial[i] = carry + ex;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0 && grand_total) *grand_total = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(uint32_t* __restrict__ a, uint64_t n,
                                                          const uint32_t* __restrict__ partial) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    // each thread scans kScanItems consecutive elements
    uint32_t v[kScanItems];
    uint32_t s = 0;
    const uint64_t tb = base + (uint64_t)threadIdx.x * kScanItems;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; k++) {
        uint64_t i = tb + k;
        v[k] = i < n ? a[i] : 0;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<kScanBlock>(s, scratch, &total) + partial[blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; k++) {
        uint64_t i = tb + k;
        if (i < n) a[i] = run;
        run += v[k];
    }
}

// ---------------------------------------------------------------------------
// Scatter: load the tile, rank every tuple inside its bin with LDS atomics,
// stage the tile bin-sorted in LDS, then stream each bin's run to its global
// offset.  Consecutive lanes write consecutive slots of the same run.

__global__ __launch_bounds__(kPartBlock) void k_scatter(const uint2* __restrict__ in,
                                                        uint2* __restrict__ out,
                                                        const TileDesc* __restrict__ tiles,
                                                        const SegDesc* __restrict__ segs,
                                                        const uint32_t* __restrict__ ntiles_dev,
                                                        uint32_t tile_base,
                                                        const uint32_t* __restrict__ scanned,
                                                        BinSpec bs, uint32_t nbins) {
    const uint32_t t = blockIdx.x + tile_base;
    if (t >= *ntiles_dev) return;
    __shared__ uint2 stage[kTile];
    __shared__ uint16_t sbin[kTile];
    __shared__ uint32_t loff[kMaxBins];
    __shared__ uint32_t goff[kMaxBins];
    __shared__ uint32_t scratch[kPartBlock / TBJ_WAVE + 1];
    extern __shared__ uint8_t dyn_lds[];
    LdsModel& lm = *reinterpret_cast<LdsModel*>(dyn_lds);
    for (uint32_t i = threadIdx.x; i < nbins; i += kPartBlock) loff[i] = 0;
    BinFn bin = make_binfn(bs, &lm);
    __syncthreads();
    const TileDesc td = tiles[t];
    uint2 v[kPartItems];
    uint32_t b[kPartItems];
    uint32_t rank[kPartItems];
#pragma unroll
    for (uint32_t k = 0; k < kPartItems; k++) {
        uint32_t j = k * kPartBlock + threadIdx.x;
        b[k] = TBJ_DROP;
        if (j < td.len) v[k] = in[td.in_off + j];
    }
#pragma unroll
    for (uint32_t k = 0; k < kPartItems; k++) {
        uint32_t j = k * kPartBlock + threadIdx.x;
        if (j < td.len) {
            b[k] = bin(v[k].x);
            if (b[k] != TBJ_DROP) rank[k] = atomicAdd(&loff[b[k]], 1u);
        }
    }
    __syncthreads();
    const uint32_t nvalid = block_scan_lds<kPartBlock>(loff, nbins, scratch);
    const SegDesc sd = segs[td.seg];
    const uint32_t tis = t - sd.first_tile;
    for (uint32_t i = threadIdx.x; i < nbins; i += kPartBlock)
        goff[i] = scanned[sd.hist_base + (uint64_t)i * sd.ntiles + tis] - loff[i];
#pragma unroll
    for (uint32_t k = 0; k < kPartItems; k++) {
        if (b[k] != TBJ_DROP) {
            uint32_t slot = loff[b[k]] + rank[k];
            stage[slot] = v[k];
            sbin[slot] = (uint16_t)b[k];
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nvalid; j += kPartBlock) {
        uint32_t bb = sbin[j];
        out[goff[bb] + j] = stage[j];
    }
}

// ---------------------------------------------------------------------------
// Second-pass work descriptors: each (relation, first-pass partition) becomes
// a segment cut into tiles.  One workgroup; npart1 <= kMaxBins.

__global__ __launch_bounds__(kScanBlock) void k_make_segments(
    const uint32_t* __restrict__ scanned1, const SegDesc* __restrict__ segs1, uint32_t nseg1,
    uint32_t nbins1, uint32_t nbins2, SegDesc* __restrict__ segs2, TileDesc* __restrict__ tiles2,
    uint32_t* __restrict__ ntiles2_dev, uint32_t max_tiles2) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    __shared__ uint32_t carry_s;
    const uint32_t nseg2 = nseg1 * nbins1;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nseg2; base += kScanBlock) {
        uint32_t s2 = base + threadIdx.x;
        uint32_t st = 0, n = 0, nt = 0;
        if (s2 < nseg2) {
            uint32_t s1 = s2 / nbins1, b = s2 % nbins1;
            const SegDesc sd = segs1[s1];
            st = scanned1[sd.hist_base + (uint64_t)b * sd.ntiles];
            uint32_t en = scanned1[sd.hist_base + (uint64_t)(b + 1) * sd.ntiles];
            n = en - st;
            nt = (n + kTile - 1) / kTile;
        }
        uint32_t total;
        uint32_t ex = block_exclusive_scan<kScanBlock>(nt, scratch, &total);
        uint32_t first = carry_s + ex;
        if (s2 < nseg2) {
            SegDesc d;
            d.in_off = st;
            d.n = n;
            d.ntiles = nt;
            d.first_tile = first;
            d.hist_base = (uint64_t)first * nbins2;
            segs2[s2] = d;
            for (uint32_t k = 0; k < nt && first + k < max_tiles2; k++) {
                TileDesc td;
                td.in_off = st + k * kTile;
                td.len = min((uint32_t)kTile, n - k * kTile);
                td.seg = s2;
                td.pad = 0;
                tiles2[first + k] = td;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) carry_s += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *ntiles2_dev = carry_s;
}

// ---------------------------------------------------------------------------
// Learned partitioner: strided key sample of R, then a one-workgroup fit of
// an equal-mass piecewise-linear CDF (knots at sample quantiles).

__global__ void k_sample_keys(const uint2* __restrict__ in, uint64_t n, uint32_t m,
                              uint32_t* __restrict__ sample) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    // evenly spaced positions with a per-slot offset inside the stride
    uint64_t stride = n / m;
    uint64_t pos = (uint64_t)i * stride + ((uint64_t)i * 2654435761u) % (stride ? stride : 1);
    if (pos >= n) pos = n - 1;
    sample[i] = in[pos].x;
}

// Bitonic sort of m (power of two, <= kFitMax) keys in LDS, then knots.
// Segment count nseg <= kLdsKnots.  Partition space: 2^bits partitions with
// frac fixed-point bits (bits + frac == 32).
__global__ __launch_bounds__(kFitBlock) void k_fit_cdf(const uint32_t* __restrict__ sample,
                                                       uint32_t m, uint32_t nseg_req,
                                                       uint32_t bits, CdfModel* __restrict__ model) {
    __shared__ uint32_t a[kFitMax];
    __shared__ uint32_t knots[kLdsKnots + 1];
    __shared__ uint32_t nk_s;
    for (uint32_t i = threadIdx.x; i < m; i += kFitBlock) a[i] = sample[i];
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < m; i += kFitBlock) {
                uint32_t ixj = i ^ j;
                if (ixj > i) {
                    bool up = (i & k) == 0;
                    uint32_t x = a[i], y = a[ixj];
                    if ((x > y) == up) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        // distinct quantile knots; the first knot is the sample minimum
        uint32_t nk = 0;
        for (uint32_t j = 0; j < nseg_req; j++) {
            uint32_t q = a[(uint64_t)j * m / nseg_req];
            if (nk == 0 || q > knots[nk - 1]) knots[nk++] = q;
        }
        nk_s = nk;
    }
    __syncthreads();
    const uint32_t nseg = nk_s;
    const uint32_t frac = 32 - bits;
    const uint64_t span = (uint64_t)1 << 32;  // y spans [0, 2^32): 2^bits partitions
    for (uint32_t i = threadIdx.x; i <= nseg; i += kFitBlock) {
        uint64_t xi = i < nseg ? (uint64_t)knots[i] : ((uint64_t)a[m - 1] + 1);
        uint64_t yi = (span * i) / nseg;
        if (yi > 0xFFFFFFFFull) yi = 0xFFFFFFFFull;
        model->x[i] = xi;
        model->y[i] = (uint32_t)yi;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nseg; i += kFitBlock) {
        uint64_t dx = model->x[i + 1] - model->x[i];
        uint64_t dy = (uint64_t)model->y[i + 1] - model->y[i];
        model->slope[i] = dx ? ((dy << 32) / dx) : 0;
    }
    if (threadIdx.x == 0) {
        model->nseg = nseg;
        model->frac = frac;
    }
}

// Fold the exact R key range found by the histogram pass into the model.
__global__ void k_set_bounds(CdfModel* model, const uint32_t* minmax, uint32_t drop) {
    model->kmin = minmax[0];
    model->kmax = minmax[1];
    model->drop_outside = drop;
}

// First-pass work descriptors: segment 0 = R, segment 1 = S, tiles of R
// first.  Tile offsets are relative to the owning relation's base pointer.
__global__ void k_make_tiles1(TileDesc* __restrict__ tiles, SegDesc* __restrict__ segs,
                              uint32_t nR, uint32_t nS, uint32_t nbins,
                              uint32_t* __restrict__ ntiles_dev) {
    const uint32_t tR = (nR + kTile - 1) / kTile, tS = (nS + kTile - 1) / kTile;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tR) {
        TileDesc d;
        d.in_off = i * kTile;
        d.len = min((uint32_t)kTile, nR - i * kTile);
        d.seg = 0;
        d.pad = 0;
        tiles[i] = d;
    } else if (i < tR + tS) {
        const uint32_t k = i - tR;
        TileDesc d;
        d.in_off = k * kTile;
        d.len = min((uint32_t)kTile, nS - k * kTile);
        d.seg = 1;
        d.pad = 0;
        tiles[i] = d;
    }
    if (i == 0) {
        SegDesc a, b;
        a.hist_base = 0;
        a.in_off = 0;
        a.n = nR;
        a.ntiles = tR;
        a.first_tile = 0;
        b.hist_base = (uint64_t)nbins * tR;
        b.in_off = 0;
        b.n = nS;
        b.ntiles = tS;
        b.first_tile = tR;
        segs[0] = a;
        segs[1] = b;
        *ntiles_dev = tR + tS;
    }
}

// Zero hist[i] for start <= i < n where start = mult * (*count_dev).
__global__ void k_zero_tail(uint32_t* __restrict__ hist, uint64_t n,
                            const uint32_t* __restrict__ count_dev, uint32_t mult) {
    const uint64_t start = (uint64_t)mult * (*count_dev);
    const uint64_t i = start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hist[i] = 0;
}

// ---------------------------------------------------------------------------
// Host launchers

size_t lds_model_bytes() { return sizeof(LdsModel); }

static inline size_t model_lds(const BinSpec& bs) {
    return bs.kind == kBinLearned ? sizeof(LdsModel) : 0;
}

void launch_hist(hipStream_t st, uint32_t grid, const uint2* in, const TileDesc* tiles,
                 const SegDesc* segs, const uint32_t* ntiles_dev, uint32_t tile_base,
                 uint32_t* hist, const BinSpec& bs, uint32_t nbins, uint32_t* minmax) {
    if (!grid) return;
    hipLaunchKernelGGL(k_hist, dim3(grid), dim3(kPartBlock), model_lds(bs), st, in, tiles, segs,
                       ntiles_dev, tile_base, hist, bs, nbins, minmax);
}

void launch_scatter(hipStream_t st, uint32_t grid, const uint2* in, uint2* out,
                    const TileDesc* tiles, const SegDesc* segs, const uint32_t* ntiles_dev,
                    uint32_t tile_base, const uint32_t* scanned, const BinSpec& bs,
                    uint32_t nbins) {
    if (!grid) return;
    hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(kPartBlock), model_lds(bs), st, in, out, tiles,
                       segs, ntiles_dev, tile_base, scanned, bs, nbins);
}

uint32_t scan_partials(uint64_t n) { return (uint32_t)((n + kScanChunk - 1) / kScanChunk) + 1; }

void launch_scan(hipStream_t st, uint32_t* a, uint64_t n, uint32_t* partial,
                 uint32_t* grand_total) {
    const uint32_t nb = (uint32_t)((n + kScanChunk - 1) / kScanChunk);
    if (nb == 0) return;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, st, a, n, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanBlock), 0, st, partial, nb,
                       grand_total);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanBlock), 0, st, a, n, partial);
}

void launch_make_segments(hipStream_t st, const uint32_t* scanned1, const SegDesc* segs1,
                          uint32_t nseg1, uint32_t nbins1, uint32_t nbins2, SegDesc* segs2,
                          TileDesc* tiles2, uint32_t* ntiles2_dev, uint32_t max_tiles2) {
    hipLaunchKernelGGL(k_make_segments, dim3(1), dim3(kScanBlock), 0, st, scanned1, segs1, nseg1,
                       nbins1, nbins2, segs2, tiles2, ntiles2_dev, max_tiles2);
}

void launch_sample(hipStream_t st, const uint2* in, uint64_t n, uint32_t m, uint32_t* sample) {
    hipLaunchKernelGGL(k_sample_keys, dim3((m + 255) / 256), dim3(256), 0, st, in, n, m, sample);
}

void launch_fit(hipStream_t st, const uint32_t* sample, uint32_t m, uint32_t nseg, uint32_t bits,
                CdfModel* model) {
    hipLaunchKernelGGL(k_fit_cdf, dim3(1), dim3(kFitBlock), 0, st, sample, m, nseg, bits, model);
}

void launch_set_bounds(hipStream_t st, CdfModel* model, const uint32_t* minmax, uint32_t drop) {
    hipLaunchKernelGGL(k_set_bounds, dim3(1), dim3(1), 0, st, model, minmax, drop);
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_make_tiles1(hipStream_t st, TileDesc* tiles, SegDesc* segs, uint32_t nR, uint32_t nS,
                        uint32_t nbins, uint32_t* ntiles_dev) {
    const uint32_t t = (nR + kTile - 1) / kTile + (nS + kTile - 1) / kTile;
    const uint32_t grid = t ? (t + 255) / 256 : 1;
    hipLaunchKernelGGL(k_make_tiles1, dim3(grid), dim3(256), 0, st, tiles, segs, nR, nS, nbins,
                       ntiles_dev);
}

void launch_zero_tail(hipStream_t st, uint32_t* hist, uint64_t n, const uint32_t* count_dev,
                      uint32_t mult, uint64_t max_tail) {
    if (!max_tail) return;
    hipLaunchKernelGGL(k_zero_tail, dim3((uint32_t)((max_tail + 255) / 256)), dim3(256), 0, st,
                       hist, n, count_dev, mult);
}

void launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

}  // namespace tbj
