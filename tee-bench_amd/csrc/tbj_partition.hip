// Partitioning kernels: tile histograms, device-wide scan, LDS-staged
// scatter, and the learned CDF model (sample + fit).
//
// The first pass restates parallel_radix_partition (radix_join.c:629-703):
// per-chunk histograms, exclusive offsets, scatter.  Here a chunk is a tile
// of TILE tuples owned by one workgroup; each tile writes its bin counts as
// one contiguous row, and the offsets of all (bin, tile) pairs follow the
// bin-major order of a device-wide scan over 32-tile group sums, so every
// partition is one contiguous run and a tile writes each bin as one
// coalesced run staged in LDS (the role of the reference's software
// write-combining buffers, radix_join.c:733-827).  The second pass restates serial_radix_partition
// (radix_join.c:573-619) with every first-pass partition as its own segment.
#include "tbj_device.h"
#include "tbj_kernels.h"

namespace tbj {

// ---------------------------------------------------------------------------
// Partition function

// LDS copy of a CdfModel's cell table.  Within cell c a key at offset d
// from the cell start maps to cells[c] + (d * (cells[c+1] - cells[c]) >>
// cshift) (1/256 partitions): d < 2^cshift keeps the value below
// cells[c+1], so the map is monotone inside and across cells, and R and S
// go through the same code, so equal keys always land in the same partition.
struct LdsModel {
    uint32_t cells[kCells + 1];
};

__device__ __forceinline__ void load_model(LdsModel* lm, const CdfModel* m) {
    for (uint32_t i = threadIdx.x; i <= m->ncells; i += blockDim.x) lm->cells[i] = m->cells[i];
}

struct BinFn {
    BinSpec bs;
    const LdsModel* lm;
    uint32_t kmin, kmax, cshift, ncells, last;
    uint32_t dlo, dhi;  // keys outside [dlo, dhi] are dropped (bs.drop)
    uint32_t lin;       // the linear model: ((key - kmin) * lmul) >> 32
    uint64_t lmul;

    __device__ __forceinline__ uint32_t linear(uint32_t key) const {
        return key < kmin ? 0 : key > kmax ? last
                                           : (uint32_t)(((uint64_t)(key - kmin) * lmul) >> 32);
    }

    // The key's partition among all 2^bits (both passes), or TBJ_DROP.
    __device__ __forceinline__ uint32_t full(uint32_t key) const {
        if (bs.kind == kBinRadix) return bs.bits >= 32 ? key : (key & ((1u << bs.bits) - 1));
        if (bs.drop && (key < dlo || key > dhi)) return TBJ_DROP;
        if (lin) return linear(key);
        if (key < kmin) return 0;
        const uint32_t d = key - kmin, c = d >> cshift;
        if (c >= ncells) return last;
        const uint32_t y0 = lm->cells[c], y1 = lm->cells[c + 1];
        const uint32_t off = d & ((1u << cshift) - 1);
        return (y0 + (uint32_t)(((uint64_t)off * (y1 - y0)) >> cshift)) >> 8;
    }

    __device__ __forceinline__ uint32_t operator()(uint32_t key) const {
        const uint32_t f = full(key);
        return f == TBJ_DROP ? TBJ_DROP : (f >> bs.pass_shift) & bs.pass_mask;
    }

    // full() of N keys at once (TBJ_DROP for dropped keys and for items with
    // valid[q] false).  The learned model's cell reads of all N keys are
    // issued before any is used (global loads when GMODEL), so their round
    // trips overlap; branch-free, the same values as full().
    template <bool GMODEL, int N>
    __device__ __forceinline__ void fulls(const uint32_t (&key)[N], const bool (&valid)[N],
                                          uint32_t (&out)[N]) const {
        // the map is the same for every key: chosen once, then straight-line
        // code for all N keys (not N copies of full()'s branches)
        if (bs.kind == kBinRadix) {
            const uint32_t mask = bs.bits >= 32 ? 0xFFFFFFFFu : (1u << bs.bits) - 1;
#pragma unroll
            for (int q = 0; q < N; q++) out[q] = valid[q] ? (key[q] & mask) : TBJ_DROP;
            return;
        }
        if (lin) {
#pragma unroll
            for (int q = 0; q < N; q++) {
                const uint32_t k = key[q];
                const bool drop = !valid[q] || (bs.drop && (k < dlo || k > dhi));
                out[q] = drop ? TBJ_DROP : linear(k);
            }
            return;
        }
        typedef __attribute__((address_space(1))) const uint32_t gu32;
        const uint32_t cmask = (1u << cshift) - 1;
        uint32_t c[N], y0[N], y1[N];
#pragma unroll
        for (int q = 0; q < N; q++) {
            const uint32_t d = key[q] - kmin;
            c[q] = key[q] < kmin ? 0 : min(d >> cshift, ncells - 1);
        }
#pragma unroll
        for (int q = 0; q < N; q++) {
            if (GMODEL) {
                gu32* g = (gu32*)lm->cells;
                y0[q] = g[c[q]];
                y1[q] = g[c[q] + 1];
            } else {
                y0[q] = lm->cells[c[q]];
                y1[q] = lm->cells[c[q] + 1];
            }
        }
#pragma unroll
        for (int q = 0; q < N; q++) {
            const uint32_t k = key[q], d = k - kmin;
            uint32_t f = (y0[q] + (uint32_t)(((uint64_t)(d & cmask) * (y1[q] - y0[q])) >> cshift)) >> 8;
            f = k < kmin ? 0 : (d >> cshift) >= ncells ? last : f;
            const bool drop = !valid[q] || (bs.drop && (k < dlo || k > dhi));
            out[q] = drop ? TBJ_DROP : f;
        }
    }

    // The pass's bins of N keys (fulls, then this pass's bits).
    template <bool GMODEL, int N>
    __device__ __forceinline__ void bins(const uint32_t (&key)[N], const bool (&valid)[N],
                                         uint32_t (&out)[N]) const {
        fulls<GMODEL>(key, valid, out);
#pragma unroll
        for (int q = 0; q < N; q++)
            out[q] = out[q] == TBJ_DROP ? TBJ_DROP : (out[q] >> bs.pass_shift) & bs.pass_mask;
    }
};

// GMODEL = false: loads the cell table into LDS (no barrier: the caller's
// next barrier covers it).  GMODEL = true: reads it in place from the model
// (global memory, cached), so that one tile per workgroup needs no copy.
// A template parameter, so that each kernel's model reads keep one address
// space (LDS reads stay ds_read).
template <bool GMODEL = false>
__device__ __forceinline__ BinFn make_binfn(const BinSpec& bs, LdsModel* lm) {
    BinFn f;
    f.bs = bs;
    f.lm = GMODEL && bs.kind == kBinLearned ? reinterpret_cast<const LdsModel*>(bs.model->cells)
                                            : lm;
    f.kmin = f.kmax = f.cshift = f.ncells = f.last = f.dlo = f.dhi = f.lin = 0;
    f.lmul = 0;
    if (bs.kind == kBinLearned) {
        f.lin = bs.model->linear;
        f.lmul = bs.model->lmul;
        if (!GMODEL && !f.lin) load_model(lm, bs.model);
        f.kmin = bs.model->kmin;
        f.kmax = bs.model->kmax;
        f.cshift = bs.model->cshift;
        f.ncells = bs.model->ncells;
        f.last = bs.bits >= 32 ? 0xFFFFFFFFu : (1u << bs.bits) - 1;
        f.dlo = bs.range ? bs.range[0] : f.kmin;
        f.dhi = bs.range ? bs.range[1] : f.kmax;
    }
    return f;
}

// ---------------------------------------------------------------------------
// Histogram: one workgroup per tile, LDS-privatised counts, written as the
// tile's row of u16 counts.

// Persistent: each workgroup walks tiles tile_base + blockIdx.x, + gridDim.x,
// ... below tile_end and *ntiles_dev, loading the partition function (the
// learned model) into LDS once.  Tile t writes its counts to row t of hist
// (nbins u16: a tile holds at most kMaxTile tuples).  IT: tuples per thread
// (the tile holds kPartBlock * IT).
template <int IT>
__global__ __launch_bounds__(kPartBlock) void k_hist(const uint2* __restrict__ in,
                                                     const TileDesc* __restrict__ tiles,
                                                     const uint32_t* __restrict__ ntiles_dev,
                                                     uint32_t tile_base, uint32_t tile_end,
                                                     uint16_t* __restrict__ hist, BinSpec bs,
                                                     uint32_t nbins) {
    __shared__ uint32_t cnt[kMaxBins];
    extern __shared__ uint8_t dyn_lds[];
    LdsModel& lm = *reinterpret_cast<LdsModel*>(dyn_lds);
    BinFn bin = make_binfn(bs, &lm);
    const uint32_t end = min(tile_end, *ntiles_dev);
    for (uint32_t t = tile_base + blockIdx.x; t < end; t += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < nbins; i += kPartBlock) cnt[i] = 0;
        __syncthreads();
        const TileDesc td = tiles[t];
        uint32_t key[IT];
#pragma unroll
        for (uint32_t k = 0; k < IT; k++) {
            const uint32_t j = k * kPartBlock + threadIdx.x;
            key[k] = j < td.len ? in[td.in_off + j].x : 0;
        }
#pragma unroll
        for (uint32_t k = 0; k < IT; k++) {
            const uint32_t j = k * kPartBlock + threadIdx.x;
            if (j < td.len) {
                const uint32_t b = bin(key[k]);
                if (b != TBJ_DROP) atomicAdd(&cnt[b], 1u);
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nbins; i += kPartBlock)
            hist[(uint64_t)t * nbins + i] = (uint16_t)cnt[i];
        __syncthreads();
    }
}

// Exact key range of a relation: minmax[0] = min, minmax[1] = max (must start
// at {UINT_MAX, 0}).  Grid-stride over pairs of tuples (16-byte loads when
// the base is 16-byte aligned, which the host checks), a block reduction,
// then one atomic pair per workgroup.
__global__ __launch_bounds__(256) void k_minmax(const uint2* __restrict__ in, uint64_t n,
                                                uint32_t* __restrict__ minmax) {
    __shared__ uint32_t red[2][256 / TBJ_WAVE];
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    const uint64_t npair = n / 2;
    const uint4* in4 = reinterpret_cast<const uint4*>(in);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < npair; i += 4 * stride) {  // four 16-byte loads in flight
        uint4 v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = in4[i + q * stride];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            mn = min(mn, min(v[q].x, v[q].z));
            mx = max(mx, max(v[q].x, v[q].z));
        }
    }
    for (; i < npair; i += stride) {
        const uint4 v = in4[i];
        mn = min(mn, min(v.x, v.z));
        mx = max(mx, max(v.x, v.z));
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        mn = min(mn, in[n - 1].x);
        mx = max(mx, in[n - 1].x);
    }
    for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, d, TBJ_WAVE));
        mx = max(mx, (uint32_t)__shfl_xor(mx, d, TBJ_WAVE));
    }
    if (lane_id() == 0) {
        red[0][wave_id()] = mn;
        red[1][wave_id()] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / TBJ_WAVE; w++) {
            mn = min(mn, red[0][w]);
            mx = max(mx, red[1][w]);
        }
        atomicMin(&minmax[0], mn);
        atomicMax(&minmax[1], mx);
    }
}

// minmax[0] = least of the n tiles' minima, minmax[1] = greatest of their
// maxima (one workgroup).
__global__ __launch_bounds__(1024) void k_range_reduce(const uint2* __restrict__ r, uint32_t n,
                                                       uint32_t* __restrict__ minmax) {
    __shared__ uint2 red[1024 / TBJ_WAVE];
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 1024 * 8) {  // eight loads in flight
        uint2 v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t i = i0 + k * 1024 + threadIdx.x;
            v[k] = i < n ? r[i] : make_uint2(0xFFFFFFFFu, 0);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            mn = min(mn, v[k].x);
            mx = max(mx, v[k].y);
        }
    }
    for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, d, TBJ_WAVE));
        mx = max(mx, (uint32_t)__shfl_xor(mx, d, TBJ_WAVE));
    }
    if (lane_id() == 0) red[wave_id()] = make_uint2(mn, mx);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 1024 / TBJ_WAVE; w++) {
            mn = min(mn, red[w].x);
            mx = max(mx, red[w].y);
        }
        minmax[0] = mn;
        minmax[1] = mx;
    }
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

// Small arrays (n <= kScanSingleMax): the whole exclusive scan in one
// workgroup, in rounds of kScanSingleBlock * 8 words carried forward.
constexpr int kScanSingleBlock = 1024;
constexpr uint32_t kScanSingleItems = 8;
constexpr uint32_t kScanSingleRound = kScanSingleBlock * kScanSingleItems;
constexpr uint64_t kScanSingleMax = 4ull * kScanSingleRound;

__global__ __launch_bounds__(kScanSingleBlock) void k_scan_single(uint32_t* __restrict__ a,
                                                                  uint64_t n,
                                                                  uint32_t* __restrict__ grand_total) {
    __shared__ uint32_t scratch[kScanSingleBlock / TBJ_WAVE + 1];
    __shared__ uint32_t buf[kScanSingleRound + kScanSingleRound / 32];  // padded
    uint32_t carry = 0;
    for (uint64_t base = 0; base < n; base += kScanSingleRound) {
#pragma unroll
        for (uint32_t k = 0; k < kScanSingleItems; k++) {
            const uint32_t j = k * kScanSingleBlock + threadIdx.x;
            buf[j + (j >> 5)] = base + j < n ? a[base + j] : 0;
        }
        __syncthreads();
        uint32_t v[kScanSingleItems], s = 0;
#pragma unroll
        for (uint32_t k = 0; k < kScanSingleItems; k++) {
            const uint32_t j = threadIdx.x * kScanSingleItems + k;
            v[k] = buf[j + (j >> 5)];
            s += v[k];
        }
        uint32_t total;
        uint32_t run = carry + block_exclusive_scan<kScanSingleBlock>(s, scratch, &total);
#pragma unroll
        for (uint32_t k = 0; k < kScanSingleItems; k++) {
            const uint32_t j = threadIdx.x * kScanSingleItems + k;
            buf[j + (j >> 5)] = run;
            run += v[k];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kScanSingleItems; k++) {
            const uint32_t j = k * kScanSingleBlock + threadIdx.x;
            if (base + j < n) a[base + j] = buf[j + (j >> 5)];
        }
        carry += total;
        __syncthreads();  // buf is refilled by the next round
    }
    if (threadIdx.x == 0 && grand_total) *grand_total = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(uint32_t* __restrict__ a, uint64_t n,
                                                          const uint32_t* __restrict__ partial) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    __shared__ uint32_t buf[kScanChunk + kScanChunk / 32];  // padded against bank conflicts
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    // coalesced load into LDS; each thread then scans kScanItems consecutive words
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; k++) {
        const uint32_t j = k * kScanBlock + threadIdx.x;
        const uint64_t i = base + j;
        buf[j + (j >> 5)] = i < n ? a[i] : 0;
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; k++) {
        const uint32_t j = threadIdx.x * kScanItems + k;
        v[k] = buf[j + (j >> 5)];
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan<kScanBlock>(s, scratch, &total) + partial[blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; k++) {
        const uint32_t j = threadIdx.x * kScanItems + k;
        buf[j + (j >> 5)] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kScanItems; k++) {
        const uint32_t j = k * kScanBlock + threadIdx.x;
        const uint64_t i = base + j;
        if (i < n) a[i] = buf[j + (j >> 5)];
    }
}

// ---------------------------------------------------------------------------
// Scatter: load the tile, rank every tuple inside its bin with LDS atomics,
// stage the tile bin-sorted in LDS, then stream each bin's run to its global
// offset.  Consecutive lanes write consecutive slots of the same run.
//
// One tile per workgroup, blocks mapped to tiles by xcd_tile.

// Dynamic LDS of k_scatter: the learned model (if any), then the bin
// offsets, sized by the pass's bin count so that a small pass leaves room
// for a third workgroup per CU.
__host__ __device__ constexpr size_t scatter_model_bytes(bool learned) {
    return learned ? (sizeof(LdsModel) + 15) / 16 * 16 : 0;
}

// Where a claiming scatter puts tile t's run of bin i: region r = cb0 +
// tiles[t].aux * cmul + i * cstride owns [starts[r], starts[r + 1]); the run
// takes its place with one atomicAdd on cursors[r].  A run that would pass
// the region's end sets *overflow and writes only what fits (the caller then
// redoes the join with exact histograms).
struct Claims {
    const uint32_t* starts;
    uint32_t* cursors;
    uint32_t* overflow;
    uint32_t cb0, cmul, cstride;
    uint32_t xmul;  // r += xmul * the tile's XCD eighth (block index mod 8)
    const uint32_t* agg_flag;  // AGG only while *agg_flag (NULL: always)
    uint2* tile_range;  // non-null: {min, max} key of each tile, by tile - tile_base
};

// CLAIM = false: offs row t = the output offset of each bin's run of tile t.
// CLAIM = true: the runs claim space in sample-sized regions (cl).
// The global words a tile needs per bin (its offset row, or its regions'
// bounds) are loaded together with the tile, and the claims are issued
// before the LDS scan and staging and consumed after them, so that neither
// round trip sits between the tile's load and its writes.
constexpr int kScatterBinsPerThread = kMaxBins / kPartBlock;

// IT: tuples per thread (the tile holds kPartBlock * IT).  With 4096-tuple
// tiles three workgroups per CU (six waves per SIMD) is what the LDS stage
// allows, and the register budget is held to it; 8192-tuple tiles fit one
// workgroup per CU.
template <bool GMODEL, bool AGG, bool CLAIM, int IT = kPartItems>
__global__ __launch_bounds__(kPartBlock) __attribute__((amdgpu_waves_per_eu(IT > kPartItems ? 2 : 6))) void k_scatter(const uint2* __restrict__ in,
                                                        uint2* __restrict__ out,
                                                        const TileDesc* __restrict__ tiles,
                                                        const uint32_t* __restrict__ ntiles_dev,
                                                        uint32_t tile_base, uint32_t tile_end,
                                                        const uint32_t* __restrict__ offs,
                                                        Claims cl, BinSpec bs, uint32_t nbins) {
    __shared__ uint2 stage[kPartBlock * IT];
    __shared__ uint16_t sbin[kPartBlock * IT];
    __shared__ uint32_t scratch[kPartBlock / TBJ_WAVE + 1];
    __shared__ uint2 wrange[CLAIM ? kPartBlock / TBJ_WAVE : 1];
    extern __shared__ uint8_t dyn_lds[];
    LdsModel& lm = *reinterpret_cast<LdsModel*>(dyn_lds);
    uint32_t* loff = reinterpret_cast<uint32_t*>(
        dyn_lds + scatter_model_bytes(!GMODEL && bs.kind == kBinLearned));
    uint32_t* goff = loff + nbins;
    uint32_t* glim = goff + nbins;  // CLAIM only
    BinFn bin = make_binfn<GMODEL>(bs, &lm);
    // the launch's tiles that exist (the grid may be an upper bound)
    const uint32_t end = min(tile_end, *ntiles_dev);
    const uint32_t n = end > tile_base ? end - tile_base : 0;
    if (blockIdx.x < n) {
        const uint32_t t = tile_base + xcd_tile(blockIdx.x, n);
        for (uint32_t i = threadIdx.x; i < nbins; i += kPartBlock) loff[i] = 0;
        const TileDesc td = tiles[t];
        uint2 v[IT];
        uint32_t b[IT];
        uint32_t rank[IT];
#pragma unroll
        for (uint32_t k = 0; k < IT; k++) {
            const uint32_t j = k * kPartBlock + threadIdx.x;
            b[k] = TBJ_DROP;
            if (j < td.len) v[k] = in[td.in_off + j];
        }
        // per bin of this thread: the tile's offset (exact) or its region's
        // start and end (claims)
        const uint32_t r0 =
            CLAIM ? cl.cb0 + td.aux * cl.cmul + cl.xmul * (blockIdx.x % kNumXcd) : 0;
        uint32_t w0[kScatterBinsPerThread], w1[kScatterBinsPerThread];
#pragma unroll
        for (int q = 0; q < kScatterBinsPerThread; q++) {
            const uint32_t i = threadIdx.x + q * kPartBlock;
            w0[q] = w1[q] = 0;
            if (i < nbins) {
                if (CLAIM) {
                    const uint32_t r = r0 + i * cl.cstride;
                    w0[q] = cl.starts[r];
                    w1[q] = cl.starts[r + 1];
                } else {
                    w0[q] = offs[(uint64_t)t * nbins + i];
                }
            }
        }
        __syncthreads();
        {
            uint32_t keys[IT];
            bool ok[IT];
#pragma unroll
            for (uint32_t k = 0; k < IT; k++) {
                keys[k] = v[k].x;
                ok[k] = k * kPartBlock + threadIdx.x < td.len;
            }
            bin.template bins<GMODEL>(keys, ok, b);
            if (CLAIM && cl.tile_range) {  // the tile's key range, per wave
                uint32_t mn = 0xFFFFFFFFu, mx = 0;
#pragma unroll
                for (uint32_t k = 0; k < IT; k++)
                    if (ok[k]) {
                        mn = min(mn, keys[k]);
                        mx = max(mx, keys[k]);
                    }
                for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) {
                    mn = min(mn, (uint32_t)__shfl_xor(mn, d, TBJ_WAVE));
                    mx = max(mx, (uint32_t)__shfl_xor(mx, d, TBJ_WAVE));
                }
                if (lane_id() == 0) wrange[wave_id()] = make_uint2(mn, mx);
            }
        }
        // the aggregated ranks cost a ballot per item: a sample-sized pass
        // takes them only when its sample saw a hot bin
        const bool agg = AGG && (!CLAIM || !cl.agg_flag || *cl.agg_flag);
        if (agg) {
#pragma unroll
            for (uint32_t k = 0; k < IT; k++) rank[k] = agg_inc(loff, b[k]);
        } else {
#pragma unroll
            for (uint32_t k = 0; k < IT; k++)
                if (b[k] != TBJ_DROP) rank[k] = atomicAdd(&loff[b[k]], 1u);
        }
        __syncthreads();
        if (CLAIM && cl.tile_range && threadIdx.x == 0) {
            uint2 r = wrange[0];
            for (int w = 1; w < kPartBlock / TBJ_WAVE; w++) {
                r.x = min(r.x, wrange[w].x);
                r.y = max(r.y, wrange[w].y);
            }
            cl.tile_range[t - tile_base] = r;
        }
        // claims: one atomic per non-empty bin, issued now, awaited after the
        // scan and the staging
        uint32_t cnt[kScatterBinsPerThread], base[kScatterBinsPerThread];
#pragma unroll
        for (int q = 0; q < kScatterBinsPerThread; q++) {
            const uint32_t i = threadIdx.x + q * kPartBlock;
            cnt[q] = i < nbins ? loff[i] : 0;
            base[q] = 0;
            if (CLAIM && cnt[q]) base[q] = atomicAdd(&cl.cursors[r0 + i * cl.cstride], cnt[q]);
        }
        const uint32_t nvalid = block_scan_lds<kPartBlock>(loff, nbins, scratch);
#pragma unroll
        for (uint32_t k = 0; k < IT; k++) {
            if (b[k] != TBJ_DROP) {
                const uint32_t slot = loff[b[k]] + rank[k];
                stage[slot] = v[k];
                sbin[slot] = (uint16_t)b[k];
            }
        }
#pragma unroll
        for (int q = 0; q < kScatterBinsPerThread; q++) {
            const uint32_t i = threadIdx.x + q * kPartBlock;
            if (i < nbins) {
                if (CLAIM) {
                    const uint32_t cap = w1[q] - w0[q];
                    if (cnt[q] && base[q] + cnt[q] > cap) atomicOr(cl.overflow, 1u);
                    goff[i] = w0[q] + base[q] - loff[i];
                    glim[i] = w0[q] + cap;
                } else {
                    goff[i] = w0[q] - loff[i];
                }
            }
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < nvalid; j += kPartBlock) {
            const uint32_t bb = sbin[j];
            const uint32_t pos = goff[bb] + j;
            if (!CLAIM || pos < glim[bb]) out[pos] = stage[j];
        }
    }
}

// ---------------------------------------------------------------------------
// Sample-sized passes: no histogram read.  One kernel reads 1/16 of the
// input's 16-tuple groups (exactly one group in every 16, at a hashed
// position) and counts, per first-pass tile row, the sampled tuples of each
// first-pass bin (sizing that row's regions of the first pass) and, per
// relation, the sampled tuples of each (first-pass bin, second-pass bin)
// pair (sizing the second pass's regions inside each first-pass partition).


// One workgroup per sample chunk: chunk c of tile row `row` (rows of R first,
// then S; kChunksPerRow chunks per row).  cnt1 (zeroed) accumulates the
// row's sampled counts per first-pass bin at cnt1[pbase[rel] + row * F1 +
// b1] (the first-pass regions are laid out row by row, so that a tile's
// claims touch consecutive words); fine (zeroed, 2 * F1 * F2 words)
// accumulates (rel, b1, b2) counts.  Those only apportion each exact
// first-pass partition size among its second-pass bins, so a quarter of the
// sample serves: a chunk counts the (b1, b2) bins of one quarter of the b1
// values, quarter (row + chunk) mod 4, so that every b1 takes its counts
// from a like share of the chunks and the LDS histogram is a quarter of
// F1 * F2 bins (32 KB: several workgroups per CU).  The counts live in LDS
// as 16-bit halves of 32-bit words: a chunk holds at most kSampleTiles *
// kTile / 16 < 2^16 sampled tuples.
constexpr uint32_t kFineParts = 4;
constexpr uint32_t kFineLds = (1u << 16) / kFineParts;  // (F1 / parts) * F2 at most
constexpr int kSampleBlock = 512;

__global__ __launch_bounds__(kSampleBlock) void k_sample_rows(
    const uint2* __restrict__ R, uint32_t nR, const uint2* __restrict__ S, uint32_t nS,
    uint32_t nrowR, uint32_t nrowS, uint32_t row0, BinSpec bsr, BinSpec bss, BinSpec bs2,
    uint32_t F1, uint32_t F2, uint32_t* __restrict__ cnt1, uint32_t* __restrict__ fine,
    uint32_t* __restrict__ hot1) {
    __shared__ uint32_t h[kFineLds / 2];
    __shared__ uint32_t c1[kMaxBins];
    const uint32_t rowg = row0 + blockIdx.x / kChunksPerRow, chunk = blockIdx.x % kChunksPerRow;
    const uint32_t rel = rowg < nrowR ? 0 : 1;
    const uint32_t row = rel ? rowg - nrowR : rowg;
    const uint2* in = rel ? S : R;
    const uint32_t n = rel ? nS : nR;
    const uint32_t ntiles = (n + kTile - 1) / kTile;
    const uint32_t rpe = rows_per_eighth(ntiles);
    const uint32_t x = row / rpe, c = row % rpe;
    const uint32_t e0 = eighth_start(x, ntiles), e1 = eighth_start(x + 1, ntiles);
    const uint32_t r0 = min(e0 + c * kRowTiles, e1), r1 = min(r0 + kRowTiles, e1);
    const uint32_t t0 = min(r0 + chunk * kSampleTiles, r1), t1 = min(t0 + kSampleTiles, r1);
    if (t0 == t1) return;  // whole workgroup
    const uint32_t nf = F1 * F2;
    // this chunk's quarter of the b1 values (F1 / nq of them)
    const uint32_t nq = F1 < kFineParts ? F1 : kFineParts;
    const uint32_t fq = F1 / nq, qlo = ((rowg + chunk) % nq) * fq;
    const uint32_t nfq = fq * F2;
    for (uint32_t i = threadIdx.x; i < nfq / 2 + 1 && i < kFineLds / 2; i += kSampleBlock) h[i] = 0;
    for (uint32_t i = threadIdx.x; i < F1; i += kSampleBlock) c1[i] = 0;
    __syncthreads();
    // one evaluation of the partition function per key gives both passes'
    // bins (and the drop decision of the first pass)
    BinFn b1f = make_binfn<true>(rel ? bss : bsr, nullptr);
    const uint32_t s1 = b1f.bs.pass_shift, m1 = b1f.bs.pass_mask;
    const uint32_t s2 = bs2.pass_shift, m2 = bs2.pass_mask;
    const uint32_t lo = t0 * kTile, hi = min(t1 * kTile, n);
    // blocks of 16 groups = 256 tuples, one sampled group (128 B) each.  Eight
    // lanes share a block's group, two tuples a lane, so that a load
    // instruction of a wave touches eight lines rather than 64; a wave takes
    // 64 blocks per round and a lane gathers its 16 keys from eight of them.
    const uint32_t blo = lo / 256, bhi = (hi + 255) / 256;
    const uint32_t lane = threadIdx.x % TBJ_WAVE, wv = threadIdx.x / TBJ_WAVE;
    constexpr uint32_t kRound = kSampleBlock / TBJ_WAVE * 64;  // blocks per workgroup round
    for (uint32_t wb = blo + wv * 64; wb < bhi; wb += kRound) {
        uint32_t key[16], full[16];
        bool ok[16];
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) {
            const uint32_t blk = wb + q * 8 + lane / 8;
            const uint32_t i = blk * 256 + sampled_group(blk) * 16 + (lane % 8) * 2;
#pragma unroll
            for (uint32_t e = 0; e < 2; e++) {
                const uint32_t k = 2 * q + e;
                ok[k] = blk < bhi && i + e >= lo && i + e < hi;
                key[k] = in[ok[k] ? i + e : lo].x;
            }
        }
        b1f.template fulls<true>(key, ok, full);
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) {
            if (full[k] != TBJ_DROP) {
                const uint32_t b1 = (full[k] >> s1) & m1;
                if (b1 - qlo < fq) {
                    const uint32_t f = (b1 - qlo) * F2 + ((full[k] >> s2) & m2);
                    atomicAdd(&h[f >> 1], 1u << ((f & 1) * 16));
                }
                atomicAdd(&c1[b1], 1u);
            }
        }
    }
    __syncthreads();
    const uint32_t pbase = rel ? F1 * nrowR : 0;
    // a bin with over 1/8 of the chunk's sampled positions (hi - lo) / 16
    const uint32_t hot_min = max(8u, (hi - lo) / 128);
    for (uint32_t i = threadIdx.x; i < F1; i += kSampleBlock)
        if (c1[i]) {
            atomicAdd(&cnt1[pbase + row * F1 + i], c1[i]);
            if (hot1 && c1[i] > hot_min) atomicOr(hot1, 1u);
        }
    uint32_t* fr = fine + (uint64_t)rel * nf + (uint64_t)qlo * F2;
    for (uint32_t f = threadIdx.x; f < nfq; f += kSampleBlock) {
        const uint32_t v = (h[f >> 1] >> ((f & 1) * 16)) & 0xFFFFu;
        if (v) atomicAdd(&fr[f], v);
    }
}

// Region capacity (tuples) from an estimate e = 16 * sampled count: room for
// 10 % over e, 12 standard deviations of a 1/16 sample of independent tuples
// (48 sqrt(e)), and 640 more -- a region of 640 tuples shows no sampled tuple
// with probability (15/16)^640 < 1e-17, so an empty sample is no evidence of
// an empty region.
constexpr double kCapFloor = 640.0;
__device__ __forceinline__ uint32_t region_cap(double e) {
    return (uint32_t)min(4.0e9, e * 1.1 + 48.0 * sqrt(e) + kCapFloor);
}

// caps[i] = region_cap(16 * cnt[i]) for the n first-pass regions; caps[n] =
// 0 (the scan's sentinel).
__global__ void k_row_caps(const uint32_t* __restrict__ cnt, uint32_t n,
                           uint32_t* __restrict__ caps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) caps[i] = region_cap(16.0 * cnt[i]);
    else if (i == n) caps[i] = 0;
}

// Exact size of each first-pass partition p = rel * F1 + b1 (the sum of its
// rows' filled regions): one wave per partition, its lanes over the rows;
// *total (zeroed) receives the sum over all partitions.
__global__ void k_part_sizes(const uint32_t* __restrict__ starts,
                             const uint32_t* __restrict__ cursors, uint32_t F1, uint32_t nrowR,
                             uint32_t nrowS, uint32_t* __restrict__ npart,
                             uint32_t* __restrict__ total) {
    const uint32_t p = (blockIdx.x * blockDim.x + threadIdx.x) / TBJ_WAVE;
    if (p >= 2 * F1) return;  // whole waves
    const uint32_t rel = p / F1, b1 = p % F1;
    const uint32_t nrow = rel ? nrowS : nrowR;
    const uint32_t r0 = (rel ? F1 * nrowR : 0) + b1;
    uint64_t tot = 0;
    for (uint32_t k = lane_id(); k < nrow; k += TBJ_WAVE) {
        const uint32_t r = r0 + k * F1;
        tot += min(cursors[r], starts[r + 1] - starts[r]);
    }
    tot = wave_reduce_u64(tot);
    if (lane_id() == 0) {
        npart[p] = (uint32_t)tot;
        if (tot) atomicAdd(total, (uint32_t)tot);
    }
}

// Estimated largest first-pass R partition (16 times its sampled count over
// the rows) into stats[0] (zeroed): one thread per R bin.
__global__ void k_est_part_max(const uint32_t* __restrict__ cnt1, uint32_t F1, uint32_t nrowR,
                               uint32_t* __restrict__ stats) {
    const uint32_t b1 = blockIdx.x * blockDim.x + threadIdx.x;
    if (b1 >= F1) return;
    uint32_t tot = 0;
    for (uint32_t r = 0; r < nrowR; r++) tot += cnt1[r * F1 + b1];
    atomicMax(&stats[0], 16 * tot);
}

// Second-pass region caps inside first-pass partition s (one workgroup per
// partition): the ratio estimate of bin b2's share of the exact partition
// size from the sampled (b1, b2) counts; bins with few samples are shrunk
// toward an even split by 8 pseudo-samples, the larger estimate wins.  Each
// region is split into 8 sub-regions, one per XCD eighth of the second pass's
// tiles (so that the claims of a partition whose tiles span several XCDs --
// a hot key's -- are spread over several counters and never cross an XCD),
// each sized by the share of the partition's tiles in that eighth: region
// (s, b2) sub-region x is caps[(s * 8 + x) * nbins2 + b2] (a tile's claims
// then hit consecutive words).  first: the first tile of each first-pass
// region in partition order (scanned), *ntiles2 the tile count.
// caps[8 * nsub] = 0 is the scan's sentinel.
__global__ __launch_bounds__(256) void k_region_caps(const uint32_t* __restrict__ fine,
                                                     const uint32_t* __restrict__ npart,
                                                     uint32_t nbins2,
                                                     const uint32_t* __restrict__ first,
                                                     const uint32_t* __restrict__ ntiles2,
                                                     uint32_t F1, uint32_t nrowR, uint32_t nrowS,
                                                     uint32_t* __restrict__ caps,
                                                     uint32_t nsub, uint32_t* __restrict__ hot) {
    __shared__ uint32_t red[256 / TBJ_WAVE];
    const uint32_t s = blockIdx.x;
    const uint32_t* f = fine + (uint64_t)s * nbins2;
    uint32_t tot = 0;
    for (uint32_t b = threadIdx.x; b < nbins2; b += 256) tot += f[b];
    for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) tot += __shfl_xor(tot, d, TBJ_WAVE);
    if (lane_id() == 0) red[wave_id()] = tot;
    __syncthreads();
    tot = 0;
    for (int w = 0; w < 256 / TBJ_WAVE; w++) tot += red[w];
    // the partition's tiles [ft, et) and their share in each eighth
    const uint32_t rel = s / F1, b1 = s % F1;
    const uint32_t nrow = rel ? nrowS : nrowR;
    const uint32_t pos0 = (rel ? F1 * nrowR : 0) + b1 * nrow;
    const uint32_t ft = first[pos0], et = first[pos0 + nrow], T = *ntiles2;
    double share[8];
#pragma unroll
    for (uint32_t x = 0; x < 8; x++) {
        const uint32_t a = max(ft, eighth_start(x, T)), b = min(et, eighth_start(x + 1, T));
        share[x] = et > ft && b > a ? (double)(b - a) / (double)(et - ft) : 0.0;
    }
    const double n = (double)npart[s];
    for (uint32_t b = threadIdx.x; b < nbins2; b += 256) {
        // a second-pass bin with over 1/8 of its partition's sample is hot
        if (f[b] >= 8 && 8ull * f[b] > tot) atomicOr(hot, 1u);
        const double raw = tot ? n * (double)f[b] / (double)tot : 0.0;
        const double shrunk = n * ((double)f[b] + 8.0) / ((double)tot + 8.0 * nbins2);
        const double e = raw > shrunk ? raw : shrunk;
#pragma unroll
        for (uint32_t x = 0; x < 8; x++) {
            const uint32_t cap = n > 0 && share[x] > 0 ? (uint32_t)min(n, (double)region_cap(e * share[x])) : 0;
            caps[((uint64_t)s * 8 + x) * nbins2 + b] = cap;
        }
    }
    if (s == 0 && threadIdx.x == 0) caps[8 * (uint64_t)nsub] = 0;
}

// Second-pass tiles over the first pass's filled regions.  Region i is
// (rel, row, b1) in memory order; the second pass visits the regions in
// partition order (rel, b1, row), so that consecutive tiles -- one XCD's
// share -- fill the regions of one partition.  ntiles[pos] = region i's tile
// count at its position pos in that order (scanned afterwards into first
// tiles); ntiles[n] = 0.
__device__ __forceinline__ uint32_t region_order(uint32_t i, uint32_t F1, uint32_t nrowR,
                                                 uint32_t nrowS, uint32_t* part) {
    const uint32_t nR1 = F1 * nrowR;
    const uint32_t rel = i < nR1 ? 0 : 1, j = rel ? i - nR1 : i;
    const uint32_t nrow = rel ? nrowS : nrowR;
    const uint32_t row = j / F1, b1 = j % F1;
    *part = rel * F1 + b1;
    return (rel ? nR1 : 0) + b1 * nrow + row;
}

__global__ void k_region_tiles(const uint32_t* __restrict__ starts,
                               const uint32_t* __restrict__ cursors, uint32_t n, uint32_t F1,
                               uint32_t nrowR, uint32_t nrowS, uint32_t* __restrict__ ntiles,
                               uint32_t tile) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t len = min(cursors[i], starts[i + 1] - starts[i]);
        uint32_t part;
        ntiles[region_order(i, F1, nrowR, nrowS, &part)] = (len + tile - 1) / tile;
    } else if (i == n) {
        ntiles[i] = 0;
    }
}

// One thread per region: its tiles (from first, in partition order), each
// tagged with the region's first-pass partition rel * F1 + b1.
__global__ void k_region_tile_descs(const uint32_t* __restrict__ starts,
                                    const uint32_t* __restrict__ cursors,
                                    const uint32_t* __restrict__ first, uint32_t n, uint32_t F1,
                                    uint32_t nrowR, uint32_t nrowS,
                                    TileDesc* __restrict__ tiles, uint32_t tile,
                                    uint32_t seg_is_part) {
    // one wave per region, its lanes over the region's tiles (a hot key's
    // region holds hundreds)
    const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / TBJ_WAVE;
    if (i >= n) return;  // whole waves
    const uint32_t st = starts[i], len = min(cursors[i], starts[i + 1] - st);
    uint32_t part;
    const uint32_t t0 = first[region_order(i, F1, nrowR, nrowS, &part)];
    for (uint32_t q = lane_id(); (uint64_t)q * tile < len; q += TBJ_WAVE) {
        const uint32_t k = q * tile;
        TileDesc d;
        d.in_off = st + k;
        d.len = min(tile, len - k);
        d.seg = seg_is_part ? part : i;
        d.aux = part;
        tiles[t0 + q] = d;
    }
}

// Join bounds of the per-XCD sub-regions: sub-region ((rel * F1 + b1) * 8 +
// x) * F2 + b2 is piece x of side rel of partition p = b1 * F2 + b2,
// bounds[(rel * F1 * F2 + p) * 8 + x] = {start, start + min(cursor, cap)}.
__global__ void k_bounds_x8(const uint32_t* __restrict__ starts,
                            const uint32_t* __restrict__ cursors, uint32_t F1, uint32_t F2,
                            uint2* __restrict__ bounds) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= 16 * F1 * F2) return;
    const uint32_t b2 = r % F2, x = (r / F2) % 8, s = r / (8 * F2);  // s = rel * F1 + b1
    const uint32_t st = starts[r], cap = starts[r + 1] - st;
    bounds[((uint64_t)s * F2 + b2) * 8 + x] = make_uint2(st, st + min(cursors[r], cap));
}

// bounds[i] = {start, start + min(cursor, cap)} for the 2 * nparts partitions.
__global__ void k_bounds_est(const uint32_t* __restrict__ starts,
                             const uint32_t* __restrict__ cursors, uint32_t n,
                             uint2* __restrict__ bounds) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t st = starts[i], cap = starts[i + 1] - st;
    bounds[i] = make_uint2(st, st + min(cursors[i], cap));
}

// ---------------------------------------------------------------------------
// Second-pass work descriptors: each (relation, first-pass partition) becomes
// a segment cut into tiles.  One workgroup; npart1 <= kMaxBins.

__global__ __launch_bounds__(kScanBlock) void k_make_segments(
    const uint32_t* __restrict__ scanned1, const SegDesc* __restrict__ segs1, uint32_t nseg1,
    uint32_t nbins1, uint32_t nbins2, SegDesc* __restrict__ segs2,
    uint32_t* __restrict__ ntiles2_dev, uint32_t* __restrict__ ngroups2_dev, uint32_t tile) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    __shared__ uint32_t carry_s, gcarry_s;
    const uint32_t nseg2 = nseg1 * nbins1;
    if (threadIdx.x == 0) carry_s = gcarry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nseg2; base += kScanBlock) {
        const uint32_t s2 = base + threadIdx.x;
        uint32_t st = 0, n = 0, nt = 0;
        if (s2 < nseg2) {
            const uint32_t s1 = s2 / nbins1, b = s2 % nbins1;
            const SegDesc sd = segs1[s1];
            st = scanned1[sd.hist_base + (uint64_t)b * sd.ngroups];
            const uint32_t en = scanned1[sd.hist_base + (uint64_t)(b + 1) * sd.ngroups];
            n = en - st;
            nt = (n + tile - 1) / tile;
        }
        uint32_t total, gtotal;
        const uint32_t ft = carry_s + block_exclusive_scan<kScanBlock>(nt, scratch, &total);
        const uint32_t ng = seg_groups(ft, nt);
        __syncthreads();
        const uint32_t gex = block_exclusive_scan<kScanBlock>(ng, scratch, &gtotal);
        if (s2 < nseg2) {
            SegDesc d;
            d.in_off = st;
            d.n = n;
            d.ntiles = nt;
            d.first_tile = ft;
            d.ngroups = ng;
            d.aux = 0;
            d.hist_base = (uint64_t)(gcarry_s + gex) * nbins2;
            segs2[s2] = d;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            carry_s += total;
            gcarry_s += gtotal;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *ntiles2_dev = carry_s;
        if (ngroups2_dev) *ngroups2_dev = gcarry_s;
    }
}

// One thread per second-pass tile: find the owning segment by binary search
// over first_tile, then describe the tile.
__global__ void k_make_tiles2(const SegDesc* __restrict__ segs2, uint32_t nseg2,
                              const uint32_t* __restrict__ ntiles2_dev,
                              TileDesc* __restrict__ tiles2, uint32_t tile) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *ntiles2_dev) return;
    uint32_t lo = 0, hi = nseg2;  // last segment with first_tile <= t and ntiles > 0
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (segs2[mid].first_tile <= t) lo = mid;
        else hi = mid;
    }
    // empty segments share first_tile with their successor: step forward
    while (lo + 1 < nseg2 &&
           (segs2[lo].ntiles == 0 || t >= segs2[lo].first_tile + segs2[lo].ntiles))
        lo++;
    const SegDesc sd = segs2[lo];
    const uint32_t k = t - sd.first_tile;
    TileDesc td;
    td.in_off = sd.in_off + k * tile;
    td.len = min(tile, sd.n - k * tile);
    td.seg = lo;
    td.aux = 0;
    tiles2[t] = td;
}

// ---------------------------------------------------------------------------
// Learned partitioner: strided key sample of R, then a one-workgroup fit of
// an equal-mass piecewise-linear CDF (knots at sample quantiles).

__global__ void k_sample_keys(const uint2* __restrict__ in, uint64_t n, uint32_t m,
                              uint32_t* __restrict__ sample) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    // slot i covers positions [i n / m, (i + 1) n / m) -- the slots tile all
    // of the input, its tail included -- and takes one at a hashed offset
    const uint64_t a = (uint64_t)i * n / m, b = (uint64_t)(i + 1) * n / m;
    uint64_t pos = a + ((uint64_t)i * 2654435761u) % (b > a ? b - a : 1);
    if (pos >= n) pos = n - 1;
    sample[i] = in[pos].x;
}

// The linear model when the sample (m keys, any order) is near-linear;
// model->linear says which.  Its key range is the sample's, widened on each
// side by 8 expected gaps between sampled keys (span / m), where the R keys
// beyond the sample's extremes lie; it needs span + 1 >= 2^bits so that
// lmul fits the product's 64 bits.
__global__ __launch_bounds__(kFitBlock) void k_fit_linear(const uint32_t* __restrict__ sample,
                                                          uint32_t m, uint32_t bits,
                                                          CdfModel* __restrict__ model,
                                                          uint32_t* __restrict__ lin_out) {
    __shared__ unsigned long long lds[kLinBuckets + 5];
    uint32_t kmin, kmax;
    const bool lin = sample_is_linear(sample, m, lds, &kmin, &kmax);
    if (threadIdx.x != 0) return;
    const uint64_t gap = 8ull * ((uint64_t)(kmax - kmin) / m);
    const uint64_t lo = kmin > gap ? kmin - gap : 0;
    const uint64_t hi = min((uint64_t)kmax + gap, (uint64_t)0xFFFFFFFFu);
    const uint64_t span1 = hi - lo + 1;
    const bool ok = lin && span1 >= (1ull << bits);
    model->linear = ok;
    *lin_out = ok;
    if (!ok) return;
    model->lmul = (1ull << (32 + bits)) / span1;  // bits <= 22: 2^(32 + bits) fits
    model->kmin = (uint32_t)lo;
    model->kmax = (uint32_t)hi;
    model->drop_outside = 1;
    model->nseg = 0;
    model->frac = 32 - bits;
    model->cshift = 0;
    model->ncells = 0;
}

// Bitonic sort of m (power of two, <= kFitMax) sampled keys in LDS, in
// place.  It needs only the sample, so it runs beside the exact key range.
__global__ __launch_bounds__(kFitBlock) void k_fit_sort(uint32_t* __restrict__ sample,
                                                        uint32_t m,
                                                        const CdfModel* __restrict__ model) {
    __shared__ uint32_t a[kFitMax];
    if (model && model->linear) return;  // no knots to place
    for (uint32_t i = threadIdx.x; i < m; i += kFitBlock) a[i] = sample[i];
    __syncthreads();
    // one compare-exchange per pair index q < m / 2 (every lane busy)
    for (uint32_t k = 2; k <= m; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t q = threadIdx.x; q < m / 2; q += kFitBlock) {
                const uint32_t i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                const uint32_t ixj = i | j;
                const bool up = (i & k) == 0;
                const uint32_t x = a[i], y = a[ixj];
                if ((x > y) == up) {
                    a[i] = y;
                    a[ixj] = x;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < m; i += kFitBlock) sample[i] = a[i];
}

// Knots and cell table from the sorted sample.  Segment count nseg <=
// kLdsKnots.  Partition space: 2^bits partitions with frac fixed-point bits
// (bits + frac == 32).
__global__ __launch_bounds__(kFitBlock) void k_fit_cdf(const uint32_t* __restrict__ sorted,
                                                       uint32_t m, uint32_t nseg_req,
                                                       uint32_t bits,
                                                       const uint32_t* __restrict__ minmax,
                                                       CdfModel* __restrict__ model) {
    if (model->linear) return;  // the linear model is complete
    __shared__ uint32_t quant[kLdsKnots];
    __shared__ uint32_t knots[kLdsKnots + 1];
    __shared__ uint64_t kx[kLdsKnots + 1];
    __shared__ uint32_t ky[kLdsKnots + 1];
    __shared__ uint32_t nk_s;
    for (uint32_t j = threadIdx.x; j < nseg_req; j += kFitBlock)
        quant[j] = sorted[(uint64_t)j * m / nseg_req];
    __syncthreads();
    // the model's key range: R's exact one, or without it the sample's
    const uint32_t kmin = minmax ? minmax[0] : sorted[0], kmax = minmax ? minmax[1] : sorted[m - 1];
    if (threadIdx.x == 0) {
        // first knot at the range's minimum, then distinct sample quantiles
        uint32_t nk = 0;
        knots[nk++] = kmin;
        for (uint32_t j = 1; j < nseg_req; j++) {
            const uint32_t q = quant[j];
            if (q > knots[nk - 1] && q <= kmax) knots[nk++] = q;
        }
        nk_s = nk;
    }
    __syncthreads();
    const uint32_t nseg = nk_s;
    const uint32_t frac = 32 - bits;
    const uint64_t span = (uint64_t)1 << 32;  // y spans [0, 2^32): 2^bits partitions
    for (uint32_t i = threadIdx.x; i <= nseg; i += kFitBlock) {
        // the last knot closes the model at R's exact maximum
        const uint64_t xi = i < nseg ? (uint64_t)knots[i] : ((uint64_t)kmax + 1);
        uint64_t yi = (span * i) / nseg;
        if (yi > 0xFFFFFFFFull) yi = 0xFFFFFFFFull;
        model->x[i] = xi;
        model->y[i] = (uint32_t)yi;
        kx[i] = xi;
        ky[i] = (uint32_t)yi;
    }
    __syncthreads();
    // cell table: the model at equal-width cell starts, in 1/256 partitions
    uint32_t cshift = 0;
    while ((((uint64_t)kmax - kmin) >> cshift) >= (uint64_t)kCells) cshift++;
    const uint32_t ncells = (uint32_t)((((uint64_t)kmax - kmin) >> cshift) + 1);
    const double to_cells = 1.0 / (double)(1ull << (frac - 8));  // 2^-frac -> 2^-8 units
    for (uint32_t c = threadIdx.x; c <= ncells; c += kFitBlock) {
        const uint64_t key = (uint64_t)kmin + ((uint64_t)c << cshift);
        double y;
        if (c == ncells || key > kmax) {
            y = 4294967296.0;
        } else {
            uint32_t lo = 0, hi = nseg;  // largest i with x[i] <= key
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (kx[mid] <= key) lo = mid;
                else hi = mid;
            }
            const double dx = (double)(kx[lo + 1] - kx[lo]);
            const double dy = (double)ky[lo + 1] - (double)ky[lo];
            y = (double)ky[lo] + (double)(key - kx[lo]) * dy / dx;
        }
        model->cells[c] = (uint32_t)(y * to_cells);
    }
    if (threadIdx.x == 0) {
        model->nseg = nseg;
        model->frac = frac;
        model->kmin = kmin;
        model->kmax = kmax;
        model->drop_outside = 1;
        model->cshift = cshift;
        model->ncells = ncells;
    }
}


// First-pass work descriptors: segment 0 = R, segment 1 = S, tiles of R
// first.  Tile offsets are relative to the owning relation's base pointer.
// Descriptor of first-pass tile i (R's tiles, then S's); i == 0 also writes
// the two segments and the tile count.
__device__ __forceinline__ void make_tile1(uint32_t i, TileDesc* __restrict__ tiles,
                                           SegDesc* __restrict__ segs, uint32_t nR, uint32_t nS,
                                           uint32_t nbins, uint32_t* __restrict__ ntiles_dev) {
    const uint32_t tR = (nR + kTile - 1) / kTile, tS = (nS + kTile - 1) / kTile;
    if (i < tR) {
        TileDesc d;
        d.in_off = i * kTile;
        d.len = min((uint32_t)kTile, nR - i * kTile);
        d.seg = 0;
        d.aux = tile_row(i, tR);
        tiles[i] = d;
    } else if (i < tR + tS) {
        const uint32_t k = i - tR;
        TileDesc d;
        d.in_off = k * kTile;
        d.len = min((uint32_t)kTile, nS - k * kTile);
        d.seg = 1;
        d.aux = tile_row(k, tS);
        tiles[i] = d;
    }
    if (i == 0) {
        SegDesc a, b;
        a.hist_base = 0;
        a.in_off = 0;
        a.n = nR;
        a.ntiles = tR;
        a.first_tile = 0;
        a.ngroups = seg_groups(0, tR);
        a.aux = 0;
        b.hist_base = (uint64_t)nbins * a.ngroups;
        b.in_off = 0;
        b.n = nS;
        b.ntiles = tS;
        b.first_tile = tR;
        b.ngroups = seg_groups(tR, tS);
        b.aux = 0;
        segs[0] = a;
        segs[1] = b;
        *ntiles_dev = tR + tS;
    }
}

__global__ void k_make_tiles1(TileDesc* __restrict__ tiles, SegDesc* __restrict__ segs,
                              uint32_t nR, uint32_t nS, uint32_t nbins,
                              uint32_t* __restrict__ ntiles_dev) {
    make_tile1(blockIdx.x * blockDim.x + threadIdx.x, tiles, segs, nR, nS, nbins, ntiles_dev);
}

// The first launch of a sample-sized call: the arrays of z cleared and the
// first pass's tiles described, grid-stride (z must not hold *ntiles_dev).
__global__ void k_start_tiles1(ZeroList z, TileDesc* __restrict__ tiles,
                               SegDesc* __restrict__ segs, uint32_t nR, uint32_t nS,
                               uint32_t nbins, uint32_t* __restrict__ ntiles_dev) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t a = 0; a < z.k; a++)
        for (uint64_t i = t0; i < z.n[a]; i += step) z.p[a][i] = 0;
    const uint64_t nt = (uint64_t)(nR + kTile - 1) / kTile + (nS + kTile - 1) / kTile;
    for (uint64_t i = t0; i < (nt ? nt : 1); i += step)
        make_tile1((uint32_t)i, tiles, segs, nR, nS, nbins, ntiles_dev);
}

// ---------------------------------------------------------------------------
// Offsets from per-tile counts.  One workgroup per group of kGroupTiles
// tiles, one thread per bin: the group's rows are loaded at once (each load
// a contiguous row), then walked in tile order; a tile of another segment
// than its predecessor starts a new sum (segments may be shorter than a
// group).

// gsum[(segment, bin, group)] = the counts of the group's tiles of that
// segment.
__global__ __launch_bounds__(256) void k_group_sums(const uint16_t* __restrict__ counts,
                                                    const TileDesc* __restrict__ tiles,
                                                    const SegDesc* __restrict__ segs,
                                                    const uint32_t* __restrict__ ntiles_dev,
                                                    uint32_t nbins, uint32_t* __restrict__ gsum) {
    const uint32_t g = blockIdx.x, t0 = g * kGroupTiles, nt = *ntiles_dev;
    if (t0 >= nt) return;
    const uint32_t len = min((uint32_t)kGroupTiles, nt - t0);
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) {
        uint32_t c[kGroupTiles];
#pragma unroll
        for (uint32_t k = 0; k < kGroupTiles; k++)
            c[k] = counts[(uint64_t)(t0 + min(k, len - 1)) * nbins + b];
        uint32_t seg = tiles[t0].seg, acc = 0;
#pragma unroll
        for (uint32_t k = 0; k < kGroupTiles; k++) {
            if (k < len) {
                const uint32_t sk = tiles[t0 + k].seg;
                if (sk != seg) {
                    const SegDesc sd = segs[seg];
                    gsum[sd.hist_base + (uint64_t)b * sd.ngroups + g - sd.first_tile / kGroupTiles] =
                        acc;
                    seg = sk;
                    acc = 0;
                }
                acc += c[k];
            }
        }
        const SegDesc sd = segs[seg];
        gsum[sd.hist_base + (uint64_t)b * sd.ngroups + g - sd.first_tile / kGroupTiles] = acc;
    }
}

// Per-tile offsets: a tile's bin offset is its group's scanned sum plus the
// counts of the segment's earlier tiles in the group.
__global__ __launch_bounds__(256) void k_group_offsets(const uint16_t* __restrict__ counts,
                                                       uint32_t* __restrict__ offs,
                                                       const TileDesc* __restrict__ tiles,
                                                       const SegDesc* __restrict__ segs,
                                                       const uint32_t* __restrict__ ntiles_dev,
                                                       uint32_t nbins,
                                                       const uint32_t* __restrict__ gsum) {
    const uint32_t g = blockIdx.x, t0 = g * kGroupTiles, nt = *ntiles_dev;
    if (t0 >= nt) return;
    const uint32_t len = min((uint32_t)kGroupTiles, nt - t0);
    for (uint32_t b = threadIdx.x; b < nbins; b += blockDim.x) {
        uint32_t c[kGroupTiles];
#pragma unroll
        for (uint32_t k = 0; k < kGroupTiles; k++)
            c[k] = counts[(uint64_t)(t0 + min(k, len - 1)) * nbins + b];
        uint32_t seg = ~0u, run = 0;
#pragma unroll
        for (uint32_t k = 0; k < kGroupTiles; k++) {
            if (k < len) {
                const uint32_t sk = tiles[t0 + k].seg;
                if (sk != seg) {
                    const SegDesc sd = segs[sk];
                    run = gsum[sd.hist_base + (uint64_t)b * sd.ngroups + g -
                               sd.first_tile / kGroupTiles];
                    seg = sk;
                }
                offs[(uint64_t)(t0 + k) * nbins + b] = run;
                run += c[k];
            }
        }
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

static inline size_t model_lds(const BinSpec& bs) {
    return bs.kind == kBinLearned ? sizeof(LdsModel) : 0;
}

// Persistent grid of the learned histogram: six workgroups per CU (each
// loads the 16 KB model once; measured on C3 against 2, 4, 5, 8 and 12 per
// CU, 6 was fastest), a multiple of the XCD count.
static uint32_t persistent_grid(uint32_t ntiles) {
    static uint32_t cap = 0;
    if (!cap) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        cap = (uint32_t)(cus > 0 ? cus : 256) * 6;
        cap = (cap + 7) / 8 * 8;
    }
    return ntiles < cap ? ntiles : cap;
}

void launch_hist(hipStream_t st, uint32_t ntiles, const uint2* in, const TileDesc* tiles,
                 const uint32_t* ntiles_dev, uint32_t tile_base, uint16_t* hist,
                 const BinSpec& bs, uint32_t nbins, uint32_t tile) {
    if (!ntiles) return;
    // the learned model is loaded into LDS once per workgroup: persistent
    // grid; radix: one tile per workgroup
    const uint32_t grid = bs.kind == kBinLearned ? persistent_grid(ntiles) : ntiles;
    const auto k = tile == kTile / 2 ? k_hist<kPartItems / 2>
                   : tile == kMaxTile ? k_hist<2 * kPartItems> : k_hist<kPartItems>;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kPartBlock), model_lds(bs), st, in, tiles,
                       ntiles_dev, tile_base, tile_base + ntiles, hist, bs, nbins);
}

void launch_scatter(hipStream_t st, uint32_t ntiles, const uint2* in, uint2* out,
                    const TileDesc* tiles, const uint32_t* ntiles_dev, uint32_t tile_base,
                    const uint32_t* offs, const BinSpec& bs, uint32_t nbins, bool last_pass,
                    const ClaimSpec* claims, uint32_t tile) {
    if (!ntiles) return;
    // one tile per workgroup; the learned model is read in place (cached)
    // rather than copied into LDS per tile
    const bool claim = claims != nullptr;
    const size_t lds = (claim ? 3 : 2) * sizeof(uint32_t) * nbins;
    const bool learned = bs.kind == kBinLearned;
    Claims cl{};
    if (claim) {
        cl.starts = claims->starts;
        cl.cursors = claims->cursors;
        cl.overflow = claims->overflow;
        cl.cb0 = claims->cb0;
        cl.cmul = claims->cmul;
        cl.cstride = claims->cstride;
        cl.xmul = claims->xmul;
        cl.agg_flag = claims->agg_flag;
        cl.tile_range = claims->tile_range;
    }
    decltype(&k_scatter<false, false, false>) k;
    const bool agg = last_pass || (claim && claims->agg_flag);
    if (claim)
        k = learned ? (agg ? k_scatter<true, true, true> : k_scatter<true, false, true>)
                    : (agg ? k_scatter<false, true, true> : k_scatter<false, false, true>);
    else if (tile == kTile / 2 && !learned)  // other tile shapes: radix passes without claims
        k = last_pass ? k_scatter<false, true, false, kPartItems / 2>
                      : k_scatter<false, false, false, kPartItems / 2>;
    else if (tile == kMaxTile && !learned)
        k = last_pass ? k_scatter<false, true, false, 2 * kPartItems>
                      : k_scatter<false, false, false, 2 * kPartItems>;
    else
        k = learned ? (last_pass ? k_scatter<true, true, false> : k_scatter<true, false, false>)
                    : (last_pass ? k_scatter<false, true, false> : k_scatter<false, false, false>);
    hipLaunchKernelGGL(k, dim3(ntiles), dim3(kPartBlock), lds, st, in, out, tiles, ntiles_dev,
                       tile_base, tile_base + ntiles, offs, cl, bs, nbins);
}

__global__ void k_fold_bins(const uint32_t* __restrict__ in, uint32_t rows, uint32_t F2,
                            uint32_t* __restrict__ out) {
    const uint32_t h = F2 / 2;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)rows * h) return;
    const uint32_t r = (uint32_t)(i / h), j = (uint32_t)(i % h);
    out[i] = in[(uint64_t)r * F2 + j] + in[(uint64_t)r * F2 + j + h];
}

void launch_fold_bins(hipStream_t st, const uint32_t* in, uint32_t rows, uint32_t F2,
                      uint32_t* out) {
    const uint64_t n = (uint64_t)rows * (F2 / 2);
    if (n) hipLaunchKernelGGL(k_fold_bins, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                              in, rows, F2, out);
}

void launch_range_reduce(hipStream_t st, const uint2* tile_range, uint32_t n, uint32_t* minmax) {
    hipLaunchKernelGGL(k_range_reduce, dim3(1), dim3(1024), 0, st, tile_range, n, minmax);
}

void launch_tile_offsets(hipStream_t st, uint32_t max_tiles, const uint16_t* counts,
                         uint32_t* offs, const TileDesc* tiles, const SegDesc* segs,
                         const uint32_t* ntiles_dev, uint32_t nbins, uint32_t* gsum,
                         uint64_t glen, uint32_t* partial, uint32_t* grand_total) {
    const uint32_t ng = (max_tiles + kGroupTiles - 1) / kGroupTiles;
    if (!ng) return;
    hipLaunchKernelGGL(k_group_sums, dim3(ng), dim3(256), 0, st, counts, tiles, segs, ntiles_dev,
                       nbins, gsum);
    launch_scan(st, gsum, glen, partial, grand_total);
    hipLaunchKernelGGL(k_group_offsets, dim3(ng), dim3(256), 0, st, counts, offs, tiles, segs,
                       ntiles_dev, nbins, gsum);
}

uint32_t scan_partials(uint64_t n) { return (uint32_t)((n + kScanChunk - 1) / kScanChunk) + 1; }

void launch_scan(hipStream_t st, uint32_t* a, uint64_t n, uint32_t* partial,
                 uint32_t* grand_total) {
    const uint32_t nb = (uint32_t)((n + kScanChunk - 1) / kScanChunk);
    if (nb == 0) return;
    if (n <= kScanSingleMax) {  // one launch instead of three
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanSingleBlock), 0, st, a, n,
                           grand_total);
        return;
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, st, a, n, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanBlock), 0, st, partial, nb,
                       grand_total);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanBlock), 0, st, a, n, partial);
}

void launch_make_segments(hipStream_t st, const uint32_t* scanned1, const SegDesc* segs1,
                          uint32_t nseg1, uint32_t nbins1, uint32_t nbins2, SegDesc* segs2,
                          TileDesc* tiles2, uint32_t* ntiles2_dev, uint32_t max_tiles2,
                          uint32_t* ngroups2_dev, uint32_t tile) {
    hipLaunchKernelGGL(k_make_segments, dim3(1), dim3(kScanBlock), 0, st, scanned1, segs1, nseg1,
                       nbins1, nbins2, segs2, ntiles2_dev, ngroups2_dev, tile);
    if (max_tiles2)
        hipLaunchKernelGGL(k_make_tiles2, dim3((max_tiles2 + 255) / 256), dim3(256), 0, st, segs2,
                           nseg1 * nbins1, ntiles2_dev, tiles2, tile);
}

void launch_tiles_from_segs(hipStream_t st, const SegDesc* segs, uint32_t nseg,
                            const uint32_t* ntiles_dev, TileDesc* tiles, uint32_t max_tiles,
                            uint32_t tile) {
    if (max_tiles)
        hipLaunchKernelGGL(k_make_tiles2, dim3((max_tiles + 255) / 256), dim3(256), 0, st, segs,
                           nseg, ntiles_dev, tiles, tile);
}

void launch_sample(hipStream_t st, const uint2* in, uint64_t n, uint32_t m, uint32_t* sample) {
    hipLaunchKernelGGL(k_sample_keys, dim3((m + 255) / 256), dim3(256), 0, st, in, n, m, sample);
}

void launch_fit_sort(hipStream_t st, uint32_t* sample, uint32_t m, const CdfModel* model) {
    hipLaunchKernelGGL(k_fit_sort, dim3(1), dim3(kFitBlock), 0, st, sample, m, model);
}

void launch_fit_linear(hipStream_t st, const uint32_t* sample, uint32_t m, uint32_t bits,
                       CdfModel* model, uint32_t* lin_out) {
    hipLaunchKernelGGL(k_fit_linear, dim3(1), dim3(kFitBlock), 0, st, sample, m, bits, model,
                       lin_out);
}

void launch_fit(hipStream_t st, const uint32_t* sorted, uint32_t m, uint32_t nseg, uint32_t bits,
                const uint32_t* minmax, CdfModel* model) {
    hipLaunchKernelGGL(k_fit_cdf, dim3(1), dim3(kFitBlock), 0, st, sorted, m, nseg, bits, minmax,
                       model);
}

void launch_minmax(hipStream_t st, const uint2* in, uint64_t n, uint32_t* minmax) {
    if (!n) return;
    if (reinterpret_cast<uintptr_t>(in) % 16) {  // pair loads need 16-byte alignment
        launch_minmax(st, in + 1, n - 1, minmax);
        if (n == 1) return;
        hipLaunchKernelGGL(k_minmax, dim3(1), dim3(256), 0, st, in, (uint64_t)1, minmax);
        return;
    }
    uint64_t g = (n / 2 + 255) / 256;
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_minmax, dim3((uint32_t)g), dim3(256), 0, st, in, n, minmax);
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Zeroes up to kZeroArrays arrays in one launch (grid-stride over each).
__global__ void k_zero_multi(ZeroList z) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint32_t a = 0; a < z.k; a++)
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < z.n[a]; i += step)
            z.p[a][i] = 0;
}

void launch_start_tiles1(hipStream_t st, const ZeroList& z, TileDesc* tiles, SegDesc* segs,
                         uint32_t nR, uint32_t nS, uint32_t nbins, uint32_t* ntiles_dev) {
    uint64_t most = (uint64_t)(nR + kTile - 1) / kTile + (nS + kTile - 1) / kTile;
    for (uint32_t a = 0; a < z.k; a++) most = z.n[a] > most ? z.n[a] : most;
    const uint64_t grid = std::min<uint64_t>((most + 255) / 256, 4096);
    hipLaunchKernelGGL(k_start_tiles1, dim3((uint32_t)(grid ? grid : 1)), dim3(256), 0, st, z,
                       tiles, segs, nR, nS, nbins, ntiles_dev);
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

void launch_sample_rows(hipStream_t st, const uint2* R, uint32_t nR, const uint2* S,
                        uint32_t nS, uint32_t nrowR, uint32_t nrowS, const BinSpec& bsr,
                        const BinSpec& bss, const BinSpec& bs2, uint32_t F1, uint32_t F2,
                        uint32_t* cnt1, uint32_t* fine, int which, uint32_t* hot1) {
    const uint32_t row0 = which == 1 ? nrowR : 0;
    const uint32_t rows = which == 0 ? nrowR : which == 1 ? nrowS : nrowR + nrowS;
    if (rows == 0) return;
    hipLaunchKernelGGL(k_sample_rows, dim3(rows * kChunksPerRow), dim3(kSampleBlock), 0, st, R, nR,
                       S, nS, nrowR, nrowS, row0, bsr, bss, bs2, F1, F2, cnt1, fine, hot1);
}

__global__ void k_add_base(uint32_t* __restrict__ a, uint32_t n, const uint32_t* __restrict__ base) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += *base;
}

void launch_add_base(hipStream_t st, uint32_t* a, uint32_t n, const uint32_t* base) {
    if (n) hipLaunchKernelGGL(k_add_base, dim3((n + 255) / 256), dim3(256), 0, st, a, n, base);
}

void launch_row_caps(hipStream_t st, const uint32_t* cnt, uint32_t n, uint32_t* caps) {
    hipLaunchKernelGGL(k_row_caps, dim3((n + 1 + 255) / 256), dim3(256), 0, st, cnt, n, caps);
}

void launch_part_sizes(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                       uint32_t F1, uint32_t nrowR, uint32_t nrowS, uint32_t* npart,
                       uint32_t* total) {
    const uint64_t threads = 2ull * F1 * TBJ_WAVE;
    hipLaunchKernelGGL(k_part_sizes, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, st,
                       starts, cursors, F1, nrowR, nrowS, npart, total);
}

void launch_est_part_max(hipStream_t st, const uint32_t* cnt1, uint32_t F1, uint32_t nrowR,
                         uint32_t* stats) {
    hipLaunchKernelGGL(k_est_part_max, dim3((F1 + 255) / 256), dim3(256), 0, st, cnt1, F1, nrowR,
                       stats);
}

void launch_region_caps(hipStream_t st, const uint32_t* fine, const uint32_t* npart,
                        uint32_t nseg2, uint32_t nbins2, const uint32_t* first,
                        const uint32_t* ntiles2, uint32_t F1, uint32_t nrowR, uint32_t nrowS,
                        uint32_t* caps, uint32_t* hot) {
    launch_fill_u32(st, hot, 1, 0);
    hipLaunchKernelGGL(k_region_caps, dim3(nseg2), dim3(256), 0, st, fine, npart, nbins2, first,
                       ntiles2, F1, nrowR, nrowS, caps, nseg2 * nbins2, hot);
}

void launch_region_tiles(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                         uint32_t n, uint32_t F1, uint32_t nrowR, uint32_t nrowS,
                         uint32_t* ntiles, uint32_t tile) {
    hipLaunchKernelGGL(k_region_tiles, dim3((n + 1 + 255) / 256), dim3(256), 0, st, starts,
                       cursors, n, F1, nrowR, nrowS, ntiles, tile);
}

void launch_region_tile_descs(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                              const uint32_t* first, uint32_t n, uint32_t F1, uint32_t nrowR,
                              uint32_t nrowS, TileDesc* tiles, uint32_t tile, bool seg_is_part) {
    if (!n) return;
    const uint64_t threads = (uint64_t)n * TBJ_WAVE;
    hipLaunchKernelGGL(k_region_tile_descs, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0,
                       st, starts, cursors, first, n, F1, nrowR, nrowS, tiles, tile,
                       (uint32_t)seg_is_part);
}

void launch_bounds_x8(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                      uint32_t F1, uint32_t F2, uint2* bounds) {
    const uint32_t n = 16 * F1 * F2;
    hipLaunchKernelGGL(k_bounds_x8, dim3((n + 255) / 256), dim3(256), 0, st, starts, cursors, F1,
                       F2, bounds);
}

void launch_bounds_est(hipStream_t st, const uint32_t* starts, const uint32_t* cursors,
                       uint32_t n, uint2* bounds) {
    hipLaunchKernelGGL(k_bounds_est, dim3((n + 255) / 256), dim3(256), 0, st, starts, cursors, n,
                       bounds);
}

void launch_fill_u32(hipStream_t st, uint32_t* p, uint64_t n, uint32_t v) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

void launch_zero_multi(hipStream_t st, const ZeroList& z) {
    uint64_t most = 0;
    for (uint32_t a = 0; a < z.k; a++) most = z.n[a] > most ? z.n[a] : most;
    if (!most) return;
    const uint64_t grid = std::min<uint64_t>((most + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero_multi, dim3((uint32_t)grid), dim3(256), 0, st, z);
}

}  // namespace tbj
