// 64-bit keys, 16-byte tuples {u64 key, u64 payload}: partitioned sort-merge
// join whose materialised output is ordered by key.
//
// The reference hard-wires 8-byte tuples (Include/data-types.h:32-47), so
// this path has no reference counterpart; it follows the structure of RSM
// (partition both inputs, then sort-merge each partition pair:
// Joins/radix_sortmerge_join.c:74-105) and the sort-then-merge of PSM
// (Joins/psm/parallel_sortmerge_join.h:18-111), at 16-byte tuples:
//   1. partition R and S in one or two passes (radix on the low key bits, or
//      the learned range model fitted on an R sample, whose partitions are
//      key ranges in key order);
//   2. count: each task sorts its R chunk (at most kSm64Cap tuples) by
//      (key, payload) in LDS (bucket sort; the chunk's first task also writes
//      the sorted chunk out), indexes the first position of every key run in
//      an LDS hash table, and every S tuple finds its run there; per chunk
//      the pairs are counted, and for a chunk probed by several tasks also
//      the S tuples per run (one global counter per sorted R position).  R
//      partitions above one chunk are instead sorted beforehand (chunks, then
//      merge passes);
//   3. a scan of the per-chunk pair counts gives each chunk's output offset,
//      and for shared chunks one pass turns the run counts into run output
//      offsets (scan of count x run length in key order);
//   4. emit: each S tuple finds its run again and writes run-length pairs at
//      its slot: a chunk's only task lays its runs out from the chunk offset
//      itself, tasks of shared chunks claim slot blocks in each run.
// So pairs of one key are contiguous, and with the learned partitioner the
// whole output is in key order (radix: in key order inside each partition).
#include <algorithm>

#include <mutex>

#include "tbj_device.h"
#include "tbj_kernels.h"
#include "tbj_sm64.h"

namespace tbj {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
// agg_inc / agg_add (tbj_device.h) skip lanes holding TBJ_DROP
static_assert(kNone == TBJ_DROP, "one no-bin marker");

__device__ __forceinline__ bool tuple_less(const ulonglong2& a, const ulonglong2& b) {
    return a.x < b.x || (a.x == b.x && a.y < b.y);
}

// Exclusive prefix sum of 64-bit values across a block (BLOCK / 64 + 1 words
// of scratch); two barriers, every thread must call it.
template <int BLOCK>
__device__ __forceinline__ unsigned long long block_exclusive_scan_u64(
    unsigned long long v, unsigned long long* scratch, unsigned long long* total) {
    constexpr int NW = BLOCK / TBJ_WAVE;
    const uint32_t lane = lane_id(), wid = wave_id();
    const unsigned long long inc = wave_inclusive_scan_u64(v);
    if (lane == TBJ_WAVE - 1) scratch[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const unsigned long long w = lane < NW ? scratch[lane] : 0;
        const unsigned long long wi = wave_inclusive_scan_u64(w);
        if (lane < NW) scratch[lane] = wi - w;
        if (lane == NW - 1) scratch[NW] = wi;
    }
    __syncthreads();
    *total = scratch[NW];
    return scratch[wid] + inc - v;
}

// ---------------------------------------------------------------------------
// Partition function

struct LdsModel64 {
    uint32_t cellv[kSm64Cells + 1];
};

// The model's exact piecewise-linear value at a key (segment a's line,
// clamped to the segment's partitions); below x[0] segment 0's line runs
// down to partition 0.
__device__ __forceinline__ uint32_t model64_exact(const uint64_t* x, const double* base,
                                                  const double* slope, const uint32_t* plo,
                                                  const uint32_t* phi, uint32_t nseg,
                                                  uint64_t key) {
    if (key < x[0]) {
        const double v = base[0] - (double)(x[0] - key) * slope[0];
        return min(v <= 0.0 ? 0u : (uint32_t)v, phi[0]);
    }
    uint32_t a = 0, b = nseg;  // last segment whose first knot is <= key
    while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (x[mid] <= key) a = mid;
        else b = mid;
    }
    const double v = fma((double)(key - x[a]), slope[a], base[a]);
    uint32_t p = v <= 0.0 ? 0u : (v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v);
    p = max(p, plo[a]);
    return min(p, phi[a]);
}

// Linear interpolation inside cell c: cellv[c] + off * (cellv[c+1] - cellv[c]) / 2^gshift.
__device__ __forceinline__ uint32_t cell_interp(uint32_t y0, uint32_t y1, uint64_t off,
                                                uint32_t gshift) {
    return gshift ? y0 + (uint32_t)__umul64hi(off << (64 - gshift), (uint64_t)(y1 - y0)) : y0;
}

struct Bin64Fn {
    uint32_t kind, pass_shift, pass_mask, radix_mask;
    const LdsModel64* lm;
    uint32_t gshift, ymax;
    uint64_t kmin, kmax, lo, hi;
    uint32_t lin, last;  // the linear model: min(umulhi(key - kmin, lmul), last)
    uint64_t lmul;       // (lin == 2: the key's top bits, key >> hshift)
    uint32_t hshift;
    // the two end segments, for keys outside [kmin, kmax]
    uint64_t x0, xl;
    double base0, slope0, basel, slopel;
    uint32_t phi0, plol, phil;
    bool drop;

    __device__ __forceinline__ uint32_t full(uint64_t key) const {
        if (kind == kBinRadix) return (uint32_t)key & radix_mask;
        if (lin == 2) return (uint32_t)(key >> hshift);
        if (lin)
            return key < kmin ? 0u : key > kmax ? last
                                                : min((uint32_t)__umul64hi(key - kmin, lmul), last);
        if (key < kmin) {  // below the sample: segment 0's line, down to 0
            const double v = base0 - (double)(x0 - key) * slope0;
            return min(v <= 0.0 ? 0u : (uint32_t)v, phi0);
        }
        if (key > kmax) {  // above: the last segment's line, from ymax up
            const double v = fma((double)(key - xl), slopel, basel);
            uint32_t p = v <= 0.0 ? 0u : (v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v);
            p = min(max(p, plol), phil);
            return max(p, ymax);
        }
        const uint64_t d = key - kmin;
        const uint32_t c = (uint32_t)(d >> gshift);
        return cell_interp(lm->cellv[c], lm->cellv[c + 1], d & ((1ull << gshift) - 1), gshift);
    }
    __device__ __forceinline__ uint32_t operator()(uint64_t key) const {
        if (drop && (key < lo || key > hi)) return kNone;
        return (full(key) >> pass_shift) & pass_mask;
    }

    // The pass's bins of N keys (kNone where valid is false or the key is
    // dropped).  The choice of map is the same for every key, so it is made
    // once: the radix map and the top-bits map (the linear model over all
    // 64-bit keys) then run as straight-line code, not as N copies of
    // full()'s branches.
    template <int N>
    __device__ __forceinline__ void bins(const uint64_t (&key)[N], const bool (&valid)[N],
                                         uint32_t (&out)[N]) const {
        bool keep[N];
#pragma unroll
        for (int q = 0; q < N; q++)
            keep[q] = valid[q] && !(drop && (key[q] < lo || key[q] > hi));
        if (kind == kBinRadix) {
#pragma unroll
            for (int q = 0; q < N; q++)
                out[q] = keep[q] ? ((((uint32_t)key[q] & radix_mask) >> pass_shift) & pass_mask)
                                 : kNone;
        } else if (lin == 2) {
#pragma unroll
            for (int q = 0; q < N; q++)
                out[q] = keep[q] ? (((uint32_t)(key[q] >> hshift) >> pass_shift) & pass_mask)
                                 : kNone;
        } else {
#pragma unroll
            for (int q = 0; q < N; q++)
                out[q] = keep[q] ? ((full(key[q]) >> pass_shift) & pass_mask) : kNone;
        }
    }
};

// Loads the model into LDS (no barrier: the caller's next barrier covers it).
__device__ __forceinline__ Bin64Fn make_bin64(const Bin64& b, LdsModel64* lm) {
    Bin64Fn f;
    f.kind = b.kind;
    f.pass_shift = b.pass_shift;
    f.pass_mask = b.pass_mask;
    f.radix_mask = b.bits >= 32 ? 0xFFFFFFFFu : (1u << b.bits) - 1;
    f.lm = lm;
    f.gshift = f.ymax = 0;
    f.kmin = f.kmax = f.x0 = f.xl = 0;
    f.base0 = f.slope0 = f.basel = f.slopel = 0.0;
    f.phi0 = f.plol = f.phil = 0;
    f.lin = 0;
    f.hshift = 64 - b.bits;
    f.last = b.bits >= 32 ? 0xFFFFFFFFu : (1u << b.bits) - 1;
    f.lmul = 0;
    f.drop = b.range != nullptr && (b.gate == nullptr || *b.gate > b.gate_min);
    f.lo = f.drop ? b.range[0] : 0;
    f.hi = f.drop ? b.range[1] : ~0ull;
    if (b.kind == kBinLearned && b.model->linear) {
        f.lin = b.model->linear;
        f.kmin = b.model->kmin;
        f.kmax = b.model->kmax;
        f.lmul = b.model->lmul;
    } else if (b.kind == kBinLearned) {
        const Model64* m = b.model;
        for (uint32_t i = threadIdx.x; i <= m->ncells; i += blockDim.x) lm->cellv[i] = m->cellv[i];
        const uint32_t l = m->nseg - 1;
        f.gshift = m->gshift;
        f.ymax = m->ymax;
        f.kmin = m->kmin;
        f.kmax = m->kmax;
        f.x0 = m->x[0];
        f.base0 = m->base[0];
        f.slope0 = m->slope[0];
        f.phi0 = m->phi[0];
        f.xl = m->x[l];
        f.basel = m->base[l];
        f.slopel = m->slope[l];
        f.plol = m->plo[l];
        f.phil = m->phi[l];
    }
    return f;
}

// ---------------------------------------------------------------------------
// Learned model: strided sample of R, then one workgroup sorts it and places
// equal-mass knots.

__global__ void k_sample64(const ulonglong2* __restrict__ in, uint64_t n, uint32_t m,
                           uint64_t* __restrict__ sample) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    // slot i covers positions [i n / m, (i + 1) n / m) (m <= n: the slots
    // tile all of R, its tail included) and takes one at a hashed offset
    const uint64_t a = (uint64_t)i * n / m, b = (uint64_t)(i + 1) * n / m;
    uint64_t pos = a + ((uint64_t)i * 2654435761u) % (b > a ? b - a : 1);
    if (pos >= n) pos = n - 1;
    sample[i] = in[pos].x;
}

__global__ __launch_bounds__(1024) void k_fit64(const uint64_t* __restrict__ sample, uint32_t m,
                                                uint32_t nseg_req, uint32_t bits,
                                                Model64* __restrict__ model,
                                                uint32_t* __restrict__ lin_out) {
    __shared__ uint64_t a[kSm64FitMax];
    __shared__ uint64_t knots[kSm64Knots];
    __shared__ uint32_t krank[kSm64Knots];
    __shared__ double sbase[kSm64Knots], sslope[kSm64Knots];
    __shared__ uint32_t splo[kSm64Knots], sphi[kSm64Knots];
    __shared__ uint32_t scell[kSm64Cells + 1];
    __shared__ uint32_t nk_s;
    __shared__ unsigned long long lin_lds[kLinBuckets + 5];
    for (uint32_t i = threadIdx.x; i < m; i += 1024) a[i] = sample[i];
    __syncthreads();
    {
        // a near-linear sample needs no sort, knots or cells: the linear
        // model over its range widened by 8 expected gaps per side (span + 1
        // >= 2^(bits + 1) keeps lmul below 2^64)
        uint64_t smin, smax;
        const bool lin = sample_is_linear(a, m, lin_lds, &smin, &smax);
        const uint64_t gap = 8 * ((smax - smin) / m);
        const uint64_t lo = smin > gap ? smin - gap : 0;
        const uint64_t hi = smax < ~0ull - gap ? smax + gap : ~0ull;
        const double span1 = (double)(hi - lo) + 1.0;
        if (lin && span1 >= ldexp(1.0, (int)bits + 1)) {
            if (threadIdx.x == 0) {
                *lin_out = 1;
                // a range widened to all 64-bit keys maps a key to its top
                // `bits` bits (lmul = 2^bits): a shift
                model->linear = lo == 0 && hi == ~0ull ? 2 : 1;
                model->lmul = (uint64_t)(ldexp(1.0, 64 + (int)bits) / span1);
                model->kmin = lo;
                model->kmax = hi;
                model->nseg = model->gshift = model->ncells = model->ymax = 0;
            }
            return;
        }
        if (threadIdx.x == 0) {
            model->linear = 0;
            *lin_out = 0;
        }
    }
    // bitonic sort, one compare-exchange per pair index q < m / 2 (every
    // lane busy): pair q joins i = its low bits below j and the rest shifted
    // up one place, with i ^ j
    for (uint32_t k = 2; k <= m; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t q = threadIdx.x; q < m / 2; q += 1024) {
                const uint32_t i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                const uint32_t ixj = i | j;
                const bool up = (i & k) == 0;
                const uint64_t x = a[i], y = a[ixj];
                if ((x > y) == up) {
                    a[i] = y;
                    a[ixj] = x;
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        uint32_t nk = 1;
        knots[0] = a[0];
        krank[0] = 0;
        for (uint32_t j = 1; j < nseg_req; j++) {
            const uint32_t idx = (uint32_t)((uint64_t)j * m / nseg_req);
            if (a[idx] > knots[nk - 1]) {
                knots[nk] = a[idx];
                krank[nk] = idx;
                nk++;
            }
        }
        nk_s = nk;
    }
    __syncthreads();
    const uint32_t nseg = nk_s;
    const uint64_t kmin = a[0], kmax = a[m - 1];
    const double parts = (double)(1ull << bits);
    const uint32_t last = (uint32_t)((1ull << bits) - 1);
    for (uint32_t i = threadIdx.x; i < nseg; i += 1024) {
        // knot at sample rank r: 2^bits * (r + 8) / (m + 15).  The R keys
        // beyond the sample's extremes number about n / (m + 1) per side,
        // with an exponential tail; each side keeps 8 times that room.  The
        // last segment ends at the sample maximum and continues beyond it.
        const double y0 = parts * (krank[i] + 8.0) / (m + 15.0);
        const double y1 = parts * ((i + 1 < nseg ? krank[i + 1] : m - 1) + 8.0) / (m + 15.0);
        const uint64_t x0 = knots[i];
        const uint64_t x1 = i + 1 < nseg ? knots[i + 1] : kmax;
        const double dx = x1 > x0 ? (double)(x1 - x0) : 1.0;
        const double sl = (y1 - y0) / dx;
        const uint32_t lo = i == 0 ? 0u : min((uint32_t)floor(y0), last);
        uint32_t hi = (uint32_t)ceil(y1);
        hi = hi > 0 ? hi - 1 : 0;
        const uint32_t ph = i + 1 < nseg ? min(max(hi, lo), last) : last;
        model->x[i] = x0;
        model->base[i] = y0;
        model->slope[i] = sl;
        model->plo[i] = lo;
        model->phi[i] = ph;
        sbase[i] = y0;
        sslope[i] = sl;
        splo[i] = lo;
        sphi[i] = ph;
    }
    if (threadIdx.x == 0) model->x[nseg] = kmax;
    __syncthreads();
    // cell table: the model at kmin + (c << gshift), c <= ncells (the last
    // cell start lies above kmax; saturated at 2^64 - 1)
    uint32_t gshift = 0;
    while (((kmax - kmin) >> gshift) >= (uint64_t)kSm64Cells) gshift++;
    const uint32_t ncells = (uint32_t)(((kmax - kmin) >> gshift) + 1);
    for (uint32_t c = threadIdx.x; c <= ncells; c += 1024) {
        // c << gshift itself can pass 2^64 (the last cell start of a span
        // near 2^64)
        const bool wide = gshift > 0 && ((uint64_t)c >> (64 - gshift)) != 0;
        const uint64_t off = (uint64_t)c << gshift;
        const uint64_t key = wide || off > ~0ull - kmin ? ~0ull : kmin + off;
        const uint32_t v = model64_exact(knots, sbase, sslope, splo, sphi, nseg, key);
        scell[c] = v;
        model->cellv[c] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t d = kmax - kmin;
        const uint32_t c = (uint32_t)(d >> gshift);
        model->ymax = cell_interp(scell[c], scell[c + 1], d & ((1ull << gshift) - 1), gshift);
        model->nseg = nseg;
        model->kmin = kmin;
        model->kmax = kmax;
        model->gshift = gshift;
        model->ncells = ncells;
    }
}

// ---------------------------------------------------------------------------
// Partition passes: per histogram tile, a row of bin counts; the counts
// become per-tile offsets (group sums, scan, in-place offsets: the 8-byte
// path's launch_tile_offsets); then each tile scatters its stages through
// LDS.  First pass: segment 0 is R and segment 1 is S; second pass: one
// segment per (relation, first-pass bin).

__global__ __launch_bounds__(kSm64Block) void k_hist64(
    const ulonglong2* __restrict__ in, const TileDesc* __restrict__ tiles,
    const uint32_t* __restrict__ ntiles_dev, uint32_t tile_base, uint16_t* __restrict__ hist,
    Bin64 b, uint32_t nbins, unsigned long long* __restrict__ mm) {
    __shared__ uint32_t cnt[kSm64MaxBins];
    __shared__ unsigned long long bmm[2];
    extern __shared__ __align__(16) uint8_t dyn_lds[];
    const uint32_t t = tile_base + blockIdx.x;
    if (t >= *ntiles_dev) return;
    // the tile's key loads go out before the partition function's set-up,
    // which waits on loads of its own
    const TileDesc td = tiles[t];
    const uint64_t* keys = reinterpret_cast<const uint64_t*>(in + td.in_off);
    uint64_t mn = ~0ull, mx = 0;
    constexpr int kHistItems = kSm64Span / kSm64Block;
    uint64_t kv[kHistItems];
#pragma unroll
    for (int q = 0; q < kHistItems; q++) {  // all loads in flight before any use
        const uint32_t j = q * kSm64Block + threadIdx.x;
        kv[q] = j < td.len ? keys[2 * (uint64_t)j] : 0;
    }
    const Bin64Fn bin = make_bin64(b, reinterpret_cast<LdsModel64*>(dyn_lds));
    for (uint32_t i = threadIdx.x; i < nbins; i += kSm64Block) cnt[i] = 0;
    if (threadIdx.x == 0) {
        bmm[0] = ~0ull;
        bmm[1] = 0;
    }
    __syncthreads();
    uint32_t hb[kHistItems];
    {
        bool ok[kHistItems];
#pragma unroll
        for (int q = 0; q < kHistItems; q++) ok[q] = q * kSm64Block + threadIdx.x < td.len;
        bin.bins(kv, ok, hb);
    }
#pragma unroll
    for (int q = 0; q < kHistItems; q++) {
        const uint32_t j = q * kSm64Block + threadIdx.x;
        if (j < td.len) {
            mn = min(mn, kv[q]);
            mx = max(mx, kv[q]);
        }
        agg_add(cnt, hb[q]);
    }
    if (mm) {  // tile min / max: waves -> LDS -> one global atomic pair
        wave_minmax_u64(mn, mx);
        if (lane_id() == 0) {
            atomicMin(&bmm[0], (unsigned long long)mn);
            atomicMax(&bmm[1], (unsigned long long)mx);
        }
    }
    __syncthreads();
    if (mm && threadIdx.x == 0) {
        // the stored range only narrows toward the result: a stale read can
        // only let an unneeded atomic through, and most tiles need none (one
        // atomic pair per tile on two words serialises)
        const volatile unsigned long long* vmm = mm;
        if (bmm[0] < vmm[0]) atomicMin(&mm[0], bmm[0]);
        if (bmm[1] > vmm[1]) atomicMax(&mm[1], bmm[1]);
    }
    for (uint32_t i = threadIdx.x; i < nbins; i += kSm64Block)
        hist[(uint64_t)t * nbins + i] = (uint16_t)cnt[i];
}

__host__ __device__ constexpr size_t sm64_model_bytes_aligned(uint32_t kind) {
    return kind == kBinLearned ? (sizeof(LdsModel64) + 15) / 16 * 16 : 0;
}

// One tile per workgroup, blocks mapped to tiles by xcd_tile; offs row t =
// tile t's output offset of each bin.  The whole tile is staged at once
// (8192 tuples, 128 KB of LDS, one workgroup per CU): measured on C5, runs of
// a full tile per bin beat two 4096-tuple stages at two workgroups per CU
// (scatter 4.8 -> 3.5 ms radix).  sbin keeps one byte per staged tuple,
// bin >> sh (sh = log2(nbins) - 8 above 256 bins); the low sh bits come
// back from the staged bin starts.
// CLAIM: the first pass of sample-sized regions: tile t's run of bin i goes
// to region cl.cb0 + tiles[t].aux * cl.cmul + i (a row-major (row, bin)
// layout, so a tile's claims are contiguous words), claimed from
// cl.cursors; a run past its region's end sets *cl.overflow and is not
// written.  cl.mm (optional, {2^64 - 1, 0} beforehand): the exact key range
// of the launch's tuples.
struct Claims64 {
    const uint32_t* starts;
    uint32_t* cursors;
    uint32_t* overflow;
    uint32_t cb0, cmul;
    unsigned long long* mm;  // the launch's key range, while !gate || *gate > gate_min
    const uint32_t* gate;
    uint32_t gate_min;
};

template <bool CLAIM>
__global__ __launch_bounds__(kSm64ScBlock) void k_scatter64(
    const ulonglong2* __restrict__ in, ulonglong2* __restrict__ out,
    const TileDesc* __restrict__ tiles, const uint32_t* __restrict__ ntiles_dev,
    uint32_t tile_base, uint32_t tile_end, const uint32_t* __restrict__ offs, Bin64 b,
    uint32_t nbins, Claims64 cl) {
    static_assert(kSm64MaxBins <= kSm64ScBlock, "one bin per thread");
    __shared__ ulonglong2 stage[kSm64ScTile];
    __shared__ uint8_t sbin[kSm64ScTile];
    __shared__ uint32_t scratch[kSm64ScBlock / TBJ_WAVE + 1];
    __shared__ unsigned long long bmm[2];
    // dynamic LDS: the learned model (if any), then the bin starts, the
    // bins' global offsets and (CLAIM) their region ends (nbins words each)
    extern __shared__ __align__(16) uint8_t dyn_lds[];
    uint32_t* loff = reinterpret_cast<uint32_t*>(dyn_lds + sm64_model_bytes_aligned(b.kind));
    uint32_t* goff = loff + nbins;
    uint32_t* glim = goff + nbins;
    const uint32_t end = min(tile_end, *ntiles_dev);
    const uint32_t n = end > tile_base ? end - tile_base : 0;
    if (blockIdx.x >= n) return;
    const uint32_t t = tile_base + xcd_tile(blockIdx.x, n);
    // the tile's loads go out first: the partition function's set-up waits
    // on loads of its own (a learned model's fields, the drop gate), and one
    // workgroup per CU leaves nothing else to cover that wait
    const TileDesc td = tiles[t];
    ulonglong2 v[kSm64ScItems];
#pragma unroll
    for (uint32_t k = 0; k < kSm64ScItems; k++) {
        const uint32_t j = k * kSm64ScBlock + threadIdx.x;
        if (j < td.len) v[k] = in[td.in_off + j];
    }
    // this thread's bin (one per thread): its region's bounds, loaded early
    const uint32_t r0 = CLAIM ? cl.cb0 + td.aux * cl.cmul : 0;
    uint32_t w0 = 0, w1 = 0;
    if (CLAIM && threadIdx.x < nbins) {
        w0 = cl.starts[r0 + threadIdx.x];
        w1 = cl.starts[r0 + threadIdx.x + 1];
    }
    const Bin64Fn bin = make_bin64(b, reinterpret_cast<LdsModel64*>(dyn_lds));
    const bool measure = CLAIM && cl.mm && (!cl.gate || *cl.gate > cl.gate_min);
    for (uint32_t i = threadIdx.x; i < nbins; i += kSm64ScBlock) loff[i] = 0;
    if (measure && threadIdx.x == 0) {
        bmm[0] = ~0ull;
        bmm[1] = 0;
    }
    // the stored range, read now so that the filter below does not wait for
    // it (it only narrows: a stale value lets an unneeded atomic through)
    unsigned long long smin = 0, smax = ~0ull;
    if (measure && threadIdx.x == 0) {
        const volatile unsigned long long* vmm = cl.mm;
        smin = vmm[0];
        smax = vmm[1];
    }
    __syncthreads();
    uint32_t bb[kSm64ScItems], rank[kSm64ScItems];
    {
        uint64_t keys[kSm64ScItems];
        bool ok[kSm64ScItems];
#pragma unroll
        for (uint32_t k = 0; k < kSm64ScItems; k++) {
            keys[k] = v[k].x;
            ok[k] = k * kSm64ScBlock + threadIdx.x < td.len;
        }
        bin.bins(keys, ok, bb);
    }
#pragma unroll
    for (uint32_t k = 0; k < kSm64ScItems; k++) rank[k] = agg_inc(loff, bb[k]);
    if (measure) {  // tile min / max: waves -> LDS -> one filtered atomic pair
        uint64_t mn = ~0ull, mx = 0;
#pragma unroll
        for (uint32_t k = 0; k < kSm64ScItems; k++)
            if (k * kSm64ScBlock + threadIdx.x < td.len) {
                mn = min(mn, (uint64_t)v[k].x);
                mx = max(mx, (uint64_t)v[k].x);
            }
        wave_minmax_u64(mn, mx);
        if (lane_id() == 0) {
            atomicMin(&bmm[0], (unsigned long long)mn);
            atomicMax(&bmm[1], (unsigned long long)mx);
        }
    }
    __syncthreads();
    if (measure && threadIdx.x == 0) {
        if (bmm[0] < smin) atomicMin(&cl.mm[0], bmm[0]);
        if (bmm[1] > smax) atomicMax(&cl.mm[1], bmm[1]);
    }
    // claims: one atomic per non-empty bin, awaited after the scan
    const uint32_t cnt = CLAIM && threadIdx.x < nbins ? loff[threadIdx.x] : 0;
    const uint32_t cbase = CLAIM && cnt ? atomicAdd(&cl.cursors[r0 + threadIdx.x], cnt) : 0;
    const uint32_t nvalid = block_scan_lds<kSm64ScBlock>(loff, nbins, scratch);
    if (CLAIM) {
        if (threadIdx.x < nbins) {
            if (cnt && cbase + cnt > w1 - w0) atomicOr(cl.overflow, 1u);
            goff[threadIdx.x] = w0 + cbase - loff[threadIdx.x];
            glim[threadIdx.x] = w1;
        }
    } else {
        for (uint32_t i = threadIdx.x; i < nbins; i += kSm64ScBlock)
            goff[i] = offs[(uint64_t)t * nbins + i] - loff[i];
    }
    const uint32_t sh = nbins > 256 ? 23 - __clz(nbins) : 0;  // log2(nbins) - 8
#pragma unroll
    for (uint32_t k = 0; k < kSm64ScItems; k++) {
        if (bb[k] != kNone) {
            const uint32_t slot = loff[bb[k]] + rank[k];
            stage[slot] = v[k];
            sbin[slot] = (uint8_t)(bb[k] >> sh);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nvalid; j += kSm64ScBlock) {
        // bin = the last of the 2^sh candidates whose start is <= j
        const uint32_t base = (uint32_t)sbin[j] << sh;
        uint32_t bj = base;
        for (uint32_t k = 1; k < (1u << sh); k++) bj += loff[base + k] <= j ? 1u : 0u;
        const uint32_t pos = goff[bj] + j;
        if (!CLAIM || pos < glim[bj]) out[pos] = stage[j];
    }
}

// Sample of the first pass of sample-sized regions: one workgroup per chunk
// of kS64SampleTiles tiles of one row (rows of R, then of S); exactly one
// 16-tuple group of every 256-tuple block is read (sampled_group), binned
// with the pass's function (S without the drop test: R's range is not known
// yet) and counted per (row, bin) into cnt1[(rel ? F1 * nrowR : 0) + row *
// F1 + bin].
constexpr uint32_t kS64SampleTiles = 112;
constexpr uint32_t kS64ChunksPerRow = (kRowTiles + kS64SampleTiles - 1) / kS64SampleTiles;

__global__ __launch_bounds__(256) void k_sample_rows64(
    const ulonglong2* __restrict__ R, uint32_t nR, const ulonglong2* __restrict__ S, uint32_t nS,
    uint32_t nrowR, uint32_t nrowS, Bin64 bsr, Bin64 bss, uint32_t F1,
    uint32_t* __restrict__ cnt1, uint32_t* __restrict__ s_out) {
    __shared__ uint32_t c1[kSm64MaxBins];
    extern __shared__ __align__(16) uint8_t dyn_lds[];
    const uint32_t rowg = blockIdx.x / kS64ChunksPerRow, chunk = blockIdx.x % kS64ChunksPerRow;
    const uint32_t rel = rowg < nrowR ? 0 : 1;
    const uint32_t row = rel ? rowg - nrowR : rowg;
    const ulonglong2* in = rel ? S : R;
    const uint32_t n = rel ? nS : nR;
    const uint32_t ntiles = (n + kSm64Span - 1) / kSm64Span;
    const uint32_t rpe = rows_per_eighth(ntiles);
    const uint32_t x = row / rpe, c = row % rpe;
    const uint32_t e0 = eighth_start(x, ntiles), e1 = eighth_start(x + 1, ntiles);
    const uint32_t r0 = min(e0 + c * kRowTiles, e1), r1 = min(r0 + kRowTiles, e1);
    const uint32_t t0 = min(r0 + chunk * kS64SampleTiles, r1), t1 = min(t0 + kS64SampleTiles, r1);
    if (t0 == t1) return;  // whole workgroup
    const Bin64Fn bin = make_bin64(rel ? bss : bsr, reinterpret_cast<LdsModel64*>(dyn_lds));
    for (uint32_t i = threadIdx.x; i < F1; i += 256) c1[i] = 0;
    __syncthreads();
    const uint32_t lo = t0 * kSm64Span, hi = min(t1 * kSm64Span, n);
    // S against the learned model's key range (the block-uniform test keeps
    // the wave reduction below with every lane)
    const bool count_out = rel && s_out && bss.kind == kBinLearned;
    uint32_t nout = 0;
    for (uint32_t blk = lo / 256 + threadIdx.x; blk < (hi + 255) / 256; blk += 256) {
        const uint32_t first = blk * 256 + sampled_group(blk) * 16;
        uint64_t key[16];
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) {  // all loads in flight before any use
            const uint32_t i = first + k;
            key[k] = in[i >= lo && i < hi ? i : lo].x;
        }
        bool ok[16];
        uint32_t bb[16];
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) ok[k] = first + k >= lo && first + k < hi;
        bin.bins(key, ok, bb);
#pragma unroll
        for (uint32_t k = 0; k < 16; k++) {
            if (bb[k] != kNone) atomicAdd(&c1[bb[k]], 1u);
            nout += ok[k] && (key[k] < bin.kmin || key[k] > bin.kmax);
        }
    }
    if (count_out) {
        const uint64_t w = wave_reduce_u64(nout);
        if (lane_id() == 0 && w) atomicAdd(s_out, (uint32_t)w);
    }
    __syncthreads();
    const uint32_t pbase = rel ? F1 * nrowR : 0;
    for (uint32_t i = threadIdx.x; i < F1; i += 256)
        if (c1[i]) atomicAdd(&cnt1[pbase + row * F1 + i], c1[i]);
}

// Second-pass segments of sample-sized first-pass regions: segment p = first-
// pass partition (rel, b1) = (p / F1, p % F1), whose tiles are the tiles of
// its rows' regions, [first[pos0], first[pos0 + nrow]) with pos0 = the
// partition's first region in partition order (k_region_tiles' scan).
__global__ __launch_bounds__(kScanBlock) void k_part_segments64(
    const uint32_t* __restrict__ first, uint32_t F1, uint32_t F2, uint32_t nrowR, uint32_t nrowS,
    SegDesc* __restrict__ segs2) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    __shared__ uint32_t gcarry_s;
    if (threadIdx.x == 0) gcarry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < 2 * F1; base += kScanBlock) {
        const uint32_t p = base + threadIdx.x;
        uint32_t ft = 0, nt = 0;
        if (p < 2 * F1) {
            const uint32_t rel = p / F1, b1 = p % F1, nrow = rel ? nrowS : nrowR;
            const uint32_t pos0 = (rel ? F1 * nrowR : 0) + b1 * nrow;
            ft = first[pos0];
            nt = first[pos0 + nrow] - ft;
        }
        const uint32_t ng = seg_groups(ft, nt);
        uint32_t gtotal;
        const uint32_t gex = block_exclusive_scan<kScanBlock>(ng, scratch, &gtotal);
        if (p < 2 * F1) {
            SegDesc d;
            d.in_off = 0;  // the tiles carry their own offsets
            d.n = 0;
            d.ntiles = nt;
            d.first_tile = ft;
            d.ngroups = ng;
            d.aux = p;
            d.hist_base = (uint64_t)(gcarry_s + gex) * F2;
            segs2[p] = d;
        }
        __syncthreads();
        if (threadIdx.x == 0) gcarry_s += gtotal;
        __syncthreads();
    }
}

__global__ void k_make_tiles64(TileDesc* __restrict__ tiles, SegDesc* __restrict__ segs,
                               uint32_t nR, uint32_t nS, uint32_t nbins,
                               uint32_t* __restrict__ ntiles_dev, uint32_t rows) {
    const uint32_t tR = (nR + kSm64Span - 1) / kSm64Span, tS = (nS + kSm64Span - 1) / kSm64Span;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tR + tS) {
        const bool r = i < tR;
        const uint32_t k = r ? i : i - tR, n = r ? nR : nS;
        TileDesc d;
        d.in_off = k * kSm64Span;
        d.len = min((uint32_t)kSm64Span, n - k * kSm64Span);
        d.seg = r ? 0 : 1;
        d.aux = rows ? tile_row(k, r ? tR : tS) : 0;  // sample-sized: the tile's row
        tiles[i] = d;
    }
    if (i == 0) {
        SegDesc a{}, c{};
        a.n = nR;
        a.ntiles = tR;
        a.ngroups = seg_groups(0, tR);
        c.hist_base = (uint64_t)nbins * a.ngroups;
        c.n = nS;
        c.ntiles = tS;
        c.first_tile = tR;
        c.ngroups = seg_groups(tR, tS);
        segs[0] = a;
        segs[1] = c;
        *ntiles_dev = tR + tS;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_make_segments64(
    const uint32_t* __restrict__ scanned1, const SegDesc* __restrict__ segs1, uint32_t nbins1,
    uint32_t nbins2, SegDesc* __restrict__ segs2, uint32_t* __restrict__ ntiles2_dev) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    __shared__ uint32_t carry_s, gcarry_s;
    const uint32_t nseg2 = 2 * nbins1;
    if (threadIdx.x == 0) carry_s = gcarry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nseg2; base += kScanBlock) {
        const uint32_t s2 = base + threadIdx.x;
        uint32_t st = 0, n = 0, nt = 0;
        if (s2 < nseg2) {
            const SegDesc sd = segs1[s2 / nbins1];
            const uint32_t bb = s2 % nbins1;
            st = scanned1[sd.hist_base + (uint64_t)bb * sd.ngroups];
            n = scanned1[sd.hist_base + (uint64_t)(bb + 1) * sd.ngroups] - st;
            nt = (n + kSm64Span - 1) / kSm64Span;
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
    if (threadIdx.x == 0) *ntiles2_dev = carry_s;
}

__global__ void k_make_tiles64_2(const SegDesc* __restrict__ segs2, uint32_t nseg2,
                                 const uint32_t* __restrict__ ntiles2_dev,
                                 TileDesc* __restrict__ tiles2) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *ntiles2_dev) return;
    // last segment whose first tile is <= t; step back over empty segments
    // sharing that first tile (the segment owning tile t has tiles)
    uint32_t lo = 0, hi = nseg2;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (segs2[mid].first_tile <= t) lo = mid;
        else hi = mid;
    }
    while (lo > 0 && (segs2[lo].ntiles == 0 || t >= segs2[lo].first_tile + segs2[lo].ntiles)) lo--;
    const SegDesc sd = segs2[lo];
    const uint32_t k = t - sd.first_tile;
    TileDesc td;
    td.in_off = sd.in_off + k * kSm64Span;
    td.len = min((uint32_t)kSm64Span, sd.n - k * kSm64Span);
    td.seg = lo;
    td.aux = 0;
    tiles2[t] = td;
}

// Final partition bounds: bounds[rel * nparts + p] = {start, end}.
__global__ void k_bounds64(const uint32_t* __restrict__ scanned, const SegDesc* __restrict__ segs,
                           uint32_t outer, uint32_t inner, uint32_t passes, uint32_t nR,
                           uint32_t nS, uint2* __restrict__ bounds) {
    const uint32_t nparts = outer * inner;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * nparts) return;
    const uint32_t rel = i / nparts, p = i % nparts;
    if (passes == 0) {
        bounds[i] = rel ? make_uint2(nR, nR + nS) : make_uint2(0, nR);
        return;
    }
    const uint32_t seg = passes == 1 ? rel : rel * outer + p / inner;
    const uint32_t bin = passes == 1 ? p : p % inner;
    const SegDesc sd = segs[seg];
    bounds[i] = make_uint2(scanned[sd.hist_base + (uint64_t)bin * sd.ngroups],
                           scanned[sd.hist_base + (uint64_t)(bin + 1) * sd.ngroups]);
}

// Per partition: R chunks and join tasks (R chunk x S chunk); word nparts is
// zero (scan sentinel); max_r receives the largest R partition.
__global__ __launch_bounds__(256) void k_task_counts64(const uint2* __restrict__ bounds,
                                                       uint32_t nparts,
                                                       uint32_t* __restrict__ rcnt,
                                                       uint32_t* __restrict__ jcnt,
                                                       uint32_t* __restrict__ max_r) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nr = 0;
    if (p < nparts) {
        const uint2 r = bounds[p], s = bounds[nparts + p];
        nr = r.y - r.x;
        const uint32_t ns = s.y - s.x;
        const uint32_t rc = (nr + kSm64Cap - 1) / kSm64Cap;
        rcnt[p] = rc;
        jcnt[p] = ns ? rc * ((ns + kSm64SChunk - 1) / kSm64SChunk) : 0;
    } else if (p == nparts) {
        rcnt[p] = 0;
        jcnt[p] = 0;
    }
    // one atomic per block (per-wave atomics on one word serialise)
    __shared__ uint32_t red[256 / TBJ_WAVE];
    for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) nr = max(nr, (uint32_t)__shfl_xor(nr, d, TBJ_WAVE));
    if (lane_id() == 0) red[wave_id()] = nr;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / TBJ_WAVE; w++) nr = max(nr, red[w]);
        if (nr) atomicMax(max_r, nr);
    }
}

// Task t of a list lies in the partition p with off[p] <= t < off[p + 1].
__device__ __forceinline__ uint32_t owner_part(const uint32_t* off, uint32_t nparts, uint32_t t) {
    uint32_t lo = 0, hi = nparts;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One thread per task: threads [0, nrtasks) write the R chunks, threads
// [nrtasks, nrtasks + njtasks) the join tasks (R chunk li / nsc, S chunk
// li % nsc of their partition).
__global__ void k_fill_tasks64(const uint2* __restrict__ bounds, uint32_t nparts,
                               const uint32_t* __restrict__ roff, const uint32_t* __restrict__ joff,
                               uint32_t nrtasks, uint32_t njtasks, SmTask* __restrict__ rtasks,
                               SmTask* __restrict__ jtasks) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrtasks) {
        const uint32_t p = owner_part(roff, nparts, i);
        const uint2 r = bounds[p];
        const uint2 s = bounds[nparts + p];
        SmTask t{};
        t.r_off = r.x + (i - roff[p]) * kSm64Cap;
        t.r_len = min((uint32_t)kSm64Cap, r.y - t.r_off);
        t.rchunk = i;
        t.multi = s.y - s.x > (uint32_t)kSm64SChunk;  // probed by more than one task
        rtasks[i] = t;
    } else if (i < nrtasks + njtasks) {
        const uint32_t jt = i - nrtasks;
        const uint32_t p = owner_part(joff, nparts, jt);
        const uint2 r = bounds[p], s = bounds[nparts + p];
        const uint32_t nsc = (s.y - s.x + kSm64SChunk - 1) / kSm64SChunk;
        const uint32_t li = jt - joff[p], rc = li / nsc, sc = li % nsc;
        SmTask t{};
        t.r_off = r.x + rc * kSm64Cap;
        t.r_len = min((uint32_t)kSm64Cap, r.y - t.r_off);
        t.s_off = s.x + sc * kSm64SChunk;
        t.s_len = min((uint32_t)kSm64SChunk, s.y - t.s_off);
        t.rchunk = roff[p] + rc;
        t.first = sc == 0;
        t.multi = nsc > 1;
        jtasks[jt] = t;
    }
}

// ---------------------------------------------------------------------------
// R chunk sort in LDS, by (key, payload) like the merge passes that may
// follow.  Buckets: n equal-width key ranges between the chunk's min and max
// key (a monotone map, so bucket order is key order); each bucket is then
// insertion-sorted by one thread.  A bucket above
// kBucketMax tuples (clustered keys) sends the whole chunk to a bitonic sort
// instead, and so does a chunk of one key (ordered by payload there).

constexpr uint32_t kSortItems = kSm64Cap / kSm64Block;
constexpr uint32_t kBucketMax = 32;

struct SortLds {
    ulonglong2 a[kSm64Cap];
    uint32_t cnt[kSm64Cap];
    uint32_t scratch[kSm64Block / TBJ_WAVE + 1];
    unsigned long long mm[2];
    uint32_t big;
};

// Bucket stage of chunk_sort for a chunk whose keys span lo .. lo + span
// (span > 0): leaves L.a[0, n) sorted by (key, payload), or sets L.big when
// a bucket is too large for its insertion sort.  Every thread must call it.
__device__ void bucket_sort(const ulonglong2 (&v)[kSortItems], uint32_t n, uint64_t lo,
                            uint64_t span, SortLds& L) {
    const uint32_t tid = threadIdx.x;
    const double scale = (double)n / ((double)span + 1.0);
    uint32_t bk[kSortItems], rk[kSortItems];
#pragma unroll
    for (uint32_t k = 0; k < kSortItems; k++) {
        if (k * kSm64Block + tid < n) {
            bk[k] = min(n - 1, (uint32_t)((double)(v[k].x - lo) * scale));
            rk[k] = atomicAdd(&L.cnt[bk[k]], 1u);
        }
    }
    __syncthreads();
    block_scan_lds<kSm64Block>(L.cnt, n, L.scratch);
#pragma unroll
    for (uint32_t k = 0; k < kSortItems; k++)
        if (k * kSm64Block + tid < n) L.a[L.cnt[bk[k]] + rk[k]] = v[k];
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kSm64Block) {
        const uint32_t s0 = L.cnt[i], s1 = i + 1 < n ? L.cnt[i + 1] : n;
        if (s1 - s0 > kBucketMax) {
            L.big = 1;
        } else {
            for (uint32_t x = s0 + 1; x < s1; x++) {
                const ulonglong2 t = L.a[x];
                uint32_t y = x;
                while (y > s0 && tuple_less(t, L.a[y - 1])) {
                    L.a[y] = L.a[y - 1];
                    y--;
                }
                L.a[y] = t;
            }
        }
    }
}

// v: item k of thread t is tuple k * kSm64Block + t (valid below n).  Leaves
// the chunk sorted by (key, payload) in L.a[0, n); every thread must call
// it; ends with a barrier.
__device__ void chunk_sort(const ulonglong2 (&v)[kSortItems], uint32_t n, SortLds& L) {
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        L.mm[0] = ~0ull;
        L.mm[1] = 0;
        L.big = 0;
    }
    for (uint32_t i = tid; i < n; i += kSm64Block) L.cnt[i] = 0;
    uint64_t mn = ~0ull, mx = 0;
#pragma unroll
    for (uint32_t k = 0; k < kSortItems; k++) {
        if (k * kSm64Block + tid < n) {
            mn = min(mn, (uint64_t)v[k].x);
            mx = max(mx, (uint64_t)v[k].x);
        }
    }
    wave_minmax_u64(mn, mx);
    __syncthreads();
    if (lane_id() == 0) {
        atomicMin(&L.mm[0], (unsigned long long)mn);
        atomicMax(&L.mm[1], (unsigned long long)mx);
    }
    __syncthreads();
    const uint64_t lo = L.mm[0], span = L.mm[1] - L.mm[0];
    if (span == 0) {  // one key: order by payload in the network below
#pragma unroll
        for (uint32_t k = 0; k < kSortItems; k++)
            if (k * kSm64Block + tid < n) L.a[k * kSm64Block + tid] = v[k];
        if (tid == 0) L.big = 1;
    } else {
        bucket_sort(v, n, lo, span, L);
    }
    __syncthreads();
    if (L.big) {
        const uint32_t np2 = next_pow2_u32(n);
        for (uint32_t j = n + tid; j < np2; j += kSm64Block) L.a[j] = make_ulonglong2(~0ull, ~0ull);
        __syncthreads();
        for (uint32_t k = 2; k <= np2; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = tid; i < np2; i += kSm64Block) {
                    const uint32_t ixj = i ^ j;
                    if (ixj > i) {
                        const bool up = (i & k) == 0;
                        const ulonglong2 x = L.a[i], y = L.a[ixj];
                        if (tuple_less(y, x) == up) {
                            L.a[i] = y;
                            L.a[ixj] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

__device__ __forceinline__ void load_chunk(const ulonglong2* __restrict__ src, uint32_t n,
                                           ulonglong2 (&v)[kSortItems]) {
#pragma unroll
    for (uint32_t k = 0; k < kSortItems; k++) {
        const uint32_t j = k * kSm64Block + threadIdx.x;
        if (j < n) v[k] = src[j];
    }
}

// load_chunk for n >= 1 with every load issued: items past n re-read the
// last tuple.
__device__ __forceinline__ void load_chunk_clamped(const ulonglong2* __restrict__ src, uint32_t n,
                                                   ulonglong2 (&v)[kSortItems]) {
#pragma unroll
    for (uint32_t k = 0; k < kSortItems; k++)
        v[k] = src[min(k * kSm64Block + threadIdx.x, n - 1)];
}

// S tuples j0 + q * kSm64Block + threadIdx.x (q < N) of a range of len >= 1
// tuples, every load issued (ok[q] = inside the range).
template <int N>
__device__ __forceinline__ void load_s_group(const ulonglong2* __restrict__ s, uint32_t len,
                                             uint32_t j0, uint64_t (&sk)[N], uint64_t (&sp)[N],
                                             bool (&ok)[N]) {
#pragma unroll
    for (int q = 0; q < N; q++) {
        const uint32_t j = j0 + q * kSm64Block + threadIdx.x;
        ok[q] = j < len;
        const ulonglong2 sv = s[min(j, len - 1)];
        sk[q] = sv.x;
        sp[q] = sv.y;
    }
}

// Sorts the R chunks that several join tasks probe (multi) from src into dst
// at the same offsets, so that the count kernel sorts each of them once
// rather than once per task (a hot key of a skewed S spreads its partition
// over thousands of tasks).  One workgroup per kSm64Block task descriptors:
// the multi ones are listed in LDS, then sorted one after the other.
__global__ __launch_bounds__(kSm64Block) void k_sort_shared64(const ulonglong2* __restrict__ src,
                                                              ulonglong2* __restrict__ dst,
                                                              const SmTask* __restrict__ rtasks,
                                                              uint32_t nrtasks) {
    __shared__ SortLds L;
    __shared__ uint32_t list[kSm64Block];
    __shared__ uint32_t nlist;
    if (threadIdx.x == 0) nlist = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * kSm64Block + threadIdx.x;
    if (i < nrtasks && rtasks[i].multi) list[atomicAdd(&nlist, 1u)] = i;
    __syncthreads();
    const uint32_t m = nlist;
    for (uint32_t q = 0; q < m; q++) {
        const SmTask tk = rtasks[list[q]];
        ulonglong2 v[kSortItems];
        load_chunk(src + tk.r_off, tk.r_len, v);
        chunk_sort(v, tk.r_len, L);
        for (uint32_t j = threadIdx.x; j < tk.r_len; j += kSm64Block) dst[tk.r_off + j] = L.a[j];
        __syncthreads();
    }
}

// Sorts every R chunk in place (partitions above one chunk, before merging).
__global__ __launch_bounds__(kSm64Block) void k_sort_r64(ulonglong2* __restrict__ Rb,
                                                         const SmTask* __restrict__ rtasks,
                                                         uint32_t nrtasks) {
    __shared__ SortLds L;
    for (uint32_t t = blockIdx.x; t < nrtasks; t += gridDim.x) {
        const SmTask tk = rtasks[t];
        ulonglong2 v[kSortItems];
        load_chunk(Rb + tk.r_off, tk.r_len, v);
        chunk_sort(v, tk.r_len, L);
        for (uint32_t j = threadIdx.x; j < tk.r_len; j += kSm64Block) Rb[tk.r_off + j] = L.a[j];
        __syncthreads();
    }
}

// One merge pass over R partitions longer than a chunk: sorted runs of
// `width` tuples are merged pairwise (merge path: each thread finds how
// many of its first output's predecessors come from the left run, then
// merges kSm64MergeItems outputs).  A run pair whose right run is empty is
// copied.
__global__ __launch_bounds__(kSm64Block) void k_merge64(const ulonglong2* __restrict__ src,
                                                        ulonglong2* __restrict__ dst,
                                                        const MergeBlock* __restrict__ blocks,
                                                        uint64_t width) {
    const MergeBlock mb = blocks[blockIdx.x];
    const uint64_t o = (uint64_t)mb.out_begin + (uint64_t)threadIdx.x * kSm64MergeItems;
    const uint64_t len = mb.seg_len;
    if (o >= len) return;
    const ulonglong2* s = src + mb.seg_off;
    ulonglong2* d = dst + mb.seg_off;
    const uint64_t q0 = o / (2 * width) * (2 * width);
    const uint64_t na = min(width, len - q0);
    const uint64_t b0 = q0 + na;
    const uint64_t nb = b0 < len ? min(width, len - b0) : 0;
    const ulonglong2* A = s + q0;
    const ulonglong2* B = s + b0;
    const uint64_t diag = o - q0;
    uint64_t lo = diag > nb ? diag - nb : 0, hi = min(diag, na);
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (!tuple_less(B[diag - 1 - mid], A[mid])) lo = mid + 1;
        else hi = mid;
    }
    uint64_t i = lo, j = diag - lo;
    for (uint32_t e = 0; e < kSm64MergeItems && i + j < na + nb; e++) {
        const bool take_a = j >= nb || (i < na && !tuple_less(B[j], A[i]));
        d[q0 + i + j] = take_a ? A[i] : B[j];
        if (take_a) i++;
        else j++;
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Length of the key run starting at sorted position j.
__device__ __forceinline__ uint32_t run_len(const uint64_t* keys, uint32_t n, uint32_t j) {
    uint32_t len = 1;
    while (j + len < n && keys[j + len] == keys[j]) len++;
    return len;
}

// Bucket of a 64-bit key in a table of 2^hb buckets (multiplicative hash:
// the high bits of key * 2^64/phi, so keys that share their low bits, as in
// a radix partition, still spread).
__device__ __forceinline__ uint32_t hash64(uint64_t key, uint32_t hb) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - hb));
}

__device__ __forceinline__ uint32_t hash_bits(uint32_t n) {
    uint32_t b = 1;
    while ((1u << b) < n) b++;
    return b;
}

// LDS index of the key runs of a sorted chunk of n keys: each run start j
// (its key differs from key j - 1) is chained under hash64(key) (1-based,
// 0 ends a chain).  head must be zero; a barrier must follow.
template <typename KeyAt>
__device__ __forceinline__ void index_runs(KeyAt key_at, uint32_t n, uint32_t hb, uint32_t* head,
                                           uint16_t* nxt) {
    for (uint32_t j = threadIdx.x; j < n; j += kSm64Block) {
        const uint64_t k = key_at(j);
        if (j == 0 || key_at(j - 1) != k) nxt[j] = (uint16_t)atomicExch(&head[hash64(k, hb)], j + 1);
    }
}

// Run starts of N keys at once through index_runs' table (lo[q] = first
// sorted position of key q, n when absent or invalid): the N chains are
// walked in lockstep, so every step issues N independent LDS reads.
template <int N, typename KeyAt>
__device__ __forceinline__ void find_runs(KeyAt key_at, uint32_t n, uint32_t hb,
                                          const uint32_t* head, const uint16_t* nxt,
                                          const uint64_t (&k)[N], const bool (&valid)[N],
                                          uint32_t (&lo)[N]) {
    uint32_t e[N];
#pragma unroll
    for (int q = 0; q < N; q++) {
        e[q] = valid[q] ? head[hash64(k[q], hb)] : 0;
        lo[q] = n;
    }
    for (;;) {
        bool any = false;
#pragma unroll
        for (int q = 0; q < N; q++) {
            if (e[q]) {
                const uint32_t i = e[q] - 1;
                if (key_at(i) == k[q]) {
                    lo[q] = i;
                    e[q] = 0;
                } else {
                    e[q] = nxt[i];
                }
                any |= e[q] != 0;
            }
        }
        if (!any) break;
    }
}

// A chunk probed by its task alone with at most kSmallS S tuples is not
// sorted: every R tuple is chained in an LDS hash table as loaded, its S
// tuples count their matches along the chains, and the emit kernel orders
// the few S tuples by key instead (so pairs still come out in key order).
constexpr uint32_t kSmallS = kSm64Block;

__device__ __forceinline__ bool small_single(const SmTask& t) {
    return !t.multi && t.s_len <= kSmallS;
}

// LDS index of every tuple of an unsorted chunk of n keys, chained under
// hash64(key) (1-based, 0 ends a chain).  head must be zero; a barrier must
// follow.
template <typename KeyAt>
__device__ __forceinline__ void index_all(KeyAt key_at, uint32_t n, uint32_t hb, uint32_t* head,
                                          uint16_t* nxt) {
    for (uint32_t j = threadIdx.x; j < n; j += kSm64Block)
        nxt[j] = (uint16_t)atomicExch(&head[hash64(key_at(j), hb)], j + 1);
}

// Matches of N keys along index_all's chains, walked in lockstep (N
// independent LDS reads per step); cks += R.payload + S.payload per match.
template <int N, typename KeyAt, typename PayAt>
__device__ __forceinline__ void count_all(KeyAt key_at, PayAt pay_at, uint32_t hb,
                                          const uint32_t* head, const uint16_t* nxt,
                                          const uint64_t (&k)[N], const uint64_t (&sp)[N],
                                          const bool (&valid)[N], uint32_t (&cnt)[N],
                                          uint64_t& cks) {
    uint32_t e[N];
#pragma unroll
    for (int q = 0; q < N; q++) {
        e[q] = valid[q] ? head[hash64(k[q], hb)] : 0;
        cnt[q] = 0;
    }
    for (;;) {
        bool any = false;
#pragma unroll
        for (int q = 0; q < N; q++) {
            if (e[q]) {
                const uint32_t i = e[q] - 1;
                if (key_at(i) == k[q]) {
                    cnt[q]++;
                    cks += pay_at(i) + sp[q];
                }
                e[q] = nxt[i];
                any |= e[q] != 0;
            }
        }
        if (!any) break;
    }
}

// Staging of a small chunk's pairs (count kernel, materialising).  The
// task's pairs are laid out as k_emit64 would lay them out -- S tuples in
// (key, position) order, each followed by its matches along the chains --
// in the task's own S positions of the free buffer stage (s_off .. s_off +
// s_len: the sorted R copies there use R positions only); pairs past s_len
// are not written.  When all of them fit (tm <= s_len: R keys unique, or
// nearly), the count kernel sets runs[r_off], which no task of a single-task
// chunk uses otherwise, to tm + 1, and k_emit64 copies the staged pairs to
// the chunk's output range instead of reading the R chunk again.  Every
// thread of the block calls it with its S tuple (ok: one of the task's) and
// that tuple's matches.  A task of at most one wave of S tuples has them all
// in wave 0, which ranks them with lane reads alone; a larger one goes
// through L.cnt (S keys and counts) and one barrier.
static_assert(kSmallS * (sizeof(uint64_t) + sizeof(uint32_t)) <= sizeof(SortLds::cnt),
              "the small chunk's S keys and counts fit the count array");
__device__ __forceinline__ void stage_small(const SmTask& tk, bool ok, uint64_t key, uint64_t pay,
                                            uint32_t cnt, uint32_t hb, SortLds& L,
                                            const uint32_t* head, const uint16_t* nxt,
                                            ulonglong2* __restrict__ stage) {
    const uint32_t ns = tk.s_len;
    uint32_t off = 0;  // pairs of the task's S tuples ordered before this one
    if (ns <= (uint32_t)TBJ_WAVE) {
        if (wave_id() != 0) return;
        const uint32_t lane = lane_id();
        for (uint32_t j = 0; j < ns; j++) {
            const uint64_t kj =
                ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(key >> 32), (int)j) << 32) |
                (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, (int)j);
            const uint32_t cj = (uint32_t)__builtin_amdgcn_readlane((int)cnt, (int)j);
            if (kj < key || (kj == key && j < lane)) off += cj;
        }
    } else {
        uint64_t* sks = reinterpret_cast<uint64_t*>(L.cnt);
        uint32_t* scnt = reinterpret_cast<uint32_t*>(sks + kSmallS);
        if (ok) {
            sks[threadIdx.x] = key;
            scnt[threadIdx.x] = cnt;
        }
        __syncthreads();
        if (ok)
            for (uint32_t j = 0; j < ns; j++) {
                const uint64_t kj = sks[j];
                if (kj < key || (kj == key && j < threadIdx.x)) off += scnt[j];
            }
    }
    if (!ok || !cnt) return;
    ulonglong2* o = stage + tk.s_off;
    for (uint32_t e = head[hash64(key, hb)]; e && off < ns; e = nxt[e - 1])
        if (L.a[e - 1].x == key) o[off++] = make_ulonglong2(L.a[e - 1].y, pay);
}

// Count.  counters[0] += matches, counters[1] += sum of (R.payload +
// S.payload).  MAT: runs[g] += S tuples matching the key run starting at
// sorted R position g; chunk_vol[rchunk] += pairs of the task.
template <bool MAT, bool PRESORTED>
__global__ __launch_bounds__(kSm64Block) __attribute__((amdgpu_waves_per_eu(6))) void k_count64(
    const ulonglong2* __restrict__ Rs, ulonglong2* __restrict__ Rout,
    const ulonglong2* __restrict__ Sb, const SmTask* __restrict__ tasks,
    const uint32_t* __restrict__ ntasks_dev, unsigned long long* __restrict__ runs,
    unsigned long long* __restrict__ chunk_vol, unsigned long long* __restrict__ counters) {
    __shared__ SortLds L;
    __shared__ uint32_t head[kSm64Cap];
    __shared__ uint16_t nxt[kSm64Cap];
    __shared__ unsigned long long tm;
    __shared__ unsigned long long red_m[kSm64Block / TBJ_WAVE], red_c[kSm64Block / TBJ_WAVE];
    uint64_t matches = 0, cks = 0;
    const uint32_t ntasks = *ntasks_dev;
    if (blockIdx.x >= ntasks) return;  // no task: nothing to count or flush
    for (uint32_t i = threadIdx.x; i < kSm64Cap; i += kSm64Block) head[i] = 0;
    const auto key_at = [&](uint32_t m) { return (uint64_t)L.a[m].x; };
    constexpr int G = 4;  // S tuples looked up in lockstep per thread
    // Software pipeline: a task's first S group is loaded before its R chunk
    // is sorted, and the next task's R chunk is loaded before this task's
    // probes (clamped, unconditional loads, so that nothing waits for them
    // early).
    // chunks that several tasks probe were sorted into Rout beforehand
    // (k_sort_shared64) unless the whole input is presorted
    const auto chunk_src = [&](const SmTask& x) {
        return (!PRESORTED && x.multi ? (const ulonglong2*)Rout : Rs) + x.r_off;
    };
    SmTask tk = tasks[blockIdx.x];
    ulonglong2 v[kSortItems];
    load_chunk_clamped(chunk_src(tk), tk.r_len, v);
    for (uint32_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const uint32_t n = tk.r_len;
        const uint32_t hb = hash_bits(n);
        uint64_t sk[G], sp[G];
        bool ok[G];
        load_s_group<G>(Sb + tk.s_off, tk.s_len, 0, sk, sp, ok);
        const bool small = !PRESORTED && small_single(tk);
        if (PRESORTED || tk.multi || small) {
            // presorted, or not sorted at all (small); slots past n (re-read
            // copies of the last tuple) are never read
#pragma unroll
            for (uint32_t k = 0; k < kSortItems; k++) L.a[k * kSm64Block + threadIdx.x] = v[k];
            __syncthreads();
        } else {
            chunk_sort(v, n, L);
            if (MAT && tk.first)
                for (uint32_t j = threadIdx.x; j < n; j += kSm64Block) Rout[tk.r_off + j] = L.a[j];
        }
        const uint32_t tn = t + gridDim.x;
        const SmTask tkn = tasks[tn < ntasks ? tn : t];
        load_chunk_clamped(chunk_src(tkn), tkn.r_len, v);
        if (small) index_all(key_at, n, hb, head, nxt);
        else index_runs(key_at, n, hb, head, nxt);
        if (MAT)
            for (uint32_t j = threadIdx.x; j < n; j += kSm64Block) L.cnt[j] = 0;
        if (threadIdx.x == 0) tm = 0;
        __syncthreads();
        uint64_t mt = 0;
        if (small) {  // one group: s_len <= kSmallS <= kSm64Block
            uint32_t c[G];
            count_all<G>(key_at, [&](uint32_t m) { return (uint64_t)L.a[m].y; }, hb, head, nxt,
                         sk, sp, ok, c, cks);
#pragma unroll
            for (int q = 0; q < G; q++) mt += c[q];
            if (MAT) stage_small(tk, ok[0], sk[0], sp[0], c[0], hb, L, head, nxt, Rout);
        }
        for (uint32_t j0 = 0; !small;) {
            uint32_t lo[G];
            find_runs<G>(key_at, n, hb, head, nxt, sk, ok, lo);
#pragma unroll
            for (int q = 0; q < G; q++) {
                const bool hit = lo[q] < n;
                if (hit) {
                    for (uint32_t e = lo[q]; e < n && L.a[e].x == sk[q]; e++) {
                        mt++;
                        cks += L.a[e].y + sp[q];
                    }
                }
                if (MAT) agg_add(L.cnt, hit ? lo[q] : kNone);
            }
            j0 += G * kSm64Block;
            if (j0 >= tk.s_len) break;
            load_s_group<G>(Sb + tk.s_off, tk.s_len, j0, sk, sp, ok);
        }
        matches += mt;
        if (MAT) {
            mt = wave_reduce_u64(mt);
            if (lane_id() == 0 && mt) atomicAdd(&tm, (unsigned long long)mt);
            __syncthreads();
            // a chunk shared by several tasks accumulates its run counts for
            // k_run_offsets64; a chunk probed by this task alone needs only its
            // volume (k_emit64 recounts its runs)
            if (tk.multi) {
                for (uint32_t j = threadIdx.x; j < n; j += kSm64Block)
                    if (L.cnt[j]) atomicAdd(&runs[tk.r_off + j], (unsigned long long)L.cnt[j]);
                if (threadIdx.x == 0 && tm) atomicAdd(&chunk_vol[tk.rchunk], tm);
            } else if (threadIdx.x == 0) {
                chunk_vol[tk.rchunk] = tm;
                if (small) runs[tk.r_off] = tm <= tk.s_len ? tm + 1 : 0;  // stage_small
            }
        }
        __syncthreads();
        // (the next task's sort or load ends with a barrier before its index)
        for (uint32_t i = threadIdx.x; i < (1u << hb); i += kSm64Block) head[i] = 0;
        tk = tkn;
    }
    matches = wave_reduce_u64(matches);
    cks = wave_reduce_u64(cks);
    if (lane_id() == 0) {
        red_m[wave_id()] = matches;
        red_c[wave_id()] = cks;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = 0, c = 0;
        for (int w = 0; w < kSm64Block / TBJ_WAVE; w++) {
            m += red_m[w];
            c += red_c[w];
        }
        if (m) atomicAdd(&counters[0], m);
        if (c) atomicAdd(&counters[1], c);
    }
}

// Device-wide exclusive scan of u64 values in three launches: per-block
// sums of 8192-value chunks, one scan of those sums (*total = grand total),
// then each block scans its chunk from its carry.
constexpr uint32_t kScan64Chunk = 1024 * 8;

__global__ __launch_bounds__(1024) void k_scan64_reduce(const unsigned long long* __restrict__ a,
                                                        uint32_t n,
                                                        unsigned long long* __restrict__ part) {
    __shared__ unsigned long long red[1024 / TBJ_WAVE];
    const uint64_t base = (uint64_t)blockIdx.x * kScan64Chunk;
    unsigned long long s = 0;
    for (uint32_t i = threadIdx.x; i < kScan64Chunk; i += 1024)
        if (base + i < n) s += a[base + i];
    s = wave_reduce_u64(s);
    if (lane_id() == 0) red[wave_id()] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 1024 / TBJ_WAVE; w++) t += red[w];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(1024) void k_scan64_partials(unsigned long long* __restrict__ part,
                                                          uint32_t nb,
                                                          unsigned long long* __restrict__ total) {
    __shared__ unsigned long long scratch[1024 / TBJ_WAVE + 1];
    unsigned long long carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const unsigned long long v = i < nb ? part[i] : 0;
        unsigned long long tot;
        const unsigned long long ex = block_exclusive_scan_u64<1024>(v, scratch, &tot);
        if (i < nb) part[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void k_scan64_down(unsigned long long* __restrict__ a,
                                                      uint32_t n,
                                                      const unsigned long long* __restrict__ part) {
    __shared__ unsigned long long scratch[1024 / TBJ_WAVE + 1];
    __shared__ unsigned long long tile[kScan64Chunk];
    const uint64_t base = (uint64_t)blockIdx.x * kScan64Chunk;
    for (uint32_t i = threadIdx.x; i < kScan64Chunk; i += 1024)
        tile[i] = base + i < n ? a[base + i] : 0;
    __syncthreads();
    unsigned long long v[8], s = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        v[q] = tile[threadIdx.x * 8 + q];
        s += v[q];
    }
    unsigned long long tot;
    unsigned long long run = part[blockIdx.x] + block_exclusive_scan_u64<1024>(s, scratch, &tot);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        tile[threadIdx.x * 8 + q] = run;
        run += v[q];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kScan64Chunk; i += 1024)
        if (base + i < n) a[base + i] = tile[i];
}

// Per chunk with output: runs[g] = chunk_off + sum of count x run length over
// the runs before g, for the run starts g with a nonzero count.  Keys are
// read only where a run starts (count > 0), to measure its length.
__global__ __launch_bounds__(kSm64Block) void k_run_offsets64(
    const ulonglong2* __restrict__ Rsorted, const SmTask* __restrict__ rtasks, uint32_t nrtasks,
    const unsigned long long* __restrict__ chunk_off, const unsigned long long* __restrict__ total,
    unsigned long long* __restrict__ runs) {
    __shared__ unsigned long long scratch[kSm64Block / TBJ_WAVE + 1];
    constexpr uint32_t kPer = kSm64Cap / kSm64Block;
    for (uint32_t c = blockIdx.x; c < nrtasks; c += gridDim.x) {
        const SmTask tk = rtasks[c];
        if (!tk.multi) continue;  // k_emit64 places a single task's runs itself
        const unsigned long long base = chunk_off[c];
        const unsigned long long end = c + 1 < nrtasks ? chunk_off[c + 1] : *total;
        if (end == base) continue;  // no output: the chunk may be unsorted
        const uint32_t n = tk.r_len;
        const ulonglong2* Rc = Rsorted + tk.r_off;
        uint32_t cnt[kPer];
        unsigned long long vol[kPer], vsum = 0;
#pragma unroll
        for (uint32_t q = 0; q < kPer; q++) {
            const uint32_t j = threadIdx.x * kPer + q;
            cnt[q] = j < n ? (uint32_t)runs[tk.r_off + j] : 0;
            vol[q] = 0;
            if (cnt[q]) {
                const uint64_t k = Rc[j].x;
                uint32_t len = 1;
                while (j + len < n && Rc[j + len].x == k) len++;
                vol[q] = (unsigned long long)cnt[q] * len;
            }
            vsum += vol[q];
        }
        unsigned long long tot;
        unsigned long long run = base + block_exclusive_scan_u64<kSm64Block>(vsum, scratch, &tot);
#pragma unroll
        for (uint32_t q = 0; q < kPer; q++) {
            if (cnt[q]) runs[tk.r_off + threadIdx.x * kPer + q] = run;
            run += vol[q];
        }
        __syncthreads();
    }
}

// Clears the run counters of the R chunks that several tasks share (the
// chunk's first task does it, before any task adds to them).
__global__ void k_clear_shared_runs(const SmTask* __restrict__ tasks, uint32_t ntasks,
                                    unsigned long long* __restrict__ runs) {
    for (uint32_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const SmTask tk = tasks[t];
        if (!(tk.multi && tk.first)) continue;
        for (uint32_t j = threadIdx.x; j < tk.r_len; j += blockDim.x) runs[tk.r_off + j] = 0;
    }
}

// Emit.  runs[g] holds the next free output slot of the run starting at g;
// the S tuples of a task claim run-length blocks from it.  A small chunk
// whose pairs the count kernel staged (runs[r_off] = pairs + 1) is copied
// from its S positions of Rsorted's buffer (the free buffer when small_ok).
__global__ __launch_bounds__(kSm64Block) __attribute__((amdgpu_waves_per_eu(6))) void k_emit64(
    const ulonglong2* __restrict__ Rsorted, const ulonglong2* __restrict__ Sb,
    const SmTask* __restrict__ tasks, const uint32_t* __restrict__ ntasks_dev,
    unsigned long long* __restrict__ runs, const unsigned long long* __restrict__ chunk_off,
    ulonglong2* __restrict__ pairs, uint64_t cap, uint32_t small_ok) {
    __shared__ uint64_t rk[kSm64Cap];
    __shared__ uint32_t lcnt[kSm64Cap];
    __shared__ uint32_t scratch[kSm64Block / TBJ_WAVE + 1];
    __shared__ unsigned long long lbase[kSm64Cap];
    __shared__ uint32_t head[kSm64Cap];
    __shared__ uint16_t nxt[kSm64Cap];
    const uint32_t ntasks = *ntasks_dev;
    for (uint32_t i = threadIdx.x; i < kSm64Cap; i += kSm64Block) head[i] = 0;
    const auto key_at = [&](uint32_t m) { return rk[m]; };
    for (uint32_t t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const SmTask tk = tasks[t];
        const uint32_t n = tk.r_len;
        const uint32_t hb = hash_bits(n);
        if (small_ok && small_single(tk)) {
            const unsigned long long staged = runs[tk.r_off];
            if (staged) {  // laid out by the count kernel (stage_small)
                const unsigned long long o = chunk_off[tk.rchunk] + threadIdx.x;
                if (threadIdx.x < staged - 1 && o < cap) pairs[o] = Rsorted[tk.s_off + threadIdx.x];
                continue;
            }
            // unsorted R (the count kernel did not sort it): every tuple
            // chained, one S tuple per thread ranked by (key, position), the
            // pairs laid out in that order from the chunk's offset
            const ulonglong2* Ru = Sb + tk.r_off;  // R's partitions share the buffer
            for (uint32_t j = threadIdx.x; j < n; j += kSm64Block) rk[j] = Ru[j].x;
            const bool ok = threadIdx.x < tk.s_len;
            const ulonglong2 sv = ok ? Sb[tk.s_off + threadIdx.x] : make_ulonglong2(0, 0);
            __syncthreads();
            index_all(key_at, n, hb, head, nxt);
            uint64_t* sks = reinterpret_cast<uint64_t*>(lbase);
            sks[threadIdx.x] = sv.x;
            __syncthreads();
            uint32_t m = 0;
            if (ok)
                for (uint32_t e = head[hash64(sv.x, hb)]; e; e = nxt[e - 1]) m += rk[e - 1] == sv.x;
            uint32_t rank = 0;
            if (ok)
                for (uint32_t j = 0; j < tk.s_len; j++) {
                    const uint64_t kj = sks[j];
                    rank += kj < sv.x || (kj == sv.x && j < threadIdx.x);
                }
            if (ok) lcnt[rank] = m;
            __syncthreads();
            block_scan_lds<kSm64Block>(lcnt, tk.s_len, scratch);
            if (ok && m) {
                unsigned long long o = chunk_off[tk.rchunk] + lcnt[rank];
                for (uint32_t e = head[hash64(sv.x, hb)]; e; e = nxt[e - 1])
                    if (rk[e - 1] == sv.x) {
                        if (o < cap) pairs[o] = make_ulonglong2(Ru[e - 1].y, sv.y);
                        o++;
                    }
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < (1u << hb); i += kSm64Block) head[i] = 0;
            continue;
        }
        const ulonglong2* Rc = Rsorted + tk.r_off;
        for (uint32_t j = threadIdx.x; j < n; j += kSm64Block) {
            rk[j] = Rc[j].x;
            lcnt[j] = 0;
        }
        __syncthreads();
        index_runs(key_at, n, hb, head, nxt);
        __syncthreads();
        uint64_t sp[kSm64SItems];
        uint32_t slo[kSm64SItems], srank[kSm64SItems];
        {
            uint64_t sk[kSm64SItems];
            bool ok[kSm64SItems];
#pragma unroll
            for (uint32_t k = 0; k < kSm64SItems; k++) {
                const uint32_t j = k * kSm64Block + threadIdx.x;
                ok[k] = j < tk.s_len;
                const ulonglong2 s = ok[k] ? Sb[tk.s_off + j] : make_ulonglong2(0, 0);
                sk[k] = s.x;
                sp[k] = s.y;
            }
            find_runs<kSm64SItems>(key_at, n, hb, head, nxt, sk, ok, slo);
#pragma unroll
            for (uint32_t k = 0; k < kSm64SItems; k++) {
                if (slo[k] >= n) slo[k] = kNone;
                srank[k] = agg_inc(lcnt, slo[k]);
            }
        }
        __syncthreads();
        if (tk.multi) {  // claim slots inside each run's range (k_run_offsets64)
            for (uint32_t j = threadIdx.x; j < n; j += kSm64Block) {
                const uint32_t c = lcnt[j];
                if (c)
                    lbase[j] = atomicAdd(&runs[tk.r_off + j],
                                         (unsigned long long)c * run_len(rk, n, j));
            }
        } else {  // the chunk's only task: runs laid out in key order from its offset
            // (a run's volume is at most 4096 S tuples x 2048 R tuples: 32 bits)
            for (uint32_t j = threadIdx.x; j < n; j += kSm64Block) {
                const uint32_t c = lcnt[j];
                lcnt[j] = c ? c * run_len(rk, n, j) : 0;
            }
            __syncthreads();
            block_scan_lds<kSm64Block>(lcnt, n, scratch);
            const unsigned long long base = chunk_off[tk.rchunk];
            for (uint32_t j = threadIdx.x; j < n; j += kSm64Block) lbase[j] = base + lcnt[j];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < kSm64SItems; k++) {
            if (slo[k] != kNone) {
                const uint32_t lo = slo[k];
                const uint32_t len = run_len(rk, n, lo);
                unsigned long long o = lbase[lo] + (unsigned long long)srank[k] * len;
                for (uint32_t e = 0; e < len; e++, o++)
                    if (o < cap) pairs[o] = make_ulonglong2(Rc[lo + e].y, sp[k]);
            }
        }
        __syncthreads();
        // (the next task's key load ends with a barrier before its index)
        for (uint32_t i = threadIdx.x; i < (1u << hb); i += kSm64Block) head[i] = 0;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers

size_t sm64_model_lds(const Bin64& b) { return b.kind == kBinLearned ? sizeof(LdsModel64) : 0; }

void sm64_fit(hipStream_t st, const ulonglong2* R, uint64_t nR, uint32_t m, uint32_t nseg,
              uint32_t bits, uint64_t* sample, Model64* model, uint32_t* lin_out) {
    hipLaunchKernelGGL(k_sample64, dim3((m + 255) / 256), dim3(256), 0, st, R, nR, m, sample);
    hipLaunchKernelGGL(k_fit64, dim3(1), dim3(1024), 0, st, sample, m, nseg, bits, model,
                       lin_out);
}

void sm64_make_tiles(hipStream_t st, TileDesc* tiles, SegDesc* segs, uint32_t nR, uint32_t nS,
                     uint32_t nbins, uint32_t* ntiles_dev, bool rows) {
    const uint32_t t = (nR + kSm64Span - 1) / kSm64Span + (nS + kSm64Span - 1) / kSm64Span;
    hipLaunchKernelGGL(k_make_tiles64, dim3(t ? (t + 255) / 256 : 1), dim3(256), 0, st, tiles,
                       segs, nR, nS, nbins, ntiles_dev, (uint32_t)rows);
}

void sm64_sample_rows(hipStream_t st, const ulonglong2* R, uint32_t nR, const ulonglong2* S,
                      uint32_t nS, uint32_t nrowR, uint32_t nrowS, const Bin64& bsr,
                      const Bin64& bss, uint32_t F1, uint32_t* cnt1, uint32_t* s_out) {
    const uint32_t rows = nrowR + nrowS;
    if (!rows) return;
    hipLaunchKernelGGL(k_sample_rows64, dim3(rows * kS64ChunksPerRow), dim3(256),
                       sm64_model_lds(bsr), st, R, nR, S, nS, nrowR, nrowS, bsr, bss, F1, cnt1,
                       s_out);
}

void sm64_part_segments(hipStream_t st, const uint32_t* first, uint32_t F1, uint32_t F2,
                        uint32_t nrowR, uint32_t nrowS, SegDesc* segs2) {
    hipLaunchKernelGGL(k_part_segments64, dim3(1), dim3(kScanBlock), 0, st, first, F1, F2, nrowR,
                       nrowS, segs2);
}

void sm64_hist(hipStream_t st, uint32_t grid, const ulonglong2* in, const TileDesc* tiles,
               const uint32_t* ntiles_dev, uint32_t tile_base, uint16_t* hist, const Bin64& b,
               uint32_t nbins, unsigned long long* mm) {
    if (!grid) return;
    hipLaunchKernelGGL(k_hist64, dim3(grid), dim3(kSm64Block), sm64_model_lds(b), st, in, tiles,
                       ntiles_dev, tile_base, hist, b, nbins, mm);
}

void sm64_scatter(hipStream_t st, uint32_t grid, const ulonglong2* in, ulonglong2* out,
                  const TileDesc* tiles, const uint32_t* ntiles_dev, uint32_t tile_base,
                  const uint32_t* offs, const Bin64& b, uint32_t nbins,
                  const Sm64Claims* claims) {
    if (!grid) return;
    const size_t lds = sm64_model_bytes_aligned(b.kind) + 3 * sizeof(uint32_t) * nbins;
    Claims64 cl{};
    if (claims) {
        cl.starts = claims->starts;
        cl.cursors = claims->cursors;
        cl.overflow = claims->overflow;
        cl.cb0 = claims->cb0;
        cl.cmul = claims->cmul;
        cl.mm = claims->mm;
        cl.gate = claims->gate;
        cl.gate_min = claims->gate_min;
        hipLaunchKernelGGL(k_scatter64<true>, dim3(grid), dim3(kSm64ScBlock), lds, st, in, out,
                           tiles, ntiles_dev, tile_base, tile_base + grid, offs, b, nbins, cl);
    } else {
        hipLaunchKernelGGL(k_scatter64<false>, dim3(grid), dim3(kSm64ScBlock), lds, st, in, out,
                           tiles, ntiles_dev, tile_base, tile_base + grid, offs, b, nbins, cl);
    }
}

void sm64_segments(hipStream_t st, const uint32_t* scanned1, const SegDesc* segs1, uint32_t nbins1,
                   uint32_t nbins2, SegDesc* segs2, TileDesc* tiles2, uint32_t* ntiles2_dev,
                   uint32_t max_tiles2) {
    hipLaunchKernelGGL(k_make_segments64, dim3(1), dim3(kScanBlock), 0, st, scanned1, segs1,
                       nbins1, nbins2, segs2, ntiles2_dev);
    hipLaunchKernelGGL(k_make_tiles64_2, dim3((max_tiles2 + 255) / 256), dim3(256), 0, st, segs2,
                       2 * nbins1, ntiles2_dev, tiles2);
}

void sm64_bounds(hipStream_t st, const uint32_t* scanned, const SegDesc* segs, uint32_t outer,
                 uint32_t inner, uint32_t passes, uint32_t nR, uint32_t nS, uint2* bounds) {
    const uint32_t n = 2 * outer * inner;
    hipLaunchKernelGGL(k_bounds64, dim3((n + 255) / 256), dim3(256), 0, st, scanned, segs, outer,
                       inner, passes, nR, nS, bounds);
}

void sm64_task_counts(hipStream_t st, const uint2* bounds, uint32_t nparts, uint32_t* rcnt,
                      uint32_t* jcnt, uint32_t* max_r) {
    hipLaunchKernelGGL(k_task_counts64, dim3((nparts + 1 + 255) / 256), dim3(256), 0, st, bounds,
                       nparts, rcnt, jcnt, max_r);
}

void sm64_fill_tasks(hipStream_t st, const uint2* bounds, uint32_t nparts, const uint32_t* roff,
                     const uint32_t* joff, uint32_t nrtasks, uint32_t njtasks, SmTask* rtasks,
                     SmTask* jtasks) {
    const uint64_t n = (uint64_t)nrtasks + njtasks;
    if (n)
        hipLaunchKernelGGL(k_fill_tasks64, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st,
                           bounds, nparts, roff, joff, nrtasks, njtasks, rtasks, jtasks);
}

// The persistent kernels below stride their task lists by gridDim.x, so a
// grid larger than what stays resident leaves the surplus workgroups to
// start late and run a tail: clamp to the kernel's resident capacity.
template <typename K>
static uint32_t resident_grid(K kernel, uint32_t want) {
    // capacity per kernel, looked up once (kernels of one signature share K)
    static const void* keys[8] = {};
    static uint32_t caps[8] = {};
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    uint32_t i = 0;
    while (i < 8 && keys[i] && keys[i] != (const void*)kernel) i++;
    if (i == 8) return want;
    if (!keys[i]) {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kSm64Block, 0) !=
                hipSuccess ||
            per < 1)
            per = 1;
        caps[i] = (uint32_t)(cus > 0 ? cus : 256) * (uint32_t)per;
        keys[i] = (const void*)kernel;
    }
    return want < caps[i] ? want : caps[i];
}

void sm64_sort_r(hipStream_t st, uint32_t grid, ulonglong2* Rb, const SmTask* rtasks,
                 uint32_t nrtasks) {
    grid = resident_grid(k_sort_r64, grid);
    if (nrtasks && grid)
        hipLaunchKernelGGL(k_sort_r64, dim3(grid), dim3(kSm64Block), 0, st, Rb, rtasks, nrtasks);
}

void sm64_sort_shared(hipStream_t st, const ulonglong2* src, ulonglong2* dst,
                      const SmTask* rtasks, uint32_t nrtasks) {
    if (!nrtasks) return;
    hipLaunchKernelGGL(k_sort_shared64, dim3((nrtasks + kSm64Block - 1) / kSm64Block),
                       dim3(kSm64Block), 0, st, src, dst, rtasks, nrtasks);
}

void sm64_merge(hipStream_t st, const ulonglong2* src, ulonglong2* dst, const MergeBlock* blocks,
                uint32_t nblocks, uint64_t width) {
    if (nblocks)
        hipLaunchKernelGGL(k_merge64, dim3(nblocks), dim3(kSm64Block), 0, st, src, dst, blocks,
                           width);
}

void sm64_fill_u64(hipStream_t st, unsigned long long* p, uint64_t n, unsigned long long v) {
    if (!n) return;
    uint64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)g), dim3(256), 0, st, p, n, v);
}

void sm64_count(hipStream_t st, uint32_t grid, bool mat, bool presorted, const ulonglong2* Rs,
                ulonglong2* Rout, const ulonglong2* Sb, const SmTask* tasks,
                const uint32_t* ntasks_dev, unsigned long long* runs,
                unsigned long long* chunk_vol, unsigned long long* counters) {
    const dim3 b(kSm64Block);
    if (mat && presorted)
        hipLaunchKernelGGL((k_count64<true, true>), dim3(resident_grid(k_count64<true, true>, grid)),
                           b, 0, st, Rs, Rout, Sb, tasks, ntasks_dev, runs, chunk_vol, counters);
    else if (mat)
        hipLaunchKernelGGL((k_count64<true, false>),
                           dim3(resident_grid(k_count64<true, false>, grid)), b, 0, st, Rs, Rout,
                           Sb, tasks, ntasks_dev, runs, chunk_vol, counters);
    else if (presorted)
        hipLaunchKernelGGL((k_count64<false, true>),
                           dim3(resident_grid(k_count64<false, true>, grid)), b, 0, st, Rs, Rout,
                           Sb, tasks, ntasks_dev, runs, chunk_vol, counters);
    else
        hipLaunchKernelGGL((k_count64<false, false>),
                           dim3(resident_grid(k_count64<false, false>, grid)), b, 0, st, Rs, Rout,
                           Sb, tasks, ntasks_dev, runs, chunk_vol, counters);
}

void sm64_clear_shared_runs(hipStream_t st, const SmTask* tasks, uint32_t ntasks,
                            unsigned long long* runs) {
    if (ntasks)
        hipLaunchKernelGGL(k_clear_shared_runs, dim3(std::min<uint32_t>(ntasks, 4096)), dim3(256),
                           0, st, tasks, ntasks, runs);
}

uint32_t sm64_scan_partials(uint32_t n) { return (n + kScan64Chunk - 1) / kScan64Chunk + 1; }

void sm64_scan_u64(hipStream_t st, unsigned long long* a, uint32_t n, unsigned long long* part,
                   unsigned long long* total) {
    const uint32_t nb = (n + kScan64Chunk - 1) / kScan64Chunk;
    if (!nb) return;
    hipLaunchKernelGGL(k_scan64_reduce, dim3(nb), dim3(1024), 0, st, a, n, part);
    hipLaunchKernelGGL(k_scan64_partials, dim3(1), dim3(1024), 0, st, part, nb, total);
    hipLaunchKernelGGL(k_scan64_down, dim3(nb), dim3(1024), 0, st, a, n, part);
}

void sm64_run_offsets(hipStream_t st, uint32_t grid, const ulonglong2* Rsorted,
                      const SmTask* rtasks, uint32_t nrtasks,
                      const unsigned long long* chunk_off, const unsigned long long* total,
                      unsigned long long* runs) {
    if (nrtasks)
        hipLaunchKernelGGL(k_run_offsets64, dim3(resident_grid(k_run_offsets64, grid)),
                           dim3(kSm64Block), 0, st, Rsorted, rtasks, nrtasks, chunk_off, total,
                           runs);
}

void sm64_emit(hipStream_t st, uint32_t grid, const ulonglong2* Rsorted, const ulonglong2* Sb,
               const SmTask* tasks, const uint32_t* ntasks_dev, unsigned long long* runs,
               const unsigned long long* chunk_off, ulonglong2* pairs, uint64_t cap,
               bool presorted) {
    hipLaunchKernelGGL(k_emit64, dim3(resident_grid(k_emit64, grid)), dim3(kSm64Block), 0, st,
                       Rsorted, Sb, tasks, ntasks_dev, runs, chunk_off, pairs, cap,
                       (uint32_t)!presorted);
}

}  // namespace tbj
