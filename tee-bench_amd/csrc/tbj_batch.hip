// Batches of joins: partition bounds over many joins at once, a persistent
// count-only LDS hash join whose tasks are decoded from a per-partition
// prefix (no task list, no host round trip), the bulk build of the PHT
// joins' tables (a block of buckets per workgroup) and their probe in one
// launch, and the device generator of the batch workload.
//
// The join semantics are RHO's (radix partition, then a per-partition
// bucket-chained hash join: Joins/radix_join.c:373-440, 1557-1561) and PHT's
// (one shared hash table per join: Joins/no_partitioning_join.h:543-623);
// the batch itself has no reference counterpart.
#include "tbj_batch.h"
#include "tbj_device.h"
#include "tbj_kernels.h"
#include "tbj_pht.h"

namespace tbj {

namespace {

// Partition bounds of a group of nj joins.  Partition q = j * F1 * F2 + p
// of join j; bounds[q] = R range, bounds[NP + q] = S range.  passes 0: the
// joins' own arena ranges (segs[2j], segs[2j+1]); passes 1: first-pass bins
// of segs = (R_0, S_0, R_1, S_1, ...); passes 2: second-pass segments,
// segment (2j + rel) * F1 + p / F2, bin p % F2.
__global__ void k_bbounds(const uint32_t* __restrict__ scanned, const SegDesc* __restrict__ segs,
                          uint32_t nj, uint32_t F1, uint32_t F2, uint32_t passes,
                          uint2* __restrict__ bounds) {
    const uint32_t P = F1 * F2, NP = nj * P;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * NP) return;
    const uint32_t rel = i / NP, q = i % NP, j = q / P, p = q % P;
    if (passes == 0) {
        const SegDesc sd = segs[2 * j + rel];
        bounds[i] = make_uint2(sd.in_off, sd.in_off + sd.n);
        return;
    }
    const uint32_t seg = passes == 1 ? 2 * j + rel : (2 * j + rel) * F1 + p / F2;
    const uint32_t bin = passes == 1 ? p : p % F2;
    const SegDesc sd = segs[seg];
    bounds[i] = make_uint2(scanned[sd.hist_base + (uint64_t)bin * sd.ngroups],
                           scanned[sd.hist_base + (uint64_t)(bin + 1) * sd.ngroups]);
}

struct TaskView {
    uint32_t r_off, r_len, s_off, s_len, join;
};

// Decodes task t (all 64 lanes of one wave): a 64-ary search for the last
// partition q with task_off[q] <= t (each step one load per lane, so
// log64(NP) round trips), then the task's R and S chunks inside q.  Lane 0
// returns the view; join = ~0 past the last task.
__device__ __forceinline__ TaskView decode_task(uint32_t t, uint32_t ntasks,
                                                const uint32_t* __restrict__ task_off, uint32_t NP,
                                                uint32_t P, const uint2* __restrict__ bounds,
                                                uint32_t s_chunk,
                                                const uint32_t* __restrict__ join_ids) {
    TaskView v{};
    v.join = 0xFFFFFFFFu;
    if (t >= ntasks) return v;
    const uint32_t lane = lane_id();
    uint32_t lo = 0, hi = NP;
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + TBJ_WAVE - 1) / TBJ_WAVE;
        const uint32_t idx = lo + lane * step;
        const bool le = idx < hi && task_off[idx] <= t;
        const uint64_t m = __ballot(le);  // lanes 0..L hold (task_off is non-decreasing)
        const uint32_t L = 63 - __clzll((unsigned long long)m);
        const uint32_t nlo = lo + L * step;
        hi = min(hi, nlo + step);
        lo = nlo;
    }
    const uint2 r = bounds[lo], s = bounds[NP + lo];
    const uint32_t nsc = (s.y - s.x + s_chunk - 1) / s_chunk;
    const uint32_t li = t - task_off[lo];
    v.r_off = r.x + (li / nsc) * kJoinCap;
    v.r_len = min((uint32_t)kJoinCap, r.y - v.r_off);
    v.s_off = s.x + (li % nsc) * s_chunk;
    v.s_len = min(s_chunk, s.y - v.s_off);
    v.join = join_ids[lo / P];
    return v;
}

// Persistent count-only join.  Task t lies in the partition q with
// task_off[q] <= t < task_off[q + 1]; inside it, task li covers R chunk
// li / nsc and S chunk li % nsc (nsc = S chunks of the partition).
// counts[join_ids[q / P]] += matches of the task.  Wave 0 claims and decodes
// the next task while the current one runs.
__global__ __launch_bounds__(kJoinBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_bjoin(
    const uint2* __restrict__ Rb, const uint2* __restrict__ Sb, const uint2* __restrict__ bounds,
    uint32_t NP, uint32_t P, const uint32_t* __restrict__ task_off,
    const uint32_t* __restrict__ ntasks_dev, uint32_t* __restrict__ task_ctr, uint32_t s_chunk,
    uint32_t hshift, const uint32_t* __restrict__ join_ids, unsigned long long* __restrict__ counts) {
    constexpr int RITEMS = kJoinCap / kJoinBlock;
    constexpr int SITEMS = kJoinSItems;
    __shared__ uint32_t rkey[kJoinCap];
    __shared__ uint32_t head[kJoinCap];
    __shared__ uint16_t nxt[kJoinCap];
    __shared__ TaskView tv;
    __shared__ unsigned long long tm;
    const uint32_t ntasks = *ntasks_dev;
    const bool w0 = threadIdx.x < TBJ_WAVE;
    for (uint32_t i = threadIdx.x; i < kJoinCap; i += kJoinBlock) head[i] = 0;
    if (w0) {
        uint32_t t = 0;
        if (threadIdx.x == 0) t = atomicAdd(task_ctr, 1u);
        t = __shfl(t, 0, TBJ_WAVE);
        const TaskView v = decode_task(t, ntasks, task_off, NP, P, bounds, s_chunk, join_ids);
        if (threadIdx.x == 0) {
            tv = v;
            tm = 0;
        }
    }
    __syncthreads();
    for (;;) {
        const TaskView tk = tv;
        if (tk.join == 0xFFFFFFFFu) break;
        const uint32_t n = tk.r_len;
        uint2 rv[RITEMS], sv[SITEMS];
#pragma unroll
        for (int k = 0; k < RITEMS; k++) {
            const uint32_t j = k * kJoinBlock + threadIdx.x;
            if (j < n) rv[k] = Rb[tk.r_off + j];
        }
#pragma unroll
        for (int k = 0; k < SITEMS; k++) {
            const uint32_t j = k * kJoinBlock + threadIdx.x;
            if (j < tk.s_len) sv[k] = Sb[tk.s_off + j];
        }
        const uint32_t hmask = next_pow2_u32(n) - 1;
#pragma unroll
        for (int k = 0; k < RITEMS; k++) {
            const uint32_t j = k * kJoinBlock + threadIdx.x;
            if (j < n) {
                rkey[j] = rv[k].x;
                nxt[j] = (uint16_t)atomicExch(&head[(rv[k].x >> hshift) & hmask], j + 1);
            }
        }
        __syncthreads();
        // the task's S range in sub-chunks of SITEMS * kJoinBlock against the
        // one build (sub-chunk 0 was loaded with R); chains of G S tuples
        // walked in lockstep: G independent LDS reads per step
        uint32_t c = 0;
        constexpr int G = 8;
        for (uint32_t s0 = 0; s0 < tk.s_len; s0 += SITEMS * kJoinBlock) {
            if (s0) {
#pragma unroll
                for (int k = 0; k < SITEMS; k++) {
                    const uint32_t j = s0 + k * kJoinBlock + threadIdx.x;
                    if (j < tk.s_len) sv[k] = Sb[tk.s_off + j];
                }
            }
#pragma unroll
            for (int k0 = 0; k0 < SITEMS; k0 += G) {
                uint32_t e[G];
#pragma unroll
                for (int q = 0; q < G; q++) {
                    const uint32_t j = s0 + (k0 + q) * kJoinBlock + threadIdx.x;
                    e[q] = j < tk.s_len ? head[(sv[k0 + q].x >> hshift) & hmask] : 0;
                }
                for (;;) {
                    uint32_t rk[G], nx[G];
                    bool any = false;
#pragma unroll
                    for (int q = 0; q < G; q++) {
                        const uint32_t i = e[q] ? e[q] - 1 : 0;
                        rk[q] = rkey[i];
                        nx[q] = nxt[i];
                    }
#pragma unroll
                    for (int q = 0; q < G; q++) {
                        if (e[q]) {
                            c += rk[q] == sv[k0 + q].x;
                            e[q] = nx[q];
                            any |= e[q] != 0;
                        }
                    }
                    if (!any) break;
                }
            }
        }
        // the next task, decoded by wave 0 while the other waves finish probing
        TaskView nv{};
        if (w0) {
            uint32_t t = 0;
            if (threadIdx.x == 0) t = atomicAdd(task_ctr, 1u);
            t = __shfl(t, 0, TBJ_WAVE);
            nv = decode_task(t, ntasks, task_off, NP, P, bounds, s_chunk, join_ids);
        }
        const uint64_t wc = wave_reduce_u64(c);
        if (lane_id() == 0 && wc) atomicAdd(&tm, (unsigned long long)wc);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i <= hmask; i += kJoinBlock) head[i] = 0;
        if (threadIdx.x == 0) {
            if (tm) atomicAdd(&counts[tk.join], tm);
            tv = nv;
            tm = 0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// PHT joins of a wave.  Chunk c of a launch belongs to the last job whose
// first chunk is <= c (jobs without chunks share their first chunk with the
// next job, so the search never stops on one).

__device__ __forceinline__ uint32_t pht_job_of(const PhtJob* __restrict__ jobs, uint32_t njobs,
                                               uint32_t c) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].sc0 <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

constexpr int kPhtBlock = 256;
constexpr int kPhtProbeItems = 8;  // bucket loads in flight per thread

// Bulk build of the tables of a wave's PHT joins, one block of 2^hbits
// buckets per workgroup (pht_place_block).  Partitioned: a part of joins with
// the same pbits, whose R ranges were partitioned in one pass on the low pbits
// bits into src; block k is partition k & (2^pbits - 1) of the part's segment
// (join) k >> pbits, its bounds taken from the scanned group sums.  DIRECT
// (pbits = 0, tables of at most 2^kPhtPlaceHBits buckets): block k is all of
// join k's R, read from the arena.  jobs: the part's joins.
template <bool DIRECT>
__global__ __launch_bounds__(kPhtPlaceBlock) void k_bpht_place(
    const uint2* __restrict__ src, const PhtJob* __restrict__ jobs,
    const SegDesc* __restrict__ segs, const uint32_t* __restrict__ gsum, uint32_t pbits,
    PhtBucket* __restrict__ tab, OvNode* __restrict__ pool, uint32_t* __restrict__ pool_ctr) {
    const uint32_t k = blockIdx.x;
    const uint32_t s = k >> pbits, b = k & ((1u << pbits) - 1);
    const PhtJob jb = jobs[s];
    if (DIRECT) {
        pht_place_block(src + jb.r_off, 0, jb.nR, 0, jb.hbits, tab + jb.tab, pool, pool_ctr);
    } else {
        const SegDesc sd = segs[s];
        const uint32_t lo = gsum[sd.hist_base + (uint64_t)b * sd.ngroups];
        const uint32_t hi = gsum[sd.hist_base + (uint64_t)(b + 1) * sd.ngroups];
        pht_place_block(src, lo, hi, pbits, jb.hbits, tab + jb.tab + ((uint64_t)b << jb.hbits),
                        pool, pool_ctr);
    }
}

__global__ __launch_bounds__(kPhtBlock) void k_bpht_probe(
    const uint2* __restrict__ arena, const PhtJob* __restrict__ jobs, uint32_t njobs,
    uint32_t nchunks, const PhtBucket* __restrict__ tab, const OvNode* __restrict__ pool,
    unsigned long long* __restrict__ counts) {
    for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const PhtJob jb = jobs[pht_job_of(jobs, njobs, c)];
        const PhtBucket* t0 = tab + jb.tab;
        const uint32_t base = (c - jb.sc0) * kPhtProbeChunk + threadIdx.x;
        uint32_t m = 0;
        for (uint32_t r = 0; r < kPhtProbeChunk; r += kPhtBlock * kPhtProbeItems) {
            uint2 sv[kPhtProbeItems];
            PhtBucket bk[kPhtProbeItems];
#pragma unroll
            for (int k = 0; k < kPhtProbeItems; k++) {
                const uint32_t i = base + r + k * kPhtBlock;
                sv[k] = i < jb.nS ? arena[jb.s_off + i] : make_uint2(0, 0);
            }
            // every bucket load issued before any is used
#pragma unroll
            for (int k = 0; k < kPhtProbeItems; k++) {
                if (base + r + k * kPhtBlock < jb.nS) {
                    bk[k] = t0[pht_slot(sv[k].x & jb.mask, jb.pbits, jb.hbits)];
                } else {
                    bk[k].count = 0;
                    bk[k].ovf = 0;
                }
            }
#pragma unroll
            for (int k = 0; k < kPhtProbeItems; k++) m += pht_matches(bk[k], sv[k].x, pool);
        }
        const uint64_t wm = wave_reduce_u64(m);
        if (lane_id() == 0 && wm) atomicAdd(&counts[jb.id], (unsigned long long)wm);
    }
}

__global__ void k_fill_u64b(unsigned long long* __restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ __launch_bounds__(256) void k_sum_u64(const unsigned long long* __restrict__ a,
                                                 uint32_t n, unsigned long long* __restrict__ out) {
    __shared__ unsigned long long red[256 / TBJ_WAVE];
    unsigned long long s = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) s += a[i];
    s = wave_reduce_u64(s);
    if (lane_id() == 0) red[wave_id()] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 256 / TBJ_WAVE; w++) t += red[w];
        *out = t;
    }
}

// ---------------------------------------------------------------------------
// Batch generator

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// Bijection of [0, n): a 4-round Feistel network on 2h-bit values (2^2h >= n)
// with cycle walking back into [0, n).
__device__ __forceinline__ uint32_t feistel_perm(uint32_t x, uint32_t n, uint32_t h,
                                                 uint64_t seed) {
    const uint32_t mask = (1u << h) - 1;
    do {
        uint32_t L = x >> h, R = x & mask;
#pragma unroll
        for (uint32_t r = 0; r < 4; r++) {
            const uint32_t f = mix32(R ^ (uint32_t)(seed >> (16 * r)) ^ (r * 0x9E3779B9u)) & mask;
            const uint32_t t = L ^ f;
            L = R;
            R = t;
        }
        x = (L << h) | R;
    } while (x >= n);
    return x;
}

__device__ __forceinline__ double zipf_h_integral(double x, double s) {
    const double lx = log(x), t = (1.0 - s) * lx;
    const double h2 = fabs(t) > 1e-8 ? expm1(t) / t : 1.0 + t * 0.5 * (1.0 + t / 3.0 * (1.0 + 0.25 * t));
    return h2 * lx;
}

__device__ __forceinline__ double zipf_h_integral_inv(double x, double s) {
    double t = x * (1.0 - s);
    if (t < -1.0) t = -1.0;
    const double h1 = fabs(t) > 1e-8 ? log1p(t) / t : 1.0 - t * (0.5 - t * (1.0 / 3.0 - 0.25 * t));
    return exp(h1 * x);
}

// Rejection-inversion (Hormann and Derflinger 1996) for P(k) ~ k^-s, k in
// [1, n]; hx1 = H(1.5) - 1, hn = H(n + 1/2), sp = 2 - Hinv(H(2.5) - 2^-s).
__device__ __forceinline__ uint32_t zipf_sample(uint64_t& st, uint32_t n, double s, double hx1,
                                                double hn, double sp) {
    for (;;) {
        st += 0x9E3779B97F4A7C15ull;
        const double u01 = (double)(mix64(st) >> 11) * (1.0 / 9007199254740992.0);
        const double u = hn + u01 * (hx1 - hn);
        const double x = zipf_h_integral_inv(u, s);
        double kd = floor(x + 0.5);
        if (kd < 1.0) kd = 1.0;
        if (kd > (double)n) kd = (double)n;
        if (kd - x <= sp || u >= zipf_h_integral(kd + 0.5, s) - exp(-s * log(kd)))
            return (uint32_t)kd;
    }
}

__global__ __launch_bounds__(256) void k_gen_batch(uint2* __restrict__ arena,
                                                   const GenJob* __restrict__ jobs, uint32_t njobs,
                                                   uint64_t total) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = njobs;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].start <= g) lo = mid;
            else hi = mid;
        }
        const GenJob jb = jobs[lo];
        const uint64_t local = g - jb.start;
        if (local < jb.nR) {
            const uint32_t i = (uint32_t)local;
            arena[jb.r_off + i] = make_uint2(feistel_perm(i, (uint32_t)jb.nR, jb.fbits, jb.seed) + 1, i);
        } else {
            const uint32_t i = (uint32_t)(local - jb.nR);
            uint64_t st = jb.seed ^ ((uint64_t)(i + 1) * 0xD1B54A32D192ED03ull);
            uint32_t key;
            if (jb.z == 0.0) {
                st += 0x9E3779B97F4A7C15ull;
                key = (uint32_t)((mix64(st) >> 32) * jb.nR >> 32) + 1;
            } else {
                key = zipf_sample(st, (uint32_t)jb.nR, jb.z, jb.hx1, jb.hn, jb.sp);
            }
            arena[jb.s_off + i] = make_uint2(key, i);
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers

void batch_bounds(hipStream_t st, const uint32_t* scanned, const SegDesc* segs, uint32_t nj,
                  uint32_t F1, uint32_t F2, uint32_t passes, uint2* bounds) {
    const uint64_t n = 2ull * nj * F1 * F2;
    hipLaunchKernelGGL(k_bbounds, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, scanned,
                       segs, nj, F1, F2, passes, bounds);
}

void batch_join(hipStream_t st, uint32_t grid, const uint2* Rb, const uint2* Sb,
                const uint2* bounds, uint32_t NP, uint32_t P, const uint32_t* task_off,
                const uint32_t* ntasks_dev, uint32_t* task_ctr, uint32_t s_chunk, uint32_t hshift,
                const uint32_t* join_ids, unsigned long long* counts) {
    hipLaunchKernelGGL(k_bjoin, dim3(grid), dim3(kJoinBlock), 0, st, Rb, Sb, bounds, NP, P,
                       task_off, ntasks_dev, task_ctr, s_chunk, hshift, join_ids, counts);
}

void batch_pht_place(hipStream_t st, const uint2* src, const PhtJob* jobs, uint32_t njobs,
                     const SegDesc* segs, const uint32_t* gsum, uint32_t pbits, void* tab,
                     void* pool, uint32_t* pool_ctr) {
    if (!njobs) return;
    const uint32_t nblk = njobs << pbits;
    if (pbits == 0)
        hipLaunchKernelGGL(k_bpht_place<true>, dim3(nblk), dim3(kPhtPlaceBlock), 0, st, src, jobs,
                           segs, gsum, 0u, reinterpret_cast<PhtBucket*>(tab),
                           reinterpret_cast<OvNode*>(pool), pool_ctr);
    else
        hipLaunchKernelGGL(k_bpht_place<false>, dim3(nblk), dim3(kPhtPlaceBlock), 0, st, src,
                           jobs, segs, gsum, pbits, reinterpret_cast<PhtBucket*>(tab),
                           reinterpret_cast<OvNode*>(pool), pool_ctr);
}

void batch_pht_probe(hipStream_t st, uint32_t grid, const uint2* arena, const PhtJob* jobs,
                     uint32_t njobs, uint32_t nsc, const void* tab, const void* pool,
                     unsigned long long* counts) {
    if (njobs && nsc)
        hipLaunchKernelGGL(k_bpht_probe, dim3(min(grid, nsc)), dim3(kPhtBlock), 0, st, arena, jobs,
                           njobs, nsc, reinterpret_cast<const PhtBucket*>(tab),
                           reinterpret_cast<const OvNode*>(pool), counts);
}

void batch_zero_u64(hipStream_t st, unsigned long long* p, uint64_t n) {
    if (!n) return;
    uint64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_u64b, dim3((uint32_t)g), dim3(256), 0, st, p, n);
}

void batch_sum(hipStream_t st, const unsigned long long* a, uint32_t n, unsigned long long* out) {
    hipLaunchKernelGGL(k_sum_u64, dim3(1), dim3(256), 0, st, a, n, out);
}

void batch_generate(hipStream_t st, uint2* arena, const GenJob* jobs, uint32_t njobs,
                    uint64_t total) {
    if (!total) return;
    uint64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_gen_batch, dim3((uint32_t)g), dim3(256), 0, st, arena, jobs, njobs,
                       total);
}

}  // namespace tbj
