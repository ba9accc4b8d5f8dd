// Per-partition build/probe: partition bounds, join tasks and the
// LDS-resident hash-table join.
//
// One task pairs an R range of at most kJoinCap tuples (it fits the LDS hash
// table) with a chunk of S.  Large S partitions are split across tasks, each
// of which rebuilds the same R table (the reference's skew handling splits
// S the same way, radix_join.c:1085-1108).  Persistent workgroups pop tasks
// from an atomic counter, as the reference's threads pop join tasks
// (radix_join.c:1184-1191, task_queue.h:123-137).
//
// kModeChain restates bucket_chaining_join (radix_join.c:186-244): heads of
// nextpow2(|R_p|) buckets on key bits above the partition bits, chains
// through next[].  kModeHist restates histogram_join (radix_join.c:373-440):
// R_p reordered by bucket, S probes a contiguous bucket range.
#include "tbj_device.h"
#include "tbj_kernels.h"

namespace tbj {

// ---------------------------------------------------------------------------
// Final partition bounds.  passes == 2: segment rel * outer + p1, bin p2;
// passes == 1: segment rel, bin p; passes == 0: the whole relation.
__global__ void k_part_bounds(const uint32_t* __restrict__ scanned,
                              const SegDesc* __restrict__ segs, uint32_t outer, uint32_t inner,
                              uint32_t passes, uint32_t nR, uint32_t nS,
                              uint2* __restrict__ bounds) {
    const uint32_t nparts = outer * inner;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * nparts) return;
    const uint32_t rel = i / nparts, p = i % nparts;
    if (passes == 0) {
        bounds[i] = make_uint2(0, rel ? nS : nR);
        return;
    }
    uint32_t seg, bin;
    if (passes == 1) {
        seg = rel;
        bin = p;
    } else {
        seg = rel * outer + p / inner;
        bin = p % inner;
    }
    const SegDesc sd = segs[seg];
    const uint32_t st = scanned[sd.hist_base + (uint64_t)bin * sd.ntiles];
    const uint32_t en = scanned[sd.hist_base + (uint64_t)(bin + 1) * sd.ntiles];
    bounds[i] = make_uint2(st, en);
}

__device__ __forceinline__ uint32_t tasks_of(uint2 r, uint2 s, uint32_t cap, uint32_t chunk) {
    const uint32_t nr = r.y - r.x, ns = s.y - s.x;
    if (!nr || !ns) return 0;
    return ((nr + cap - 1) / cap) * ((ns + chunk - 1) / chunk);
}

// Per-partition task counts -> exclusive offsets (one workgroup).
// stats[0] = largest R partition, stats[1] = largest S partition.
__global__ __launch_bounds__(kScanBlock) void k_count_tasks(const uint2* __restrict__ bounds,
                                                            uint32_t nparts, uint32_t cap,
                                                            uint32_t chunk,
                                                            uint32_t* __restrict__ task_off,
                                                            uint32_t* __restrict__ ntasks_dev,
                                                            uint32_t* __restrict__ stats) {
    __shared__ uint32_t scratch[kScanBlock / TBJ_WAVE + 1];
    __shared__ uint32_t carry_s, rmax_s, smax_s;
    if (threadIdx.x == 0) carry_s = rmax_s = smax_s = 0;
    __syncthreads();
    uint32_t rmax = 0, smax = 0;
    for (uint32_t base = 0; base < nparts; base += kScanBlock) {
        const uint32_t p = base + threadIdx.x;
        uint32_t c = 0;
        if (p < nparts) {
            const uint2 r = bounds[p], s = bounds[nparts + p];
            c = tasks_of(r, s, cap, chunk);
            rmax = max(rmax, r.y - r.x);
            smax = max(smax, s.y - s.x);
        }
        uint32_t total;
        const uint32_t ex = block_exclusive_scan<kScanBlock>(c, scratch, &total);
        if (p < nparts) task_off[p] = carry_s + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry_s += total;
        __syncthreads();
    }
    atomicMax(&rmax_s, rmax);
    atomicMax(&smax_s, smax);
    __syncthreads();
    if (threadIdx.x == 0) {
        *ntasks_dev = carry_s;
        stats[0] = rmax_s;
        stats[1] = smax_s;
    }
}

// Larger partitions first (a partition's tasks are contiguous; the task
// counter hands them out in order).
__global__ void k_fill_tasks(const uint2* __restrict__ bounds, uint32_t nparts, uint32_t cap,
                             uint32_t chunk, const uint32_t* __restrict__ task_off,
                             JoinTask* __restrict__ tasks) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nparts) return;
    const uint2 r = bounds[p], s = bounds[nparts + p];
    const uint32_t n = tasks_of(r, s, cap, chunk);
    uint32_t o = task_off[p];
    for (uint32_t ro = r.x; ro < r.y && n; ro += cap) {
        for (uint32_t so = s.x; so < s.y; so += chunk) {
            JoinTask t;
            t.r_off = ro;
            t.r_len = min(cap, r.y - ro);
            t.s_off = so;
            t.s_len = min(chunk, s.y - so);
            tasks[o++] = t;
        }
    }
}

// ---------------------------------------------------------------------------
// LDS build/probe join.  counters[0] = matches, [1] = checksum of
// (R.payload + S.payload), [2] = output cursor.

template <int MODE, bool MAT>
__global__ __launch_bounds__(kJoinBlock) void k_join(const uint2* __restrict__ Rb,
                                                     const uint2* __restrict__ Sb,
                                                     const JoinTask* __restrict__ tasks,
                                                     const uint32_t* __restrict__ ntasks_dev,
                                                     uint32_t* __restrict__ task_ctr,
                                                     uint32_t hshift, uint2* __restrict__ pairs,
                                                     uint64_t cap,
                                                     unsigned long long* __restrict__ counters) {
    constexpr int RITEMS = kJoinCap / kJoinBlock;
    constexpr int NHEAD = MODE == kModeChain ? kJoinCap : kJoinCap / 2 + 1;
    __shared__ uint32_t rkey[kJoinCap];
    __shared__ uint32_t rpay[kJoinCap];
    __shared__ uint32_t head[NHEAD];
    __shared__ uint16_t nxt[MODE == kModeChain ? kJoinCap : 1];
    __shared__ uint32_t scratch[kJoinBlock / TBJ_WAVE + 1];
    __shared__ uint32_t task_s;
    __shared__ unsigned long long obase_s;
    __shared__ unsigned long long red_m[kJoinBlock / TBJ_WAVE], red_c[kJoinBlock / TBJ_WAVE];

    const uint32_t ntasks = *ntasks_dev;
    uint64_t matches = 0, cks = 0;
    for (;;) {
        if (threadIdx.x == 0) task_s = atomicAdd(task_ctr, 1u);
        __syncthreads();
        const uint32_t t = task_s;
        if (t >= ntasks) break;
        const JoinTask tk = tasks[t];
        const uint32_t n = tk.r_len;
        uint32_t nb;
        if (MODE == kModeChain) {
            nb = next_pow2_u32(n);
        } else {
            nb = next_pow2_u32(n) >> 1;
            if (nb == 0) nb = 1;
        }
        const uint32_t hmask = nb - 1;

        // ---- build
        if (MODE == kModeChain) {
            for (uint32_t i = threadIdx.x; i < nb; i += kJoinBlock) head[i] = 0;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RITEMS; k++) {
                const uint32_t j = k * kJoinBlock + threadIdx.x;
                if (j < n) {
                    const uint2 v = Rb[tk.r_off + j];
                    rkey[j] = v.x;
                    rpay[j] = v.y;
                    const uint32_t h = (v.x >> hshift) & hmask;
                    nxt[j] = (uint16_t)atomicExch(&head[h], j + 1);
                }
            }
            __syncthreads();
        } else {
            for (uint32_t i = threadIdx.x; i <= nb; i += kJoinBlock) head[i] = 0;
            __syncthreads();
            uint2 v[RITEMS];
            uint32_t rk[RITEMS];
#pragma unroll
            for (int k = 0; k < RITEMS; k++) {
                const uint32_t j = k * kJoinBlock + threadIdx.x;
                if (j < n) {
                    v[k] = Rb[tk.r_off + j];
                    rk[k] = atomicAdd(&head[(v[k].x >> hshift) & hmask], 1u);
                }
            }
            __syncthreads();
            block_scan_lds<kJoinBlock>(head, nb + 1, scratch);
#pragma unroll
            for (int k = 0; k < RITEMS; k++) {
                const uint32_t j = k * kJoinBlock + threadIdx.x;
                if (j < n) {
                    const uint32_t pos = head[(v[k].x >> hshift) & hmask] + rk[k];
                    rkey[pos] = v[k].x;
                    rpay[pos] = v[k].y;
                }
            }
            __syncthreads();
        }

        // ---- probe
        for (uint32_t base = 0; base < tk.s_len; base += kJoinBlock * kJoinProbeItems) {
            uint2 sv[kJoinProbeItems];
            uint32_t c = 0;
#pragma unroll
            for (int k = 0; k < kJoinProbeItems; k++) {
                const uint32_t j = base + k * kJoinBlock + threadIdx.x;
                sv[k] = j < tk.s_len ? Sb[tk.s_off + j] : make_uint2(0, 0);
            }
#pragma unroll
            for (int k = 0; k < kJoinProbeItems; k++) {
                const uint32_t j = base + k * kJoinBlock + threadIdx.x;
                if (j >= tk.s_len) continue;
                const uint32_t key = sv[k].x;
                const uint32_t h = (key >> hshift) & hmask;
                if (MODE == kModeChain) {
                    for (uint32_t e = head[h]; e; e = nxt[e - 1]) {
                        if (rkey[e - 1] == key) {
                            c++;
                            cks += (uint64_t)rpay[e - 1] + sv[k].y;
                        }
                    }
                } else {
                    const uint32_t en = head[h + 1];
                    for (uint32_t e = head[h]; e < en; e++) {
                        if (rkey[e] == key) {
                            c++;
                            cks += (uint64_t)rpay[e] + sv[k].y;
                        }
                    }
                }
            }
            matches += c;
            if (MAT) {
                uint32_t total;
                const uint32_t ex = block_exclusive_scan<kJoinBlock>(c, scratch, &total);
                if (threadIdx.x == 0) obase_s = total ? atomicAdd(&counters[2], (unsigned long long)total) : 0;
                __syncthreads();
                uint64_t o = obase_s + ex;
                if (c) {
#pragma unroll
                    for (int k = 0; k < kJoinProbeItems; k++) {
                        const uint32_t j = base + k * kJoinBlock + threadIdx.x;
                        if (j >= tk.s_len) continue;
                        const uint32_t key = sv[k].x;
                        const uint32_t h = (key >> hshift) & hmask;
                        if (MODE == kModeChain) {
                            for (uint32_t e = head[h]; e; e = nxt[e - 1]) {
                                if (rkey[e - 1] == key) {
                                    if (o < cap) pairs[o] = make_uint2(rpay[e - 1], sv[k].y);
                                    o++;
                                }
                            }
                        } else {
                            const uint32_t en = head[h + 1];
                            for (uint32_t e = head[h]; e < en; e++) {
                                if (rkey[e] == key) {
                                    if (o < cap) pairs[o] = make_uint2(rpay[e], sv[k].y);
                                    o++;
                                }
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    // block reduction of the counters
    matches = wave_reduce_u64(matches);
    cks = wave_reduce_u64(cks);
    if (lane_id() == 0) {
        red_m[wave_id()] = matches;
        red_c[wave_id()] = cks;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = 0, c = 0;
        for (int w = 0; w < kJoinBlock / TBJ_WAVE; w++) {
            m += red_m[w];
            c += red_c[w];
        }
        if (m) atomicAdd(&counters[0], m);
        if (c) atomicAdd(&counters[1], c);
    }
}

// ---------------------------------------------------------------------------
// Host launchers

void launch_part_bounds(hipStream_t st, const uint32_t* scanned, const SegDesc* segs,
                        uint32_t outer, uint32_t inner, uint32_t passes, uint32_t nR,
                        uint32_t nS, uint2* bounds) {
    const uint32_t n = 2 * outer * inner;
    hipLaunchKernelGGL(k_part_bounds, dim3((n + 255) / 256), dim3(256), 0, st, scanned, segs,
                       outer, inner, passes, nR, nS, bounds);
}

void launch_count_tasks(hipStream_t st, const uint2* bounds, uint32_t nparts, uint32_t cap,
                        uint32_t s_chunk, uint32_t* task_off, uint32_t* partial,
                        uint32_t* ntasks_dev, uint32_t* stats) {
    (void)partial;
    hipLaunchKernelGGL(k_count_tasks, dim3(1), dim3(kScanBlock), 0, st, bounds, nparts, cap,
                       s_chunk, task_off, ntasks_dev, stats);
}

void launch_fill_tasks(hipStream_t st, const uint2* bounds, uint32_t nparts, uint32_t cap,
                       uint32_t s_chunk, const uint32_t* task_off, JoinTask* tasks) {
    hipLaunchKernelGGL(k_fill_tasks, dim3((nparts + 255) / 256), dim3(256), 0, st, bounds, nparts,
                       cap, s_chunk, task_off, tasks);
}

template <int MODE, bool MAT>
static void launch_join_t(hipStream_t st, uint32_t grid, const uint2* Rb, const uint2* Sb,
                          const JoinTask* tasks, const uint32_t* ntasks_dev, uint32_t* task_ctr,
                          uint32_t hshift, uint2* pairs, uint64_t cap,
                          unsigned long long* counters) {
    hipLaunchKernelGGL((k_join<MODE, MAT>), dim3(grid), dim3(kJoinBlock), 0, st, Rb, Sb, tasks,
                       ntasks_dev, task_ctr, hshift, pairs, cap, counters);
}

void launch_join(hipStream_t st, int mode, bool materialize, uint32_t grid, const uint2* Rb,
                 const uint2* Sb, const JoinTask* tasks, const uint32_t* ntasks_dev,
                 uint32_t* task_ctr, uint32_t hshift, uint2* pairs, uint64_t cap,
                 unsigned long long* counters) {
    if (mode == kModeChain) {
        if (materialize)
            launch_join_t<kModeChain, true>(st, grid, Rb, Sb, tasks, ntasks_dev, task_ctr, hshift,
                                            pairs, cap, counters);
        else
            launch_join_t<kModeChain, false>(st, grid, Rb, Sb, tasks, ntasks_dev, task_ctr, hshift,
                                             pairs, cap, counters);
    } else {
        if (materialize)
            launch_join_t<kModeHist, true>(st, grid, Rb, Sb, tasks, ntasks_dev, task_ctr, hshift,
                                           pairs, cap, counters);
        else
            launch_join_t<kModeHist, false>(st, grid, Rb, Sb, tasks, ntasks_dev, task_ctr, hshift,
                                            pairs, cap, counters);
    }
}

int join_occupancy(int mode, bool materialize) {
    int blocks = 0;
    hipError_t e;
    if (mode == kModeChain)
        e = materialize ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                              &blocks, k_join<kModeChain, true>, kJoinBlock, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                              &blocks, k_join<kModeChain, false>, kJoinBlock, 0);
    else
        e = materialize ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                              &blocks, k_join<kModeHist, true>, kJoinBlock, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                              &blocks, k_join<kModeHist, false>, kJoinBlock, 0);
    if (e != hipSuccess || blocks < 1) blocks = 1;
    return blocks;
}

}  // namespace tbj
