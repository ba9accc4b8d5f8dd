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
#include <algorithm>

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
    const uint32_t st = scanned[sd.hist_base + (uint64_t)bin * sd.ngroups];
    const uint32_t en = scanned[sd.hist_base + (uint64_t)(bin + 1) * sd.ngroups];
    bounds[i] = make_uint2(st, en);
}

// Builds (or chunks) of one partition side: its npieces ranges from `b`,
// each cut into pieces of at most `cap` tuples; *len = its tuple count.
__device__ __forceinline__ uint32_t side_chunks(const uint2* __restrict__ b, uint32_t npieces,
                                                uint32_t cap, uint32_t* len) {
    uint32_t c = 0, n = 0;
    for (uint32_t x = 0; x < npieces; x++) {
        const uint2 r = b[x];
        const uint32_t l = r.y - r.x;
        c += (l + cap - 1) / cap;
        n += l;
    }
    *len = n;
    return c;
}

// Range of chunk k of a side (k < its chunk count).
__device__ __forceinline__ void side_chunk(const uint2* __restrict__ b, uint32_t npieces,
                                           uint32_t cap, uint32_t k, uint32_t* off,
                                           uint32_t* len) {
    for (uint32_t x = 0; x < npieces; x++) {
        const uint2 r = b[x];
        const uint32_t c = (r.y - r.x + cap - 1) / cap;
        if (k < c) {
            *off = r.x + k * cap;
            *len = min(cap, r.y - *off);
            return;
        }
        k -= c;
    }
    *off = 0;
    *len = 0;
}

// Per-partition task counts (task_cnt[nparts] is a zero sentinel so that the
// exclusive scan leaves the total there): builds times chunks, zero when a
// side is empty.  stats[0] / stats[1] = largest R / S partition.
__global__ __launch_bounds__(256) void k_task_counts(const uint2* __restrict__ bounds,
                                                     uint32_t nparts, uint32_t npieces,
                                                     uint32_t cap, uint32_t chunk,
                                                     uint32_t* __restrict__ task_cnt,
                                                     uint32_t* __restrict__ stats) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t rn = 0, sn = 0;
    if (p < nparts) {
        const uint32_t rc = side_chunks(bounds + (uint64_t)p * npieces, npieces, cap, &rn);
        const uint32_t sc =
            side_chunks(bounds + ((uint64_t)nparts + p) * npieces, npieces, chunk, &sn);
        task_cnt[p] = rc * sc;
    } else if (p == nparts) {
        task_cnt[p] = 0;
    }
    // one atomic pair per block: per-wave atomics on the two words
    // serialise (tens of microseconds at 2^16 partitions)
    __shared__ uint32_t red[2][256 / TBJ_WAVE];
    for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) {
        rn = max(rn, (uint32_t)__shfl_xor(rn, d, TBJ_WAVE));
        sn = max(sn, (uint32_t)__shfl_xor(sn, d, TBJ_WAVE));
    }
    if (lane_id() == 0) {
        red[0][wave_id()] = rn;
        red[1][wave_id()] = sn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / TBJ_WAVE; w++) {
            rn = max(rn, red[0][w]);
            sn = max(sn, red[1][w]);
        }
        if (rn | sn) {
            atomicMax(&stats[0], rn);
            atomicMax(&stats[1], sn);
        }
    }
}

// Task t of partition p (task_off[p] <= t < task_off[p + 1]): build li / nsc
// and chunk li % nsc of it (li = t - task_off[p], nsc = S chunks of p).  A
// partition's tasks are contiguous in partition order.
__device__ __forceinline__ JoinTask task_in_part(const TaskSrc& ts, uint32_t p, uint32_t t) {
    const uint2* rb = ts.bounds + (uint64_t)p * ts.npieces;
    const uint2* sb = ts.bounds + ((uint64_t)ts.nparts + p) * ts.npieces;
    uint32_t sn;
    const uint32_t nsc = side_chunks(sb, ts.npieces, ts.chunk, &sn);
    const uint32_t li = t - ts.task_off[p];
    JoinTask tk;
    side_chunk(rb, ts.npieces, ts.cap, li / nsc, &tk.r_off, &tk.r_len);
    side_chunk(sb, ts.npieces, ts.chunk, li % nsc, &tk.s_off, &tk.s_len);
    return tk;
}

// Partition of task t: binary search over task_off.
__device__ __forceinline__ uint32_t task_part(const TaskSrc& ts, uint32_t t) {
    uint32_t lo = 0, hi = ts.nparts;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ts.task_off[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The descriptor of task t < ntasks: from the task list below task_cap, else
// decoded from the per-partition task offsets (only partitions whose R side
// spans several builds and whose S side several chunks at once can push the
// count past the list, which the host sizes without knowing the count).
__device__ __forceinline__ JoinTask get_task(const TaskSrc& ts, uint32_t t) {
    if (t < ts.task_cap) return ts.tasks[t];
    return task_in_part(ts, task_part(ts, t), t);
}

// One thread per list entry t < min(ntasks, task_cap).
__global__ void k_fill_tasks(TaskSrc ts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= min(*ts.ntasks_dev, ts.task_cap)) return;
    ts.tasks[t] = task_in_part(ts, task_part(ts, t), t);
}

// ---------------------------------------------------------------------------
// Output claims of the sort-merge join (the hash joins claim directly from
// counters[2]).  Returning atomics on one address retire at ~83 M/s on this
// chip, so a join claiming once per group of 2048 S tuples spends ~3 ms of a
// 2^29-pair output waiting on the counter.  In chunk mode (OutBuf::chunk > 0) a workgroup claims whole
// chunks, one atomic each, and hands out slots from its current chunk; the
// unused tail of each workgroup's last chunk is a gap that launch_out_fix
// fills afterwards from the pairs above the total.

// Slot of the off-th pair of a group: o1 + off below split, else o2 + off -
// split (a group that overflows its workgroup's chunk continues in new ones).
struct OutMap {
    unsigned long long o1, o2;
    uint32_t split, pad;
};

__device__ __forceinline__ unsigned long long out_slot(const OutMap& m, uint64_t off) {
    return off < m.split ? m.o1 + off : m.o2 + (off - m.split);
}

// Slots below cap are the caller's; chunk mode keeps side_cap more in the
// side buffer; beyond that pairs are dropped (pairs_written reports what the
// caller's buffer holds).
__device__ __forceinline__ void out_put(const OutBuf& b, unsigned long long p, uint2 v) {
    if (p < b.cap) b.pairs[p] = v;
    else if (p - b.cap < b.side_cap) b.side[p - b.cap] = v;
}

// One thread of the workgroup: tot slots for its current group.  cur is the
// workgroup's current chunk (LDS; used == chunk before its first claim).
__device__ __forceinline__ OutMap out_claim(const OutBuf& b, OutTail& cur, uint32_t tot,
                                            unsigned long long* ctr) {
    OutMap m;
    m.o2 = 0;
    m.split = 0xFFFFFFFFu;
    m.pad = 0;
    if (!b.chunk) {
        m.o1 = tot ? atomicAdd(ctr, (unsigned long long)tot) : 0;
        return m;
    }
    const uint32_t rem = b.chunk - cur.used;
    m.o1 = cur.base + cur.used;
    if (tot <= rem) {
        cur.used += tot;
        return m;
    }
    const uint32_t need = tot - rem, k = (need + b.chunk - 1) / b.chunk;
    m.o2 = atomicAdd(ctr, (unsigned long long)k * b.chunk);
    m.split = rem;
    cur.base = m.o2 + (unsigned long long)(k - 1) * b.chunk;
    cur.used = need - (k - 1) * b.chunk;
    return m;
}

// ---------------------------------------------------------------------------
// LDS build/probe join.  counters[0] = matches, [1] = checksum of
// (R.payload + S.payload), [2] = output cursor.

// Looks one S tuple up in the LDS table; EMIT=false counts matches and sums
// the checksum, EMIT=true writes them from output slot o on: (R.payload,
// S.payload) pairs, or with rel the (key, R.payload) tuples of the
// to_relation format.
template <int MODE, bool EMIT>
__device__ __forceinline__ void lds_probe(uint2 sv, uint32_t hshift, uint32_t hmask,
                                          const uint32_t* rkey, const uint32_t* rpay,
                                          const uint32_t* head, const uint16_t* nxt,
                                          uint32_t& c, uint64_t& cks, uint32_t& first,
                                          uint2* __restrict__ pairs, uint64_t cap, uint64_t& o,
                                          bool rel = false) {
    const uint32_t key = sv.x;
    const uint32_t h = (key >> hshift) & hmask;
    if (MODE == kModeChain) {
        for (uint32_t e = head[h]; e; e = nxt[e - 1]) {
            if (rkey[e - 1] == key) {
                if (EMIT) {
                    if (o < cap) pairs[o] = rel ? make_uint2(key, rpay[e - 1]) : make_uint2(rpay[e - 1], sv.y);
                    o++;
                } else {
                    first = rpay[e - 1];
                    c++;
                    cks += (uint64_t)first + sv.y;
                }
            }
        }
    } else {
        const uint32_t en = head[h + 1];
        for (uint32_t e = head[h]; e < en; e++) {
            if (rkey[e] == key) {
                if (EMIT) {
                    if (o < cap) pairs[o] = rel ? make_uint2(key, rpay[e]) : make_uint2(rpay[e], sv.y);
                    o++;
                } else {
                    first = rpay[e];
                    c++;
                    cks += (uint64_t)first + sv.y;
                }
            }
        }
    }
}

// Counts the matches of N S tuples at once (sv[k0 .. k0 + N), item k at S
// position s0 + k * kJoinBlock + threadIdx.x, valid below s_len): the N
// chains (or ranges) are walked in lockstep, so each step issues N
// independent LDS reads instead of one dependent read per step.  ck[q] =
// matches of item q, fp[q] = R.payload of its (last) match.
template <int MODE, int N, int SI>
__device__ __forceinline__ void lds_count_batch(const uint2 (&sv)[SI], int k0, uint32_t s0,
                                                uint32_t s_len, uint32_t hshift, uint32_t hmask,
                                                const uint32_t* rkey, const uint32_t* rpay,
                                                const uint32_t* head, const uint16_t* nxt,
                                                uint32_t (&ck)[N], uint32_t (&fp)[N],
                                                uint64_t& cks) {
    uint32_t e[N], en[N];
#pragma unroll
    for (int q = 0; q < N; q++) {
        const bool ok = s0 + (k0 + q) * kJoinBlock + threadIdx.x < s_len;
        const uint32_t h = (sv[k0 + q].x >> hshift) & hmask;
        ck[q] = 0;
        fp[q] = 0;
        if (MODE == kModeChain) {
            e[q] = ok ? head[h] : 0;  // 1-based, 0 = end of chain
            en[q] = 0;
        } else {
            e[q] = ok ? head[h] : 0;
            en[q] = ok ? head[h + 1] : 0;
        }
    }
    for (;;) {
        bool any = false;
        uint32_t rk[N], nx[N];
#pragma unroll
        for (int q = 0; q < N; q++) {
            const bool live = MODE == kModeChain ? e[q] != 0 : e[q] < en[q];
            const uint32_t i = MODE == kModeChain ? (live ? e[q] - 1 : 0) : (live ? e[q] : 0);
            rk[q] = rkey[i];
            nx[q] = MODE == kModeChain ? nxt[i] : 0;
        }
#pragma unroll
        for (int q = 0; q < N; q++) {
            const bool live = MODE == kModeChain ? e[q] != 0 : e[q] < en[q];
            if (live) {
                const uint32_t i = MODE == kModeChain ? e[q] - 1 : e[q];
                if (rk[q] == sv[k0 + q].x) {
                    fp[q] = rpay[i];
                    ck[q]++;
                    cks += (uint64_t)fp[q] + sv[k0 + q].y;
                }
                e[q] = MODE == kModeChain ? nx[q] : e[q] + 1;
                any |= MODE == kModeChain ? e[q] != 0 : e[q] < en[q];
            }
        }
        if (!any) break;
    }
}

// Loads tuples first + k * kJoinBlock + threadIdx.x (k < N) of a range of
// len >= 1 tuples; positions past the end re-read the last tuple, so every
// lane issues the same loads and no load waits on a branch.
template <int N>
__device__ __forceinline__ void load_range(uint2 (&v)[N], const uint2* __restrict__ base,
                                           uint32_t off, uint32_t len, uint32_t first) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        const uint32_t j = first + k * kJoinBlock + threadIdx.x;
        v[k] = base[off + min(j, len - 1)];
    }
}

// Builds the LDS table of an R range of n tuples (rv[k] = tuple k *
// kJoinBlock + threadIdx.x) into zeroed heads.  kModeChain: chains through
// nxt (1-based, 0 ends a chain); kModeHist: R reordered by bucket, head[b]
// = first tuple of bucket b.  The caller's barrier completes the table.
template <int MODE, int RITEMS>
__device__ __forceinline__ void lds_build(const uint2 (&rv)[RITEMS], uint32_t n, uint32_t hshift,
                                          uint32_t hmask, uint32_t nb, uint32_t* rkey,
                                          uint32_t* rpay, uint32_t* head, uint16_t* nxt,
                                          uint32_t* scratch) {
    if (MODE == kModeChain) {
#pragma unroll
        for (int k = 0; k < RITEMS; k++) {
            const uint32_t j = k * kJoinBlock + threadIdx.x;
            if (j < n) {
                rkey[j] = rv[k].x;
                rpay[j] = rv[k].y;
                nxt[j] = (uint16_t)atomicExch(&head[(rv[k].x >> hshift) & hmask], j + 1);
            }
        }
    } else {
        uint32_t rk[RITEMS];
#pragma unroll
        for (int k = 0; k < RITEMS; k++) {
            const uint32_t j = k * kJoinBlock + threadIdx.x;
            if (j < n) rk[k] = atomicAdd(&head[(rv[k].x >> hshift) & hmask], 1u);
        }
        __syncthreads();
        block_scan_lds<kJoinBlock>(head, nb + 1, scratch);
#pragma unroll
        for (int k = 0; k < RITEMS; k++) {
            const uint32_t j = k * kJoinBlock + threadIdx.x;
            if (j < n) {
                const uint32_t pos = head[(rv[k].x >> hshift) & hmask] + rk[k];
                rkey[pos] = rv[k].x;
                rpay[pos] = rv[k].y;
            }
        }
    }
}

__device__ __forceinline__ uint32_t table_buckets(int mode, uint32_t n) {
    if (mode == kModeChain) return next_pow2_u32(n);
    const uint32_t nb = next_pow2_u32(n) >> 1;
    return nb ? nb : 1;
}

// Block reduction of (matches, checksum) into counters[0..1].
__device__ __forceinline__ void flush_counters(uint64_t matches, uint64_t cks,
                                               unsigned long long* red_m,
                                               unsigned long long* red_c,
                                               unsigned long long* __restrict__ counters) {
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

// Counting join, one task per loop trip: the R range and the first S
// sub-chunk are loaded together (one memory phase), the R range is built into
// LDS once, and the task's S range is probed in sub-chunks of SITEMS *
// kJoinBlock tuples.
template <int MODE, int SITEMS = kJoinSItems>
__global__ __launch_bounds__(kJoinBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_join(
    const uint2* __restrict__ Rb, const uint2* __restrict__ Sb, TaskSrc ts,
    uint32_t* __restrict__ task_ctr, uint32_t hshift, unsigned long long* __restrict__ counters) {
    constexpr int RITEMS = kJoinCap / kJoinBlock;
    constexpr int NHEAD = MODE == kModeChain ? kJoinCap : kJoinCap / 2 + 1;
    __shared__ uint32_t rkey[kJoinCap];
    __shared__ uint32_t rpay[kJoinCap];
    __shared__ uint32_t head[NHEAD];
    __shared__ uint16_t nxt[MODE == kModeChain ? kJoinCap : 1];
    __shared__ uint32_t scratch[kJoinBlock / TBJ_WAVE + 1];
    __shared__ uint32_t task_s;
    __shared__ unsigned long long red_m[kJoinBlock / TBJ_WAVE], red_c[kJoinBlock / TBJ_WAVE];
    __shared__ JoinTask tk_s;

    const uint32_t ntasks = *ts.ntasks_dev;
    uint64_t matches = 0, cks = 0;
    if (threadIdx.x == 0) {
        task_s = atomicAdd(task_ctr, 1u);
        if (task_s < ntasks) tk_s = get_task(ts, task_s);
    }
    for (uint32_t i = threadIdx.x; i < NHEAD; i += kJoinBlock) head[i] = 0;
    __syncthreads();
    for (;;) {
        const uint32_t t = task_s;
        if (t >= ntasks) break;
        const JoinTask tk = tk_s;
        const uint32_t n = tk.r_len;
        // the next task's index is claimed now and its descriptor read after
        // the build, so neither round trip sits between two tasks
        uint32_t next_t = 0;
        JoinTask nd{};
        if (threadIdx.x == 0) next_t = atomicAdd(task_ctr, 1u);
        uint2 rv[RITEMS], sv[SITEMS];
        load_range(rv, Rb, tk.r_off, n, 0);
        load_range(sv, Sb, tk.s_off, tk.s_len, 0);
        const uint32_t nb = table_buckets(MODE, n);
        const uint32_t hmask = nb - 1;
        lds_build<MODE>(rv, n, hshift, hmask, nb, rkey, rpay, head, nxt, scratch);
        __syncthreads();

        if (threadIdx.x == 0 && next_t < ntasks) nd = get_task(ts, next_t);
        for (uint32_t s0 = 0; s0 < tk.s_len; s0 += SITEMS * kJoinBlock) {
            if (s0) load_range(sv, Sb, tk.s_off, tk.s_len, s0);
            constexpr int G = MODE == kModeChain ? 8 : 4;
#pragma unroll
            for (int k0 = 0; k0 < SITEMS; k0 += G) {
                uint32_t ck[G], fp[G];
                lds_count_batch<MODE, G>(sv, k0, s0, tk.s_len, hshift, hmask, rkey, rpay, head,
                                         nxt, ck, fp, cks);
#pragma unroll
                for (int q = 0; q < G; q++) matches += ck[q];
            }
        }
        __syncthreads();
        // clear the heads for the next task and fetch its index
        for (uint32_t i = threadIdx.x; i <= nb && i < (uint32_t)NHEAD; i += kJoinBlock) head[i] = 0;
        if (threadIdx.x == 0) {
            task_s = next_t;
            tk_s = nd;
        }
        __syncthreads();
    }
    flush_counters(matches, cks, red_m, red_c, counters);
}

// Materialising join, software-pipelined: each S sub-chunk (SI * kJoinBlock
// tuples) is counted, its output range claimed with one atomic, and its
// pairs written; the loads of the next sub-chunk -- or, after a task's last
// sub-chunk, of the next task's R range and first sub-chunk -- are issued
// before the claim is awaited, so they travel while the claim returns and the
// pairs are written.  Output slots go in item-major order (all threads' item
// 0, then item 1, ...) so that a wave's stores of one item are contiguous.
// rel: (key, R.payload) tuples instead of (R.payload, S.payload) pairs.
template <int MODE>
__global__ __launch_bounds__(kJoinBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_join_mat(
    const uint2* __restrict__ Rb, const uint2* __restrict__ Sb, TaskSrc ts,
    uint32_t* __restrict__ task_ctr, uint32_t hshift, uint2* __restrict__ pairs, uint64_t cap,
    bool rel, unsigned long long* __restrict__ counters) {
    constexpr int SI = kJoinSItems / 2;
    constexpr int SUB = SI * kJoinBlock;
    constexpr int RITEMS = kJoinCap / kJoinBlock;
    constexpr int NHEAD = MODE == kModeChain ? kJoinCap : kJoinCap / 2 + 1;
    constexpr int NW = kJoinBlock / TBJ_WAVE;
    __shared__ uint32_t rkey[kJoinCap];
    __shared__ uint32_t rpay[kJoinCap];
    __shared__ uint32_t head[NHEAD];
    __shared__ uint16_t nxt[MODE == kModeChain ? kJoinCap : 1];
    __shared__ uint32_t scratch[NW + 1];
    // per-wave match totals of each item, and each wave's output offset for
    // each item; double-buffered by sub-chunk parity (a wave may count the
    // next sub-chunk while others still write this one)
    __shared__ uint32_t wtot[2][NW][SI];
    __shared__ uint32_t woff[2][NW][SI];
    __shared__ uint32_t first_s, next_s;
    __shared__ JoinTask tk_s;
    __shared__ unsigned long long obase_s;
    __shared__ unsigned long long red_m[NW], red_c[NW];

    const uint32_t ntasks = *ts.ntasks_dev;
    const uint32_t lane = lane_id(), w = wave_id();
    uint64_t matches = 0, cks = 0;
    if (threadIdx.x == 0) {
        const uint32_t t = atomicAdd(task_ctr, 1u);
        first_s = t;
        if (t < ntasks) tk_s = get_task(ts, t);
    }
    for (uint32_t i = threadIdx.x; i < NHEAD; i += kJoinBlock) head[i] = 0;
    __syncthreads();
    bool live = first_s < ntasks;
    JoinTask tk = tk_s;
    uint2 rv[RITEMS], sv[SI];
    if (live) {
        load_range(rv, Rb, tk.r_off, tk.r_len, 0);
        load_range(sv, Sb, tk.s_off, tk.s_len, 0);
    }
    uint32_t ph = 0;
    while (live) {
        const uint32_t n = tk.r_len;
        const uint32_t nb = table_buckets(MODE, n);
        const uint32_t hmask = nb - 1;
        // the next task is popped during the build; every thread then reads
        // its descriptor (one broadcast load per wave), which travels while
        // the first sub-chunk is probed
        uint32_t nt = 0;
        if (threadIdx.x == 0) nt = atomicAdd(task_ctr, 1u);
        lds_build<MODE>(rv, n, hshift, hmask, nb, rkey, rpay, head, nxt, scratch);
        if (threadIdx.x == 0) next_s = nt;
        __syncthreads();
        const uint32_t ntu = next_s;
        const bool has_next = ntu < ntasks;
        const JoinTask nd = get_task(ts, has_next ? ntu : 0);
        for (uint32_t s0 = 0;; s0 += SUB) {
            uint32_t ck[SI], fp[SI];
#pragma unroll
            for (int g0 = 0; g0 < SI; g0 += 4) {
                uint32_t c4[4], f4[4];
                lds_count_batch<MODE, 4>(sv, g0, s0, tk.s_len, hshift, hmask, rkey, rpay, head,
                                         nxt, c4, f4, cks);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    ck[g0 + q] = c4[q];
                    fp[g0 + q] = f4[q];
                }
            }
#pragma unroll
            for (int q = 0; q < SI; q++) {
                matches += ck[q];
                const uint32_t incl = wave_inclusive_scan(ck[q]);
                if (lane == TBJ_WAVE - 1) wtot[ph][w][q] = incl;
            }
            __syncthreads();
            const bool last = s0 + SUB >= tk.s_len;
            // the S payloads (keys for rel) are kept for the writes; an item
            // with several matches re-reads its tuple (rare: R keys repeat)
            uint32_t cpay[SI];
#pragma unroll
            for (int k = 0; k < SI; k++) cpay[k] = rel ? sv[k].x : sv[k].y;
            if (!last) {
                load_range(sv, Sb, tk.s_off, tk.s_len, s0 + SUB);
            } else if (has_next) {
                load_range(sv, Sb, nd.s_off, nd.s_len, 0);
            }
            // wave 0: output offsets (item q's slots follow items < q; inside
            // an item, wave order) and the claim, issued after the loads so
            // that its round trip overlaps theirs
            if (w == 0) {
                uint32_t allq = 0;
                if (lane < SI)
                    for (int ww = 0; ww < NW; ww++) allq += wtot[ph][ww][lane];
                const uint32_t inc = wave_inclusive_scan(allq);
                const uint32_t tot = __shfl(inc, TBJ_WAVE - 1, TBJ_WAVE);
                if (lane < SI) {
                    uint32_t run = inc - allq;
                    for (int ww = 0; ww < NW; ww++) {
                        const uint32_t v = wtot[ph][ww][lane];
                        woff[ph][ww][lane] = run;
                        run += v;
                    }
                }
                if (lane == 0) obase_s = tot ? atomicAdd(&counters[2], (unsigned long long)tot) : 0;
            }
            __syncthreads();
            const uint64_t obase = obase_s;
#pragma unroll
            for (int q = 0; q < SI; q++) {
                const uint32_t incl = wave_inclusive_scan(ck[q]);
                uint64_t o = obase + woff[ph][w][q] + incl - ck[q];
                if (ck[q] == 1) {
                    if (o < cap) pairs[o] = rel ? make_uint2(cpay[q], fp[q]) : make_uint2(fp[q], cpay[q]);
                } else if (ck[q] > 1) {
                    uint32_t c2 = 0, f2 = 0;
                    uint64_t k2 = 0;
                    const uint2 v = Sb[tk.s_off + s0 + q * kJoinBlock + threadIdx.x];
                    lds_probe<MODE, true>(v, hshift, hmask, rkey, rpay, head, nxt, c2, k2, f2,
                                          pairs, cap, o, rel);
                }
            }
            ph ^= 1;
            if (last) break;
        }
        if (has_next) load_range(rv, Rb, nd.r_off, nd.r_len, 0);
        __syncthreads();  // the table is no longer read
        for (uint32_t i = threadIdx.x; i <= nb && i < (uint32_t)NHEAD; i += kJoinBlock) head[i] = 0;
        live = has_next;
        tk = nd;
        __syncthreads();
    }
    flush_counters(matches, cks, red_m, red_c, counters);
}

// ---------------------------------------------------------------------------
// Sort-merge task (RSM: radix_sortmerge_join.c:74-100; PSM's merge:
// parallel_sortmerge_join.h:18-51).  The R range is sorted in LDS by
// (key, payload) -- a bucket sort over the range's key span, each bucket
// insertion-sorted by one thread, with a bitonic network when a bucket is
// too large (clustered keys); every S tuple then finds its equal-key run with
// a lower-bound search inside its bucket of the sort (several searches in
// lockstep per thread) and walks it.
// The merge is bounded on both sides (the reference's unbounded inner loops
// overrun the arrays and miscount, parallel_sortmerge_join.h:26-46).  A task
// sorts its R range once and probes its S range in sub-chunks.

constexpr uint32_t kSortBucketMax = 32;

// Bucket of key among n buckets spread evenly over [lo, lo + span):
// floor((key - lo) * n / span), monotone in key.  The floor of the double
// quotient is exact: (key - lo) * n < 2^44 and span <= 2^32 are exact
// doubles, and a quotient below n <= 4096 that is not an integer lies at
// least 1 / span >= 2^-32 below the next integer, while its rounding error
// is at most 2^12 * 2^-53.
__device__ __forceinline__ uint32_t sort_bucket(uint32_t key, uint32_t lo, uint32_t n,
                                                uint64_t span) {
    return (uint32_t)((double)((uint64_t)(key - lo) * n) / (double)span);
}

// Sorts the n values v (item k of the thread at k * kJoinBlock + threadIdx.x)
// into srt[0, n) ascending; cnt (n words), scratch and mm/big are LDS work
// space.  Every thread must call it; ends with a barrier.  Afterwards mm holds
// the least and greatest key and cnt[b] the first sorted position of bucket b
// (sort_bucket over span = mm[1] - mm[0] + 1): a directory of the sorted keys.
template <int ITEMS>
__device__ void lds_sort_u64(const unsigned long long (&v)[ITEMS], uint32_t n,
                             unsigned long long* srt, uint32_t* cnt, uint32_t* scratch,
                             uint32_t* mm, uint32_t* big) {
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        mm[0] = 0xFFFFFFFFu;
        mm[1] = 0;
        *big = 0;
    }
    for (uint32_t i = tid; i < n; i += kJoinBlock) cnt[i] = 0;
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        if (k * kJoinBlock + tid < n) {
            const uint32_t key = (uint32_t)(v[k] >> 32);
            mn = min(mn, key);
            mx = max(mx, key);
        }
    }
    for (int d = TBJ_WAVE / 2; d > 0; d >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor(mn, d, TBJ_WAVE));
        mx = max(mx, (uint32_t)__shfl_xor(mx, d, TBJ_WAVE));
    }
    __syncthreads();
    if (lane_id() == 0) {
        atomicMin(&mm[0], mn);
        atomicMax(&mm[1], mx);
    }
    __syncthreads();
    const uint32_t lo = mm[0];
    const uint64_t span = (uint64_t)mm[1] - lo + 1;  // keys map monotonically to n buckets
    uint32_t bk[ITEMS], rk[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        if (k * kJoinBlock + tid < n) {
            bk[k] = sort_bucket((uint32_t)(v[k] >> 32), lo, n, span);
            rk[k] = atomicAdd(&cnt[bk[k]], 1u);
        }
    }
    __syncthreads();
    block_scan_lds<kJoinBlock>(cnt, n, scratch);
#pragma unroll
    for (int k = 0; k < ITEMS; k++)
        if (k * kJoinBlock + tid < n) srt[cnt[bk[k]] + rk[k]] = v[k];
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kJoinBlock) {
        const uint32_t s0 = cnt[i], s1 = i + 1 < n ? cnt[i + 1] : n;
        if (s1 - s0 > kSortBucketMax) {
            *big = 1;
        } else {
            for (uint32_t x = s0 + 1; x < s1; x++) {
                const unsigned long long t = srt[x];
                uint32_t y = x;
                while (y > s0 && srt[y - 1] > t) {
                    srt[y] = srt[y - 1];
                    y--;
                }
                srt[y] = t;
            }
        }
    }
    __syncthreads();
    if (*big) {
        const uint32_t np2 = next_pow2_u32(n);
        for (uint32_t j = n + tid; j < np2; j += kJoinBlock) srt[j] = ~0ull;
        __syncthreads();
        for (uint32_t k = 2; k <= np2; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = tid; i < np2; i += kJoinBlock) {
                    const uint32_t ixj = i ^ j;
                    if (ixj > i) {
                        const bool up = (i & k) == 0;
                        const unsigned long long a = srt[i], b = srt[ixj];
                        if ((a > b) == up) {
                            srt[i] = b;
                            srt[ixj] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
}

__global__ __launch_bounds__(kJoinBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_join_sort(const uint2* __restrict__ Rb,
                                                          const uint2* __restrict__ Sb, TaskSrc ts,
                                                          uint32_t* __restrict__ task_ctr,
                                                          OutBuf out, bool mat, bool rel,
                                                          unsigned long long* __restrict__ counters) {
    constexpr int RITEMS = kJoinCap / kJoinBlock;
    constexpr int G = kJoinProbeItems;  // S tuples searched in lockstep per thread
    __shared__ unsigned long long srt[kJoinCap];
    __shared__ uint32_t cnt[kJoinCap];
    __shared__ uint32_t scratch[kJoinBlock / TBJ_WAVE + 1];
    __shared__ uint32_t mm[2], big;
    __shared__ uint32_t task_s;
    __shared__ JoinTask tk_s;
    __shared__ OutMap omap_s;
    __shared__ OutTail cur_s;  // the workgroup's current output chunk
    __shared__ unsigned long long red_m[kJoinBlock / TBJ_WAVE], red_c[kJoinBlock / TBJ_WAVE];
    // materialising: per-wave match totals of each item and each wave's
    // output offset for each item, double-buffered by group parity (a wave
    // may count the next group while others still write this one)
    constexpr int NW = kJoinBlock / TBJ_WAVE;
    __shared__ uint32_t wtot[2][NW][G];
    __shared__ uint32_t woff[2][NW][G];
    const uint32_t lane = lane_id(), w = wave_id();
    uint32_t ph = 0;
    const uint32_t ntasks = *ts.ntasks_dev;
    uint64_t matches = 0, cks = 0;
    if (threadIdx.x == 0) {
        cur_s.base = 0;
        cur_s.used = out.chunk;
    }
    for (;;) {
        if (threadIdx.x == 0) {
            task_s = atomicAdd(task_ctr, 1u);
            if (task_s < ntasks) tk_s = get_task(ts, task_s);
        }
        __syncthreads();
        const uint32_t t = task_s;
        if (t >= ntasks) break;
        const JoinTask tk = tk_s;
        const uint32_t n = tk.r_len;
        {
            unsigned long long v[RITEMS];
#pragma unroll
            for (int k = 0; k < RITEMS; k++) {
                const uint32_t j = k * kJoinBlock + threadIdx.x;
                const uint2 r = j < n ? Rb[tk.r_off + j] : make_uint2(0, 0);
                v[k] = ((unsigned long long)r.x << 32) | r.y;
            }
            lds_sort_u64<RITEMS>(v, n, srt, cnt, scratch, mm, &big);
        }
        const uint32_t kmin = mm[0], kmax = mm[1];
        const uint64_t span = (uint64_t)kmax - kmin + 1;
        // S groups are loaded one ahead (clamped, unconditional loads), so a
        // group's loads fly during the previous group's searches and claim
        const uint32_t slast = tk.s_len ? tk.s_len - 1 : 0;
        uint2 sv[G];
#pragma unroll
        for (int k = 0; k < G; k++) sv[k] = Sb[tk.s_off + min(k * kJoinBlock + threadIdx.x, slast)];
        for (uint32_t base = 0; base < tk.s_len; base += kJoinBlock * G) {
            uint2 sn[G];
#pragma unroll
            for (int k = 0; k < G; k++)
                sn[k] = Sb[tk.s_off + min(base + kJoinBlock * G + k * kJoinBlock + threadIdx.x, slast)];
            uint32_t lo[G], hi[G];
            // first sorted entry with key >= S key, searched inside the S
            // key's bucket [cnt[b], cnt[b + 1]) (keys outside [kmin, kmax]
            // and S tuples past the range: n, no match)
#pragma unroll
            for (int k = 0; k < G; k++) {
                const uint32_t j = base + k * kJoinBlock + threadIdx.x;
                const uint32_t x = sv[k].x;
                lo[k] = hi[k] = n;
                if (j < tk.s_len && x >= kmin && x <= kmax) {
                    const uint32_t bk = sort_bucket(x, kmin, n, span);
                    lo[k] = cnt[bk];
                    hi[k] = bk + 1 < n ? cnt[bk + 1] : n;
                }
            }
            for (;;) {
                bool more = false;
#pragma unroll
                for (int k = 0; k < G; k++) {
                    if (lo[k] < hi[k]) {
                        const uint32_t m = (lo[k] + hi[k]) >> 1;
                        if (srt[m] < ((unsigned long long)sv[k].x << 32)) lo[k] = m + 1;
                        else hi[k] = m;
                        more |= lo[k] < hi[k];
                    }
                }
                if (!more) break;
            }
            uint32_t c[G], fpay[G];  // matches of each item, the first one's R payload
#pragma unroll
            for (int k = 0; k < G; k++) {
                c[k] = 0;
                fpay[k] = 0;
                for (uint32_t e = lo[k]; e < n && (uint32_t)(srt[e] >> 32) == sv[k].x; e++) {
                    if (!c[k]) fpay[k] = (uint32_t)srt[e];
                    c[k]++;
                    cks += (uint64_t)(uint32_t)srt[e] + sv[k].y;
                }
                matches += c[k];
            }
            if (mat) {
                // output slots item-major (item q's pairs follow those of
                // items < q; inside an item, wave order, then lane order), so
                // that a wave's stores of one item are contiguous; wave 0
                // lays the items out and claims the group's range
#pragma unroll
                for (int k = 0; k < G; k++) {
                    const uint32_t incl = wave_inclusive_scan(c[k]);
                    if (lane == TBJ_WAVE - 1) wtot[ph][w][k] = incl;
                }
                __syncthreads();
                if (w == 0) {
                    uint32_t allq = 0;
                    if (lane < (uint32_t)G)
                        for (int ww = 0; ww < NW; ww++) allq += wtot[ph][ww][lane];
                    const uint32_t inc = wave_inclusive_scan(allq);
                    const uint32_t tot = __shfl(inc, TBJ_WAVE - 1, TBJ_WAVE);
                    if (lane < (uint32_t)G) {
                        uint32_t run = inc - allq;
                        for (int ww = 0; ww < NW; ww++) {
                            const uint32_t v = wtot[ph][ww][lane];
                            woff[ph][ww][lane] = run;
                            run += v;
                        }
                    }
                    if (lane == 0) omap_s = out_claim(out, cur_s, tot, &counters[2]);
                }
                __syncthreads();
                const OutMap om = omap_s;
#pragma unroll
                for (int k = 0; k < G; k++) {
                    // (the scan again: cheaper than keeping it live across the barriers)
                    const uint32_t incl = wave_inclusive_scan(c[k]);
                    if (!c[k]) continue;
                    uint64_t off = woff[ph][w][k] + incl - c[k];  // in the group's claim
                    if (c[k] == 1) {
                        out_put(out, out_slot(om, off),
                                rel ? make_uint2(sv[k].x, fpay[k]) : make_uint2(fpay[k], sv[k].y));
                        continue;
                    }
                    for (uint32_t e = lo[k]; e < n && (uint32_t)(srt[e] >> 32) == sv[k].x; e++) {
                        out_put(out, out_slot(om, off),
                                rel ? make_uint2(sv[k].x, (uint32_t)srt[e])
                                    : make_uint2((uint32_t)srt[e], sv[k].y));
                        off++;
                    }
                }
                ph ^= 1;
            }
#pragma unroll
            for (int k = 0; k < G; k++) sv[k] = sn[k];
        }
        __syncthreads();
    }
    if (out.chunk && threadIdx.x == 0) out.tails[blockIdx.x] = cur_s;
    flush_counters(matches, cks, red_m, red_c, counters);
}

// ---------------------------------------------------------------------------
// Gap fill after a chunk-mode join.  The gaps are the unused tails of the
// workgroups' last chunks; every other claimed slot holds a pair.  With
// W = min(matches, cap), the pairs in non-gap slots at or above W (below
// cap + side_cap: the rest were dropped) move, in slot order, into the gap
// slots below W, in slot order; there are exactly as many of them as gap
// slots below W when nothing was dropped, and at least as many otherwise
// (the side buffer holds as many slots as all gaps together).  k_out_plan
// (one workgroup) sorts the gaps and lays out both lists with their prefix
// lengths; k_out_move copies.

constexpr int kFixBlock = 1024;
constexpr uint32_t kFixMaxTails = 2048;

// The plan in device memory: total pairs to move, gap count m, then the
// destination list (m entries) and the source list (m + 1 entries: the
// stretch before each gap and the one after the last), each as start slots
// and exclusive prefix lengths with the total appended.
struct FixPlan {
    unsigned long long total;
    uint32_t m, pad;
};
struct FixLists {
    unsigned long long* dst_start;
    uint32_t* dst_pre;  // m + 1
    unsigned long long* src_start;
    uint32_t* src_pre;  // m + 2
};

__host__ __device__ inline FixLists fix_lists(void* work, uint32_t ntails) {
    uint8_t* p = (uint8_t*)work + 64;
    FixLists l;
    l.dst_start = (unsigned long long*)p;
    p += sizeof(unsigned long long) * (ntails + 2);
    l.src_start = (unsigned long long*)p;
    p += sizeof(unsigned long long) * (ntails + 2);
    l.dst_pre = (uint32_t*)p;
    p += sizeof(uint32_t) * (ntails + 2);
    l.src_pre = (uint32_t*)p;
    return l;
}

__global__ __launch_bounds__(kFixBlock) void k_out_plan(OutBuf out, uint32_t ntails,
                                                        const unsigned long long* __restrict__ counters,
                                                        void* __restrict__ work) {
    __shared__ unsigned long long gs[kFixMaxTails];  // gap starts, sorted
    __shared__ uint32_t gl[kFixMaxTails];            // their lengths
    __shared__ uint32_t dl[kFixMaxTails + 1];        // destination lengths, then prefix
    __shared__ uint32_t sl[kFixMaxTails + 2];        // source lengths, then prefix
    __shared__ uint32_t scratch[kFixBlock / TBJ_WAVE + 1];
    __shared__ uint32_t m_s;
    __shared__ unsigned long long keys[kFixMaxTails];
    const uint32_t chunk = out.chunk;
    const unsigned long long W = min((unsigned long long)counters[0], (unsigned long long)out.cap);
    const unsigned long long E = min((unsigned long long)counters[2],
                                     (unsigned long long)(out.cap + out.side_cap));
    if (threadIdx.x == 0) m_s = 0;
    // a tail's gap start, or ~0 without a gap
    for (uint32_t b = threadIdx.x; b < ntails; b += kFixBlock) {
        const OutTail t = out.tails[b];
        keys[b] = t.used < chunk ? t.base + t.used : ~0ull;
        dl[b] = chunk - t.used;  // (lengths, until the ranks are taken)
    }
    __syncthreads();
    // rank sort: gap starts are distinct (each in its own chunk); tails
    // without a gap sort last, by index
    uint32_t rk[(kFixMaxTails + kFixBlock - 1) / kFixBlock];
    uint32_t q = 0;
    for (uint32_t b = threadIdx.x; b < ntails; b += kFixBlock, q++) {
        const unsigned long long k = keys[b];
        uint32_t r = 0;
        for (uint32_t x = 0; x < ntails; x++) {
            const unsigned long long kx = keys[x];
            r += kx < k || (kx == k && x < b);
        }
        rk[q] = r;
    }
    q = 0;
    for (uint32_t b = threadIdx.x; b < ntails; b += kFixBlock, q++) {
        if (keys[b] != ~0ull) {
            gs[rk[q]] = keys[b];
            gl[rk[q]] = dl[b];
            atomicAdd(&m_s, 1u);
        }
    }
    __syncthreads();
    const uint32_t m = m_s;
    for (uint32_t i = threadIdx.x; i <= m; i += kFixBlock) {
        if (i < m) {  // gap i below W
            const unsigned long long a = gs[i], e = a + gl[i];
            dl[i] = a < W ? (uint32_t)(min(e, W) - a) : 0;
        }
        // the stretch before gap i (after the last gap when i == m), inside [W, E)
        const unsigned long long prev = i ? gs[i - 1] + gl[i - 1] : 0;
        const unsigned long long next = i < m ? gs[i] : E;
        const unsigned long long lo = max(prev, W), hi = min(next, E);
        sl[i] = hi > lo ? (uint32_t)(hi - lo) : 0;
    }
    __syncthreads();
    FixLists L = fix_lists(work, ntails);
    for (uint32_t i = threadIdx.x; i <= m; i += kFixBlock) {
        if (i < m) L.dst_start[i] = gs[i];
        const unsigned long long prev = i ? gs[i - 1] + gl[i - 1] : 0;
        L.src_start[i] = max(prev, W);
    }
    const uint32_t dtot = block_scan_lds<kFixBlock>(dl, m, scratch);
    const uint32_t stot = block_scan_lds<kFixBlock>(sl, m + 1, scratch);
    for (uint32_t i = threadIdx.x; i < m; i += kFixBlock) L.dst_pre[i] = dl[i];
    for (uint32_t i = threadIdx.x; i <= m; i += kFixBlock) L.src_pre[i] = sl[i];
    if (threadIdx.x == 0) {
        L.dst_pre[m] = dtot;
        L.src_pre[m + 1] = stot;
        FixPlan* plan = (FixPlan*)work;
        plan->total = min(dtot, stot);
        plan->m = m;
    }
}

// Last index i of pre[0, n] with pre[i] <= t (pre non-decreasing, pre[0] = 0 <= t).
__device__ __forceinline__ uint32_t fix_find(const uint32_t* pre, uint32_t n, uint32_t t) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_out_move(OutBuf out, uint32_t ntails,
                                                  const void* __restrict__ work) {
    const FixPlan plan = *(const FixPlan*)work;
    const FixLists L = fix_lists((void*)work, ntails);
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < plan.total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t tt = (uint32_t)t;
        const uint32_t i = fix_find(L.dst_pre, plan.m, tt);
        const uint32_t j = fix_find(L.src_pre, plan.m + 1, tt);
        const unsigned long long dst = L.dst_start[i] + (tt - L.dst_pre[i]);
        const unsigned long long src = L.src_start[j] + (tt - L.src_pre[j]);
        out.pairs[dst] = src < out.cap ? out.pairs[src] : out.side[src - out.cap];
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
                        uint32_t* ntasks_dev, uint32_t* stats, uint32_t npieces) {
    hipLaunchKernelGGL(k_task_counts, dim3((nparts + 1 + 255) / 256), dim3(256), 0, st, bounds,
                       nparts, npieces, cap, s_chunk, task_off, stats);
    launch_scan(st, task_off, (uint64_t)nparts + 1, partial, ntasks_dev);
}

void launch_fill_tasks(hipStream_t st, const TaskSrc& ts) {
    if (ts.task_cap)
        hipLaunchKernelGGL(k_fill_tasks, dim3((ts.task_cap + 255) / 256), dim3(256), 0, st, ts);
}

template <int MODE>
static void launch_join_t(hipStream_t st, bool mat, bool rel, uint32_t grid, const uint2* Rb,
                          const uint2* Sb, const TaskSrc& ts, uint32_t* task_ctr, uint32_t hshift,
                          const OutBuf& out, unsigned long long* counters) {
    if (mat)
        hipLaunchKernelGGL((k_join_mat<MODE>), dim3(grid), dim3(kJoinBlock), 0, st, Rb, Sb, ts,
                           task_ctr, hshift, out.pairs, out.cap, rel, counters);
    else
        hipLaunchKernelGGL((k_join<MODE>), dim3(grid), dim3(kJoinBlock), 0, st, Rb, Sb, ts,
                           task_ctr, hshift, counters);
}

void launch_join(hipStream_t st, int mode, bool materialize, bool rel, uint32_t grid,
                 const uint2* Rb, const uint2* Sb, const TaskSrc& ts, uint32_t* task_ctr,
                 uint32_t hshift, const OutBuf& out, unsigned long long* counters) {
    if (mode == kModeSort) {
        hipLaunchKernelGGL(k_join_sort, dim3(grid), dim3(kJoinBlock), 0, st, Rb, Sb, ts, task_ctr,
                           out, materialize, rel, counters);
        return;
    }
    if (mode == kModeChain)
        launch_join_t<kModeChain>(st, materialize, rel, grid, Rb, Sb, ts, task_ctr, hshift, out,
                                  counters);
    else
        launch_join_t<kModeHist>(st, materialize, rel, grid, Rb, Sb, ts, task_ctr, hshift, out,
                                 counters);
}

// The hash joins claim once per 4096 S tuples: RHO's moves its pairs at
// streaming bandwidth anyway, and RHT's spills with the chunk logic compiled
// in (C2: 3.04 against 2.43 ms).  The sort-merge join, claiming once per
// 2048 S tuples, is bound by the counter (C2: 3.36 ms, 2.21 with the claims
// taken out) and claims chunks of up to 8192 slots -- about a quarter of a
// workgroup's share of S, so that gaps stay a few percent of the output and
// small joins keep a small side buffer.
uint32_t out_chunk(int mode, uint64_t nS, uint32_t grid) {
    if (mode != kModeSort || !grid || grid > kFixMaxTails) return 0;
    uint64_t c = nS / (4ull * grid);
    uint32_t chunk = 256;
    while (chunk < 8192 && 2ull * chunk <= c) chunk *= 2;
    return chunk;
}

size_t out_fix_bytes(uint32_t ntails) {
    return 64 + (sizeof(unsigned long long) * 2 + sizeof(uint32_t) * 2) * ((size_t)ntails + 2);
}

void launch_out_fix(hipStream_t st, const OutBuf& out, uint32_t ntails,
                    const unsigned long long* counters, void* work) {
    if (!out.chunk || !ntails || ntails > kFixMaxTails) return;
    hipLaunchKernelGGL(k_out_plan, dim3(1), dim3(kFixBlock), 0, st, out, ntails, counters, work);
    // the moves are at most the gaps below W: ntails * chunk slots
    const uint64_t most = (uint64_t)ntails * out.chunk;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((most + 255) / 256, 2048);
    hipLaunchKernelGGL(k_out_move, dim3(grid), dim3(256), 0, st, out, ntails, work);
}

template <typename K>
static int blocks_per_cu(K kernel) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kJoinBlock, 0) !=
            hipSuccess ||
        blocks < 1)
        blocks = 1;
    return blocks;
}

int join_occupancy(int mode, bool materialize) {
    if (mode == kModeSort) return blocks_per_cu(k_join_sort);
    if (mode == kModeChain)
        return materialize ? blocks_per_cu(k_join_mat<kModeChain>)
                           : blocks_per_cu(k_join<kModeChain>);
    return materialize ? blocks_per_cu(k_join_mat<kModeHist>) : blocks_per_cu(k_join<kModeHist>);
}

}  // namespace tbj
