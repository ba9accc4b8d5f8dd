// Orchestration of the partitioned joins (RHO, RHT, RJ, RHO_atomic): up to
// two partition passes over HBM, then persistent LDS build/probe tasks.
//
// Restates join_init_run / prj_thread (radix_join.c:839-1396) for one GPU:
// the pthread barriers become kernel boundaries, the per-thread histograms
// become per-tile histograms, and the task queues become a device task list
// consumed through an atomic counter.
#include <string.h>

#include <algorithm>

#include "tbj_kernels.h"
#include "tbj_runtime.h"

namespace tbj {

namespace {

struct Small {  // device scalars, one allocation
    uint32_t ntiles1, ntiles2, ntasks, task_ctr;
    uint32_t stats[2];
    uint32_t minmax[2];
    uint32_t total1, total2;
    uint32_t pad[6];
    unsigned long long counters[4];
};

uint32_t ceil_log2(uint64_t v) {
    uint32_t b = 0;
    while (((uint64_t)1 << b) < v) b++;
    return b;
}

int choose_bits(uint64_t nR, const tbj_options& o) {
    if (o.radix_bits > 0) return std::min(o.radix_bits, 2 * kMaxPassBits);
    // average R partition at half the LDS capacity
    const uint64_t target = kJoinCap / 2;
    if (nR <= target) return 0;
    return (int)std::min<uint32_t>(ceil_log2((nR + target - 1) / target), 2 * kMaxPassBits);
}

struct PassPlan {
    uint32_t B, passes, b1, b2, F1, F2;
};

PassPlan plan_passes(uint32_t B, int req_passes) {
    PassPlan p{};
    p.B = B;
    if (B == 0) {
        p.passes = 0;
    } else if (req_passes == 1 && B <= (uint32_t)kMaxPassBits) {
        p.passes = 1;
    } else if (req_passes == 2 || B > 8) {
        p.passes = 2;
    } else {
        p.passes = 1;
    }
    if (p.passes == 1) {
        p.b1 = B;
        p.b2 = 0;
    } else if (p.passes == 2) {
        p.b1 = (B + 1) / 2;
        p.b2 = B - p.b1;
        if (p.b1 > (uint32_t)kMaxPassBits) p.b1 = kMaxPassBits;
        p.b2 = B - p.b1;
    }
    p.F1 = 1u << p.b1;
    p.F2 = 1u << p.b2;
    return p;
}

BinSpec radix_spec(const PassPlan& p, int pass) {
    BinSpec s{};
    s.kind = kBinRadix;
    s.bits = p.B;
    s.pass_shift = pass == 1 ? 0 : p.b1;
    s.pass_mask = (pass == 1 ? p.F1 : p.F2) - 1;
    return s;
}

BinSpec learned_spec(const PassPlan& p, int pass, const CdfModel* m, uint32_t drop) {
    BinSpec s{};
    s.kind = kBinLearned;
    s.bits = p.B;
    s.pass_shift = pass == 1 ? p.b2 : 0;
    s.pass_mask = (pass == 1 ? p.F1 : p.F2) - 1;
    s.model = m;
    s.drop = drop;
    return s;
}

}  // namespace

int run_radix_join(int algo, const tbj_tuple* R, uint64_t nR64, const tbj_tuple* S,
                   uint64_t nS64, tbj_pair* pairs, uint64_t pairs_cap, const tbj_options& opt,
                   tbj_result* res, hipStream_t st) {
    Context& C = ctx();
    Workspace& W = C.ws;
    if (nR64 + nS64 >= (1ull << 32) - (1ull << 20)) return TBJ_ERR_TOO_LARGE;
    const uint32_t nR = (uint32_t)nR64, nS = (uint32_t)nS64;
    const bool mat = opt.materialize && pairs && pairs_cap;
    const int mode = algo == TBJ_RHT ? kModeHist : kModeChain;
    bool learned = opt.partitioner == TBJ_PART_LEARNED;

    Small* sm = nullptr;
    TBJ_TRY(W.get("small", sizeof(Small), (void**)&sm));
    TBJ_HIP(hipEventRecord(C.ev[0], st));
    TBJ_HIP(hipMemsetAsync(sm, 0, sizeof(Small), st));

    if (nR == 0 || nS == 0) {
        TBJ_HIP(hipEventRecord(C.ev[1], st));
        TBJ_HIP(hipEventRecord(C.ev[2], st));
        TBJ_HIP(hipEventSynchronize(C.ev[2]));
        res->partitioner_used = learned ? TBJ_PART_LEARNED : TBJ_PART_RADIX;
        return TBJ_OK;
    }

    const PassPlan plan = plan_passes((uint32_t)choose_bits(nR, opt), opt.passes);
    if (plan.passes == 0) learned = false;
    const uint32_t nparts = plan.passes ? plan.F1 * plan.F2 : 1;
    res->radix_bits = (int)plan.B;
    res->passes = (int)plan.passes;
    res->num_partitions = nparts;

    const uint2* Rb = (const uint2*)R;
    const uint2* Sb = (const uint2*)S;
    uint2* bounds = nullptr;
    TBJ_TRY(W.get("bounds", sizeof(uint2) * 2 * nparts, (void**)&bounds));

    CdfModel* model = nullptr;
    if (plan.passes == 0) {
        launch_part_bounds(st, nullptr, nullptr, 1, 1, 0, nR, nS, bounds);
    } else {
        const uint64_t ntot = (uint64_t)nR + nS;
        uint2 *C1 = nullptr, *C2 = nullptr;
        TBJ_TRY(W.get("c1", sizeof(uint2) * ntot, (void**)&C1));
        const uint32_t tR = (nR + kTile - 1) / kTile, tS = (nS + kTile - 1) / kTile;
        const uint32_t T1 = tR + tS;
        TileDesc* tiles1 = nullptr;
        SegDesc* segs1 = nullptr;
        uint32_t *hist1 = nullptr, *partial = nullptr;
        const uint64_t L1 = (uint64_t)plan.F1 * T1 + 1;
        TBJ_TRY(W.get("tiles1", sizeof(TileDesc) * T1, (void**)&tiles1));
        TBJ_TRY(W.get("segs1", sizeof(SegDesc) * 2, (void**)&segs1));
        TBJ_TRY(W.get("hist1", sizeof(uint32_t) * L1, (void**)&hist1));
        const uint32_t T2max = plan.passes == 2 ? (uint32_t)(ntot / kTile) + 2 * plan.F1 + 2 : 0;
        const uint64_t L2max = (uint64_t)plan.F2 * T2max + 1;
        TBJ_TRY(W.get("partial", sizeof(uint32_t) * scan_partials(std::max(L1, L2max)),
                      (void**)&partial));

        BinSpec s1r, s1s, s2;
        if (learned) {
            TBJ_TRY(W.get("model", sizeof(CdfModel), (void**)&model));
            uint32_t m = 2;
            while (m * 2 <= (uint32_t)std::min<uint64_t>(nR, kFitMax)) m *= 2;
            if (opt.learned_samples > 0)
                while (m > (uint32_t)opt.learned_samples && m > 2) m >>= 1;
            uint32_t* sample = nullptr;
            TBJ_TRY(W.get("sample", sizeof(uint32_t) * m, (void**)&sample));
            const uint32_t nseg = std::max<uint32_t>(1, std::min<uint32_t>(kLdsKnots, m / 16));
            launch_sample(st, Rb, nR, m, sample);
            launch_fit(st, sample, m, nseg, plan.B, model);
            s1r = learned_spec(plan, 1, model, 0);
            s1s = learned_spec(plan, 1, model, 1);
            s2 = learned_spec(plan, 2, model, 1);
        } else {
            s1r = s1s = radix_spec(plan, 1);
            s2 = radix_spec(plan, 2);
        }

        // ---- pass 1
        launch_make_tiles1(st, tiles1, segs1, nR, nS, plan.F1, &sm->ntiles1);
        launch_fill_u32(st, sm->minmax, 1, 0xFFFFFFFFu);
        launch_hist(st, tR, Rb, tiles1, segs1, &sm->ntiles1, 0, hist1, s1r, plan.F1,
                    learned ? sm->minmax : nullptr);
        if (learned) launch_set_bounds(st, model, sm->minmax, 1);
        launch_hist(st, tS, Sb, tiles1, segs1, &sm->ntiles1, tR, hist1, s1s, plan.F1, nullptr);
        launch_fill_u32(st, hist1 + (L1 - 1), 1, 0);
        launch_scan(st, hist1, L1, partial, &sm->total1);

        if (learned) {
            // Judge the model on the first-pass R partitions; fall back to radix
            // when one would overflow the per-partition bound.
            launch_part_bounds(st, hist1, segs1, 1, plan.F1, 1, nR, nS, bounds);
            uint32_t* toff = nullptr;
            TBJ_TRY(W.get("task_off", sizeof(uint32_t) * std::max(nparts, plan.F1), (void**)&toff));
            launch_count_tasks(st, bounds, plan.F1, kJoinCap, 1u << 30, toff, nullptr,
                               &sm->ntasks, sm->stats);
            Small h;
            TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small), hipMemcpyDeviceToHost, st));
            TBJ_HIP(hipStreamSynchronize(st));
            const uint64_t per = opt.learned_bound ? opt.learned_bound : (uint64_t)kJoinCap * 85 / 100;
            const uint64_t bound1 = per * plan.F2;
            if (h.stats[0] > bound1) {
                learned = false;
                res->fallback = 1;
                s1r = s1s = radix_spec(plan, 1);
                s2 = radix_spec(plan, 2);
                launch_hist(st, tR, Rb, tiles1, segs1, &sm->ntiles1, 0, hist1, s1r, plan.F1, nullptr);
                launch_hist(st, tS, Sb, tiles1, segs1, &sm->ntiles1, tR, hist1, s1s, plan.F1, nullptr);
                launch_fill_u32(st, hist1 + (L1 - 1), 1, 0);
                launch_scan(st, hist1, L1, partial, &sm->total1);
            } else {
                res->s_dropped = (uint64_t)nS - ((uint64_t)h.total1 - nR);
            }
        }
        launch_scatter(st, tR, Rb, C1, tiles1, segs1, &sm->ntiles1, 0, hist1, s1r, plan.F1);
        launch_scatter(st, tS, Sb, C1, tiles1, segs1, &sm->ntiles1, tR, hist1, s1s, plan.F1);

        if (plan.passes == 1) {
            launch_part_bounds(st, hist1, segs1, 1, plan.F1, 1, nR, nS, bounds);
            Rb = Sb = C1;
        } else {
            // ---- pass 2: every first-pass partition is a segment
            TBJ_TRY(W.get("c2", sizeof(uint2) * ntot, (void**)&C2));
            TileDesc* tiles2 = nullptr;
            SegDesc* segs2 = nullptr;
            uint32_t* hist2 = nullptr;
            TBJ_TRY(W.get("tiles2", sizeof(TileDesc) * T2max, (void**)&tiles2));
            TBJ_TRY(W.get("segs2", sizeof(SegDesc) * 2 * plan.F1, (void**)&segs2));
            TBJ_TRY(W.get("hist2", sizeof(uint32_t) * L2max, (void**)&hist2));
            launch_make_segments(st, hist1, segs1, 2, plan.F1, plan.F2, segs2, tiles2, &sm->ntiles2,
                                 T2max);
            launch_hist(st, T2max, C1, tiles2, segs2, &sm->ntiles2, 0, hist2, s2, plan.F2, nullptr);
            launch_zero_tail(st, hist2, L2max, &sm->ntiles2, plan.F2, (uint64_t)plan.F2 * (2 * plan.F1 + 2) + 1);
            launch_scan(st, hist2, L2max, partial, &sm->total2);
            launch_scatter(st, T2max, C1, C2, tiles2, segs2, &sm->ntiles2, 0, hist2, s2, plan.F2);
            launch_part_bounds(st, hist2, segs2, plan.F1, plan.F2, 2, nR, nS, bounds);
            Rb = Sb = C2;
        }
    }
    res->partitioner_used = learned ? TBJ_PART_LEARNED : TBJ_PART_RADIX;
    TBJ_HIP(hipEventRecord(C.ev[1], st));

    // ---- join tasks
    const uint32_t s_chunk = opt.s_chunk > 0 ? (uint32_t)opt.s_chunk : 16384u;
    uint32_t* task_off = nullptr;
    TBJ_TRY(W.get("task_off", sizeof(uint32_t) * nparts, (void**)&task_off));
    launch_count_tasks(st, bounds, nparts, kJoinCap, s_chunk, task_off, nullptr, &sm->ntasks,
                       sm->stats);
    Small h;
    TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    res->max_r_partition = h.stats[0];
    JoinTask* tasks = nullptr;
    TBJ_TRY(W.get("tasks", sizeof(JoinTask) * std::max<uint32_t>(1, h.ntasks), (void**)&tasks));
    launch_fill_tasks(st, bounds, nparts, kJoinCap, s_chunk, task_off, tasks);
    const uint32_t hshift = learned ? 0 : plan.B;
    const uint32_t occ = (uint32_t)join_occupancy(mode, mat);
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(h.ntasks, C.num_cus * occ));
    launch_join(st, mode, mat, grid, Rb, Sb, tasks, &sm->ntasks, &sm->task_ctr, hshift,
                (uint2*)pairs, mat ? pairs_cap : 0, sm->counters);
    TBJ_HIP(hipGetLastError());
    TBJ_HIP(hipEventRecord(C.ev[2], st));
    TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    res->matches = h.counters[0];
    res->checksum = h.counters[1];
    res->pairs_written = mat ? std::min<uint64_t>(h.counters[2], pairs_cap) : 0;
    float ms = 0;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[0], C.ev[1]));
    res->ms_partition = ms;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[1], C.ev[2]));
    res->ms_join = ms;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[0], C.ev[2]));
    res->ms_total = ms;
    return TBJ_OK;
}

}  // namespace tbj
