// Orchestration of the partitioned joins (RHO, RHT, RJ, RHO_atomic): up to
// two partition passes over HBM, then persistent LDS build/probe tasks.
//
// Restates join_init_run / prj_thread (radix_join.c:839-1396) for one GPU:
// the pthread barriers become kernel boundaries, the per-thread histograms
// become per-tile histograms, and the task queues become a device task list
// consumed through an atomic counter.
//
// Both passes are exact and deterministic (histogram, scan, scatter).  With
// tbj_options.sampled_regions the second pass instead sizes every
// sub-partition from a 1/16 sample taken during the first histogram and
// claims space with atomics, skipping its histogram read; should a
// sub-partition outgrow its region, the call is redone exactly.
#include <string.h>

#include <algorithm>

#include "tbj_kernels.h"
#include "tbj_runtime.h"

namespace tbj {

namespace {

enum Phase { kPhHist1 = 0, kPhScan1, kPhScatter1, kPhSize2, kPhScatter2, kPhTasks, kPhJoin, kPhOther };

struct Small {  // device scalars, one allocation
    uint32_t ntiles1, ntiles2, ntasks, task_ctr;
    uint32_t stats[2];
    uint32_t minmax[2];
    uint32_t total1, total2, overflow, region_total;
    uint32_t ngroups2, pad[3];
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
    if (B == 0) p.passes = 0;
    else if (req_passes == 1 && B <= (uint32_t)kMaxPassBits) p.passes = 1;
    else if (req_passes == 2 || B > 8) p.passes = 2;
    else p.passes = 1;
    if (p.passes == 1) {
        p.b1 = B;
        p.b2 = 0;
    } else if (p.passes == 2) {
        p.b1 = std::min<uint32_t>((B + 1) / 2, kMaxPassBits);
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

struct Job {
    int algo;
    const uint2* R;
    const uint2* S;
    uint32_t nR, nS;
    uint2* pairs;
    uint64_t cap;
    bool mat;
    bool rel;  // materialise (key, R.payload) tuples
    int mode;
};

// One full run; *overflow_out reports a too-small sample-sized region.  The
// host waits for the device once, at the end (and once more for the learned
// partitioner's judge).
int run_once(Context& C, const Job& J, const tbj_options& opt, bool exact2, tbj_result* res,
             hipStream_t st, bool* overflow_out) {
    Workspace& W = C.ws;
    PhaseTimer pt(C, st, opt.profile != 0);
    *overflow_out = false;
    bool learned = opt.partitioner == TBJ_PART_LEARNED;
    const uint32_t nR = J.nR, nS = J.nS;

    Small* sm = nullptr;
    TBJ_TRY(W.get("small", sizeof(Small), (void**)&sm));
    TBJ_HIP(hipEventRecord(C.ev[0], st));
    TBJ_TRY(pt.begin());
    TBJ_HIP(hipMemsetAsync(sm, 0, sizeof(Small), st));

    const PassPlan plan = plan_passes((uint32_t)choose_bits(nR, opt), opt.passes);
    if (plan.passes == 0) learned = false;
    const uint32_t nparts = plan.passes ? plan.F1 * plan.F2 : 1;
    res->radix_bits = (int)plan.B;
    res->passes = (int)plan.passes;
    res->num_partitions = nparts;

    const uint2* Rb = J.R;
    const uint2* Sb = J.S;
    uint2* bounds = nullptr;
    TBJ_TRY(W.get("bounds", sizeof(uint2) * 2 * std::max(nparts, plan.F1), (void**)&bounds));
    uint32_t* task_off = nullptr;
    TBJ_TRY(W.get("task_off", sizeof(uint32_t) * (std::max(nparts, plan.F1) + 1),
                  (void**)&task_off));
    uint32_t* tpartial = nullptr;
    TBJ_TRY(W.get("tpartial", sizeof(uint32_t) * scan_partials((uint64_t)std::max(nparts, plan.F1) + 1),
                  (void**)&tpartial));

    if (plan.passes == 0) {
        launch_part_bounds(st, nullptr, nullptr, 1, 1, 0, nR, nS, bounds);
    } else {
        const uint64_t ntot = (uint64_t)nR + nS;
        uint2* C1 = nullptr;
        TBJ_TRY(W.get("c1", sizeof(uint2) * ntot, (void**)&C1));
        const uint32_t tR = (nR + kTile - 1) / kTile, tS = (nS + kTile - 1) / kTile;
        const uint32_t T1 = tR + tS;
        TileDesc* tiles1 = nullptr;
        SegDesc* segs1 = nullptr;
        uint16_t* hist1 = nullptr;
        uint32_t *offs1 = nullptr, *gs1 = nullptr, *partial = nullptr, *fine = nullptr;
        // per-tile counts and offsets (T1 rows) and the group sums of R's and
        // S's tiles
        const uint64_t L1 = (uint64_t)plan.F1 * (seg_groups(0, tR) + seg_groups(tR, tS)) + 1;
        TBJ_TRY(W.get("tiles1", sizeof(TileDesc) * T1, (void**)&tiles1));
        TBJ_TRY(W.get("segs1", sizeof(SegDesc) * 2, (void**)&segs1));
        TBJ_TRY(W.get("hist1", sizeof(uint16_t) * plan.F1 * T1, (void**)&hist1));
        TBJ_TRY(W.get("offs1", sizeof(uint32_t) * plan.F1 * T1, (void**)&offs1));
        TBJ_TRY(W.get("gs1", sizeof(uint32_t) * L1, (void**)&gs1));
        const uint32_t T2max = plan.passes == 2 ? (uint32_t)(ntot / kTile) + 2 * plan.F1 + 2 : 0;
        // groups of the second pass: at most one more than tiles / kGroupTiles
        // per segment, plus one straddling each segment start
        const uint64_t G2max = T2max / kGroupTiles + 2ull * (2 * plan.F1) + 1;
        const uint64_t L2max = exact2 ? (uint64_t)plan.F2 * G2max + 1 : 0;
        const uint64_t nsub = 2ull * plan.F1 * plan.F2;
        TBJ_TRY(W.get("partial",
                      sizeof(uint32_t) * scan_partials(std::max({L1, L2max, nsub + 1, (uint64_t)nparts + 1})),
                      (void**)&partial));
        const bool sample_fine = plan.passes == 2 && !exact2;
        if (sample_fine) {
            TBJ_TRY(W.get("fine", sizeof(uint32_t) * nsub, (void**)&fine));
            launch_fill_u32(st, fine, nsub, 0);
        }

        CdfModel* model = nullptr;
        BinSpec s1r, s1s, s2;
        if (learned) {
            TBJ_TRY(W.get("model", sizeof(CdfModel), (void**)&model));
            uint32_t m = 2;
            while (m * 2 <= (uint32_t)std::min<uint64_t>(nR, kFitMax)) m *= 2;
            if (opt.learned_samples > 0)
                while (m > (uint32_t)opt.learned_samples && m > 2) m >>= 1;
            uint32_t* sample = nullptr;
            TBJ_TRY(W.get("sample", sizeof(uint32_t) * m, (void**)&sample));
            // 64 samples per segment: knot noise of about 1/8 of a segment
            const uint32_t nseg = std::max<uint32_t>(1, std::min<uint32_t>(kLdsKnots, m / 64));
            launch_fill_u32(st, sm->minmax, 1, 0xFFFFFFFFu);
            // the sample and its sort (one workgroup) run on the side stream
            // while the exact key range reads all of R
            TBJ_HIP(hipEventRecord(C.fork_ev, st));
            TBJ_HIP(hipStreamWaitEvent(C.side_stream, C.fork_ev, 0));
            launch_sample(C.side_stream, Rb, nR, m, sample);
            launch_fit_sort(C.side_stream, sample, m);
            TBJ_HIP(hipEventRecord(C.join_ev, C.side_stream));
            launch_minmax(st, Rb, nR, sm->minmax);
            TBJ_HIP(hipStreamWaitEvent(st, C.join_ev, 0));
            launch_fit(st, sample, m, nseg, plan.B, sm->minmax, model);
            s1r = learned_spec(plan, 1, model, 0);
            s1s = learned_spec(plan, 1, model, 1);
            s2 = learned_spec(plan, 2, model, 1);
        } else {
            s1r = s1s = radix_spec(plan, 1);
            s2 = radix_spec(plan, 2);
        }

        // ---- pass 1 (exact): histogram, scan, scatter
        launch_make_tiles1(st, tiles1, segs1, nR, nS, plan.F1, &sm->ntiles1);
        launch_hist(st, tR, Rb, tiles1, &sm->ntiles1, 0, hist1, s1r, plan.F1, fine, &s2, plan.F2);
        launch_hist(st, tS, Sb, tiles1, &sm->ntiles1, tR, hist1, s1s, plan.F1, fine, &s2, plan.F2);
        launch_fill_u32(st, gs1 + (L1 - 1), 1, 0);
        TBJ_TRY(pt.mark(kPhHist1));
        launch_tile_offsets(st, T1, hist1, offs1, tiles1, segs1, &sm->ntiles1, plan.F1, gs1, L1,
                            partial, &sm->total1);

        if (learned) {
            // Check the model on the first-pass R partitions; fall back to radix
            // when one would overflow the per-partition bound.
            launch_part_bounds(st, gs1, segs1, 1, plan.F1, 1, nR, nS, bounds);
            launch_count_tasks(st, bounds, plan.F1, kJoinCap, 1u << 30, task_off, tpartial,
                               &sm->ntasks, sm->stats);
            Small h;
            TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small), hipMemcpyDeviceToHost, st));
            TBJ_HIP(hipStreamSynchronize(st));
            const uint64_t per = opt.learned_bound ? opt.learned_bound : (uint64_t)kJoinCap * 85 / 100;
            if (h.stats[0] > per * plan.F2) {
                learned = false;
                res->fallback = 1;
                s1r = s1s = radix_spec(plan, 1);
                s2 = radix_spec(plan, 2);
                if (sample_fine) launch_fill_u32(st, fine, nsub, 0);
                launch_hist(st, tR, Rb, tiles1, &sm->ntiles1, 0, hist1, s1r, plan.F1, fine, &s2,
                            plan.F2);
                launch_hist(st, tS, Sb, tiles1, &sm->ntiles1, tR, hist1, s1s, plan.F1, fine, &s2,
                            plan.F2);
                launch_fill_u32(st, gs1 + (L1 - 1), 1, 0);
                launch_tile_offsets(st, T1, hist1, offs1, tiles1, segs1, &sm->ntiles1, plan.F1,
                                    gs1, L1, partial, &sm->total1);
            } else {
                res->s_dropped = (uint64_t)nS - ((uint64_t)h.total1 - nR);
            }
            launch_fill_u32(st, sm->stats, 2, 0);
        }
        TBJ_TRY(pt.mark(kPhScan1));
        launch_scatter(st, tR, Rb, C1, tiles1, &sm->ntiles1, 0, offs1, s1r, plan.F1,
                       plan.passes == 1);
        launch_scatter(st, tS, Sb, C1, tiles1, &sm->ntiles1, tR, offs1, s1s, plan.F1,
                       plan.passes == 1);
        TBJ_TRY(pt.mark(kPhScatter1));

        if (plan.passes == 1) {
            launch_part_bounds(st, gs1, segs1, 1, plan.F1, 1, nR, nS, bounds);
            Rb = Sb = C1;
        } else {
            // ---- pass 2: every first-pass partition is a segment
            TileDesc* tiles2 = nullptr;
            SegDesc* segs2 = nullptr;
            TBJ_TRY(W.get("tiles2", sizeof(TileDesc) * T2max, (void**)&tiles2));
            TBJ_TRY(W.get("segs2", sizeof(SegDesc) * 2 * plan.F1, (void**)&segs2));
            launch_make_segments(st, gs1, segs1, 2, plan.F1, plan.F2, segs2, tiles2,
                                 &sm->ntiles2, T2max, &sm->ngroups2);
            uint2* C2 = nullptr;
            if (exact2) {
                uint16_t* hist2 = nullptr;
                uint32_t *offs2 = nullptr, *gs2 = nullptr;
                TBJ_TRY(W.get("c2", sizeof(uint2) * ntot, (void**)&C2));
                TBJ_TRY(W.get("hist2", sizeof(uint16_t) * plan.F2 * T2max, (void**)&hist2));
                TBJ_TRY(W.get("offs2", sizeof(uint32_t) * plan.F2 * T2max, (void**)&offs2));
                TBJ_TRY(W.get("gs2", sizeof(uint32_t) * L2max, (void**)&gs2));
                launch_hist(st, T2max, C1, tiles2, &sm->ntiles2, 0, hist2, s2, plan.F2);
                // group sums past the segments' own are zero, so the scan's
                // total is the tuple count
                launch_zero_tail(st, gs2, L2max, &sm->ngroups2, plan.F2, L2max);
                launch_tile_offsets(st, T2max, hist2, offs2, tiles2, segs2, &sm->ntiles2, plan.F2,
                                    gs2, L2max, partial, &sm->total2);
                TBJ_TRY(pt.mark(kPhSize2));
                launch_scatter(st, T2max, C1, C2, tiles2, &sm->ntiles2, 0, offs2, s2, plan.F2,
                               true);
                launch_part_bounds(st, gs2, segs2, plan.F1, plan.F2, 2, nR, nS, bounds);
            } else {
                // regions sized from the 1/16 sample (ratio estimate inside each
                // first-pass partition); room for 1.1x + 6 sigma + 64 tuples
                uint32_t *starts = nullptr, *cursors = nullptr;
                TBJ_TRY(W.get("starts", sizeof(uint32_t) * (nsub + 1), (void**)&starts));
                TBJ_TRY(W.get("cursors", sizeof(uint32_t) * nsub, (void**)&cursors));
                // caps: 1.1 e + 6 sqrt(16 e) + 64 with sum(e) <= 2 ntot (raw and
                // shrunk estimates each sum to the segment size)
                const double cap_tuples = 2.2 * (double)ntot +
                                          24.0 * sqrt(2.0 * (double)ntot * (double)nsub) +
                                          64.0 * nsub + 1;
                if (cap_tuples >= 4294967295.0) return TBJ_ERR_TOO_LARGE;
                TBJ_TRY(W.get("c2", sizeof(uint2) * (uint64_t)cap_tuples, (void**)&C2));
                launch_region_caps(st, fine, segs2, 2 * plan.F1, plan.F2, starts);
                launch_scan(st, starts, nsub + 1, partial, &sm->region_total);
                launch_fill_u32(st, cursors, nsub, 0);
                TBJ_TRY(pt.mark(kPhSize2));
                launch_scatter_est(st, T2max, C1, C2, tiles2, &sm->ntiles2, starts, cursors,
                                   &sm->overflow, s2, plan.F2);
                launch_bounds_est(st, starts, cursors, (uint32_t)nsub, bounds);
            }
            TBJ_TRY(pt.mark(kPhScatter2));
            Rb = Sb = C2;
        }
    }
    res->partitioner_used = learned ? TBJ_PART_LEARNED : TBJ_PART_RADIX;
    TBJ_HIP(hipEventRecord(C.ev[1], st));

    // ---- join tasks
    const uint32_t max_chunk = (uint32_t)join_s_chunk(J.mat) * kJoinSTaskChunks;
    const uint32_t s_chunk = (opt.s_chunk > 0 && (uint32_t)opt.s_chunk < max_chunk)
                                 ? (uint32_t)opt.s_chunk : max_chunk;
    launch_count_tasks(st, bounds, nparts, kJoinCap, s_chunk, task_off, tpartial, &sm->ntasks,
                       sm->stats);
    // the task list is sized without the count: every partition's tasks are
    // listed unless one spans several builds and several S chunks at once
    // (then the join decodes the rest from task_off)
    TaskSrc ts{};
    ts.task_off = task_off;
    ts.bounds = bounds;
    ts.ntasks_dev = &sm->ntasks;
    ts.nparts = nparts;
    ts.cap = kJoinCap;
    ts.chunk = s_chunk;
    const uint64_t tcap = (uint64_t)nparts + (nR + kJoinCap - 1) / kJoinCap +
                          (nS + s_chunk - 1) / s_chunk + 64;
    ts.task_cap = (uint32_t)std::min<uint64_t>(tcap, 1ull << 31);
    TBJ_TRY(W.get("tasks", sizeof(JoinTask) * ts.task_cap, (void**)&ts.tasks));
    launch_fill_tasks(st, ts);
    TBJ_TRY(pt.mark(kPhTasks));
    const uint32_t hshift = learned ? 0 : plan.B;
    const uint32_t occ = (uint32_t)join_occupancy(J.mode, J.mat);
    const uint32_t grid = (uint32_t)C.num_cus * occ;
    launch_join(st, J.mode, J.mat, J.rel, grid, Rb, Sb, ts, &sm->task_ctr, hshift, J.pairs,
                J.mat ? J.cap : 0, sm->counters);
    TBJ_HIP(hipGetLastError());
    TBJ_TRY(pt.mark(kPhJoin));
    TBJ_HIP(hipEventRecord(C.ev[2], st));
    Small h;
    TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    TBJ_TRY(pt.finish(res));
    res->max_r_partition = h.stats[0];
    *overflow_out = h.overflow != 0;
    res->matches = h.counters[0];
    res->checksum = h.counters[1];
    res->pairs_written = J.mat ? std::min<uint64_t>(h.counters[2], J.cap) : 0;
    float ms = 0;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[0], C.ev[1]));
    res->ms_partition = ms;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[1], C.ev[2]));
    res->ms_join = ms;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[0], C.ev[2]));
    res->ms_total = ms;
    return TBJ_OK;
}

}  // namespace

int run_radix_join(Context& C, int algo, const tbj_tuple* R, uint64_t nR64, const tbj_tuple* S,
                   uint64_t nS64, void* pairs, uint64_t pairs_cap, const tbj_options& opt,
                   tbj_result* res, hipStream_t st) {
    if (nR64 + nS64 >= (1ull << 32) - (1ull << 20)) return TBJ_ERR_TOO_LARGE;
    Job J;
    J.algo = algo;
    J.R = (const uint2*)R;
    J.S = (const uint2*)S;
    J.nR = (uint32_t)nR64;
    J.nS = (uint32_t)nS64;
    J.pairs = (uint2*)pairs;
    J.cap = pairs_cap;
    J.mat = opt.materialize && pairs && pairs_cap;
    J.rel = opt.output == TBJ_OUT_RELATION;
    J.mode = algo == TBJ_RHT ? kModeHist
             : (algo == TBJ_RSM || algo == TBJ_PSM || algo == TBJ_MWAY) ? kModeSort
                                                                       : kModeChain;
    if (J.nR == 0 || J.nS == 0) {
        res->partitioner_used = opt.partitioner == TBJ_PART_LEARNED ? TBJ_PART_LEARNED : TBJ_PART_RADIX;
        return TBJ_OK;
    }
    bool overflow = false;
    TBJ_TRY(run_once(C, J, opt, opt.sampled_regions == 0, res, st, &overflow));
    if (overflow) {
        // a sample-sized region was too small: redo with exact histograms
        const float first_ms = res->ms_total;
        memset(res, 0, sizeof(*res));
        TBJ_TRY(run_once(C, J, opt, true, res, st, &overflow));
        res->redone_exact = 1;
        res->ms_total += first_ms;
    }
    return TBJ_OK;
}

}  // namespace tbj
