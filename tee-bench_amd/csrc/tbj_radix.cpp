// Orchestration of the partitioned joins (RHO, RHT, RJ, RHO_atomic): up to
// two partition passes over HBM, then persistent LDS build/probe tasks.
//
// Restates join_init_run / prj_thread (radix_join.c:839-1396) for one GPU:
// the pthread barriers become kernel boundaries, the per-thread histograms
// become per-tile histograms, and the task queues become a device task list
// consumed through an atomic counter.
//
// Two ways to place the partitions:
//  * exact: per pass a histogram read of the input, offsets from a scan,
//    then the scatter (deterministic layout);
//  * sample-sized (two-pass plans of large joins): one kernel reads 1/16 of
//    the input's 16-tuple groups and sizes, from those counts, every output
//    region of both passes; the scatters claim space in them with one atomic
//    per (tile, bin), so neither pass reads its input twice.  First-pass
//    regions are per (bin, tile row): a row is 224 consecutive tiles of one
//    XCD's eighth, so that a region is written by one XCD.  Should any region
//    turn out too small, the call is redone exactly.  The first-pass regions
//    are laid out row by row (a tile's claims hit consecutive words); the
//    second pass visits them partition by partition.
#include <string.h>

#include <algorithm>
#include <cstddef>

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
    uint32_t ngroups2, est_max;
    uint32_t lin;  // the learned fit gave the linear model
    uint32_t hot1;  // a sample chunk saw a hot first-pass bin
    uint32_t hot2;  // the sample saw a hot second-pass bin
    unsigned long long counters[4];
};

// Joins at least this large (|R| + |S|) use sample-sized regions when
// tbj_options.sampled_regions is 0 (auto).
constexpr uint64_t kSampledMin = 1ull << 24;

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

// key_shift: the plan's bits start at this key bit (0 but for
// partition_relation).
BinSpec radix_spec(const PassPlan& p, int pass, uint32_t key_shift = 0) {
    BinSpec s{};
    s.kind = kBinRadix;
    s.bits = p.B + key_shift;
    s.pass_shift = (pass == 1 ? 0 : p.b1) + key_shift;
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

// The per-call state shared by the partition paths.
struct Call {
    Context& C;
    Workspace& W;
    const Job& J;
    const tbj_options& opt;
    tbj_result* res;
    hipStream_t st;
    PhaseTimer& pt;
    Small* sm;
    PassPlan plan;
    uint32_t nparts;
    uint2* bounds;
    uint32_t* task_off;
    uint32_t* tpartial;
    bool learned;
    const uint2* Rb;  // the join's inputs: the partitioned relations
    const uint2* Sb;
    uint32_t npieces;  // ranges per partition side in bounds (1, or 8 per-XCD pieces)
    uint32_t key_shift = 0;  // radix plans: the partition bits start at this key bit
};

// The learned model, fitted on a strided key sample of R, and the partition
// functions of both passes.  exact_range: the model spans R's exact key range
// (a pass over R, beside the sample's sort); otherwise it spans the sample's
// range, and R's exact range -- which decides what S drops -- is written to
// sm->minmax by the sample-sized flow after R's scatter (from the tiles'
// ranges).  Either way the drop test reads sm->minmax.
int make_specs(Call& k, BinSpec* s1r, BinSpec* s1s, BinSpec* s2, bool exact_range) {
    const PassPlan& plan = k.plan;
    if (!k.learned) {
        *s1r = *s1s = radix_spec(plan, 1, k.key_shift);
        *s2 = radix_spec(plan, 2, k.key_shift);
        return TBJ_OK;
    }
    Context& C = k.C;
    const uint32_t nR = k.J.nR;
    CdfModel* model = nullptr;
    TBJ_TRY(k.W.get("model", sizeof(CdfModel), (void**)&model));
    uint32_t m = 2;
    while (m * 2 <= (uint32_t)std::min<uint64_t>(nR, kFitMax)) m *= 2;
    if (k.opt.learned_samples > 0)
        while (m > (uint32_t)k.opt.learned_samples && m > 2) m >>= 1;
    uint32_t* sample = nullptr;
    TBJ_TRY(k.W.get("sample", sizeof(uint32_t) * m, (void**)&sample));
    // 64 samples per segment: knot noise of about 1/8 of a segment
    const uint32_t nseg = std::max<uint32_t>(1, std::min<uint32_t>(kLdsKnots, m / 64));
    if (exact_range) {
        // {UINT_MAX, 0}: the start of the atomic min / max over R (the
        // sample-sized flow writes the range whole after R's scatter)
        launch_fill_u32(k.st, k.sm->minmax, 1, 0xFFFFFFFFu);
        launch_fill_u32(k.st, k.sm->minmax + 1, 1, 0);
        // the sample and its sort (one workgroup) run on the side stream while
        // the exact key range reads all of R
        TBJ_HIP(hipEventRecord(C.fork_ev, k.st));
        TBJ_HIP(hipStreamWaitEvent(C.side_stream, C.fork_ev, 0));
        launch_sample(C.side_stream, k.J.R, nR, m, sample);
        launch_fit_linear(C.side_stream, sample, m, plan.B, model, &k.sm->lin);
        launch_fit_sort(C.side_stream, sample, m, model);
        TBJ_HIP(hipEventRecord(C.join_ev, C.side_stream));
        launch_minmax(k.st, k.J.R, nR, k.sm->minmax);
        TBJ_HIP(hipStreamWaitEvent(k.st, C.join_ev, 0));
        launch_fit(k.st, sample, m, nseg, plan.B, k.sm->minmax, model);
    } else {
        launch_sample(k.st, k.J.R, nR, m, sample);
        launch_fit_linear(k.st, sample, m, plan.B, model, &k.sm->lin);
        launch_fit_sort(k.st, sample, m, model);
        launch_fit(k.st, sample, m, nseg, plan.B, nullptr, model);
    }
    *s1r = learned_spec(plan, 1, model, 0);
    *s1s = learned_spec(plan, 1, model, 1);
    *s2 = learned_spec(plan, 2, model, 1);
    s1s->range = s2->range = k.sm->minmax;
    return TBJ_OK;
}

// The judge of the learned partitioner: the largest first-pass R partition
// (exact, or estimated from the sample) against the per-partition bound.
// Returns true when the call must fall back to radix partitioning.
int judge_learned(Call& k, bool* fall_back) {
    Small h;
    TBJ_HIP(hipMemcpyAsync(&h, k.sm, sizeof(Small), hipMemcpyDeviceToHost, k.st));
    TBJ_HIP(hipStreamSynchronize(k.st));
    const uint64_t per = k.opt.learned_bound ? k.opt.learned_bound : (uint64_t)kJoinCap * 85 / 100;
    *fall_back = h.stats[0] > per * k.plan.F2;
    return TBJ_OK;
}

// Exact partitioning: per pass a histogram, offsets, scatter.
int partition_exact(Call& k) {
    Workspace& W = k.W;
    const PassPlan& plan = k.plan;
    hipStream_t st = k.st;
    Small* sm = k.sm;
    const uint32_t nR = k.J.nR, nS = k.J.nS;
    const uint64_t ntot = (uint64_t)nR + nS;
    uint2* C1 = nullptr;
    TBJ_TRY(W.get("c1", sizeof(uint2) * ntot, (void**)&C1));
    const uint32_t tR = (nR + kTile - 1) / kTile, tS = (nS + kTile - 1) / kTile;
    const uint32_t T1 = tR + tS;
    TileDesc* tiles1 = nullptr;
    SegDesc* segs1 = nullptr;
    uint16_t* hist1 = nullptr;
    uint32_t *offs1 = nullptr, *gs1 = nullptr, *partial = nullptr;
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
    const uint64_t L2max = plan.passes == 2 ? (uint64_t)plan.F2 * G2max + 1 : 0;
    TBJ_TRY(W.get("partial",
                  sizeof(uint32_t) * scan_partials(std::max({L1, L2max, (uint64_t)k.nparts + 1})),
                  (void**)&partial));
    BinSpec s1r, s1s, s2;
    TBJ_TRY(make_specs(k, &s1r, &s1s, &s2, true));

    // ---- pass 1: histogram, scan, scatter
    launch_make_tiles1(st, tiles1, segs1, nR, nS, plan.F1, &sm->ntiles1);
    launch_hist(st, tR, k.J.R, tiles1, &sm->ntiles1, 0, hist1, s1r, plan.F1);
    launch_hist(st, tS, k.J.S, tiles1, &sm->ntiles1, tR, hist1, s1s, plan.F1);
    launch_fill_u32(st, gs1 + (L1 - 1), 1, 0);
    TBJ_TRY(k.pt.mark(kPhHist1));
    launch_tile_offsets(st, T1, hist1, offs1, tiles1, segs1, &sm->ntiles1, plan.F1, gs1, L1,
                        partial, &sm->total1);
    if (k.learned) {
        // Check the model on the first-pass R partitions; fall back to radix
        // when one would overflow the per-partition bound.
        launch_part_bounds(st, gs1, segs1, 1, plan.F1, 1, nR, nS, k.bounds);
        launch_count_tasks(st, k.bounds, plan.F1, kJoinCap, 1u << 30, k.task_off, k.tpartial,
                           &sm->ntasks, sm->stats);
        bool fb = false;
        TBJ_TRY(judge_learned(k, &fb));
        if (fb) {
            k.learned = false;
            k.res->fallback = 1;
            s1r = s1s = radix_spec(plan, 1);
            s2 = radix_spec(plan, 2);
            launch_hist(st, tR, k.J.R, tiles1, &sm->ntiles1, 0, hist1, s1r, plan.F1);
            launch_hist(st, tS, k.J.S, tiles1, &sm->ntiles1, tR, hist1, s1s, plan.F1);
            launch_fill_u32(st, gs1 + (L1 - 1), 1, 0);
            launch_tile_offsets(st, T1, hist1, offs1, tiles1, segs1, &sm->ntiles1, plan.F1, gs1,
                                L1, partial, &sm->total1);
        }
        launch_fill_u32(st, sm->stats, 2, 0);
    }
    TBJ_TRY(k.pt.mark(kPhScan1));
    launch_scatter(st, tR, k.J.R, C1, tiles1, &sm->ntiles1, 0, offs1, s1r, plan.F1,
                   plan.passes == 1);
    launch_scatter(st, tS, k.J.S, C1, tiles1, &sm->ntiles1, tR, offs1, s1s, plan.F1,
                   plan.passes == 1);
    TBJ_TRY(k.pt.mark(kPhScatter1));

    if (plan.passes == 1) {
        launch_part_bounds(st, gs1, segs1, 1, plan.F1, 1, nR, nS, k.bounds);
        k.Rb = k.Sb = C1;
        return TBJ_OK;
    }
    // ---- pass 2: every first-pass partition is a segment
    TileDesc* tiles2 = nullptr;
    SegDesc* segs2 = nullptr;
    uint2* C2 = nullptr;
    uint16_t* hist2 = nullptr;
    uint32_t *offs2 = nullptr, *gs2 = nullptr;
    TBJ_TRY(W.get("tiles2", sizeof(TileDesc) * T2max, (void**)&tiles2));
    TBJ_TRY(W.get("segs2", sizeof(SegDesc) * 2 * plan.F1, (void**)&segs2));
    TBJ_TRY(W.get("c2", sizeof(uint2) * ntot, (void**)&C2));
    TBJ_TRY(W.get("hist2", sizeof(uint16_t) * plan.F2 * T2max, (void**)&hist2));
    TBJ_TRY(W.get("offs2", sizeof(uint32_t) * plan.F2 * T2max, (void**)&offs2));
    TBJ_TRY(W.get("gs2", sizeof(uint32_t) * L2max, (void**)&gs2));
    launch_make_segments(st, gs1, segs1, 2, plan.F1, plan.F2, segs2, tiles2, &sm->ntiles2, T2max,
                         &sm->ngroups2);
    launch_hist(st, T2max, C1, tiles2, &sm->ntiles2, 0, hist2, s2, plan.F2);
    // group sums past the segments' own are zero, so the scan's total is the
    // tuple count
    launch_zero_tail(st, gs2, L2max, &sm->ngroups2, plan.F2, L2max);
    launch_tile_offsets(st, T2max, hist2, offs2, tiles2, segs2, &sm->ntiles2, plan.F2, gs2, L2max,
                        partial, &sm->total2);
    TBJ_TRY(k.pt.mark(kPhSize2));
    launch_scatter(st, T2max, C1, C2, tiles2, &sm->ntiles2, 0, offs2, s2, plan.F2, true);
    launch_part_bounds(st, gs2, segs2, plan.F1, plan.F2, 2, nR, nS, k.bounds);
    TBJ_TRY(k.pt.mark(kPhScatter2));
    k.Rb = k.Sb = C2;
    return TBJ_OK;
}

bool sampled_fits(const PassPlan& plan) {
    return plan.passes == 2 && plan.F1 * plan.F2 <= (1u << 16);
}

// Sample-sized two-pass partitioning (no histogram read).
int partition_sampled(Call& k) {
    Workspace& W = k.W;
    const PassPlan& plan = k.plan;
    hipStream_t st = k.st;
    Small* sm = k.sm;
    const uint32_t nR = k.J.nR, nS = k.J.nS;
    const uint64_t ntot = (uint64_t)nR + nS;
    const uint32_t F1 = plan.F1;
    uint32_t F2 = plan.F2;  // a dense R may narrow the second pass (below)
    const uint32_t tR = (nR + kTile - 1) / kTile, tS = (nS + kTile - 1) / kTile;
    const uint32_t T1 = tR + tS;
    const uint32_t nrowR = 8 * rows_per_eighth(tR), nrowS = 8 * rows_per_eighth(tS);
    const uint32_t np1 = F1 * (nrowR + nrowS);  // first-pass regions
    uint32_t nsub = 2 * k.nparts;                 // second-pass regions
    uint32_t nsub8 = 8 * nsub;                    // their per-XCD sub-regions
    // Region sizes are only known on the device; the buffers take the largest
    // sum the capacity rule allows (sum of 16 * counts <= |R| + |S| plus the
    // sampled blocks' rounding; sum of sqrt <= sqrt(n * sum), Cauchy-Schwarz;
    // second-pass estimates sum to at most twice each partition).
    const double e1 = (double)ntot + 256.0 * (nrowR + nrowS);
    const double cap1 = 1.1 * e1 + 48.0 * sqrt((double)np1 * e1) + 640.0 * np1 + 1;
    // second pass: at most F2 * (2 F1 + 7) sub-regions are non-empty (a
    // partition's tiles span one eighth more per eighth boundary inside them)
    const double e2 = 2.0 * (double)ntot;
    const double nsx = (double)F2 * (2.0 * F1 + 7.0);
    const double cap2 = 1.1 * e2 + 48.0 * sqrt(nsx * e2) + 640.0 * nsx + 1;
    if (cap1 >= 4294967295.0 || cap2 >= 4294967295.0) {
        // (run_once left the clearing of sm to this path's first launch)
        TBJ_HIP(hipMemsetAsync(sm, 0, sizeof(Small), st));
        return partition_exact(k);
    }

    uint2 *C1 = nullptr, *C2 = nullptr;
    TileDesc *tiles1 = nullptr, *tiles2 = nullptr;
    SegDesc* segs1 = nullptr;
    uint32_t *cnt1 = nullptr, *starts1 = nullptr, *cursors1 = nullptr, *fine = nullptr,
             *npart = nullptr, *starts2 = nullptr, *cursors2 = nullptr, *first2 = nullptr,
             *partial = nullptr;
    const uint32_t T2max = (uint32_t)(ntot / kTile) + np1 + 1;
    TBJ_TRY(W.get("c1", sizeof(uint2) * (uint64_t)cap1, (void**)&C1));
    TBJ_TRY(W.get("c2", sizeof(uint2) * (uint64_t)cap2, (void**)&C2));
    TBJ_TRY(W.get("tiles1", sizeof(TileDesc) * T1, (void**)&tiles1));
    TBJ_TRY(W.get("segs1", sizeof(SegDesc) * 2, (void**)&segs1));
    TBJ_TRY(W.get("tiles2", sizeof(TileDesc) * T2max, (void**)&tiles2));
    TBJ_TRY(W.get("cnt1", sizeof(uint32_t) * np1, (void**)&cnt1));
    TBJ_TRY(W.get("starts1", sizeof(uint32_t) * (np1 + 1), (void**)&starts1));
    TBJ_TRY(W.get("cursors1", sizeof(uint32_t) * np1, (void**)&cursors1));
    TBJ_TRY(W.get("fine", sizeof(uint32_t) * nsub, (void**)&fine));
    TBJ_TRY(W.get("npart", sizeof(uint32_t) * 2 * F1, (void**)&npart));
    TBJ_TRY(W.get("starts2", sizeof(uint32_t) * (nsub8 + 1), (void**)&starts2));
    TBJ_TRY(W.get("cursors2", sizeof(uint32_t) * nsub8, (void**)&cursors2));
    TBJ_TRY(W.get("first2", sizeof(uint32_t) * (np1 + 1), (void**)&first2));
    TBJ_TRY(W.get("partial", sizeof(uint32_t) * scan_partials(std::max<uint64_t>(np1, nsub8) + 1),
                  (void**)&partial));
    // the call's first launch: its device scalars (sm, but ntiles1, which
    // the tile descriptors set) cleared with the counters and cursors, and
    // the first pass's tiles described -- before the learned fit, which
    // writes sm->lin
    static_assert(offsetof(Small, ntiles1) == 0 && sizeof(Small) % sizeof(uint32_t) == 0,
                  "Small clears as words after ntiles1");
    launch_start_tiles1(st,
                        ZeroList{{(uint32_t*)sm + 1, fine, cnt1, cursors1, cursors2},
                                 {sizeof(Small) / sizeof(uint32_t) - 1, nsub, np1, np1, nsub8},
                                 5},
                        tiles1, segs1, nR, nS, F1, &sm->ntiles1);
    BinSpec s1r, s1s, s2;
    TBJ_TRY(make_specs(k, &s1r, &s1s, &s2, false));
    const uint32_t nR1 = F1 * nrowR;  // R's first-pass regions (S's follow)
    uint2* tile_range = nullptr;  // learned: R's key range, per tile of R's scatter
    if (k.learned) {
        TBJ_TRY(W.get("tile_range", sizeof(uint2) * std::max<uint32_t>(1, tR), (void**)&tile_range));
        // R first: its sample, then the judge on R's estimated partitions
        launch_sample_rows(st, k.J.R, nR, k.J.S, nS, nrowR, nrowS, s1r, s1s, s2, F1, F2, cnt1,
                           fine, 0, &sm->hot1);
        launch_est_part_max(st, cnt1, F1, nrowR, sm->stats);
        bool fb = false;
        TBJ_TRY(judge_learned(k, &fb));
        launch_fill_u32(st, sm->stats, 2, 0);
        if (fb) {
            k.learned = false;
            k.res->fallback = 1;
            s1r = s1s = radix_spec(plan, 1);
            s2 = radix_spec(plan, 2);
            launch_zero_multi(st, ZeroList{{fine, cnt1, &sm->hot1}, {nsub, np1, 1}, 3});
        }
    }
    if (!k.learned) {
        // ---- sample both relations, size the first pass's regions
        launch_sample_rows(st, k.J.R, nR, k.J.S, nS, nrowR, nrowS, s1r, s1s, s2, F1, F2, cnt1,
                           fine, 2, &sm->hot1);
        TBJ_TRY(k.pt.mark(kPhHist1));
        launch_row_caps(st, cnt1, np1, starts1);
        launch_scan(st, starts1, (uint64_t)np1 + 1, partial, &sm->region_total);
        TBJ_TRY(k.pt.mark(kPhScan1));
        // ---- pass 1: claims per (tile, bin) in the tile row's regions
        ClaimSpec c1r{starts1, cursors1, &sm->overflow, 0, F1, 1, 0, &sm->hot1};
        ClaimSpec c1s{starts1, cursors1, &sm->overflow, nR1, F1, 1, 0, &sm->hot1};
        // Dense R: when R's keys span exactly |R| values (as primary keys
        // 1..|R| do), unique keys fill every partition of the plan one bit
        // narrower with exactly |R| / 2^(B-1) tuples, and if that fits one
        // build the narrower second pass writes runs twice as long and the
        // join builds half as often.  R's scatter takes R's range on the way
        // (per tile, then one reduction); the host reads it while S's
        // scatter runs.  Only plans whose first pass the narrowing keeps
        // (B even: B and B - 1 split with the same first-pass bits).
        const bool try_narrow = k.opt.radix_bits <= 0 && k.key_shift == 0 && !W.arena() &&
                                plan.B % 2 == 0 && plan.b2 >= 2 &&
                                ((uint64_t)nR >> (plan.B - 1)) <= (uint64_t)kJoinCap;
        uint2* trange = nullptr;
        if (try_narrow && !k.C.pinned) {
            void* h = nullptr;
            TBJ_HIP(hipHostMalloc(&h, 256, hipHostMallocDefault));
            k.C.pinned = static_cast<uint32_t*>(h);
        }
        if (try_narrow) {
            TBJ_TRY(W.get("tile_range", sizeof(uint2) * std::max<uint32_t>(1, tR),
                          (void**)&trange));
            c1r.tile_range = trange;
        }
        launch_scatter(st, tR, k.J.R, C1, tiles1, &sm->ntiles1, 0, nullptr, s1r, F1, false, &c1r);
        if (try_narrow) {
            launch_range_reduce(st, trange, tR, sm->minmax);
            TBJ_HIP(hipMemcpyAsync(k.C.pinned, sm->minmax, 2 * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, st));
            TBJ_HIP(hipEventRecord(k.C.host_ev, st));
        }
        launch_scatter(st, tS, k.J.S, C1, tiles1, &sm->ntiles1, tR, nullptr, s1s, F1, false,
                       &c1s);
        if (try_narrow) {
            TBJ_HIP(hipEventSynchronize(k.C.host_ev));
            const uint32_t mn = k.C.pinned[0], mx = k.C.pinned[1];
            if (nR && mx >= mn && (uint64_t)mx - mn + 1 == nR) {
                // the sample's (b1, b2) counts folded to the narrower b2
                uint32_t* fine_n = nullptr;
                TBJ_TRY(W.get("fine_narrow", sizeof(uint32_t) * (nsub / 2), (void**)&fine_n));
                launch_fold_bins(st, fine, 2 * F1, F2, fine_n);
                fine = fine_n;
                k.plan = plan_passes(plan.B - 1, 2);
                s2 = radix_spec(k.plan, 2);
                F2 = k.plan.F2;
                k.nparts = F1 * F2;
                nsub = 2 * k.nparts;
                nsub8 = 8 * nsub;
                k.res->radix_bits = (int)k.plan.B;
                k.res->num_partitions = k.nparts;
            }
        }
    } else {
        // R's regions and scatter, which also takes R's exact key range (what
        // S drops against: its reads cost nothing there, where a pass of its
        // own beside R's sample took that sample from ~40 to ~230 us); then
        // S's sample, regions and scatter
        TBJ_TRY(k.pt.mark(kPhHist1));
        launch_row_caps(st, cnt1, nR1, starts1);
        launch_scan(st, starts1, (uint64_t)nR1 + 1, partial, &sm->region_total);
        ClaimSpec c1r{starts1, cursors1, &sm->overflow, 0, F1, 1, 0, &sm->hot1};
        c1r.tile_range = tile_range;
        launch_scatter(st, tR, k.J.R, C1, tiles1, &sm->ntiles1, 0, nullptr, s1r, F1, false, &c1r);
        launch_range_reduce(st, tile_range, tR, sm->minmax);
        launch_sample_rows(st, k.J.R, nR, k.J.S, nS, nrowR, nrowS, s1r, s1s, s2, F1, F2, cnt1,
                           fine, 1, &sm->hot1);
        launch_row_caps(st, cnt1 + nR1, np1 - nR1, starts1 + nR1);
        launch_scan(st, starts1 + nR1, (uint64_t)(np1 - nR1) + 1, partial, nullptr);
        launch_add_base(st, starts1 + nR1, np1 - nR1 + 1, &sm->region_total);
        TBJ_TRY(k.pt.mark(kPhScan1));
        ClaimSpec c1s{starts1, cursors1, &sm->overflow, nR1, F1, 1, 0, &sm->hot1};
        launch_scatter(st, tS, k.J.S, C1, tiles1, &sm->ntiles1, tR, nullptr, s1s, F1, false,
                       &c1s);
    }
    TBJ_TRY(k.pt.mark(kPhScatter1));

    // ---- pass 2: every filled first-pass region is a segment of tiles
    launch_part_sizes(st, starts1, cursors1, F1, nrowR, nrowS, npart, &sm->total1);
    launch_region_tiles(st, starts1, cursors1, np1, F1, nrowR, nrowS, first2);
    launch_scan(st, first2, (uint64_t)np1 + 1, partial, &sm->ntiles2);
    launch_region_caps(st, fine, npart, 2 * F1, F2, first2, &sm->ntiles2, F1, nrowR, nrowS,
                       starts2, &sm->hot2);
    launch_scan(st, starts2, (uint64_t)nsub8 + 1, partial, &sm->total2);
    launch_region_tile_descs(st, starts1, cursors1, first2, np1, F1, nrowR, nrowS, tiles2);
    TBJ_TRY(k.pt.mark(kPhSize2));
    ClaimSpec c2{starts2, cursors2, &sm->overflow, 0, 8 * F2, 1, F2, &sm->hot2};
    launch_scatter(st, T2max, C1, C2, tiles2, &sm->ntiles2, 0, nullptr, s2, F2, true, &c2);
    // partition p's side rel is the 8 pieces bounds[(rel * nparts + p) * 8 + x]
    launch_bounds_x8(st, starts2, cursors2, F1, F2, k.bounds);
    k.npieces = 8;
    TBJ_TRY(k.pt.mark(kPhScatter2));
    k.Rb = k.Sb = C2;
    return TBJ_OK;
}

// One full run; *overflow_out reports a too-small sample-sized region.  The
// host waits for the device once, at the end (and once more for the learned
// partitioner's judge).
int run_once(Context& C, const Job& J, const tbj_options& opt, bool allow_sampled,
             tbj_result* res, hipStream_t st, bool* overflow_out) {
    Workspace& W = C.ws;
    PhaseTimer pt(C, st, opt.profile != 0);
    *overflow_out = false;
    const uint32_t nR = J.nR, nS = J.nS;

    Small* sm = nullptr;
    TBJ_TRY(W.get("small", sizeof(Small), (void**)&sm));
    Call k{C, W, J, opt, res, st, pt, sm, {}, 0, nullptr, nullptr, nullptr, false, J.R, J.S, 1};
    k.plan = plan_passes((uint32_t)choose_bits(nR, opt), opt.passes);
    const PassPlan& plan = k.plan;
    const uint64_t ntot = (uint64_t)nR + nS;
    const bool sampled = allow_sampled && sampled_fits(plan) &&
                         (opt.sampled_regions > 0 || (opt.sampled_regions == 0 && ntot >= kSampledMin));
    TBJ_HIP(hipEventRecord(C.ev[0], st));
    TBJ_TRY(pt.begin());
    // (the sample-sized path clears sm in its first launch)
    if (!(plan.passes && sampled)) TBJ_HIP(hipMemsetAsync(sm, 0, sizeof(Small), st));

    k.learned = opt.partitioner == TBJ_PART_LEARNED && plan.passes > 0;
    k.nparts = plan.passes ? plan.F1 * plan.F2 : 1;
    res->radix_bits = (int)plan.B;
    res->passes = (int)plan.passes;
    res->num_partitions = k.nparts;
    const uint32_t nb = std::max(k.nparts, plan.F1);
    TBJ_TRY(W.get("bounds", sizeof(uint2) * 2 * 8 * nb, (void**)&k.bounds));
    TBJ_TRY(W.get("task_off", sizeof(uint32_t) * (nb + 1), (void**)&k.task_off));
    TBJ_TRY(W.get("tpartial", sizeof(uint32_t) * scan_partials((uint64_t)nb + 1),
                  (void**)&k.tpartial));

    if (plan.passes == 0) {
        launch_part_bounds(st, nullptr, nullptr, 1, 1, 0, nR, nS, k.bounds);
    } else if (sampled) {
        TBJ_TRY(partition_sampled(k));
    } else {
        TBJ_TRY(partition_exact(k));
    }
    const bool learned = k.learned;
    const uint2* Rb = k.Rb;
    const uint2* Sb = k.Sb;
    const uint32_t nparts = k.nparts;
    uint2* bounds = k.bounds;
    uint32_t* task_off = k.task_off;
    uint32_t* tpartial = k.tpartial;
    res->partitioner_used = learned ? TBJ_PART_LEARNED : TBJ_PART_RADIX;
    TBJ_HIP(hipEventRecord(C.ev[1], st));

    // ---- join tasks
    const uint32_t max_chunk = (uint32_t)join_s_chunk(J.mat) * kJoinSTaskChunks;
    const uint32_t s_chunk = (opt.s_chunk > 0 && (uint32_t)opt.s_chunk < max_chunk)
                                 ? (uint32_t)opt.s_chunk : max_chunk;
    launch_count_tasks(st, bounds, nparts, kJoinCap, s_chunk, task_off, tpartial, &sm->ntasks,
                       sm->stats, k.npieces);
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
    ts.npieces = k.npieces;
    const uint64_t tcap = (uint64_t)nparts * k.npieces + (nR + kJoinCap - 1) / kJoinCap +
                          (nS + s_chunk - 1) / s_chunk + 64;
    ts.task_cap = (uint32_t)std::min<uint64_t>(tcap, 1ull << 31);
    TBJ_TRY(W.get("tasks", sizeof(JoinTask) * ts.task_cap, (void**)&ts.tasks));
    launch_fill_tasks(st, ts);
    TBJ_TRY(pt.mark(kPhTasks));
    const uint32_t hshift = learned ? 0 : plan.B;
    const uint32_t occ = (uint32_t)join_occupancy(J.mode, J.mat);
    const uint32_t grid = (uint32_t)C.num_cus * occ;
    // output claims in chunks where the join would wait on one counter (not
    // from the fixed no-allocation arena: its joins are small)
    OutBuf out{J.pairs, J.mat ? J.cap : 0, nullptr, 0, nullptr, 0, 0};
    void* fix_work = nullptr;
    out.chunk = J.mat && !W.arena() ? out_chunk(J.mode, nS, grid) : 0;
    if (out.chunk) {
        out.side_cap = (uint64_t)grid * out.chunk;
        // under a workspace cap too tight for the side buffer the join
        // claims once per group instead
        if (W.get("out_side", sizeof(uint2) * out.side_cap, (void**)&out.side) != TBJ_OK ||
            W.get("out_tails", sizeof(OutTail) * grid, (void**)&out.tails) != TBJ_OK ||
            W.get("out_fix", out_fix_bytes(grid), &fix_work) != TBJ_OK) {
            out.chunk = 0;
            out.side = nullptr;
            out.side_cap = 0;
            out.tails = nullptr;
        }
    }
    launch_join(st, J.mode, J.mat, J.rel, grid, Rb, Sb, ts, &sm->task_ctr, hshift, out,
                sm->counters);
    if (out.chunk) launch_out_fix(st, out, grid, sm->counters, fix_work);
    TBJ_HIP(hipGetLastError());
    TBJ_TRY(pt.mark(kPhJoin));
    TBJ_HIP(hipEventRecord(C.ev[2], st));
    Small h;
    TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    TBJ_TRY(pt.finish(res));
    res->max_r_partition = h.stats[0];
    if (learned) res->s_dropped = (uint64_t)nS - ((uint64_t)h.total1 - nR);
    res->learned_model = learned ? (h.lin ? 2 : 1) : 0;
    *overflow_out = h.overflow != 0;
    res->matches = h.counters[0];
    res->checksum = h.counters[1];
    res->pairs_written = J.mat ? std::min<uint64_t>(h.counters[0], J.cap) : 0;
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

int partition_relation(Context& C, const tbj_tuple* R, uint32_t nR, uint32_t bits,
                       uint32_t shift, hipStream_t st, const uint2** Rb, const uint2** bounds,
                       uint32_t* b1) {
    if (bits == 0 || bits > 2 * (uint32_t)kMaxPassBits || shift + bits > 32 || nR == 0)
        return TBJ_ERR_INVALID;
    Job J{};
    J.algo = TBJ_RHO;
    J.R = J.S = (const uint2*)R;
    J.nR = nR;
    J.nS = 0;
    J.mode = kModeChain;
    tbj_options o;
    memset(&o, 0, sizeof(o));
    o.partitioner = TBJ_PART_RADIX;
    o.radix_bits = (int32_t)bits;
    o.sampled_regions = -1;
    tbj_result res{};
    PhaseTimer pt(C, st, false);
    Small* sm = nullptr;
    TBJ_TRY(C.ws.get("prel_small", sizeof(Small), (void**)&sm));
    TBJ_HIP(hipMemsetAsync(sm, 0, sizeof(Small), st));
    Call k{C, C.ws, J, o, &res, st, pt, sm, {}, 0, nullptr, nullptr, nullptr, false, J.R, J.S, 1};
    k.plan = plan_passes(bits, 0);
    k.nparts = k.plan.F1 * k.plan.F2;
    k.key_shift = shift;
    TBJ_TRY(C.ws.get("prel_bounds", sizeof(uint2) * 2 * k.nparts, (void**)&k.bounds));
    TBJ_TRY(partition_exact(k));
    *Rb = k.Rb;
    *bounds = k.bounds;
    *b1 = k.plan.b1;
    return TBJ_OK;
}

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
    TBJ_TRY(run_once(C, J, opt, opt.sampled_regions >= 0, res, st, &overflow));
    if (overflow) {
        // a sample-sized region was too small: redo with exact histograms
        const float first_ms = res->ms_total;
        memset(res, 0, sizeof(*res));
        TBJ_TRY(run_once(C, J, opt, false, res, st, &overflow));
        res->redone_exact = 1;
        res->ms_total += first_ms;
    }
    return TBJ_OK;
}

}  // namespace tbj
