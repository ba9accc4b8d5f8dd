// Batches of joins: the planner (waves of at most tuning.wave_tuples
// tuples; inside a wave, one group per radix plan of its RHO joins and one
// group of all its PHT joins), the launch sequence (no host
// synchronisation, so it can be captured as one HIP graph), the autotuner
// and the batch generator's host side.
//
// A radix group partitions all of its joins in the same passes: the first
// pass sees one segment per relation of each join (R_0, S_0, R_1, S_1, ...),
// the second one segment per (relation, first-pass bin), exactly as one
// join's second pass does.  Per-join results come from the join kernel,
// which maps a partition back to its join.  A PHT group builds the tables
// of all of its joins in bulk -- per part (its joins with the same table
// layout) the joins' R ranges partitioned in one pass on the buckets' low
// bits, then one workgroup per block of buckets -- and probes them all in one
// launch.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/tbj_gen.h"
#include "tbj_batch.h"
#include "tbj_kernels.h"
#include "tbj_pht.h"
#include "tbj_runtime.h"

namespace {

constexpr uint64_t kWindow = (1ull << 32) - (1ull << 20);  // u32 tuple offsets inside a wave

enum GroupKind : uint32_t { kGroupRadix = 0, kGroupPht = 1 };

struct Group {
    uint32_t kind;
    uint32_t B, passes, b1, b2, F1, F2;
    uint32_t nj;
    uint64_t tuples;
    uint32_t T1, T2max;
    uint64_t L1, L2;       // group-sum lengths of the two passes
    uint32_t NP;
    uint64_t base_off;     // arena tuple offset of the wave window
    size_t segs_off;       // byte offsets into the plan's constant block
    size_t ids_off;
    uint32_t small_idx;    // 8 words each: ntiles1, ntiles2, ntasks, task_ctr, stats[2], total1, total2
                           // (PHT groups: word 0 = the overflow pool counter)
    // PHT groups
    size_t jobs_off;
    uint32_t njobs, nsc;
    uint64_t nbuckets, npool;
    uint32_t part0, part1;  // its parts in tbj_batch_plan::parts
};

// The joins of a PHT group whose tables share pbits (PhtJob): jobs [j0, j1)
// of the group.  pbits > 0: their R ranges (segments, relative to the wave
// window at base_off) are partitioned in one pass on the low pbits bits
// before the blocks are placed, in tiles of kPhtTile tuples whatever the
// radix joins' tile (up to 2^11 bins: the largest tile keeps the runs
// longest).
constexpr uint32_t kPhtTile = tbj::kMaxTile;
struct PhtPart {
    uint32_t pbits;
    uint32_t j0, j1;
    uint32_t T1;
    uint64_t L1;
    uint64_t base_off;
    size_t segs_off;
    uint32_t small_idx;  // word 0: its tile count (set at planning)
};

struct DevBuf {
    void* p = nullptr;
    int get(size_t bytes) {
        if (!bytes) bytes = 256;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return TBJ_ERR_OOM;
        }
        return TBJ_OK;
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

uint32_t ceil_log2(uint64_t v) {
    uint32_t b = 0;
    while ((1ull << b) < v) b++;
    return b;
}

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

tbj::BinSpec radix_pass(uint32_t B, uint32_t shift, uint32_t fanout) {
    tbj::BinSpec s{};
    s.kind = tbj::kBinRadix;
    s.bits = B;
    s.pass_shift = shift;
    s.pass_mask = fanout - 1;
    return s;
}

}  // namespace

struct tbj_batch_plan {
    int algo = TBJ_RHO;
    const tbj_tuple* arena = nullptr;
    uint32_t n = 0;
    unsigned long long* counts = nullptr;
    tbj_batch_tuning tune{};
    std::vector<Group> groups;
    std::vector<PhtPart> parts;
    uint32_t waves = 0;
    uint64_t tuples = 0;
    DevBuf c1, c2, tiles1, hist1, offs1, gs1, tiles2, hist2, offs2, gs2, segs2, bounds, task_off,
        partial, konst, small, total, ptab, ppool;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint32_t grid = 0, pht_grid = 0;

    ~tbj_batch_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }

    // Every launch of one run, in stream order.
    int launch(hipStream_t st) {
        using namespace tbj;
        batch_zero_u64(st, counts, n);
        for (const Group& g : groups) {
            if (g.kind == kGroupPht) {
                const PhtJob* jobs = reinterpret_cast<const PhtJob*>((char*)konst.p + g.jobs_off);
                uint32_t* pool_ctr = (uint32_t*)small.p + 8 * g.small_idx;
                const uint2* arena2 = reinterpret_cast<const uint2*>(arena);
                launch_fill_u32(st, pool_ctr, 1, 0);
                for (uint32_t k = g.part0; k < g.part1; k++) {
                    const PhtPart& pp = parts[k];
                    if (pp.pbits == 0) {
                        batch_pht_place(st, arena2, jobs + pp.j0, pp.j1 - pp.j0, nullptr, nullptr,
                                        0, ptab.p, ppool.p, pool_ctr);
                        continue;
                    }
                    const uint2* base = arena2 + pp.base_off;
                    const SegDesc* segs = reinterpret_cast<const SegDesc*>((char*)konst.p + pp.segs_off);
                    uint32_t* sm = (uint32_t*)small.p + 8 * pp.small_idx;
                    TileDesc* t1 = (TileDesc*)tiles1.p;
                    uint16_t* h1 = (uint16_t*)hist1.p;
                    uint32_t* o1 = (uint32_t*)offs1.p;
                    uint32_t* q1 = (uint32_t*)gs1.p;
                    const uint32_t F = 1u << pp.pbits;
                    const BinSpec bs = radix_pass(pp.pbits, 0, F);
                    const uint32_t nseg = pp.j1 - pp.j0;
                    launch_tiles_from_segs(st, segs, nseg, sm, t1, pp.T1, kPhtTile);
                    launch_hist(st, pp.T1, base, t1, sm, 0, h1, bs, F, kPhtTile);
                    launch_fill_u32(st, q1 + (pp.L1 - 1), 1, 0);
                    launch_tile_offsets(st, pp.T1, h1, o1, t1, segs, sm, F, q1, pp.L1,
                                        (uint32_t*)partial.p, sm + 6);
                    launch_scatter(st, pp.T1, base, (uint2*)c1.p, t1, sm, 0, o1, bs, F, true,
                                   nullptr, kPhtTile);
                    batch_pht_place(st, (const uint2*)c1.p, jobs + pp.j0, nseg, segs, q1,
                                    pp.pbits, ptab.p, ppool.p, pool_ctr);
                }
                batch_pht_probe(st, pht_grid, arena2, jobs, g.njobs, g.nsc, ptab.p, ppool.p,
                                counts);
                continue;
            }
            const uint2* base = reinterpret_cast<const uint2*>(arena) + g.base_off;
            const SegDesc* segs1 = reinterpret_cast<const SegDesc*>((char*)konst.p + g.segs_off);
            const uint32_t* ids = reinterpret_cast<const uint32_t*>((char*)konst.p + g.ids_off);
            uint32_t* sm = (uint32_t*)small.p + 8 * g.small_idx;
            launch_fill_u32(st, sm + 1, 7, 0);
            uint2* bnd = (uint2*)bounds.p;
            const uint2* Rb = base;
            if (g.passes == 0) {
                batch_bounds(st, nullptr, segs1, g.nj, 1, 1, 0, bnd);
            } else {
                TileDesc* t1 = (TileDesc*)tiles1.p;
                uint16_t* h1 = (uint16_t*)hist1.p;
                uint32_t* o1 = (uint32_t*)offs1.p;
                uint32_t* q1 = (uint32_t*)gs1.p;
                const BinSpec s1 = radix_pass(g.B, 0, g.F1);
                launch_tiles_from_segs(st, segs1, 2 * g.nj, sm, t1, g.T1, tune.tile);
                launch_hist(st, g.T1, base, t1, sm, 0, h1, s1, g.F1, tune.tile);
                launch_fill_u32(st, q1 + (g.L1 - 1), 1, 0);
                launch_tile_offsets(st, g.T1, h1, o1, t1, segs1, sm, g.F1, q1, g.L1,
                                    (uint32_t*)partial.p, sm + 6);
                launch_scatter(st, g.T1, base, (uint2*)c1.p, t1, sm, 0, o1, s1, g.F1,
                               g.passes == 1, nullptr, tune.tile);
                if (g.passes == 1) {
                    batch_bounds(st, q1, segs1, g.nj, g.F1, 1, 1, bnd);
                    Rb = (const uint2*)c1.p;
                } else {
                    TileDesc* t2 = (TileDesc*)tiles2.p;
                    uint16_t* h2 = (uint16_t*)hist2.p;
                    uint32_t* o2 = (uint32_t*)offs2.p;
                    uint32_t* q2 = (uint32_t*)gs2.p;
                    SegDesc* s2d = (SegDesc*)segs2.p;
                    const BinSpec s2 = radix_pass(g.B, g.b1, g.F2);
                    launch_make_segments(st, q1, segs1, 2 * g.nj, g.F1, g.F2, s2d, t2, sm + 1,
                                         g.T2max, nullptr, tune.tile);
                    launch_hist(st, g.T2max, (const uint2*)c1.p, t2, sm + 1, 0, h2, s2, g.F2,
                                tune.tile);
                    // group sums past the segments' own (up to g.L2) are not
                    // read: every bound lies before them
                    launch_tile_offsets(st, g.T2max, h2, o2, t2, s2d, sm + 1, g.F2, q2, g.L2,
                                        (uint32_t*)partial.p, sm + 7);
                    launch_scatter(st, g.T2max, (const uint2*)c1.p, (uint2*)c2.p, t2, sm + 1, 0,
                                   o2, s2, g.F2, true, nullptr, tune.tile);
                    batch_bounds(st, q2, s2d, g.nj, g.F1, g.F2, 2, bnd);
                    Rb = (const uint2*)c2.p;
                }
            }
            launch_count_tasks(st, bnd, g.NP, kJoinCap, tune.s_chunk, (uint32_t*)task_off.p,
                               (uint32_t*)partial.p, sm + 2, sm + 4);
            batch_join(st, grid, Rb, Rb, bnd, g.NP, g.F1 * g.F2, (const uint32_t*)task_off.p,
                       sm + 2, sm + 3, tune.s_chunk, g.B, ids, counts);
        }
        batch_sum(st, counts, n, (unsigned long long*)total.p);
        TBJ_HIP(hipGetLastError());
        return TBJ_OK;
    }
};

namespace {

int plan_build(tbj_batch_plan* P, const tbj_batch_item* items) {
    using namespace tbj;
    tbj_batch_tuning& T = P->tune;
    if (!T.target_r) T.target_r = kJoinCap / 2;
    if (!T.one_pass_bits) T.one_pass_bits = 8;
    // s_chunk: S tuples per task, probed in sub-chunks of kJoinSChunk
    const uint32_t max_span = (uint32_t)kJoinSChunk * kJoinSTaskChunks;
    if (!T.s_chunk || T.s_chunk > max_span) T.s_chunk = max_span;
    if (!T.wave_tuples) T.wave_tuples = 1u << 31;
    if (!T.tile) T.tile = kTile;
    if (!valid_tile(T.tile)) return TBJ_ERR_INVALID;
    const uint32_t tile = T.tile;
    T.one_pass_bits = std::min<uint32_t>(T.one_pass_bits, kMaxPassBits);
    const uint32_t n = P->n;
    auto algo_of = [&](uint32_t i) {
        return P->algo == TBJ_BATCH_PER_ITEM ? items[i].algo : P->algo;
    };
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    auto lo_of = [&](uint32_t i) { return std::min(items[i].r_off, items[i].s_off); };
    auto hi_of = [&](uint32_t i) {
        return std::max(items[i].r_off + items[i].nR, items[i].s_off + items[i].nS);
    };
    for (uint32_t i = 0; i < n; i++) {
        if (items[i].nR >= kWindow || items[i].nS >= kWindow || hi_of(i) - lo_of(i) > kWindow)
            return TBJ_ERR_TOO_LARGE;
        if (algo_of(i) != TBJ_RHO && algo_of(i) != TBJ_PHT) return TBJ_ERR_UNSUPPORTED;
        P->tuples += items[i].nR + items[i].nS;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return lo_of(a) < lo_of(b); });

    std::vector<char> konst;
    size_t max_tup = 0, max_T1 = 0, max_T2 = 0, max_L1 = 0, max_L2 = 0, max_s2 = 0, max_NP = 0;
    size_t max_H1 = 0, max_H2 = 0, max_bk = 0, max_pool = 0;
    uint32_t nsmall = 0;  // 8-word device slots of the groups and PHT parts
    auto align = [&](size_t a) { konst.resize((konst.size() + a - 1) / a * a); };
    size_t w0 = 0;
    while (w0 < n) {
        // ---- one wave: consecutive items (by arena offset) within the window and budget
        size_t w1 = w0;
        uint64_t wlo = lo_of(order[w0]), whi = hi_of(order[w0]), wtup = 0;
        while (w1 < n) {
            const uint32_t i = order[w1];
            const uint64_t nlo = std::min(wlo, lo_of(i)), nhi = std::max(whi, hi_of(i));
            const uint64_t ntup = wtup + items[i].nR + items[i].nS;
            if (w1 > w0 && (nhi - nlo > kWindow || ntup > T.wave_tuples)) break;
            wlo = nlo;
            whi = nhi;
            wtup = ntup;
            w1++;
        }
        P->waves++;
        // ---- groups: the wave's PHT joins, and its RHO joins with the same
        // partition bits
        std::vector<std::vector<uint32_t>> byB(2 * kMaxPassBits + 1);
        std::vector<uint32_t> pht;
        for (size_t k = w0; k < w1; k++) {
            const uint32_t i = order[k];
            if (algo_of(i) == TBJ_PHT) {
                pht.push_back(i);
                continue;
            }
            const uint64_t nR = items[i].nR;
            uint32_t B = nR <= T.target_r ? 0 : ceil_log2((nR + T.target_r - 1) / T.target_r);
            B = std::min<uint32_t>(B, 2 * kMaxPassBits);
            byB[B].push_back(i);
        }
        for (uint32_t B = 0; B < byB.size(); B++) {
            const auto& js = byB[B];
            if (js.empty()) continue;
            Group g{};
            g.kind = kGroupRadix;
            g.B = B;
            g.passes = B == 0 ? 0 : B <= T.one_pass_bits ? 1 : 2;
            g.b1 = g.passes == 2 ? (B + 1) / 2 : B;
            g.b2 = B - g.b1;
            g.F1 = 1u << g.b1;
            g.F2 = 1u << g.b2;
            g.nj = (uint32_t)js.size();
            g.base_off = wlo;
            g.NP = g.nj * g.F1 * g.F2;
            std::vector<SegDesc> segs(2 * g.nj);
            uint32_t tiles = 0;
            uint64_t groups = 0;
            for (uint32_t j = 0; j < g.nj; j++) {
                const tbj_batch_item& it = items[js[j]];
                for (int rel = 0; rel < 2; rel++) {
                    SegDesc& d = segs[2 * j + rel];
                    d.in_off = (uint32_t)((rel ? it.s_off : it.r_off) - wlo);
                    d.n = (uint32_t)(rel ? it.nS : it.nR);
                    d.ntiles = (d.n + tile - 1) / tile;
                    d.first_tile = tiles;
                    d.ngroups = seg_groups(tiles, d.ntiles);
                    d.aux = 0;
                    d.hist_base = groups * g.F1;
                    tiles += d.ntiles;
                    groups += d.ngroups;
                    g.tuples += d.n;
                }
            }
            g.T1 = tiles;
            g.L1 = (uint64_t)g.F1 * groups + 1;
            g.T2max = g.passes == 2 ? (uint32_t)(g.tuples / tile) + 2 * g.nj * g.F1 + 1 : 0;
            // second-pass groups: tiles / kGroupTiles plus two per segment
            g.L2 = g.passes == 2
                       ? (uint64_t)g.F2 * (g.T2max / kGroupTiles + 2ull * (2 * g.nj * g.F1) + 1) + 1
                       : 0;
            align(64);
            g.segs_off = konst.size();
            konst.resize(konst.size() + segs.size() * sizeof(SegDesc));
            memcpy(konst.data() + g.segs_off, segs.data(), segs.size() * sizeof(SegDesc));
            align(64);
            g.ids_off = konst.size();
            konst.resize(konst.size() + js.size() * sizeof(uint32_t));
            memcpy(konst.data() + g.ids_off, js.data(), js.size() * sizeof(uint32_t));
            g.small_idx = nsmall++;
            if (g.passes) max_tup = std::max<size_t>(max_tup, g.tuples);
            max_T1 = std::max<size_t>(max_T1, g.T1);
            max_T2 = std::max<size_t>(max_T2, g.T2max);
            max_H1 = std::max<size_t>(max_H1, (size_t)g.F1 * g.T1);
            max_H2 = std::max<size_t>(max_H2, (size_t)g.F2 * g.T2max);
            max_L1 = std::max<size_t>(max_L1, g.L1);
            max_L2 = std::max<size_t>(max_L2, g.L2);
            max_s2 = std::max<size_t>(max_s2, g.passes == 2 ? 2ull * g.nj * g.F1 : 0);
            max_NP = std::max<size_t>(max_NP, g.NP);
            P->groups.push_back(g);
        }
        if (!pht.empty()) {
            // one table of nextpow2(|R| / 2) buckets per join
            // (no_partitioning_join.h:555-556), all in one bucket array
            Group g{};
            g.kind = kGroupPht;
            g.nj = g.njobs = (uint32_t)pht.size();
            // table layout of each join: blocks of at most 2^kPhtPlaceHBits
            // buckets, one per value of the buckets' low pbits bits
            auto lg_of = [&](uint32_t i) {
                return ceil_log2(std::max<uint64_t>(1, items[i].nR / 2));
            };
            auto pbits_of = [&](uint32_t i) {
                const uint32_t lg = lg_of(i);
                return lg > kPhtPlaceHBits ? lg - kPhtPlaceHBits : 0u;
            };
            std::stable_sort(pht.begin(), pht.end(),
                             [&](uint32_t a, uint32_t b) { return pbits_of(a) < pbits_of(b); });
            std::vector<PhtJob> jobs(pht.size());
            for (uint32_t j = 0; j < g.njobs; j++) {
                const tbj_batch_item& it = items[pht[j]];
                PhtJob& jb = jobs[j];
                jb.r_off = it.r_off;
                jb.s_off = it.s_off;
                jb.nR = (uint32_t)it.nR;
                jb.nS = (uint32_t)it.nS;
                const uint32_t lg = lg_of(pht[j]);
                jb.mask = (uint32_t)((1ull << lg) - 1);
                jb.pbits = pbits_of(pht[j]);
                jb.hbits = lg - jb.pbits;
                jb.tab = g.nbuckets;
                jb.id = pht[j];
                jb.sc0 = g.nsc;
                g.nbuckets += 1ull << lg;
                // no R tuple, no match: S is not probed
                if (it.nR) g.nsc += (uint32_t)((it.nS + kPhtProbeChunk - 1) / kPhtProbeChunk);
                g.npool += it.nR;
                g.tuples += it.nR + it.nS;
            }
            // parts: runs of joins with the same pbits; those with pbits > 0
            // get the R segments of their partition pass
            g.part0 = (uint32_t)P->parts.size();
            for (uint32_t j0 = 0; j0 < g.njobs;) {
                uint32_t j1 = j0 + 1;
                while (j1 < g.njobs && jobs[j1].pbits == jobs[j0].pbits) j1++;
                PhtPart pp{};
                pp.pbits = jobs[j0].pbits;
                pp.j0 = j0;
                pp.j1 = j1;
                pp.base_off = wlo;
                if (pp.pbits) {
                    const uint32_t F = 1u << pp.pbits;
                    std::vector<SegDesc> segs(j1 - j0);
                    uint32_t tiles = 0;
                    uint64_t groups = 0, tup = 0;
                    for (uint32_t j = j0; j < j1; j++) {
                        SegDesc& d = segs[j - j0];
                        d.in_off = (uint32_t)(jobs[j].r_off - wlo);
                        d.n = jobs[j].nR;
                        d.ntiles = (d.n + kPhtTile - 1) / kPhtTile;
                        d.first_tile = tiles;
                        d.ngroups = seg_groups(tiles, d.ntiles);
                        d.aux = 0;
                        d.hist_base = groups * F;
                        tiles += d.ntiles;
                        groups += d.ngroups;
                        tup += d.n;
                    }
                    pp.T1 = tiles;
                    pp.L1 = (uint64_t)F * groups + 1;
                    align(64);
                    pp.segs_off = konst.size();
                    konst.resize(konst.size() + segs.size() * sizeof(SegDesc));
                    memcpy(konst.data() + pp.segs_off, segs.data(), segs.size() * sizeof(SegDesc));
                    pp.small_idx = nsmall++;
                    max_tup = std::max<size_t>(max_tup, tup);
                    max_T1 = std::max<size_t>(max_T1, pp.T1);
                    max_H1 = std::max<size_t>(max_H1, (size_t)F * pp.T1);
                    max_L1 = std::max<size_t>(max_L1, pp.L1);
                }
                P->parts.push_back(pp);
                j0 = j1;
            }
            g.part1 = (uint32_t)P->parts.size();
            align(64);
            g.jobs_off = konst.size();
            konst.resize(konst.size() + jobs.size() * sizeof(PhtJob));
            memcpy(konst.data() + g.jobs_off, jobs.data(), jobs.size() * sizeof(PhtJob));
            g.small_idx = nsmall++;
            max_bk = std::max<size_t>(max_bk, g.nbuckets);
            max_pool = std::max<size_t>(max_pool, g.npool);
            P->groups.push_back(g);
        }
        w0 = w1;
    }
    // ---- device buffers (owned by the plan so that a captured graph stays valid)
    TBJ_TRY(P->c1.get(sizeof(uint2) * max_tup));
    TBJ_TRY(P->c2.get(sizeof(uint2) * max_tup));
    TBJ_TRY(P->tiles1.get(sizeof(TileDesc) * max_T1));
    TBJ_TRY(P->hist1.get(sizeof(uint16_t) * max_H1));
    TBJ_TRY(P->offs1.get(sizeof(uint32_t) * max_H1));
    TBJ_TRY(P->gs1.get(sizeof(uint32_t) * max_L1));
    TBJ_TRY(P->tiles2.get(sizeof(TileDesc) * max_T2));
    TBJ_TRY(P->hist2.get(sizeof(uint16_t) * max_H2));
    TBJ_TRY(P->offs2.get(sizeof(uint32_t) * max_H2));
    TBJ_TRY(P->gs2.get(sizeof(uint32_t) * max_L2));
    TBJ_TRY(P->segs2.get(sizeof(SegDesc) * max_s2));
    TBJ_TRY(P->bounds.get(sizeof(uint2) * 2 * max_NP));
    TBJ_TRY(P->task_off.get(sizeof(uint32_t) * (max_NP + 1)));
    TBJ_TRY(P->partial.get(sizeof(uint32_t) *
                           scan_partials(std::max({max_L1, max_L2, max_NP + 1}))));
    TBJ_TRY(P->konst.get(konst.size()));
    TBJ_TRY(P->small.get(sizeof(uint32_t) * 8 * std::max<size_t>(1, nsmall)));
    TBJ_TRY(P->total.get(sizeof(unsigned long long)));
    TBJ_TRY(P->ptab.get(32 * max_bk));
    TBJ_TRY(P->ppool.get(16 * max_pool));
    if (!konst.empty())
        TBJ_HIP(hipMemcpy(P->konst.p, konst.data(), konst.size(), hipMemcpyHostToDevice));
    std::vector<uint32_t> sm(8 * std::max<size_t>(1, nsmall), 0);
    for (const Group& g : P->groups)
        if (g.kind == kGroupRadix) sm[8 * g.small_idx] = g.T1;
    for (const PhtPart& pp : P->parts)
        if (pp.pbits) sm[8 * pp.small_idx] = pp.T1;
    TBJ_HIP(hipMemcpy(P->small.p, sm.data(), sm.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    TBJ_HIP(hipEventCreate(&P->e0));
    TBJ_HIP(hipEventCreate(&P->e1));
    P->grid = (uint32_t)tbj::dev().num_cus * 4;
    P->pht_grid = (uint32_t)tbj::dev().num_cus * 8;
    return TBJ_OK;
}

}  // namespace

extern "C" {

int tbj_batch_create(int algo, const tbj_tuple* arena, const tbj_batch_item* items, uint32_t n,
                     uint64_t* counts, const tbj_batch_tuning* tuning, tbj_batch_plan** out) {
    if (!out) return TBJ_ERR_INVALID;
    *out = nullptr;
    if (algo != TBJ_RHO && algo != TBJ_PHT && algo != TBJ_BATCH_PER_ITEM) return TBJ_ERR_UNSUPPORTED;
    if (n == 0 || !items || !arena || !counts) return TBJ_ERR_INVALID;
    int s = tbj_init(tbj::dev().ready ? tbj::dev().device : 0);
    if (s != TBJ_OK) return s;
    tbj_batch_plan* P = new tbj_batch_plan();
    P->algo = algo;
    P->arena = arena;
    P->n = n;
    P->counts = (unsigned long long*)counts;
    if (tuning) P->tune = *tuning;
    s = plan_build(P, items);
    if (s != TBJ_OK) {
        delete P;
        return s;
    }
    *out = P;
    return TBJ_OK;
}

int tbj_batch_capture(tbj_batch_plan* P, void* stream) {
    if (!P) return TBJ_ERR_INVALID;
    // the legacy default stream cannot be captured: record on the library's own
    hipStream_t st = stream ? (hipStream_t)stream : tbj::dev().own_stream;
    if (P->exec) {
        TBJ_HIP(hipGraphExecDestroy(P->exec));
        P->exec = nullptr;
    }
    if (P->graph) {
        TBJ_HIP(hipGraphDestroy(P->graph));
        P->graph = nullptr;
    }
    TBJ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const int s = P->launch(st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (s != TBJ_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    TBJ_HIP(e);
    P->graph = g;
    TBJ_HIP(hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0));
    return TBJ_OK;
}

int tbj_batch_run(tbj_batch_plan* P, void* stream, tbj_batch_result* res) {
    if (!P) return TBJ_ERR_INVALID;
    hipStream_t st = (hipStream_t)stream;
    TBJ_HIP(hipEventRecord(P->e0, st));
    if (P->exec) {
        TBJ_HIP(hipGraphLaunch(P->exec, st));
    } else {
        TBJ_TRY(P->launch(st));
    }
    TBJ_HIP(hipEventRecord(P->e1, st));
    unsigned long long total = 0;
    TBJ_HIP(hipMemcpyAsync(&total, P->total.p, sizeof(total), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    if (res) {
        memset(res, 0, sizeof(*res));
        res->matches = total;
        res->tuples = P->tuples;
        res->waves = P->waves;
        res->groups = (uint32_t)P->groups.size();
        TBJ_HIP(hipEventElapsedTime(&res->ms, P->e0, P->e1));
    }
    return TBJ_OK;
}

void tbj_batch_destroy(tbj_batch_plan* P) { delete P; }

namespace {

// Median device time of `reps` runs (after one warm run) of the batch
// planned with tuning t.
int time_tuning(int algo, const tbj_tuple* arena, const tbj_batch_item* items, uint32_t n,
                uint64_t* counts, int reps, const tbj_batch_tuning& t, float* ms_out) {
    tbj_batch_plan* P = nullptr;
    TBJ_TRY(tbj_batch_create(algo, arena, items, n, counts, &t, &P));
    std::vector<float> ms;
    int s = TBJ_OK;
    for (int r = 0; r <= reps && s == TBJ_OK; r++) {
        tbj_batch_result res;
        s = tbj_batch_run(P, nullptr, &res);
        if (r) ms.push_back(res.ms);
    }
    tbj_batch_destroy(P);
    TBJ_TRY(s);
    std::sort(ms.begin(), ms.end());
    *ms_out = ms[ms.size() / 2];
    return TBJ_OK;
}

}  // namespace

int tbj_batch_autotune(int algo, const tbj_tuple* arena, const tbj_batch_item* items, uint32_t n,
                       uint64_t* counts, int reps, tbj_batch_tuning* best, float* best_ms) {
    if (!best) return TBJ_ERR_INVALID;
    if (reps < 1) reps = 1;
    float bms = 1e30f;
    tbj_batch_tuning bt{};
    bt.wave_tuples = best->wave_tuples;
    auto consider = [&](const tbj_batch_tuning& t) -> int {
        float ms = 0;
        TBJ_TRY(time_tuning(algo, arena, items, n, counts, reps, t, &ms));
        if (ms < bms) {
            bms = ms;
            bt = t;
        }
        return TBJ_OK;
    };
    // stage 1: the partition tile shape at the default plan
    const uint32_t tiles[] = {tbj::kTile / 2, tbj::kTile, tbj::kMaxTile};
    for (uint32_t tile : tiles) {
        tbj_batch_tuning t{};
        t.target_r = tbj::kJoinCap / 2;
        t.one_pass_bits = 8;
        t.s_chunk = (uint32_t)tbj::kJoinSChunk * tbj::kJoinSTaskChunks;
        t.wave_tuples = best->wave_tuples;
        t.tile = tile;
        TBJ_TRY(consider(t));
    }
    // stage 2: radix bits (target_r), pass count (one_pass_bits) and the
    // join's task shape (s_chunk) at that tile
    const uint32_t tile = bt.tile;
    const uint32_t targets[] = {1024, 2048};
    const uint32_t onepass[] = {6, 8, 10};
    const uint32_t chunks[] = {8192, 32768};
    for (uint32_t tr : targets)
        for (uint32_t op : onepass)
            for (uint32_t ch : chunks) {
                tbj_batch_tuning t{};
                t.target_r = tr;
                t.one_pass_bits = op;
                t.s_chunk = ch;
                t.wave_tuples = best->wave_tuples;
                t.tile = tile;
                TBJ_TRY(consider(t));
            }
    *best = bt;
    if (best_ms) *best_ms = bms;
    return TBJ_OK;
}

// ---------------------------------------------------------------------------
// Batch generator (host side)

uint64_t tbj_gen_batch_specs(uint64_t base_seed, uint32_t n, tbj_batch_spec* specs) {
    uint64_t off = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t x = base_seed ^ i;
        const uint64_t s = splitmix(x);
        uint64_t y = s;
        const double u = (double)(splitmix(y) >> 11) * (1.0 / 9007199254740992.0);
        const uint64_t nR = (uint64_t)floor(pow(2.0, 10.0 + 12.0 * u));
        const uint64_t k = 1 + splitmix(y) % 16;
        const double z = 1.5 * (double)(splitmix(y) >> 11) * (1.0 / 9007199254740992.0);
        const int algo = (splitmix(y) >> 63) ? TBJ_PHT : TBJ_RHO;
        tbj_batch_spec& sp = specs[i];
        sp.seed = s;
        sp.nR = nR;
        sp.nS = k * nR;
        sp.z = z;
        sp.algo = algo;
        sp.reserved = 0;
        sp.r_off = off;
        sp.s_off = off + nR;
        off += nR + sp.nS;
    }
    return off;
}

static double zipf_H(double x, double s) {
    const double lx = log(x), t = (1.0 - s) * lx;
    const double h2 = fabs(t) > 1e-8 ? expm1(t) / t : 1.0 + t * 0.5 * (1.0 + t / 3.0 * (1.0 + 0.25 * t));
    return h2 * lx;
}

static double zipf_Hinv(double x, double s) {
    double t = x * (1.0 - s);
    if (t < -1.0) t = -1.0;
    const double h1 = fabs(t) > 1e-8 ? log1p(t) / t : 1.0 - t * (0.5 - t * (1.0 / 3.0 - 0.25 * t));
    return exp(h1 * x);
}

int tbj_gen_batch_device(tbj_tuple* arena, const tbj_batch_spec* specs, uint32_t n,
                         void* stream) {
    if ((!arena && n) || (!specs && n)) return TBJ_ERR_INVALID;
    if (n == 0) return TBJ_OK;
    int s = tbj_init(tbj::dev().ready ? tbj::dev().device : 0);
    if (s != TBJ_OK) return s;
    std::vector<tbj::GenJob> jobs(n);
    uint64_t start = 0;
    for (uint32_t i = 0; i < n; i++) {
        const tbj_batch_spec& sp = specs[i];
        if (sp.nR == 0 || sp.nR >= (1ull << 32) || sp.nS >= (1ull << 32)) return TBJ_ERR_INVALID;
        tbj::GenJob& j = jobs[i];
        j.start = start;
        j.seed = sp.seed;
        j.nR = sp.nR;
        j.nS = sp.nS;
        j.r_off = sp.r_off;
        j.s_off = sp.s_off;
        j.z = sp.z;
        if (sp.z > 0) {
            j.hx1 = zipf_H(1.5, sp.z) - 1.0;
            j.hn = zipf_H((double)sp.nR + 0.5, sp.z);
            j.sp = 2.0 - zipf_Hinv(zipf_H(2.5, sp.z) - exp(-sp.z * log(2.0)), sp.z);
        }
        const uint32_t bits = ceil_log2(sp.nR);
        j.fbits = std::max<uint32_t>(1, (bits + 1) / 2);
        start += sp.nR + sp.nS;
    }
    hipStream_t st = (hipStream_t)stream;
    DevBuf dj;
    TBJ_TRY(dj.get(sizeof(tbj::GenJob) * n));
    TBJ_HIP(hipMemcpyAsync(dj.p, jobs.data(), sizeof(tbj::GenJob) * n, hipMemcpyHostToDevice, st));
    tbj::batch_generate(st, (uint2*)arena, (const tbj::GenJob*)dj.p, n, start);
    TBJ_HIP(hipGetLastError());
    TBJ_HIP(hipStreamSynchronize(st));
    return TBJ_OK;
}

}  // extern "C"
