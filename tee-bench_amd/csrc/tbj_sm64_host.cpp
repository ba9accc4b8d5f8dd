// Orchestration of the 64-bit-key sort-merge join (tbj_join64): partition
// passes, task lists, count (with the R chunk sort), chunk-offset scan, run
// offsets, emit.
//
// Phases reported in tbj_result.phase_ms: hist1, scan1, scatter1, size2,
// scatter2 as for the 8-byte path; "tasks" = task lists (+ R sort and merge
// passes when an R partition exceeds one chunk); "join" = count, offsets and
// emit.
#include <string.h>

#include <algorithm>
#include <vector>

#include "tbj_kernels.h"
#include "tbj_runtime.h"
#include "tbj_sm64.h"

namespace tbj {

namespace {

enum Phase { kPhHist1 = 0, kPhScan1, kPhScatter1, kPhSize2, kPhScatter2, kPhTasks, kPhJoin };

struct Small64 {  // device scalars, one allocation
    uint32_t ntiles1, ntiles2, nrtasks, njtasks;
    uint32_t max_r, total1, total2;
    uint32_t lin;  // the learned fit gave the linear model
    uint32_t overflow;      // a sample-sized region was too small
    uint32_t region_total;  // sample-sized first pass: region capacity in all
    uint32_t s_out;  // sampled S keys outside the learned model's key range
    uint32_t ntiles2_pad[1];
    unsigned long long range[2];     // exact R key range (learned partitioner)
    unsigned long long counters[2];  // matches, checksum
    unsigned long long volume;       // pairs over all chunks
};

uint32_t ceil_log2(uint64_t v) {
    uint32_t b = 0;
    while (((uint64_t)1 << b) < v) b++;
    return b;
}

Bin64 bin64(uint32_t kind, uint32_t B, uint32_t shift, uint32_t fanout, const Model64* m,
            const unsigned long long* range) {
    Bin64 b{};
    b.kind = kind;
    b.bits = B;
    b.pass_shift = shift;
    b.pass_mask = fanout - 1;
    b.model = m;
    b.range = range;
    return b;
}

}  // namespace

// Joins of at least this many tuples (|R| + |S|) size their first pass from
// a sample when tbj_options.sampled_regions is 0 (auto).
constexpr uint64_t kSampledMin64 = 1ull << 24;

// One run; *overflow_out: a sample-sized region was too small (the run then
// stops before the join, to be redone with exact histograms).
static int run_sm64_once(Context& C, int algo, const tbj_tuple64* R, uint64_t nR64,
                  const tbj_tuple64* S, uint64_t nS64, tbj_pair64* pairs, uint64_t pairs_cap,
                  const tbj_options& opt, bool allow_sampled, tbj_result* res, hipStream_t st,
                  bool* overflow_out) {
    *overflow_out = false;
    if (nR64 + nS64 >= (1ull << 32) - (1ull << 20)) return TBJ_ERR_TOO_LARGE;
    bool learned = opt.partitioner == TBJ_PART_LEARNED ||
                   (opt.partitioner == TBJ_PART_AUTO && algo == TBJ_PSM);
    const bool mat = opt.materialize && pairs && pairs_cap;
    const uint32_t nR = (uint32_t)nR64, nS = (uint32_t)nS64;
    if (nR == 0 || nS == 0) {
        res->partitioner_used = learned ? TBJ_PART_LEARNED : TBJ_PART_RADIX;
        return TBJ_OK;
    }
    Workspace& W = C.ws;
    PhaseTimer pt(C, st, opt.profile != 0);

    // plan: average R partition at half a chunk, up to two passes of <= 10 bits
    uint32_t B;
    if (opt.radix_bits > 0) B = std::min<uint32_t>(opt.radix_bits, 2 * kSm64PassBits);
    else B = nR <= kSm64Cap / 2 ? 0 : std::min<uint32_t>(ceil_log2((nR + kSm64Cap / 2 - 1) / (kSm64Cap / 2)), 2 * kSm64PassBits);
    uint32_t passes = B == 0 ? 0 : (B <= (uint32_t)kSm64PassBits && opt.passes != 2) ? 1 : 2;
    if (B > (uint32_t)kSm64PassBits) passes = 2;
    const uint32_t b1 = passes == 2 ? (B + 1) / 2 : B, b2 = passes == 2 ? B - b1 : 0;
    const uint32_t F1 = 1u << b1, F2 = 1u << b2;
    const uint32_t nparts = passes ? F1 * F2 : 1;
    if (passes == 0) learned = false;
    res->radix_bits = (int)B;
    res->passes = (int)passes;
    res->num_partitions = nparts;
    res->partitioner_used = learned ? TBJ_PART_LEARNED : TBJ_PART_RADIX;

    Small64* sm = nullptr;
    TBJ_TRY(W.get("s64_small", sizeof(Small64), (void**)&sm));
    const uint64_t ntot = (uint64_t)nR + nS;
    // sample-sized first pass (two-pass plans): its regions leave slack, so C1
    // takes the largest sum of capacities the rule allows (sum of 16 x counts
    // <= |R| + |S| plus the sampled blocks' rounding; sum of sqrt <=
    // sqrt(n x sum))
    const uint32_t tR = (nR + kSm64Span - 1) / kSm64Span, tS = (nS + kSm64Span - 1) / kSm64Span;
    const uint32_t nrowR = 8 * rows_per_eighth(tR), nrowS = 8 * rows_per_eighth(tS);
    const uint32_t np1 = F1 * (nrowR + nrowS);
    const double e1 = (double)ntot + 256.0 * (nrowR + nrowS);
    const double cap1 = 1.1 * e1 + 48.0 * sqrt((double)np1 * e1) + 640.0 * np1 + 1;
    const bool sampled = allow_sampled && passes == 2 && cap1 < 4294967295.0 &&
                         (opt.sampled_regions > 0 ||
                          (opt.sampled_regions == 0 && ntot >= kSampledMin64));
    ulonglong2 *C1 = nullptr, *C2 = nullptr;
    TBJ_TRY(W.get("c1", sizeof(ulonglong2) * (sampled ? (uint64_t)cap1 : ntot), (void**)&C1));
    TBJ_TRY(W.get("c2", sizeof(ulonglong2) * ntot, (void**)&C2));
    uint2* bounds = nullptr;
    TBJ_TRY(W.get("bounds", sizeof(uint2) * 2 * nparts, (void**)&bounds));

    TBJ_HIP(hipEventRecord(C.ev[0], st));
    TBJ_TRY(pt.begin());
    TBJ_HIP(hipMemsetAsync(sm, 0, sizeof(Small64), st));
    launch_fill_u32(st, (uint32_t*)&sm->range[0], 2, 0xFFFFFFFFu);

    const ulonglong2* Rin = (const ulonglong2*)R;
    const ulonglong2* Sin = (const ulonglong2*)S;
    ulonglong2* Pb = nullptr;   // partitioned R then S
    ulonglong2* Tmp = nullptr;  // free buffer of ntot tuples (merge passes)
    uint32_t* partial = nullptr;

    if (passes == 0) {
        TBJ_HIP(hipMemcpyAsync(C1, Rin, sizeof(ulonglong2) * nR, hipMemcpyDeviceToDevice, st));
        TBJ_HIP(hipMemcpyAsync(C1 + nR, Sin, sizeof(ulonglong2) * nS, hipMemcpyDeviceToDevice, st));
        sm64_bounds(st, nullptr, nullptr, 1, 1, 0, nR, nS, bounds);
        Pb = C1;
        Tmp = C2;
        TBJ_TRY(W.get("partial", sizeof(uint32_t) * scan_partials(nparts + 1), (void**)&partial));
    } else {
        const uint32_t T1 = tR + tS;
        const uint32_t T2max = passes == 2 ? (uint32_t)(ntot / kSm64Span) + 2 * F1 + 2 : 0;
        // group-sum lengths: R's and S's groups; second pass at most
        // tiles / kGroupTiles plus two per segment
        const uint64_t L1 = (uint64_t)F1 * (seg_groups(0, tR) + seg_groups(tR, tS)) + 1;
        const uint64_t L2 =
            passes == 2 ? (uint64_t)F2 * (T2max / kGroupTiles + 2ull * (2 * F1) + 1) + 1 : 0;
        TileDesc *tiles1 = nullptr, *tiles2 = nullptr;
        SegDesc *segs1 = nullptr, *segs2 = nullptr;
        uint16_t *hist1 = nullptr, *hist2 = nullptr;
        uint32_t *offs1 = nullptr, *offs2 = nullptr, *gs1 = nullptr, *gs2 = nullptr;
        TBJ_TRY(W.get("tiles1", sizeof(TileDesc) * T1, (void**)&tiles1));
        TBJ_TRY(W.get("segs1", sizeof(SegDesc) * 2, (void**)&segs1));
        TBJ_TRY(W.get("hist1", sizeof(uint16_t) * F1 * T1, (void**)&hist1));
        TBJ_TRY(W.get("offs1", sizeof(uint32_t) * F1 * T1, (void**)&offs1));
        TBJ_TRY(W.get("gs1", sizeof(uint32_t) * L1, (void**)&gs1));
        TBJ_TRY(W.get("partial", sizeof(uint32_t) * scan_partials(std::max({L1, L2, (uint64_t)nparts + 1})),
                      (void**)&partial));
        Model64* model = nullptr;
        Bin64 s1r, s1s, s2;
        if (learned) {
            TBJ_TRY(W.get("s64_model", sizeof(Model64), (void**)&model));
            uint32_t m = 2;
            while (m * 2 <= std::min<uint32_t>(nR, kSm64FitMax)) m *= 2;
            if (opt.learned_samples > 0)
                while (m > (uint32_t)opt.learned_samples && m > 2) m >>= 1;
            uint64_t* sample = nullptr;
            TBJ_TRY(W.get("s64_sample", sizeof(uint64_t) * m, (void**)&sample));
            const uint32_t nseg = std::max<uint32_t>(1, std::min<uint32_t>(kSm64Knots, m / 64));
            sm64_fit(st, Rin, nR, m, nseg, B, sample, model, &sm->lin);
            s1r = bin64(kBinLearned, B, b2, F1, model, nullptr);
            s1s = bin64(kBinLearned, B, b2, F1, model, sm->range);
            s2 = bin64(kBinLearned, B, 0, F2, model, nullptr);
        } else {
            s1r = s1s = bin64(kBinRadix, B, 0, F1, nullptr, nullptr);
            s2 = bin64(kBinRadix, B, b1, F2, nullptr, nullptr);
        }
        if (sampled) {
            // ---- pass 1 from a sample: regions per (row, bin), runs claimed
            uint32_t *cnt1 = nullptr, *starts1 = nullptr, *cursors1 = nullptr, *first2 = nullptr,
                     *npart = nullptr;
            const uint32_t T2s = (uint32_t)(ntot / kSm64Span) + np1 + 1;
            const uint64_t L2s = (uint64_t)F2 * (T2s / kGroupTiles + 2ull * (2 * F1) + 1) + 1;
            TBJ_TRY(W.get("s64_cnt1", sizeof(uint32_t) * np1, (void**)&cnt1));
            TBJ_TRY(W.get("s64_starts1", sizeof(uint32_t) * (np1 + 1), (void**)&starts1));
            TBJ_TRY(W.get("s64_cursors1", sizeof(uint32_t) * np1, (void**)&cursors1));
            TBJ_TRY(W.get("s64_first2", sizeof(uint32_t) * (np1 + 1), (void**)&first2));
            TBJ_TRY(W.get("s64_npart", sizeof(uint32_t) * 2 * F1, (void**)&npart));
            TBJ_TRY(W.get("tiles2", sizeof(TileDesc) * T2s, (void**)&tiles2));
            TBJ_TRY(W.get("segs2", sizeof(SegDesc) * 2 * F1, (void**)&segs2));
            TBJ_TRY(W.get("hist2", sizeof(uint16_t) * F2 * T2s, (void**)&hist2));
            TBJ_TRY(W.get("offs2", sizeof(uint32_t) * F2 * T2s, (void**)&offs2));
            TBJ_TRY(W.get("gs2", sizeof(uint32_t) * L2s, (void**)&gs2));
            TBJ_TRY(W.get("partial2",
                          sizeof(uint32_t) * scan_partials(std::max<uint64_t>(
                                                 std::max<uint64_t>(np1 + 1, L2s), nparts + 1)),
                          (void**)&partial));
            launch_zero_multi(st, ZeroList{{cnt1, cursors1}, {np1, np1}, 2});
            sm64_make_tiles(st, tiles1, segs1, nR, nS, F1, &sm->ntiles1, true);
            Bin64 s1s_sample = s1s;
            s1s_sample.range = nullptr;  // R's range comes with R's scatter
            sm64_sample_rows(st, Rin, nR, Sin, nS, nrowR, nrowS, s1r, s1s_sample, F1, cnt1,
                             learned ? &sm->s_out : nullptr);
            TBJ_TRY(pt.mark(kPhHist1));
            launch_row_caps(st, cnt1, np1, starts1);
            launch_scan(st, starts1, (uint64_t)np1 + 1, partial, &sm->region_total);
            TBJ_TRY(pt.mark(kPhScan1));
            // Learned: R's scatter measures R's exact range and S drops the keys
            // outside it -- when the sample saw more than 1/64 of S outside the
            // model's key range (otherwise hardly any S key could be dropped,
            // and neither pays; the join's result is the same either way).
            const uint32_t gate_min = nS / 1024;
            Sm64Claims c1r{starts1, cursors1, &sm->overflow, 0, F1, learned ? sm->range : nullptr};
            c1r.gate = &sm->s_out;
            c1r.gate_min = gate_min;
            s1s.gate = &sm->s_out;
            s1s.gate_min = gate_min;
            const Sm64Claims c1s{starts1, cursors1, &sm->overflow, F1 * nrowR, F1, nullptr};
            sm64_scatter(st, tR, Rin, C1, tiles1, &sm->ntiles1, 0, nullptr, s1r, F1, &c1r);
            sm64_scatter(st, tS, Sin, C1, tiles1, &sm->ntiles1, tR, nullptr, s1s, F1, &c1s);
            TBJ_TRY(pt.mark(kPhScatter1));
            // ---- pass 2 (exact) over the filled regions, one segment per
            // first-pass partition
            launch_part_sizes(st, starts1, cursors1, F1, nrowR, nrowS, npart, &sm->total1);
            launch_region_tiles(st, starts1, cursors1, np1, F1, nrowR, nrowS, first2, kSm64Span);
            launch_scan(st, first2, (uint64_t)np1 + 1, partial, &sm->ntiles2);
            launch_region_tile_descs(st, starts1, cursors1, first2, np1, F1, nrowR, nrowS, tiles2,
                                     kSm64Span, true);
            sm64_part_segments(st, first2, F1, F2, nrowR, nrowS, segs2);
            sm64_hist(st, T2s, C1, tiles2, &sm->ntiles2, 0, hist2, s2, F2, nullptr);
            launch_tile_offsets(st, T2s, hist2, offs2, tiles2, segs2, &sm->ntiles2, F2, gs2, L2s,
                                partial, &sm->total2);
            TBJ_TRY(pt.mark(kPhSize2));
            sm64_scatter(st, T2s, C1, C2, tiles2, &sm->ntiles2, 0, offs2, s2, F2);
            sm64_bounds(st, gs2, segs2, F1, F2, 2, nR, nS, bounds);
            TBJ_TRY(pt.mark(kPhScatter2));
            Pb = C2;
            Tmp = C1;
        } else {
        // ---- pass 1
        sm64_make_tiles(st, tiles1, segs1, nR, nS, F1, &sm->ntiles1);
        sm64_hist(st, tR, Rin, tiles1, &sm->ntiles1, 0, hist1, s1r, F1,
                  learned ? sm->range : nullptr);
        sm64_hist(st, tS, Sin, tiles1, &sm->ntiles1, tR, hist1, s1s, F1, nullptr);
        launch_fill_u32(st, gs1 + (L1 - 1), 1, 0);
        TBJ_TRY(pt.mark(kPhHist1));
        launch_tile_offsets(st, T1, hist1, offs1, tiles1, segs1, &sm->ntiles1, F1, gs1, L1,
                            partial, &sm->total1);
        TBJ_TRY(pt.mark(kPhScan1));
        sm64_scatter(st, tR, Rin, C1, tiles1, &sm->ntiles1, 0, offs1, s1r, F1);
        sm64_scatter(st, tS, Sin, C1, tiles1, &sm->ntiles1, tR, offs1, s1s, F1);
        TBJ_TRY(pt.mark(kPhScatter1));
        if (passes == 1) {
            sm64_bounds(st, gs1, segs1, 1, F1, 1, nR, nS, bounds);
            Pb = C1;
            Tmp = C2;
        } else {
            // ---- pass 2: one segment per (relation, first-pass partition)
            TBJ_TRY(W.get("tiles2", sizeof(TileDesc) * T2max, (void**)&tiles2));
            TBJ_TRY(W.get("segs2", sizeof(SegDesc) * 2 * F1, (void**)&segs2));
            TBJ_TRY(W.get("hist2", sizeof(uint16_t) * F2 * T2max, (void**)&hist2));
            TBJ_TRY(W.get("offs2", sizeof(uint32_t) * F2 * T2max, (void**)&offs2));
            TBJ_TRY(W.get("gs2", sizeof(uint32_t) * L2, (void**)&gs2));
            sm64_segments(st, gs1, segs1, F1, F2, segs2, tiles2, &sm->ntiles2, T2max);
            sm64_hist(st, T2max, C1, tiles2, &sm->ntiles2, 0, hist2, s2, F2, nullptr);
            // group sums past the segments' own (never written) only follow
            // every partition bound in the scan, so they need no clearing
            launch_tile_offsets(st, T2max, hist2, offs2, tiles2, segs2, &sm->ntiles2, F2, gs2, L2,
                                partial, &sm->total2);
            TBJ_TRY(pt.mark(kPhSize2));
            sm64_scatter(st, T2max, C1, C2, tiles2, &sm->ntiles2, 0, offs2, s2, F2);
            sm64_bounds(st, gs2, segs2, F1, F2, 2, nR, nS, bounds);
            TBJ_TRY(pt.mark(kPhScatter2));
            Pb = C2;
            Tmp = C1;
        }
        }  // exact passes
    }
    TBJ_HIP(hipEventRecord(C.ev[1], st));

    // ---- task lists
    uint32_t *rcnt = nullptr, *jcnt = nullptr;
    TBJ_TRY(W.get("s64_rcnt", sizeof(uint32_t) * (nparts + 1), (void**)&rcnt));
    TBJ_TRY(W.get("s64_jcnt", sizeof(uint32_t) * (nparts + 1), (void**)&jcnt));
    sm64_task_counts(st, bounds, nparts, rcnt, jcnt, &sm->max_r);
    launch_scan(st, rcnt, nparts + 1, partial, &sm->nrtasks);
    launch_scan(st, jcnt, nparts + 1, partial, &sm->njtasks);
    Small64 h;
    TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small64), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    if (h.overflow) {  // redo with exact histograms before any join work
        *overflow_out = true;
        return TBJ_OK;
    }
    res->max_r_partition = h.max_r;
    if (learned) res->s_dropped = (uint64_t)nS - ((uint64_t)h.total1 - nR);
    res->learned_model = learned ? (h.lin ? 2 : 1) : 0;
    SmTask *rtasks = nullptr, *jtasks = nullptr;
    TBJ_TRY(W.get("s64_rtasks", sizeof(SmTask) * std::max<uint32_t>(1, h.nrtasks), (void**)&rtasks));
    TBJ_TRY(W.get("s64_jtasks", sizeof(SmTask) * std::max<uint32_t>(1, h.njtasks), (void**)&jtasks));
    sm64_fill_tasks(st, bounds, nparts, rcnt, jcnt, h.nrtasks, h.njtasks, rtasks, jtasks);

    // ---- R partitions above one chunk: sort the chunks, then merge passes.
    // Otherwise the chunks probed by several tasks are sorted into Tmp here,
    // and the count kernel sorts each other chunk in LDS and writes it to Tmp.
    const bool presorted = h.max_r > (uint32_t)kSm64Cap;
    if (presorted) {
        const uint32_t sort_grid = std::max<uint32_t>(1, std::min<uint32_t>(h.nrtasks, C.num_cus * 3));
        sm64_sort_r(st, sort_grid, Pb, rtasks, h.nrtasks);
        std::vector<uint2> hb(nparts);
        TBJ_HIP(hipMemcpyAsync(hb.data(), bounds, sizeof(uint2) * nparts, hipMemcpyDeviceToHost, st));
        TBJ_HIP(hipStreamSynchronize(st));
        std::vector<MergeBlock> mb;
        uint32_t longest = 0;
        for (uint32_t p = 0; p < nparts; p++) {
            const uint32_t len = hb[p].y - hb[p].x;
            if (len <= (uint32_t)kSm64Cap) continue;
            longest = std::max(longest, len);
            for (uint32_t o = 0; o < len; o += kSm64MergeTile) mb.push_back({hb[p].x, len, o, 0});
        }
        uint32_t npass = 0;
        while (((uint64_t)kSm64Cap << npass) < longest) npass++;
        npass += npass & 1;  // even: the result ends in the partitioned buffer
        MergeBlock* dmb = nullptr;
        TBJ_TRY(W.get("s64_merge", sizeof(MergeBlock) * mb.size(), (void**)&dmb));
        TBJ_HIP(hipMemcpyAsync(dmb, mb.data(), sizeof(MergeBlock) * mb.size(),
                               hipMemcpyHostToDevice, st));
        for (uint32_t k = 0; k < npass; k++) {
            const ulonglong2* src = (k & 1) ? Tmp : Pb;
            ulonglong2* dst = (k & 1) ? Pb : Tmp;
            sm64_merge(st, src, dst, dmb, (uint32_t)mb.size(), (uint64_t)kSm64Cap << k);
        }
        // the host vector must outlive the copy
        TBJ_HIP(hipStreamSynchronize(st));
    } else {
        sm64_sort_shared(st, Pb, Tmp, rtasks, h.nrtasks);
    }
    const ulonglong2* Rsorted = presorted ? Pb : Tmp;
    TBJ_TRY(pt.mark(kPhTasks));

    // ---- count, offsets, emit
    const uint32_t jgrid = std::max<uint32_t>(1, std::min<uint32_t>(h.njtasks, C.num_cus * 4));
    unsigned long long *runs = nullptr, *chunk_vol = nullptr;
    if (mat) {
        TBJ_TRY(W.get("s64_runs", sizeof(unsigned long long) * nR, (void**)&runs));
        TBJ_TRY(W.get("s64_chunks", sizeof(unsigned long long) * std::max<uint32_t>(1, h.nrtasks),
                      (void**)&chunk_vol));
        sm64_clear_shared_runs(st, jtasks, h.njtasks, runs);
        sm64_fill_u64(st, chunk_vol, h.nrtasks, 0);
    }
    if (h.njtasks)
        sm64_count(st, jgrid, mat, presorted, Pb, Tmp, Pb, jtasks, &sm->njtasks, runs, chunk_vol,
                   sm->counters);
    if (mat && h.njtasks) {
        unsigned long long* vpart = nullptr;
        TBJ_TRY(W.get("s64_vpart", sizeof(unsigned long long) * sm64_scan_partials(h.nrtasks),
                      (void**)&vpart));
        sm64_scan_u64(st, chunk_vol, h.nrtasks, vpart, &sm->volume);
        const uint32_t ogrid = std::max<uint32_t>(1, std::min<uint32_t>(h.nrtasks, C.num_cus * 4));
        sm64_run_offsets(st, ogrid, Rsorted, rtasks, h.nrtasks, chunk_vol, &sm->volume, runs);
        sm64_emit(st, jgrid, Rsorted, Pb, jtasks, &sm->njtasks, runs, chunk_vol,
                  (ulonglong2*)pairs, pairs_cap, presorted);
    }
    TBJ_HIP(hipGetLastError());
    TBJ_TRY(pt.mark(kPhJoin));
    TBJ_HIP(hipEventRecord(C.ev[2], st));
    TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small64), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    TBJ_TRY(pt.finish(res));
    res->matches = h.counters[0];
    res->checksum = h.counters[1];
    res->pairs_written = mat ? std::min<uint64_t>(h.counters[0], pairs_cap) : 0;
    float ms = 0;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[0], C.ev[1]));
    res->ms_partition = ms;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[1], C.ev[2]));
    res->ms_join = ms;
    TBJ_HIP(hipEventElapsedTime(&ms, C.ev[0], C.ev[2]));
    res->ms_total = ms;
    return TBJ_OK;
}

int run_sm64(Context& C, int algo, const tbj_tuple64* R, uint64_t nR, const tbj_tuple64* S,
             uint64_t nS, tbj_pair64* pairs, uint64_t pairs_cap, const tbj_options& opt,
             tbj_result* res, hipStream_t st) {
    bool overflow = false;
    TBJ_TRY(run_sm64_once(C, algo, R, nR, S, nS, pairs, pairs_cap, opt, opt.sampled_regions >= 0,
                          res, st, &overflow));
    if (!overflow) return TBJ_OK;
    const int32_t learned_model = res->learned_model;
    memset(res, 0, sizeof(*res));
    TBJ_TRY(run_sm64_once(C, algo, R, nR, S, nS, pairs, pairs_cap, opt, false, res, st,
                          &overflow));
    res->redone_exact = 1;
    if (!res->learned_model) res->learned_model = learned_model;
    return TBJ_OK;
}

}  // namespace tbj
