// Orchestration of the no-partitioning joins: PHT / NPO_st and CHT.
#include <string.h>

#include <algorithm>

#include "tbj_kernels.h"
#include "tbj_pht.h"
#include "tbj_runtime.h"

namespace tbj {

namespace {

uint32_t next_pow2(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return (uint32_t)p;
}

struct Small {
    uint32_t pool_ctr, pad[3];
    ChtState cht;
    unsigned long long counters[4];
};

int finish(Context& C, hipStream_t st, Small* sm, bool mat, uint64_t cap, tbj_result* res) {
    Small h;
    TBJ_HIP(hipEventRecord(C.ev[2], st));
    TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    res->matches = h.counters[0];
    res->checksum = h.counters[1];
    res->pairs_written = mat ? std::min<uint64_t>(h.counters[2], cap) : 0;
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

// Relations at least this large build the PHT table in bulk when
// tbj_options.bulk_build is 0 (auto).
constexpr uint64_t kBulkMin = 1ull << 20;

// PHT (no_partitioning_join.h:543-623) and NPO_st (474-541): one shared
// table of nextpow2(|R| / 2) buckets (allocate_hashtable, 119-151), built by
// one atomic insert per R tuple or, in bulk, from R partitioned on the
// buckets' low bits (res->radix_bits and res->passes then report that
// partitioning), and probed by streaming S.
int run_pht(Context& C, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
            tbj_pair* pairs, uint64_t pairs_cap, const tbj_options& opt, tbj_result* res,
            hipStream_t st) {
    Workspace& W = C.ws;
    const bool mat = opt.materialize && pairs && pairs_cap;
    if (nR >= (1ull << 32) || nS >= (1ull << 40)) return TBJ_ERR_TOO_LARGE;
    Small* sm = nullptr;
    TBJ_TRY(W.get("pht_small", sizeof(Small), (void**)&sm));
    TBJ_HIP(hipEventRecord(C.ev[0], st));
    TBJ_HIP(hipMemsetAsync(sm, 0, sizeof(Small), st));
    res->partitioner_used = 0;
    if (nR == 0 || nS == 0) {
        TBJ_HIP(hipEventRecord(C.ev[1], st));
        return finish(C, st, sm, mat, pairs_cap, res);
    }
    const uint32_t nb = std::max<uint32_t>(1, next_pow2(nR / 2));
    uint32_t lg = 0;
    while ((1u << lg) < nb) lg++;
    const bool bulk = nR < (1ull << 31) &&
                      (opt.bulk_build > 0 || (opt.bulk_build == 0 && nR >= kBulkMin));
    void *tab = nullptr, *pool = nullptr;
    TBJ_TRY(W.get("pht_tab", pht_bucket_bytes() * nb, &tab));
    TBJ_TRY(W.get("pht_pool", pht_node_bytes() * (nR + 1), &pool));
    uint32_t pbits = 0, hbits = lg;
    if (bulk) {
        hbits = std::min<uint32_t>(lg, kPhtPlaceHBits);
        pbits = lg - hbits;
        const uint2* Rb = (const uint2*)R;
        const uint2* bounds = nullptr;
        uint32_t b1 = 0;
        if (pbits) {
            TBJ_TRY(partition_relation(C, R, (uint32_t)nR, pbits, 0, st, &Rb, &bounds, &b1));
        } else {
            // one block of buckets: all of R
            uint2* bd = nullptr;
            TBJ_TRY(W.get("pht_bound", sizeof(uint2) * 2, (void**)&bd));
            launch_part_bounds(st, nullptr, nullptr, 1, 1, 0, (uint32_t)nR, 0, bd);
            bounds = bd;
        }
        launch_pht_place(st, Rb, bounds, b1, pbits, hbits, tab, pool, &sm->pool_ctr);
        res->radix_bits = (int32_t)pbits;
        res->passes = pbits == 0 ? 0 : pbits > 8 ? 2 : 1;
    } else {
        launch_fill_u4(st, tab, pht_bucket_bytes() * nb, 0);
        launch_pht_build(st, (const uint2*)R, nR, tab, nb - 1, pool, &sm->pool_ctr);
    }
    TBJ_HIP(hipEventRecord(C.ev[1], st));
    launch_pht_probe(st, mat, (const uint2*)S, nS, tab, nb - 1, pbits, hbits, pool,
                     (uint2*)pairs, mat ? pairs_cap : 0, sm->counters);
    TBJ_HIP(hipGetLastError());
    res->num_partitions = nb;
    return finish(C, st, sm, mat, pairs_cap, res);
}

// CHT (CHTJoinWrapper.cpp:56-114 with CHTJoin<128 partitions>): bitmap of
// nextpow2(|R|) bits, partitions of bitmap/128 bits.  Built per R tuple with
// global atomics or, in bulk (tbj_options.bulk_build), per 4096-position
// segment of the bitmap in LDS from R partitioned on the positions' high
// bits (res->radix_bits and res->passes then report that partitioning).
int run_cht(Context& C, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
            tbj_pair* pairs, uint64_t pairs_cap, const tbj_options& opt, tbj_result* res,
            hipStream_t st) {
    Workspace& W = C.ws;
    const bool mat = opt.materialize && pairs && pairs_cap;
    if (nR >= (1ull << 31) || nS >= (1ull << 40)) return TBJ_ERR_TOO_LARGE;
    Small* sm = nullptr;
    TBJ_TRY(W.get("cht_small", sizeof(Small), (void**)&sm));
    TBJ_HIP(hipEventRecord(C.ev[0], st));
    TBJ_HIP(hipMemsetAsync(sm, 0, sizeof(Small), st));
    if (nR == 0 || nS == 0) {
        TBJ_HIP(hipEventRecord(C.ev[1], st));
        return finish(C, st, sm, mat, pairs_cap, res);
    }
    const uint32_t bmsize = std::max<uint32_t>(64, next_pow2(nR));
    const uint32_t psize = std::max<uint32_t>(32, bmsize / 128);
    const uint32_t nwords = bmsize / 32;
    const uint32_t nslots = (uint32_t)nR + 2;
    const uint32_t twords = (nslots + 31) / 32;
    const uint32_t swords = (uint32_t)((nR + 31) / 32);
    uint32_t *cnt = nullptr, *partial = nullptr, *taken = nullptr, *spill = nullptr,
             *spilled = nullptr;
    uint2 *words = nullptr, *arr = nullptr;
    TBJ_TRY(W.get("cht_words", 8ull * nwords, (void**)&words));
    TBJ_TRY(W.get("cht_cnt", 4ull * (nwords + 1), (void**)&cnt));
    TBJ_TRY(W.get("cht_partial", 4ull * scan_partials((uint64_t)nwords + 1), (void**)&partial));
    TBJ_TRY(W.get("cht_taken", 4ull * twords, (void**)&taken));
    TBJ_TRY(W.get("cht_arr", 8ull * nslots, (void**)&arr));
    TBJ_TRY(W.get("cht_spill", 4ull * nR, (void**)&spill));
    TBJ_TRY(W.get("cht_spilled", 4ull * swords, (void**)&spilled));
    const bool bulk = opt.bulk_build > 0 || (opt.bulk_build == 0 && nR >= kBulkMin);
    launch_fill_u32(st, taken, twords, 0);
    launch_fill_u32(st, spilled, swords, 0);
    const uint2* Rsrc = (const uint2*)R;  // what the spill list indexes
    if (bulk) {
        uint32_t lg = 0, lgp = 0;
        while ((1u << lg) < bmsize) lg++;
        while ((1u << lgp) < psize) lgp++;
        const uint32_t sbits = std::min<uint32_t>(12, lgp), pbits = lg - sbits;
        const uint2* bounds = nullptr;
        uint32_t b1 = 0;
        TBJ_TRY(partition_relation(C, R, (uint32_t)nR, pbits, sbits, st, &Rsrc, &bounds, &b1));
        uint32_t *segcnt = nullptr, *spart = nullptr;
        TBJ_TRY(W.get("cht_segcnt", 4ull * ((1ull << pbits) + 1), (void**)&segcnt));
        TBJ_TRY(W.get("cht_spart", 4ull * scan_partials((1ull << pbits) + 1), (void**)&spart));
        launch_cht_build_bulk(st, Rsrc, bounds, b1, pbits, sbits, bmsize - 1, psize - 1, words,
                              segcnt, spart, taken, arr, spill, spilled, &sm->cht);
        res->radix_bits = (int32_t)pbits;
        res->passes = pbits > 8 ? 2 : 1;
    } else {
        launch_fill_u32(st, (uint32_t*)words, 2ull * nwords, 0);
        launch_cht_build(st, (const uint2*)R, nR, bmsize - 1, psize - 1, words, cnt, partial,
                         taken, arr, spill, spilled, &sm->cht);
    }
    Small h;
    TBJ_HIP(hipMemcpyAsync(&h, sm, sizeof(Small), hipMemcpyDeviceToHost, st));
    TBJ_HIP(hipStreamSynchronize(st));
    const uint32_t nspill = h.cht.spill_ctr;
    const uint32_t ovslots = next_pow2(2ull * nspill + 2);
    uint32_t* ov_used = nullptr;
    uint2* ov = nullptr;
    TBJ_TRY(W.get("cht_ov_used", 4ull * ovslots, (void**)&ov_used));
    TBJ_TRY(W.get("cht_ov", 8ull * ovslots, (void**)&ov));
    launch_fill_u32(st, ov_used, ovslots, 0);
    launch_cht_overflow(st, Rsrc, spill, &sm->cht, nspill, ovslots - 1, ov_used, ov);
    TBJ_HIP(hipEventRecord(C.ev[1], st));
    launch_cht_probe(st, mat, h.cht.dense != 0, (const uint2*)S, nS, bmsize - 1, psize - 1,
                     ovslots - 1, words, taken, arr, nslots, ov_used, ov, &sm->cht, (uint2*)pairs,
                     mat ? pairs_cap : 0, sm->counters);
    TBJ_HIP(hipGetLastError());
    res->num_partitions = 128;
    res->s_dropped = 0;
    res->max_r_partition = nspill;  // tuples kept in the overflow table
    return finish(C, st, sm, mat, pairs_cap, res);
}

}  // namespace tbj
