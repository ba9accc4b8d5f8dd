// The reference's join operator API (include/teebench/joins.h) on top of the
// C ABI.  Each entry point keeps the reference's name, signature and C++
// linkage (Joins/radix_join.h:17-24 and the other join headers), uploads the
// relations, runs the GPU join and returns a malloc'd result_t.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <vector>

#include "../../include/tbj.h"
#include "../../include/teebench/joins.h"
#include "teebench_log.h"

namespace {

int g_materialize = 0;
int g_learned = 0;

result_t* run(int algo, relation_t* R, relation_t* S, int nthreads) {
    tbj_options opt;
    tbj_default_options(&opt);
    opt.partitioner = g_learned ? TBJ_PART_LEARNED : TBJ_PART_AUTO;
    opt.materialize = g_materialize;
    tbj_result res;
    result_t* out = (result_t*)malloc(sizeof(result_t));
    if (!out) {
        teebench_log(TB_ERROR, "Failed to allocate memory");
        exit(EXIT_FAILURE);
    }
    out->resultlist = nullptr;
    out->nthreads = nthreads;
    int rc;
    if (!g_materialize) {
        rc = tbj_join_host(algo, (const tbj_tuple*)R->tuples, R->num_tuples,
                           (const tbj_tuple*)S->tuples, S->num_tuples, nullptr, 0, &opt, &res);
    } else {
        // Join row indices instead of payloads: each pair (i, j) then names
        // R[i] and S[j], from which (key, Rpayload, Spayload) follow exactly.
        std::vector<tbj_tuple> r(R->num_tuples), s(S->num_tuples);
        for (uint64_t i = 0; i < R->num_tuples; i++) r[i] = {R->tuples[i].key, (uint32_t)i};
        for (uint64_t i = 0; i < S->num_tuples; i++) s[i] = {S->tuples[i].key, (uint32_t)i};
        // room for one match per S tuple (every foreign-key join); a join
        // with more matches runs once more with the exact room
        std::vector<tbj_pair> pairs(std::max<uint64_t>(1, S->num_tuples));
        rc = tbj_join_host(algo, r.data(), r.size(), s.data(), s.size(), pairs.data(),
                           pairs.size(), &opt, &res);
        if (rc == TBJ_OK && res.matches > res.pairs_written) {
            pairs.resize(res.matches);
            rc = tbj_join_host(algo, r.data(), r.size(), s.data(), s.size(), pairs.data(),
                               pairs.size(), &opt, &res);
        }
        if (rc == TBJ_OK) {
            out->resultlist = (threadresult_t*)malloc(sizeof(threadresult_t));
            out->nthreads = 1;
            out->resultlist[0].nresults = (int64_t)res.pairs_written;
            out->resultlist[0].threadid = 0;
            // one block of nodes, linked in order (teebench_free_result frees
            // the block through the head)
            output_list_t* nodes = nullptr;
            if (res.pairs_written) {
                nodes = (output_list_t*)malloc(sizeof(output_list_t) * res.pairs_written);
                if (!nodes) {
                    teebench_log(TB_ERROR, "Failed to allocate memory");
                    exit(EXIT_FAILURE);
                }
            }
            for (uint64_t k = 0; k < res.pairs_written; k++) {
                const tbj_pair& p = pairs[k];
                nodes[k].key = R->tuples[p.r_payload].key;
                nodes[k].Rpayload = R->tuples[p.r_payload].payload;
                nodes[k].Spayload = S->tuples[p.s_payload].payload;
                nodes[k].next = k + 1 < res.pairs_written ? &nodes[k + 1] : nullptr;
            }
            out->resultlist[0].results = nodes;
        }
    }
    if (rc != TBJ_OK) {
        teebench_log(TB_ERROR, "%s join failed: %s", tbj_algo_name(algo), tbj_status_str(rc));
        exit(EXIT_FAILURE);
    }
    out->totalresults = (int64_t)res.matches;
    // the reference's print_timing lines (radix_join.c:148-172)
    const uint64_t n = R->num_tuples + S->num_tuples;
    const double us = res.ms_total * 1000.0;
    teebench_log(TB_ENCLAVE, "Total input tuples : %lu", (unsigned long)n);
    teebench_log(TB_ENCLAVE, "Result tuples : %lu", (unsigned long)res.matches);
    teebench_log(TB_ENCLAVE, "Phase Partition (ms) : %.3f", res.ms_partition);
    teebench_log(TB_ENCLAVE, "Phase Join (ms) : %.3f", res.ms_join);
    teebench_log(TB_ENCLAVE, "Total Runtime (us) : %lu ", (unsigned long)(us + 0.5));
    teebench_log(TB_ENCLAVE, "Throughput (M rec/sec) : %.2lf", us > 0 ? n / us : 0.0);
    return out;
}

}  // namespace

result_t* RHO(relation_t* R, relation_t* S, int n) { return run(TBJ_RHO, R, S, n); }
result_t* RHT(relation_t* R, relation_t* S, int n) { return run(TBJ_RHT, R, S, n); }
result_t* RJ(relation_t* R, relation_t* S, int n) { return run(TBJ_RJ, R, S, n); }
result_t* RHO_atomic(relation_t* R, relation_t* S, int n) { return run(TBJ_RHO_ATOMIC, R, S, n); }
result_t* PHT(relation_t* R, relation_t* S, int n) { return run(TBJ_PHT, R, S, n); }
result_t* NPO_st(relation_t* R, relation_t* S, int n) { return run(TBJ_NPO_ST, R, S, n); }
result_t* CHT(relation_t* R, relation_t* S, int n) { return run(TBJ_CHT, R, S, n); }
result_t* RSM(relation_t* R, relation_t* S, int n) { return run(TBJ_RSM, R, S, n); }
result_t* PSM(relation_t* R, relation_t* S, int n) { return run(TBJ_PSM, R, S, n); }
result_t* MWAY(relation_t* R, relation_t* S, int n) { return run(TBJ_MWAY, R, S, n); }

void teebench_set_materialize(int on) { g_materialize = on ? 1 : 0; }
void teebench_set_partitioner(int learned) { g_learned = learned ? 1 : 0; }

void teebench_free_result(result_t* res) {
    if (!res) return;
    if (res->resultlist) {
        for (int t = 0; t < res->nthreads; t++) free(res->resultlist[t].results);
        free(res->resultlist);
    }
    free(res);
}

// native.cpp:34-46 with the GPU implementations (the reference's CPU-only
// nested-loop joins NL and INL are not part of this library).
algorithm_t teebench_algorithms[] = {
    {"PHT", PHT}, {"NPO_st", NPO_st}, {"RJ", RJ},   {"PSM", PSM},
    {"RHO", RHO}, {"RHT", RHT},       {"RSM", RSM}, {"CHT", CHT},
    {"MWAY", MWAY}, {"RHO_atomic", RHO_atomic}, {"", nullptr}};
