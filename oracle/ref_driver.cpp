// Test-infrastructure driver around the reference's own join and generator
// code, compiled from the reference sources in place (see oracle/Makefile).
// It replaces the reference's CLI main (native.cpp:48-130) with a C ABI that
// tests and the CPU-baseline leg of bench.py call through ctypes.  Nothing in
// the product links this file.
//
// Exports (all extern "C"):
//   ref_gen_pk / ref_gen_fk / ref_gen_fk_sel / ref_gen_zipf   -> keys only
//   ref_join_count(name, ...)                                  -> matches
//   ref_join_pairs(...)   (JOIN_MATERIALIZE build only)         -> pairs
//   ref_rand_stream(seed, n, out)                              -> glibc rand()
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <stdio.h>
#include <unistd.h>
#include <fcntl.h>

#include "data-types.h"
#include "prj_params.h"
#include "generator.h"
#include "radix_join.h"
#include "no_partitioning_join.h"
#include "CHTJoinWrapper.hpp"
#include "radix_sortmerge_join.h"
#include "parallel_sortmerge_join.h"

// The reference's Logger reads this start stamp (Logger.cpp:6); native.cpp:23
// normally defines it.
struct timespec ts_start;

namespace {

struct StdoutSilencer {
    int saved = -1;
    explicit StdoutSilencer(int quiet) {
        if (!quiet) return;
        fflush(stdout);
        saved = dup(1);
        int devnull = open("/dev/null", O_WRONLY);
        dup2(devnull, 1);
        close(devnull);
    }
    ~StdoutSilencer() {
        if (saved < 0) return;
        fflush(stdout);
        dup2(saved, 1);
        close(saved);
    }
};

void copy_keys(const table_t& t, uint32_t* out) {
    for (uint64_t i = 0; i < t.num_tuples; i++) out[i] = t.tuples[i].key;
}

// Relations handed to the reference joins carry RELATION_PADDING slack, as
// the reference generators allocate them (generator.cpp:22,53-74).
table_t make_rel(const uint32_t* keys, const uint32_t* pay, uint64_t n) {
    table_t t;
    void* mem = nullptr;
    if (posix_memalign(&mem, CACHE_LINE_SIZE, n * sizeof(row_t) + RELATION_PADDING)) abort();
    memset(mem, 0, n * sizeof(row_t) + RELATION_PADDING);
    t.tuples = (row_t*)mem;
    t.num_tuples = n;
    t.ratio_holes = 0;
    t.sorted = 0;
    for (uint64_t i = 0; i < n; i++) {
        t.tuples[i].key = keys[i];
        t.tuples[i].payload = pay ? pay[i] : 0;
    }
    return t;
}

result_t* CHT_wrap(relation_t* R, relation_t* S, int n, uint64_t* checksum) {
    join_result_t jr = CHTJ<7>(R, S, n);
    result_t* res = (result_t*)malloc(sizeof(result_t));
    res->totalresults = jr.matches;
    res->resultlist = nullptr;
    res->nthreads = n;
    if (checksum) *checksum = jr.checksum;
    return res;
}

}  // namespace

extern "C" {

int ref_rand_stream(uint32_t seed, uint64_t n, int32_t* out) {
    seed_generator(seed);
    for (uint64_t i = 0; i < n; i++) out[i] = rand();
    return 0;
}

int ref_gen_pk(uint32_t seed, uint64_t n, uint32_t* out, int quiet) {
    StdoutSilencer s(quiet);
    table_t t;
    seed_generator(seed);
    if (create_relation_pk(&t, n, 0)) return -1;
    copy_keys(t, out);
    delete_relation(&t);
    return 0;
}

int ref_gen_fk(uint32_t seed, uint64_t n, int64_t maxid, uint32_t* out, int quiet) {
    StdoutSilencer s(quiet);
    table_t t;
    seed_generator(seed);
    if (create_relation_fk(&t, n, maxid, 0)) return -1;
    copy_keys(t, out);
    delete_relation(&t);
    return 0;
}

int ref_gen_fk_sel(uint32_t seed, uint64_t n, int64_t maxid, uint32_t* out, int quiet) {
    StdoutSilencer s(quiet);
    table_t t;
    seed_generator(seed);
    if (create_relation_fk_sel(&t, n, maxid, 0)) return -1;
    copy_keys(t, out);
    delete_relation(&t);
    return 0;
}

int ref_gen_zipf(uint32_t seed, uint64_t n, int64_t alphabet, double z, uint32_t* out, int quiet) {
    StdoutSilencer s(quiet);
    table_t t;
    seed_generator(seed);
    if (create_relation_zipf(&t, n, alphabet, z, 0)) return -1;
    copy_keys(t, out);
    delete_relation(&t);
    return 0;
}

// Runs one of the reference's join entry points on copies of the given
// relations.  Returns the match count, or -1 for an unknown name.
// checksum_out receives CHT's payload checksum (left untouched otherwise).
// elapsed_us_out receives wall time around the join call.
int64_t ref_join_count(const char* name,
                       const uint32_t* rkeys, const uint32_t* rpay, uint64_t nr,
                       const uint32_t* skeys, const uint32_t* spay, uint64_t ns,
                       int nthreads, int quiet, uint64_t* checksum_out,
                       double* elapsed_us_out) {
    StdoutSilencer s(quiet);
    clock_gettime(CLOCK_MONOTONIC, &ts_start);
    table_t R = make_rel(rkeys, rpay, nr);
    table_t S = make_rel(skeys, spay, ns);
    result_t* res = nullptr;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    if (!strcmp(name, "RHO")) res = RHO(&R, &S, nthreads);
    else if (!strcmp(name, "RHT")) res = RHT(&R, &S, nthreads);
    else if (!strcmp(name, "RJ")) res = RJ(&R, &S, nthreads);
    else if (!strcmp(name, "RSM")) res = RSM(&R, &S, nthreads);
    else if (!strcmp(name, "PSM")) res = PSM(&R, &S, nthreads);
    else if (!strcmp(name, "PHT")) res = PHT(&R, &S, nthreads);
    else if (!strcmp(name, "NPO_st")) res = NPO_st(&R, &S, nthreads);
    else if (!strcmp(name, "CHT")) res = CHT_wrap(&R, &S, nthreads, checksum_out);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (elapsed_us_out)
        *elapsed_us_out = (t1.tv_sec - t0.tv_sec) * 1e6 + (t1.tv_nsec - t0.tv_nsec) * 1e-3;
    free(R.tuples);
    free(S.tuples);
    if (!res) return -1;
    int64_t m = res->totalresults;
    free(res);
    return m;
}

#ifdef JOIN_MATERIALIZE
// RHO / RHT with JOIN_MATERIALIZE: walks each thread's prepend list for
// exactly nresults nodes (the list head starts uninitialised,
// radix_join.c:1182) and writes (Rpayload, Spayload, key) triples.
// Returns the number of triples, or -1 when cap is too small.
int64_t ref_join_pairs(const char* name,
                       const uint32_t* rkeys, const uint32_t* rpay, uint64_t nr,
                       const uint32_t* skeys, const uint32_t* spay, uint64_t ns,
                       int nthreads, int quiet,
                       uint32_t* out_rpay, uint32_t* out_spay, uint32_t* out_key,
                       uint64_t cap) {
    StdoutSilencer s(quiet);
    clock_gettime(CLOCK_MONOTONIC, &ts_start);
    table_t R = make_rel(rkeys, rpay, nr);
    table_t S = make_rel(skeys, spay, ns);
    result_t* res = nullptr;
    if (!strcmp(name, "RHO")) res = RHO(&R, &S, nthreads);
    else if (!strcmp(name, "RHT")) res = RHT(&R, &S, nthreads);
    free(R.tuples);
    free(S.tuples);
    if (!res) return -1;
    uint64_t k = 0;
    int64_t rc = 0;
    for (int t = 0; t < res->nthreads; t++) {
        output_list_t* node = res->resultlist[t].results;
        for (int64_t i = 0; i < res->resultlist[t].nresults; i++) {
            if (k >= cap) { rc = -1; break; }
            out_rpay[k] = node->Rpayload;
            out_spay[k] = node->Spayload;
            out_key[k] = node->key;
            k++;
            output_list_t* nxt = node->next;
            free(node);
            node = nxt;
        }
    }
    free(res->resultlist);
    free(res);
    return rc < 0 ? rc : (int64_t)k;
}
#endif

}  // extern "C"
