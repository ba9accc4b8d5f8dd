// Test-infrastructure driver around the reference's own join and generator
// code, compiled from the reference sources in place (see oracle/Makefile).
// It replaces the reference's CLI main (native.cpp:48-130) with a C ABI that
// tests and the CPU-baseline leg of bench.py call through ctypes.  Nothing in
// the product links this file.
//
// Exports (all extern "C"):
//   ref_gen_pk / ref_gen_fk / ref_gen_fk_sel / ref_gen_zipf   -> keys only
//   ref_gen_from_file(seed, path, ...)                          -> keys only
//   ref_join_count(name, ...)                                  -> matches
//   ref_join_pairs(...)   (JOIN_MATERIALIZE build only)         -> pairs
//   ref_join_mat_timed(...) (JOIN_MATERIALIZE build only)       -> matches, join time
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
#include "sortmergejoin_multiway.h"

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

// MWAY's thread start (mway/joincommon.c:51,137) passes a pthread_attr_t it
// never initialises.  Zeroing the stack the call is about to use makes that
// attribute read as all-zero (glibc: default stack and guard size) instead of
// garbage, on which pthread_create fails with EAGAIN.
__attribute__((noinline)) void scrub_stack() {
    volatile uint8_t pad[1 << 20];
    for (size_t i = 0; i < sizeof(pad); i++) pad[i] = 0;
}

__attribute__((noinline)) result_t* MWAY_clean_stack(relation_t* R, relation_t* S, int n) {
    scrub_stack();
    return MWAY(R, S, n);
}

// create_relation_from_file hands getline an uninitialised buffer pointer
// with size 0 (generator.cpp:298-300,329); on a zeroed stack that pointer is
// NULL and getline allocates, as the reference intends.
__attribute__((noinline)) int from_file_clean_stack(table_t* t, char* path) {
    scrub_stack();
    return create_relation_from_file(t, path, 0);
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

// create_relation_from_file (generator.cpp:290-351) after seed_generator(seed),
// as native.cpp:78-84 does for -t/-u.  Returns the tuple count (keys in out,
// at most cap), or -1 when the file cannot be read.
int64_t ref_gen_from_file(uint32_t seed, const char* path, uint32_t* out, uint64_t cap,
                          int quiet) {
    StdoutSilencer s(quiet);
    table_t t;
    seed_generator(seed);
    if (from_file_clean_stack(&t, (char*)path)) return -1;
    const uint64_t n = t.num_tuples < cap ? t.num_tuples : cap;
    for (uint64_t i = 0; i < n; i++) out[i] = t.tuples[i].key;
    const int64_t total = (int64_t)t.num_tuples;
    free(t.tuples);
    return total;
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
    else if (!strcmp(name, "MWAY")) res = MWAY_clean_stack(&R, &S, nthreads);
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

// The materialising join's own time: RHO / RHT with JOIN_MATERIALIZE (one
// output-list node per match, built inside the join), timed around the join
// call alone; the lists are then walked and freed, untimed.  Returns the
// matches the lists hold, or -1.
int64_t ref_join_mat_timed(const char* name,
                           const uint32_t* rkeys, const uint32_t* rpay, uint64_t nr,
                           const uint32_t* skeys, const uint32_t* spay, uint64_t ns,
                           int nthreads, int quiet, double* elapsed_us_out) {
    StdoutSilencer s(quiet);
    clock_gettime(CLOCK_MONOTONIC, &ts_start);
    table_t R = make_rel(rkeys, rpay, nr);
    table_t S = make_rel(skeys, spay, ns);
    result_t* res = nullptr;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    if (!strcmp(name, "RHO")) res = RHO(&R, &S, nthreads);
    else if (!strcmp(name, "RHT")) res = RHT(&R, &S, nthreads);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (elapsed_us_out)
        *elapsed_us_out = (t1.tv_sec - t0.tv_sec) * 1e6 + (t1.tv_nsec - t0.tv_nsec) * 1e-3;
    free(R.tuples);
    free(S.tuples);
    if (!res) return -1;
    int64_t k = 0;
    for (int t = 0; t < res->nthreads; t++) {
        output_list_t* node = res->resultlist[t].results;
        for (int64_t i = 0; i < res->resultlist[t].nresults; i++) {
            output_list_t* nxt = node->next;
            free(node);
            node = nxt;
            k++;
        }
    }
    free(res->resultlist);
    free(res);
    return k;
}
#endif

}  // extern "C"
