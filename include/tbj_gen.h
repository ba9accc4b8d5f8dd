/*
 * tbj_gen.h — relation generators of libtbj.so (host side).
 *
 * Bit-exact with the reference's generators on the same seeds: the keys come
 * from the same glibc TYPE_3 rand() stream and the same shuffle and Zipf
 * arithmetic (App/Lib/generator.cpp:20-23,76-111,144-171,353-379,475-555,
 * 639-662; App/Lib/genzipf.c:28-158), but independent shuffled blocks and the
 * Zipf lookups run on several host threads.
 *
 * The reference never writes payloads (they are whatever the allocation
 * held, zero in practice).  payload_mode selects TBJ_PAYLOAD_ZERO (the same
 * zero payloads) or TBJ_PAYLOAD_ROWID (payload = row index, so that
 * materialised pairs identify both tuples).
 */
#ifndef TBJ_GEN_H
#define TBJ_GEN_H
#include <stdint.h>
#include "tbj.h"

#ifdef __cplusplus
extern "C" {
#endif

enum tbj_payload_mode { TBJ_PAYLOAD_ZERO = 0, TBJ_PAYLOAD_ROWID = 1 };

/* glibc rand() restated: the first n outputs after srand(seed). */
void tbj_rand_stream(uint32_t seed, uint64_t n, int32_t* out);

/* create_relation_pk: keys 1..n shuffled (generator.cpp:353-379). */
int tbj_gen_pk(uint32_t seed, uint64_t n, tbj_tuple* out, int payload_mode, int nthreads);
/* create_relation_fk: n/maxid shuffled blocks of 1..maxid (+ remainder)
 * (generator.cpp:475-514). */
int tbj_gen_fk(uint32_t seed, uint64_t n, int64_t maxid, tbj_tuple* out, int payload_mode,
               int nthreads);
/* create_relation_fk_sel (generator.cpp:516-555); maxid = 100*|R|/selectivity
 * as native.cpp:113-114 computes it. */
int tbj_gen_fk_sel(uint32_t seed, uint64_t n, int64_t maxid, tbj_tuple* out, int payload_mode,
                   int nthreads);
/* create_relation_zipf / gen_zipf (generator.cpp:639-662, genzipf.c:97-158). */
int tbj_gen_zipf(uint32_t seed, uint64_t n, uint32_t alphabet, double z, tbj_tuple* out,
                 int payload_mode, int nthreads);

/* create_relation_from_file (generator.cpp:290-351), as native.cpp:78-84 runs
 * it for -t/-u: one integer key per line (atoi), as many tuples as the file
 * has newlines, then knuth_shuffle with the rand() stream of seed.  Returns
 * the tuple count; out (cap tuples) is filled only when the count fits, so a
 * call with out = NULL sizes the relation.  TBJ_ERR_INVALID when the file
 * cannot be opened. */
int64_t tbj_gen_from_file(uint32_t seed, const char* path, tbj_tuple* out, uint64_t cap,
                          int payload_mode);
/* The same shuffle of n given keys (the keys of a loaded file). */
int tbj_gen_from_keys(uint32_t seed, const uint32_t* keys, uint64_t n, tbj_tuple* out,
                      int payload_mode);

/* Remaps every odd-indexed tuple's key to key + offset (keys outside R's
 * domain when offset >= max R key): the out-of-domain half of the skewed
 * 50%-selectivity workload.  No reference counterpart (the reference CLI
 * cannot combine -z with -l, native.cpp:107-118). */
void tbj_remap_odd(tbj_tuple* rel, uint64_t n, uint32_t offset);

/* 16-byte tuples: n unique nonzero random 64-bit keys, payload = row index.
 * Key i is splitmix64 of counter seed + (i+1)*0x9E3779B97F4A7C15; the
 * finaliser is a bijection of the counter, so distinct counters give
 * distinct keys. */
int tbj_gen_unique64(uint64_t seed, uint64_t n, tbj_tuple64* out, int nthreads);
/* n probe tuples whose keys are Zipf(z) ranks over R's keys (rank 1 = R[0]),
 * payload = row index.  Ranks are drawn by inverse-CDF lookup over a
 * double-precision Zipf table with a splitmix64 uniform stream. */
int tbj_gen_zipf_fk64(uint64_t seed, uint64_t n, const tbj_tuple64* R, uint64_t nR, double z,
                      tbj_tuple64* out, int nthreads);

/*
 * Batches of joins generated on the device (no reference counterpart: the
 * batch workload has no reference path).  Join i of a batch takes the seed
 * s_i = splitmix64(base_seed ^ i) and from its splitmix64 stream draws
 *   u ~ U[0,1): |R_i| = floor(2^(10 + 12u));
 *   k ~ U{1..16}: |S_i| = k * |R_i|;
 *   z ~ U[0, 1.5) (1.5 times a 53-bit uniform);
 *   a ~ U{0, 1}: join i is TBJ_RHO (a = 0) or TBJ_PHT (a = 1).
 * R_i: keys 1..|R_i| in a pseudo-random order (a Feistel permutation of the
 * row index, seeded by s_i).  S_i: keys in 1..|R_i|, uniform for z = 0 and
 * Zipf(z) ranks otherwise (rejection-inversion sampling, Hormann and
 * Derflinger, "Rejection-inversion to generate variates from monotone
 * discrete distributions", 1996).  Payload = row index within the relation.
 * Every S key is an R key, so join i has exactly |S_i| matches.
 */
typedef struct tbj_batch_spec {
    uint64_t seed;      /* s_i */
    uint64_t nR, nS;
    double z;
    uint64_t r_off, s_off;  /* tuple offsets in the arena */
    int32_t algo;           /* TBJ_RHO or TBJ_PHT */
    uint32_t reserved;
} tbj_batch_spec;

/* Fills specs[0..n) for joins of base_seed (offsets: R_0, S_0, R_1, S_1, ...
 * contiguous from tuple 0); returns the total tuple count. */
uint64_t tbj_gen_batch_specs(uint64_t base_seed, uint32_t n, tbj_batch_spec* specs);
/* Writes every join's R and S into the device arena (tuples at the spec
 * offsets) on the given hipStream_t (NULL = default); asynchronous. */
int tbj_gen_batch_device(tbj_tuple* arena, const tbj_batch_spec* specs, uint32_t n,
                         void* stream);

#ifdef __cplusplus
}
#endif
#endif
