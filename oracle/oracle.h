/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the reference's equi-join path (generators, RHO,
 * RHT, PHT, CHT, PSM/RSM, MWAY, the join written as a relation) used as the
 * checker for the HIP product.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * liboracle.so, and only as the checker.  The product never calls it.
 *
 * Every function cites the reference file:line it restates.  What pins each
 * function (golden vectors from the reference build, closed forms, brute
 * force) is listed in tests/test_oracle.py; no function here is unpinned.
 */
#ifndef TBJ_ORACLE_H
#define TBJ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* glibc TYPE_3 additive-feedback generator behind rand()/srand()
 * (used by generator.cpp:20,79,106 and genzipf.c:44,122). */
typedef struct orc_rng { int32_t tbl[31]; int f, r; } orc_rng;
void    orc_srand(orc_rng* g, uint32_t seed);
int32_t orc_rand(orc_rng* g);
void    orc_rand_stream(uint32_t seed, uint64_t n, int32_t* out);

/* Generators: keys only, like the reference (payload never written). */
void orc_gen_pk(uint32_t seed, uint64_t n, uint32_t* keys);
void orc_gen_fk(uint32_t seed, uint64_t n, int64_t maxid, uint32_t* keys);
void orc_gen_fk_sel(uint32_t seed, uint64_t n, int64_t maxid, uint32_t* keys);
int  orc_gen_zipf(uint32_t seed, uint64_t n, uint32_t alphabet, double z, uint32_t* keys);

/* Joins over 8-byte tuples {u32 key, u32 payload} given as split arrays.
 * pairs: if out_rpay/out_spay are non-NULL, up to cap (R.payload, S.payload)
 * pairs are written in the algorithm's natural order.  Return = matches. */
int64_t orc_rho(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                uint32_t* out_rpay, uint32_t* out_spay, uint64_t cap);
int64_t orc_rht(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                uint32_t* out_rpay, uint32_t* out_spay, uint64_t cap);
int64_t orc_pht(const uint32_t* rk, uint64_t nr, const uint32_t* sk, uint64_t ns);
int64_t orc_cht(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                uint64_t* checksum, int64_t* unresolved);
int64_t orc_sortmerge(const uint32_t* rk, uint64_t nr, const uint32_t* sk, uint64_t ns);
/* MWAY's merge join over the sorted relations (mway/joincommon.c:241-316). */
int64_t orc_mway(const uint32_t* rk, uint64_t nr, const uint32_t* sk, uint64_t ns);
/* The join as a relation of (key, R.payload) tuples, to_relation's format
 * (secure_joins.cpp:277-301), sorted; the building block of the three-way
 * join (R join S) join T (secure_joins.cpp:348-399). */
int64_t orc_join_relation(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                          const uint32_t* sk, uint64_t ns, uint32_t* out_key,
                          uint32_t* out_pay, uint64_t cap);

/* Textbook equi-join (sort both sides, multiply equal-key run lengths). */
int64_t orc_textbook_count(const uint32_t* rk, uint64_t nr, const uint32_t* sk, uint64_t ns);
int64_t orc_textbook_pairs(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                           const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                           uint32_t* out_rpay, uint32_t* out_spay, uint64_t cap);
/* 16-byte tuples {u64 key, u64 payload}: pairs sorted by (key, Rpay, Spay). */
int64_t orc_textbook_pairs64(const uint64_t* rk, const uint64_t* rp, uint64_t nr,
                             const uint64_t* sk, const uint64_t* sp, uint64_t ns,
                             uint64_t* out_rpay, uint64_t* out_spay, uint64_t cap);

/* Sorts (a[i], b[i]) pairs lexicographically, in place. */
void orc_sort_pairs32(uint32_t* a, uint32_t* b, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
