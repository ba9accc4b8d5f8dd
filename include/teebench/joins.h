/*
 * teebench/joins.h — the reference's join operator API on the MI355X.
 *
 * Same names, signature and C++ linkage as the reference's join entry
 * points (Joins/radix_join.h:17-24, Joins/no_partitioning_join.h:474,543,
 * Joins/radix_sortmerge_join.h, Joins/psm/parallel_sortmerge_join.h:61,
 * native.cpp:26-32 for CHT, rho_atomic/radix_join_atomic.cpp:1278,
 * mway/sortmergejoin_multiway.h for MWAY), so an
 * algorithm_t table like native.cpp:34-46 links against this library
 * instead of the reference's Joins objects.  Each call copies the relations
 * to HBM, runs the join through the C ABI (tbj.h) and returns a malloc'd
 * result_t.  nthreads (CPU threads in the reference) is accepted and
 * ignored.  Unlike the reference RHO/RSM/PSM the inputs are left unmodified.
 *
 * Materialisation: the reference compiles it in with -DJOIN_MATERIALIZE
 * (radix_join.c:231-233); here teebench_set_materialize(1) turns it on at
 * run time.  resultlist then holds one entry whose list has exactly
 * totalresults output_list_t nodes (key, Rpayload, Spayload), allocated as
 * one block in list order; free it with teebench_free_result().
 *
 * Errors: a failing device call logs and exits with EXIT_FAILURE, which is
 * what the reference does on allocation or thread failures (util.c:11-18).
 */
#ifndef TEEBENCH_JOINS_H
#define TEEBENCH_JOINS_H
#include "data_types.h"

result_t* RHO(relation_t* relR, relation_t* relS, int nthreads);
result_t* RHT(relation_t* relR, relation_t* relS, int nthreads);
result_t* RJ(relation_t* relR, relation_t* relS, int nthreads);
result_t* RHO_atomic(relation_t* relR, relation_t* relS, int nthreads);
result_t* PHT(relation_t* relR, relation_t* relS, int nthreads);
result_t* NPO_st(relation_t* relR, relation_t* relS, int nthreads);
result_t* CHT(relation_t* relR, relation_t* relS, int nthreads);
result_t* RSM(relation_t* relR, relation_t* relS, int nthreads);
result_t* PSM(relation_t* relR, relation_t* relS, int nthreads);
result_t* MWAY(relation_t* relR, relation_t* relS, int nthreads);

/* Runtime switches of this implementation (no reference counterpart; C
 * linkage so that FFI callers find them by name). */
#ifdef __cplusplus
extern "C" {
#endif
void teebench_set_materialize(int on);
/* 0 = radix (default), 1 = learned CDF range partitioner with radix fallback. */
void teebench_set_partitioner(int learned);
void teebench_free_result(result_t* res);
#ifdef __cplusplus
}
#endif

/* The selector table of the plain-CPU harness (native.cpp:34-46) with the
 * GPU implementations; terminated by an entry with an empty name. */
extern algorithm_t teebench_algorithms[];

#endif
