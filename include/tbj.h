/*
 * tbj.h — C ABI of the MI355X equi-join library (libtbj.so).
 *
 * This is the boundary the reference's join operator API forwards to.  The
 * reference's operator is
 *     result_t* NAME(relation_t* R, relation_t* S, int nthreads)
 * (Include/data-types.h:68-71, Joins/radix_join.h:17-24) selected by name
 * through algorithm_t tables (native.cpp:34-46); teebench/joins.h keeps
 * those C++ entry points and forwards to the functions below.  Plain
 * pointers and sizes only: no torch, no HIP types in the signatures.
 *
 * Status codes replace the reference's process exits (util.c:11-18,
 * radix_join.c:1331-1334, CHT.hpp:203-204).
 */
#ifndef TBJ_H
#define TBJ_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 8-byte tuple, layout-identical to the reference's struct row_t
 * {type_key key; type_value payload;} (Include/data-types.h:44-47). */
typedef struct tbj_tuple { uint32_t key; uint32_t payload; } tbj_tuple;
/* One materialised match (R.payload, S.payload): the two payload fields of
 * the reference's output_list_t (Include/data-types.h:49-54). */
typedef struct tbj_pair { uint32_t r_payload; uint32_t s_payload; } tbj_pair;

/* 16-byte tuples for 64-bit keys (no reference counterpart: the reference
 * hard-wires 8-byte tuples, Include/data-types.h:32-47). */
typedef struct tbj_tuple64 { uint64_t key; uint64_t payload; } tbj_tuple64;
typedef struct tbj_pair64 { uint64_t r_payload; uint64_t s_payload; } tbj_pair64;

/* Join algorithms, named like the reference's -a selector
 * (native.cpp:34-46; RHO_atomic from Enclave/TrustedJoins/secure_joins.cpp:122). */
enum tbj_algo {
    TBJ_RHO = 0,        /* radix partition + per-partition bucket join (radix_join.c:1557) */
    TBJ_RHT = 1,        /* radix partition + histogram join (radix_join.c:1563) */
    TBJ_RJ = 2,         /* single-stream radix join (radix_join.c:1398) */
    TBJ_PHT = 3,        /* no-partitioning shared hash table (no_partitioning_join.h:543) */
    TBJ_CHT = 4,        /* concise hash table (cht/CHTJoinWrapper.cpp:56) */
    TBJ_RSM = 5,        /* radix partition + sort-merge per partition (radix_sortmerge_join.c:102) */
    TBJ_PSM = 6,        /* sort both relations + merge (psm/parallel_sortmerge_join.h:61) */
    TBJ_RHO_ATOMIC = 7, /* RHO with lock-free task indices (rho_atomic/radix_join_atomic.cpp:1278) */
    TBJ_NPO_ST = 8,     /* single-table variant of PHT (no_partitioning_join.h:474) */
    TBJ_MWAY = 9,       /* partition, sort, bounded merge join (mway/sortmergejoin_multiway.c:327,
                           merge semantics of mway/joincommon.c:241-316) */
    TBJ_ALGO_COUNT = 10
};

/* What a materialising join writes per match (tbj_options.output). */
enum tbj_output {
    TBJ_OUT_PAIRS = 0,    /* tbj_pair {R.payload, S.payload} */
    TBJ_OUT_RELATION = 1  /* tbj_tuple {key, R.payload}: the tuple the reference's
                             to_relation copies out of each output_list_t node
                             (Enclave/TrustedJoins/secure_joins.cpp:277-301), so the
                             output is a relation that feeds the next join on the
                             device (radix-family algorithms only) */
};

enum tbj_partitioner {
    TBJ_PART_AUTO = 0,    /* radix */
    TBJ_PART_RADIX = 1,   /* low key bits, as prj_params.h:18-23 */
    TBJ_PART_LEARNED = 2  /* piecewise-linear key-CDF range partitioner, radix fallback */
};

enum tbj_status {
    TBJ_OK = 0,
    TBJ_ERR_INVALID = -1,      /* bad argument */
    TBJ_ERR_NO_DEVICE = -2,    /* no HIP device / library not initialised */
    TBJ_ERR_OOM = -3,          /* device allocation failed or workspace cap hit */
    TBJ_ERR_HIP = -4,          /* a HIP runtime call failed */
    TBJ_ERR_UNSUPPORTED = -5,  /* algorithm/option combination not available */
    TBJ_ERR_TOO_LARGE = -6     /* relation exceeds 2^32 - 2^20 tuples */
};

typedef struct tbj_options {
    int32_t partitioner;        /* enum tbj_partitioner */
    int32_t radix_bits;         /* total partition bits; 0 = chosen from |R| */
    int32_t passes;             /* 1 or 2 global partition passes; 0 = auto */
    int32_t materialize;        /* 1: write (R.payload, S.payload) pairs */
    uint64_t workspace_cap;     /* bytes; 0 = no cap (a cap forces TBJ_ERR_OOM) */
    uint32_t learned_bound;     /* max R tuples per learned partition before the
                                   radix fallback; 0 = the LDS capacity */
    int32_t learned_samples;    /* keys sampled to fit the CDF model; 0 = auto */
    int32_t s_chunk;            /* S tuples per join task; 0 = auto */
    int32_t sampled_regions;    /* two-pass plans: size the output regions of
                                   both passes from a 1/16 sample of the input
                                   instead of per-tile histograms, so that no
                                   pass reads its input twice (redone exactly
                                   when a region overflows).  0 = auto (for
                                   |R| + |S| >= 2^24), 1 = always, -1 = never */
    int32_t profile;            /* 1: record per-phase device times in phase_ms */
    int32_t output;             /* enum tbj_output */
    int32_t bulk_build;         /* PHT / NPO_st: build the hash table from R
                                   partitioned on the bucket's low bits (one
                                   contiguous block of buckets per partition,
                                   filled in LDS; no global atomic per tuple)
                                   instead of one atomic insert per tuple.
                                   0 = auto (|R| >= 2^20), 1 = always, -1 = never */
    int32_t reserved[1];
} tbj_options;

typedef struct tbj_result {
    uint64_t matches;           /* total matches (the reference's totalresults) */
    uint64_t checksum;          /* sum of (R.payload + S.payload) over matches, mod 2^64
                                   (CHT's join_result_t.checksum, data-types.h:86-92) */
    uint64_t pairs_written;     /* min(matches, pairs_cap) when materialising */
    int32_t partitioner_used;   /* enum tbj_partitioner actually used */
    int32_t fallback;           /* 1 if the learned partitioner fell back to radix */
    int32_t radix_bits;
    int32_t passes;
    uint32_t num_partitions;
    uint32_t max_r_partition;   /* largest R partition (tuples) */
    uint64_t s_dropped;         /* S tuples dropped as outside R's key range */
    float ms_partition;         /* device time of the partition phase */
    float ms_join;              /* device time of build/probe (or sort/merge) */
    float ms_total;             /* device time of the whole call */
    int32_t redone_exact;       /* 1 if a sample-sized region overflowed and the
                                   call was redone with exact histograms */
    int32_t learned_model;      /* the learned partitioner's model: 1 = knots and a
                                   cell table (piecewise-linear CDF), 2 = one linear
                                   map of the key range (near-linear sample); 0 = none */
    /* with opt->profile: device ms of hist1, scan1, scatter1, pass2 sizing,
       scatter2, tasks, join, other */
    float phase_ms[8];
} tbj_result;

/* Initialise on a HIP device (0-based).  Safe to call more than once.
 *
 * Concurrency: every caller stream gets its own session (device workspace,
 * events, a side stream), so calls on different streams, from different host
 * threads, run at the same time; calls on one stream serialise.  Each
 * session keeps its workspace between calls. */
int tbj_init(int device);
/* Release every session's cached device workspace. */
void tbj_shutdown(void);
/* Release the session (and workspace) of one stream; waits for a call still
 * running on it. */
int tbj_release_stream(void* stream);
void tbj_default_options(tbj_options* opt);
const char* tbj_status_str(int status);
/* Number of compiled join algorithms and their names (the -a strings). */
const char* tbj_algo_name(int algo);
int tbj_algo_from_name(const char* name);

/*
 * Equi-join of two device-resident relations.  R and S point to HBM; they
 * are not modified (unlike the reference RHO/PSM, which clobber their inputs,
 * radix_join.c:596-606, parallel_sortmerge_join.h:78-86).  pairs (device,
 * may be NULL) receives up to pairs_cap matches when opt->materialize is set,
 * in no particular order.  stream is a hipStream_t (NULL = default stream).
 * Synchronous: returns after the result is complete.
 */
int tbj_join(int algo,
             const tbj_tuple* R, uint64_t nR,
             const tbj_tuple* S, uint64_t nS,
             tbj_pair* pairs, uint64_t pairs_cap,
             const tbj_options* opt, tbj_result* res, void* stream);

/*
 * The three-way join (R join S) join T of the reference's
 * ecall_three_way_join_sealed_tables (Enclave/TrustedJoins/secure_joins.cpp:
 * 348-399) without its sealing: out1 = to_relation(join(R, S)), the
 * (key, R.payload) tuple of every match, then join(T, out1) written as
 * (key, T.payload) tuples into out (device, out_cap tuples).  out1 stays in
 * HBM between the two joins (no host round trip).  res describes the second
 * join (ms_total covers both).  algo: a radix-family selector (the
 * reference uses RHO for both joins).
 */
int tbj_join3(int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
              const tbj_tuple* T, uint64_t nT, tbj_tuple* out, uint64_t out_cap,
              const tbj_options* opt, tbj_result* res, void* stream);

/*
 * tbj_join without any device allocation (no counterpart in the reference:
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture, tight serving loops and harnesses that forbid hipMalloc).
 * The same kernels as tbj_join; their workspace comes from one fixed device
 * array of TBJ_NOALLOC_SCRATCH_MB MiB that the library holds on each GPU it
 * has been initialised on, until the process ends.  tbj_init must have run
 * first.  Limits: radix-family algorithms (RHO, RHT, RJ, RHO_atomic, RSM,
 * PSM, MWAY; others TBJ_ERR_UNSUPPORTED), nR + nS <= TBJ_NOALLOC_MAX_TUPLES
 * (TBJ_ERR_TOO_LARGE above; also when a forced radix plan needs more
 * workspace than the array holds).  opt->sampled_regions is ignored.  Every
 * call on one GPU shares that GPU's array, which keeps the last call's
 * contents (each call clears what must start at zero).  Two calls that
 * overlapped on one GPU would race on it; the calls are synchronous and the
 * library holds one lock per GPU around each, so calls from several host
 * threads run one after another.
 */
#define TBJ_NOALLOC_MAX_TUPLES (1u << 22)
#define TBJ_NOALLOC_SCRATCH_MB 80
int tbj_join_noalloc(int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
                     tbj_pair* pairs, uint64_t pairs_cap, const tbj_options* opt,
                     tbj_result* res, void* stream);

/* Same with host pointers: copies R and S to HBM, joins, copies pairs back. */
int tbj_join_host(int algo,
                  const tbj_tuple* R, uint64_t nR,
                  const tbj_tuple* S, uint64_t nS,
                  tbj_pair* pairs, uint64_t pairs_cap,
                  const tbj_options* opt, tbj_result* res);

/*
 * 64-bit keys, 16-byte tuples, sort-merge join (algo TBJ_PSM or TBJ_RSM;
 * other algos return TBJ_ERR_UNSUPPORTED).  No reference counterpart: the
 * reference's tuples are 8 bytes (Include/data-types.h:32-47).
 * TBJ_PSM (or partitioner TBJ_PART_LEARNED) range-partitions by the learned
 * key CDF: materialised pairs come out in key order.  TBJ_RSM with the radix
 * partitioner: key order inside each partition of the low key bits.  Either
 * way the pairs of one key are contiguous; their order inside that block is
 * unspecified.  Limits: nR + nS < 2^32 - 2^20 (TBJ_ERR_TOO_LARGE).
 */
int tbj_join64(int algo,
               const tbj_tuple64* R, uint64_t nR,
               const tbj_tuple64* S, uint64_t nS,
               tbj_pair64* pairs, uint64_t pairs_cap,
               const tbj_options* opt, tbj_result* res, void* stream);

/*
 * Batches of joins (no reference counterpart: the reference runs one join
 * per process, App/native.cpp:121-149).  All relations live in one device
 * arena; item i names R_i = arena[r_off, r_off + nR) and S_i = arena[s_off,
 * s_off + nS).  counts (device, n entries) receives each join's match count.
 * Count-only joins of two kinds: TBJ_RHO (radix partitioning, then LDS hash
 * tables per partition: radix_join.c:1557-1561) and TBJ_PHT (one shared
 * hash table per join in HBM: no_partitioning_join.h:543-623); both count
 * the exact equi-join.  tbj_batch_create's algo is TBJ_RHO or TBJ_PHT for
 * every item, or TBJ_BATCH_PER_ITEM to take each item's own algo field
 * (other values: TBJ_ERR_UNSUPPORTED).
 * The batch is processed in waves of at most 2^31 tuples; each item's R and
 * S must lie within one window of 2^32 - 2^20 tuples.
 *
 * A plan fixes the launch sequence for the batch's sizes; tbj_batch_run
 * launches it (no host synchronisation inside) and, after
 * tbj_batch_capture, replays it as one HIP graph.  The arena contents may
 * change between runs; the sizes and offsets may not.
 */
typedef struct tbj_batch_item {
    uint64_t r_off, nR, s_off, nS;
    int32_t algo;               /* TBJ_RHO or TBJ_PHT, read under TBJ_BATCH_PER_ITEM */
    uint32_t reserved;
} tbj_batch_item;
#define TBJ_BATCH_PER_ITEM (-1)
typedef struct tbj_batch_plan tbj_batch_plan;

typedef struct tbj_batch_tuning {
    uint32_t target_r;          /* average R tuples per partition; 0 = 2048 */
    uint32_t one_pass_bits;     /* most partition bits done in one pass; 0 = 8 */
    uint32_t s_chunk;           /* S tuples per join task (probed 8192 at a time); 0 = 32768 */
    uint32_t wave_tuples;       /* tuples per wave; 0 = 2^31 */
    uint32_t tile;              /* tuples per partition tile: 2048, 4096 or 8192; 0 = 4096 */
} tbj_batch_tuning;

typedef struct tbj_batch_result {
    uint64_t matches;           /* sum over the batch */
    uint64_t tuples;            /* sum of nR + nS */
    uint32_t waves, groups;     /* waves and (wave, radix plan) groups */
    float ms;                   /* device time of the run */
} tbj_batch_result;

int tbj_batch_create(int algo, const tbj_tuple* arena, const tbj_batch_item* items, uint32_t n,
                     uint64_t* counts, const tbj_batch_tuning* tuning, tbj_batch_plan** out);
/* stream: hipStream_t (NULL = the default stream); synchronous. */
int tbj_batch_run(tbj_batch_plan* plan, void* stream, tbj_batch_result* res);
/* Records the plan's launches into a HIP graph (on stream, or on a stream of
 * the library's own when stream is NULL); later runs replay the graph. */
int tbj_batch_capture(tbj_batch_plan* plan, void* stream);
void tbj_batch_destroy(tbj_batch_plan* plan);
/* Times candidate tunings on the batch itself (one warm and `reps` timed
 * runs each) and returns the fastest: first the tile shape (2048, 4096 or
 * 8192 tuples) at the default radix plan, then, at that tile, target_r
 * (which sets each join's radix bits) x one_pass_bits (which sets the pass
 * count) x s_chunk (the join's task shape).  best->wave_tuples is kept. */
int tbj_batch_autotune(int algo, const tbj_tuple* arena, const tbj_batch_item* items, uint32_t n,
                       uint64_t* counts, int reps, tbj_batch_tuning* best, float* best_ms);

#ifdef __cplusplus
}
#endif
#endif
