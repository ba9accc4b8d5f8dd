/*
 * teebench/data_types.h — the reference's join data model, layout-identical
 * so that code written against the reference's Include/data-types.h:32-93
 * (row_t, table_t = relation_t, result_t, threadresult_t, output_list_t,
 * algorithm_t) compiles and links against this library unchanged.
 */
#ifndef TEEBENCH_DATA_TYPES_H
#define TEEBENCH_DATA_TYPES_H
#include <stdint.h>

typedef uint32_t type_key;
typedef uint32_t type_value;

typedef struct row_t tuple_t;
typedef struct table_t relation_t;
typedef struct result_t result_t;
typedef struct threadresult_t threadresult_t;
typedef struct output_list_t output_list_t;

/* 8-byte tuple (Include/data-types.h:44-47). */
struct row_t {
    type_key key;
    type_value payload;
};

/* A relation (Include/data-types.h:49-54). */
struct table_t {
    struct row_t* tuples;
    uint64_t num_tuples;
    int ratio_holes;
    int sorted;
};

/* One materialised match (Include/data-types.h:56-61). */
struct output_list_t {
    type_key key;
    type_value Rpayload;
    type_value Spayload;
    struct output_list_t* next;
};

/* Join entry in a -a selector table (Include/data-types.h:68-71). */
typedef struct algorithm_t {
    char name[128];
    result_t* (*join)(struct table_t*, struct table_t*, int);
} algorithm_t;

/* Per-thread results (Include/data-types.h:73-78). */
struct threadresult_t {
    int64_t nresults;
    output_list_t* results;
    uint32_t threadid;
};

/* Join result (Include/data-types.h:80-85). */
struct result_t {
    int64_t totalresults;
    struct threadresult_t* resultlist;
    int nthreads;
};

#endif
