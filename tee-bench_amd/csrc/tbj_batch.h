// Batched joins: device types and launchers shared by tbj_batch.hip and the
// batch planner.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbj_part.h"

namespace tbj {

// One join of a generated batch, as the generator kernel reads it.
struct GenJob {
    uint64_t start;         // first global tuple index of the job (R then S)
    uint64_t seed;
    uint64_t nR, nS;
    uint64_t r_off, s_off;  // arena tuple offsets
    double z;               // 0: uniform keys in [1, nR]
    double hx1, hn, sp;     // rejection-inversion constants for z > 0
    uint32_t fbits;         // half width of the Feistel domain
    uint32_t pad;
};

// One PHT join of a batch: its relations, its table in the wave's bucket
// array and its first chunks in the wave's build and probe launches.
struct PhtJob {
    uint64_t r_off, s_off;  // arena tuple offsets
    uint64_t tab;           // first bucket of the join's table
    uint32_t nR, nS;
    uint32_t mask;          // table buckets - 1
    uint32_t id;            // the join's index in the batch (counts)
    uint32_t rc0, sc0;      // first build / probe chunk of the join
};
constexpr uint32_t kPhtBuildChunk = 2048;  // R tuples per build chunk
constexpr uint32_t kPhtProbeChunk = 8192;  // S tuples per probe chunk

// Count-only PHT joins of one wave: clears the njobs tables (nbuckets in
// all) and the pool counter, builds (nrc chunks), probes (nsc chunks) and
// adds each join's matches to counts[id].  pool: one node per R tuple.
void batch_pht(hipStream_t st, uint32_t grid, const uint2* arena, const PhtJob* jobs,
               uint32_t njobs, uint32_t nrc, uint32_t nsc, void* tab, uint64_t nbuckets,
               void* pool, uint32_t* pool_ctr, unsigned long long* counts);
void batch_bounds(hipStream_t st, const uint32_t* scanned, const SegDesc* segs, uint32_t nj,
                  uint32_t F1, uint32_t F2, uint32_t passes, uint2* bounds);
void batch_join(hipStream_t st, uint32_t grid, const uint2* Rb, const uint2* Sb,
                const uint2* bounds, uint32_t NP, uint32_t P, const uint32_t* task_off,
                const uint32_t* ntasks_dev, uint32_t* task_ctr, uint32_t s_chunk, uint32_t hshift,
                const uint32_t* join_ids, unsigned long long* counts);
void batch_zero_u64(hipStream_t st, unsigned long long* p, uint64_t n);
void batch_sum(hipStream_t st, const unsigned long long* a, uint32_t n, unsigned long long* out);
void batch_generate(hipStream_t st, uint2* arena, const GenJob* jobs, uint32_t njobs,
                    uint64_t total);

}  // namespace tbj
