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
// array (laid out by pht_slot: blocks of 2^hbits buckets, one per value of
// the low pbits bits) and its first chunk in the wave's probe launch.
struct PhtJob {
    uint64_t r_off, s_off;  // arena tuple offsets
    uint64_t tab;           // first bucket of the join's table
    uint32_t nR, nS;
    uint32_t mask;          // table buckets - 1
    uint32_t id;            // the join's index in the batch (counts)
    uint32_t sc0;           // first probe chunk of the join
    uint32_t pbits, hbits;  // table layout: 2^(pbits + hbits) buckets
    uint32_t pad;
};
constexpr uint32_t kPhtProbeChunk = 8192;  // S tuples per probe chunk

// Bulk build of the tables of njobs PHT joins (one part: joins with the same
// pbits).  pbits > 0: src holds their R ranges partitioned in one pass on the
// low pbits bits, segment j = jobs[j]'s R (segs, gsum: the pass's segments and
// scanned group sums).  pbits = 0: src is the arena and each join's table is
// one block built from its R directly.  Overflow nodes come from pool through
// *pool_ctr (zeroed before the wave's first part).  Every bucket is written.
void batch_pht_place(hipStream_t st, const uint2* src, const PhtJob* jobs, uint32_t njobs,
                     const SegDesc* segs, const uint32_t* gsum, uint32_t pbits, void* tab,
                     void* pool, uint32_t* pool_ctr);
// Count-only probe of a wave's PHT joins (nsc chunks in all); adds each
// join's matches to counts[id].
void batch_pht_probe(hipStream_t st, uint32_t grid, const uint2* arena, const PhtJob* jobs,
                     uint32_t njobs, uint32_t nsc, const void* tab, const void* pool,
                     unsigned long long* counts);
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
