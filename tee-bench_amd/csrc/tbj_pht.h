// The PHT table (no_partitioning_join.h:119-356): a power-of-two array of
// buckets, identity hash h = key & (buckets - 1), two tuples per bucket and
// the rest in overflow chains.  A bucket is one 32-byte record {count,
// overflow head, 2 tuples}; the reference's per-bucket latch becomes an
// atomicAdd on count, and its per-thread overflow arenas one pool claimed
// with atomicAdd.  Shared by the single join and the batch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbj {

struct __align__(32) PhtBucket {
    uint32_t count;
    uint32_t ovf;  // 1-based index into the overflow pool, 0 = none
    uint2 slot[2];
    uint32_t pad[2];
};

struct __align__(16) OvNode {
    uint2 t;
    uint32_t next;  // 1-based
    uint32_t pad;
};

// Where bucket b sits in the table.  A table built in bulk (from R
// partitioned on the bucket's low pbits bits) keeps the 2^hbits buckets of
// one partition as one contiguous block: b -> ((b & (2^pbits - 1)) << hbits)
// | (b >> pbits).  pbits = 0 is the plain array (b -> b).
__device__ __forceinline__ uint32_t pht_slot(uint32_t b, uint32_t pbits, uint32_t hbits) {
    return ((b & ((1u << pbits) - 1u)) << hbits) | (b >> pbits);
}

// Stores tuple t in bucket b, where it took position pos (the bucket's count
// before this tuple's atomicAdd): one of the two slots, else a pool node at
// the head of the bucket's overflow chain.
__device__ __forceinline__ void pht_place(PhtBucket* b, uint2 t, uint32_t pos,
                                          OvNode* __restrict__ pool,
                                          uint32_t* __restrict__ pool_ctr) {
    if (pos < 2) {
        b->slot[pos] = t;
    } else {
        const uint32_t n = atomicAdd(pool_ctr, 1u);
        pool[n].t = t;
        pool[n].next = atomicExch(&b->ovf, n + 1);
    }
}

// Matches of key in a loaded bucket and its overflow chain.
__device__ __forceinline__ uint32_t pht_matches(const PhtBucket& bk, uint32_t key,
                                                const OvNode* __restrict__ pool) {
    // slots by constant index, so the bucket stays in registers
    uint32_t c = (bk.count >= 1 && bk.slot[0].x == key) + (bk.count >= 2 && bk.slot[1].x == key);
    for (uint32_t e = bk.ovf; e; e = pool[e - 1].next) c += pool[e - 1].t.x == key;
    return c;
}

}  // namespace tbj
