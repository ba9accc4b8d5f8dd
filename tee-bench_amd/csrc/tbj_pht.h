// The PHT table (no_partitioning_join.h:119-356): a power-of-two array of
// buckets, identity hash h = key & (buckets - 1), two tuples per bucket and
// the rest in overflow chains.  A bucket is one 32-byte record {count,
// overflow head, 2 tuples}; the reference's per-bucket latch becomes an
// atomicAdd on count, and its per-thread overflow arenas one pool claimed
// with atomicAdd.  Shared by the single join and the batch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbj_device.h"

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

// Bulk build of one block of buckets: the tuples Rb[lo, hi) are every tuple
// whose bucket lies in the block (buckets sharing their low pbits bits; the
// block's 2^hbits buckets, hbits <= kPhtPlaceHBits, are told apart by bits
// [pbits, pbits + hbits) of the key) and dst is the block's first bucket.
// The workgroup (kPhtPlaceBlock threads, all of them calling) counts the
// tuples per bucket in LDS, claims pool nodes for the tuples past each
// bucket's two slots with one atomic on *pool_ctr (a bucket's overflow tuples
// take consecutive nodes, chained in order), places the tuples in an LDS copy
// of the block and writes the block out whole.  The buckets hold what per-tuple
// inserts (pht_place) would give them: the same counts and tuples, in another
// order inside a bucket.
constexpr uint32_t kPhtPlaceHBits = 10;
constexpr int kPhtPlaceBlock = 256;

__device__ __forceinline__ void pht_place_block(const uint2* __restrict__ Rb, uint32_t lo,
                                                uint32_t hi, uint32_t pbits, uint32_t hbits,
                                                PhtBucket* __restrict__ dst,
                                                OvNode* __restrict__ pool,
                                                uint32_t* __restrict__ pool_ctr) {
    __shared__ PhtBucket blk[1u << kPhtPlaceHBits];
    __shared__ uint32_t cnt[1u << kPhtPlaceHBits];
    __shared__ uint32_t ovp[1u << kPhtPlaceHBits];
    __shared__ uint32_t pos[1u << kPhtPlaceHBits];
    __shared__ uint32_t scratch[kPhtPlaceBlock / TBJ_WAVE + 1];
    __shared__ uint32_t base_s;
    const uint32_t nb = 1u << hbits, hmask = nb - 1;
    for (uint32_t i = threadIdx.x; i < nb; i += kPhtPlaceBlock) {
        cnt[i] = 0;
        pos[i] = 0;
    }
    __syncthreads();
    // counts (a block-uniform loop: every lane of a wave calls agg_add)
    for (uint32_t j0 = lo; j0 < hi; j0 += kPhtPlaceBlock) {
        const uint32_t j = j0 + threadIdx.x;
        agg_add(cnt, j < hi ? (Rb[j].x >> pbits) & hmask : TBJ_DROP);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += kPhtPlaceBlock) ovp[i] = cnt[i] > 2 ? cnt[i] - 2 : 0;
    __syncthreads();
    const uint32_t nov = block_scan_lds<kPhtPlaceBlock>(ovp, nb, scratch);
    if (threadIdx.x == 0) base_s = nov ? atomicAdd(pool_ctr, nov) : 0;
    __syncthreads();
    const uint32_t base = base_s;
    for (uint32_t i = threadIdx.x; i < nb; i += kPhtPlaceBlock) {
        PhtBucket e;
        e.count = cnt[i];
        e.ovf = cnt[i] > 2 ? base + ovp[i] + 1 : 0;
        e.slot[0] = e.slot[1] = make_uint2(0, 0);
        e.pad[0] = e.pad[1] = 0;
        blk[i] = e;
    }
    __syncthreads();
    for (uint32_t j0 = lo; j0 < hi; j0 += kPhtPlaceBlock) {
        const uint32_t j = j0 + threadIdx.x;
        uint2 t = make_uint2(0, 0);
        uint32_t b = TBJ_DROP;
        if (j < hi) {
            t = Rb[j];
            b = (t.x >> pbits) & hmask;
        }
        const uint32_t r = agg_inc(pos, b);
        if (b == TBJ_DROP) continue;
        if (r < 2) {
            blk[b].slot[r] = t;
        } else {
            // node n; the next one (1-based n + 2) unless this is the
            // bucket's last overflow tuple (r == cnt - 1)
            const uint32_t n = base + ovp[b] + (r - 2);
            OvNode nd;
            nd.t = t;
            nd.next = r + 1 < cnt[b] ? n + 2 : 0;
            nd.pad = 0;
            pool[n] = nd;
        }
    }
    __syncthreads();
    uint4* d = reinterpret_cast<uint4*>(dst);
    const uint4* src = reinterpret_cast<const uint4*>(blk);
    for (uint32_t i = threadIdx.x; i < 2 * nb; i += kPhtPlaceBlock) d[i] = src[i];
}

}  // namespace tbj
