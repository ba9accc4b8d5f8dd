// No-partitioning joins: PHT / NPO_st (shared hash table) and CHT (concise
// hash table), each built in HBM with global atomics and probed by streaming S.
//
// PHT restates no_partitioning_join.h:119-356,219-259 with nextpow2(|R|/2)
// buckets (the table of tbj_pht.h).  Large tables are built in bulk: R is
// partitioned on the buckets' low bits first, and each partition's buckets,
// one contiguous block, are filled in LDS and written whole.
//
// CHT restates CHT.hpp:26-254 / CHTJoin.hpp:46-322: a bitmap of
// nextpow2(|R|) bits in 32-bit words with a running popcount per word, a
// dense tuple array indexed by the popcount of the home bit, one linear-probe
// step inside a partition of the bitmap, and a two-slot probe.  Where the
// reference silently drops a tuple (a third colliding key, CHT.hpp:155-192)
// or aborts (a miss in the probe tail, CHT.hpp:194-206), this version keeps
// the tuple in an overflow hash table, so the result is the exact equi-join.
#include "tbj_device.h"
#include "tbj_kernels.h"
#include "tbj_pht.h"

namespace tbj {

__global__ void k_fill_u4(uint4* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(v, v, v, v);
}

constexpr int kBuildItems = 4;  // inserts per thread, their atomics in flight together

__global__ __launch_bounds__(256) void k_pht_build(const uint2* __restrict__ R, uint64_t nR,
                                                   PhtBucket* __restrict__ tab, uint32_t mask,
                                                   OvNode* __restrict__ pool,
                                                   uint32_t* __restrict__ pool_ctr) {
    const uint64_t base = (uint64_t)blockIdx.x * 256 * kBuildItems + threadIdx.x;
    uint2 t[kBuildItems];
    uint32_t pos[kBuildItems];
#pragma unroll
    for (int k = 0; k < kBuildItems; k++) {
        const uint64_t i = base + (uint64_t)k * 256;
        t[k] = R[i < nR ? i : nR - 1];
    }
#pragma unroll
    for (int k = 0; k < kBuildItems; k++) {
        const uint64_t i = base + (uint64_t)k * 256;
        pos[k] = i < nR ? atomicAdd(&tab[t[k].x & mask].count, 1u) : ~0u;
    }
#pragma unroll
    for (int k = 0; k < kBuildItems; k++) {
        if (pos[k] == ~0u) continue;
        pht_place(tab + (t[k].x & mask), t[k], pos[k], pool, pool_ctr);
    }
}

// Bulk build.  R is partitioned beforehand on the low pbits bits of the key
// (partition p holds the keys whose bits [0, b1) are p / F2 and bits [b1,
// pbits) are p % F2), and the table stores bucket b at pht_slot(b, pbits,
// hbits), so partition p's 2^hbits buckets form the block starting at its
// low bits << hbits; one workgroup per partition (pht_place_block).  Every
// bucket is written, so the table needs no clearing.
__global__ __launch_bounds__(kPhtPlaceBlock) void k_pht_place(
    const uint2* __restrict__ Rb, const uint2* __restrict__ bounds, uint32_t b1, uint32_t F2,
    uint32_t pbits, uint32_t hbits, PhtBucket* __restrict__ tab, OvNode* __restrict__ pool,
    uint32_t* __restrict__ pool_ctr) {
    const uint32_t p = blockIdx.x;
    const uint32_t low = (p / F2) | ((p % F2) << b1);
    const uint2 bd = bounds[p];
    pht_place_block(Rb, bd.x, bd.y, pbits, hbits, tab + ((uint64_t)low << hbits), pool, pool_ctr);
}

// Block-aggregated output: every round each thread handles kProbeItems S
// tuples; all their buckets are loaded before any is used (independent
// loads in flight together), matches are counted, the block claims one
// output range, and the matches are written from what each item keeps of
// its bucket (matching payloads, the overflow chain's head).  The range is
// laid out item-major (probe_offsets), so that each store of a wave writes
// consecutive slots.  Keeping whole buckets live until the writes instead
// costs a wave per SIMD of occupancy (PHT 14.4 ms instead of 12.9 ms on
// 2^29 probes of a 2 GiB table).
constexpr int kProbeBlock = 256;
constexpr int kProbeItems = 16;
constexpr int kProbeWaves = kProbeBlock / TBJ_WAVE;
static_assert(kProbeItems * kProbeWaves <= TBJ_WAVE, "one wave lays out the block's items");

// Output offsets of one round, item-major: the block's pairs of item q follow
// those of items < q, and inside item q wave order, then lane order.  c[q]:
// this thread's matches of item q; o[q] receives the offset of its first
// pair of item q from the block's claim, which wave 0 takes with one atomic
// on *ctr (*obase_s).  wtot: kProbeItems * kProbeWaves words of LDS.  Two
// barriers; every thread of the block calls it.
__device__ __forceinline__ void probe_offsets(const uint32_t (&c)[kProbeItems],
                                              uint32_t (&o)[kProbeItems], uint32_t* wtot,
                                              unsigned long long* ctr,
                                              unsigned long long* obase_s) {
    const uint32_t lane = lane_id(), w = wave_id();
    uint32_t incl[kProbeItems];
#pragma unroll
    for (int q = 0; q < kProbeItems; q++) {
        incl[q] = wave_inclusive_scan(c[q]);
        if (lane == TBJ_WAVE - 1) wtot[w * kProbeItems + q] = incl[q];
    }
    __syncthreads();
    if (w == 0) {
        // lane l = (item l / waves, wave l % waves): item-major order
        constexpr uint32_t n = kProbeItems * kProbeWaves;
        const uint32_t idx = (lane % kProbeWaves) * kProbeItems + lane / kProbeWaves;
        const uint32_t v = lane < n ? wtot[idx] : 0;
        const uint32_t inc = wave_inclusive_scan(v);
        if (lane < n) wtot[idx] = inc - v;
        if (lane == TBJ_WAVE - 1)
            *obase_s = inc ? atomicAdd(ctr, (unsigned long long)inc) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kProbeItems; q++) o[q] = wtot[w * kProbeItems + q] + incl[q] - c[q];
}

template <bool MAT>
__global__ __launch_bounds__(kProbeBlock) void k_pht_probe(
    const uint2* __restrict__ S, uint64_t nS, const PhtBucket* __restrict__ tab, uint32_t mask,
    uint32_t pbits, uint32_t hbits, const OvNode* __restrict__ pool, uint2* __restrict__ pairs,
    uint64_t cap,
    unsigned long long* __restrict__ counters) {
    __shared__ uint32_t wtot[kProbeItems * kProbeWaves];
    __shared__ unsigned long long obase_s;
    __shared__ unsigned long long red_m[kProbeBlock / TBJ_WAVE], red_c[kProbeBlock / TBJ_WAVE];
    uint64_t matches = 0, cks = 0;
    const uint64_t per_round = (uint64_t)kProbeBlock * kProbeItems;
    for (uint64_t base = (uint64_t)blockIdx.x * per_round; base < nS;
         base += (uint64_t)gridDim.x * per_round) {
        uint2 sv[kProbeItems];
        PhtBucket bk[kProbeItems];
#pragma unroll
        for (int k = 0; k < kProbeItems; k++) {
            const uint64_t j = base + k * kProbeBlock + threadIdx.x;
            sv[k] = j < nS ? S[j] : make_uint2(0, 0);
        }
#pragma unroll
        for (int k = 0; k < kProbeItems; k++) {
            const uint64_t j = base + k * kProbeBlock + threadIdx.x;
            if (j < nS) {
                bk[k] = tab[pht_slot(sv[k].x & mask, pbits, hbits)];
            } else {
                bk[k].count = 0;
                bk[k].ovf = 0;
            }
        }
        // per item, only what the writes need: the slots' matches (bits 2k,
        // 2k + 1 of hits), the first matching slot's payload (pa), the
        // second slot's (pb) and the chain's head
        uint32_t cq[kProbeItems], pa[kProbeItems], pb[kProbeItems], ov[kProbeItems];
        uint32_t hits = 0;
#pragma unroll
        for (int k = 0; k < kProbeItems; k++) {
            // slots by constant index, so the buckets stay in registers
            const bool m0 = bk[k].count >= 1 && bk[k].slot[0].x == sv[k].x;
            const bool m1 = bk[k].count >= 2 && bk[k].slot[1].x == sv[k].x;
            hits |= ((m0 ? 1u : 0u) | (m1 ? 2u : 0u)) << (2 * k);
            pa[k] = m0 ? bk[k].slot[0].y : bk[k].slot[1].y;
            pb[k] = bk[k].slot[1].y;
            ov[k] = bk[k].ovf;
            cq[k] = (uint32_t)m0 + (uint32_t)m1;
            if (m0) cks += (uint64_t)bk[k].slot[0].y + sv[k].y;
            if (m1) cks += (uint64_t)bk[k].slot[1].y + sv[k].y;
            for (uint32_t e = ov[k]; e; e = pool[e - 1].next)
                if (pool[e - 1].t.x == sv[k].x) {
                    cq[k]++;
                    cks += (uint64_t)pool[e - 1].t.y + sv[k].y;
                }
            matches += cq[k];
        }
        if (MAT) {
            uint32_t oq[kProbeItems];
            probe_offsets(cq, oq, wtot, &counters[2], &obase_s);
            const uint64_t obase = obase_s;
#pragma unroll
            for (int k = 0; k < kProbeItems; k++) {
                if (!cq[k]) continue;
                const uint32_t h = hits >> (2 * k);
                uint64_t o = obase + oq[k];
                if (h & 3u) {
                    if (o < cap) pairs[o] = make_uint2(pa[k], sv[k].y);
                    o++;
                }
                if ((h & 3u) == 3u) {
                    if (o < cap) pairs[o] = make_uint2(pb[k], sv[k].y);
                    o++;
                }
                for (uint32_t e = ov[k]; e; e = pool[e - 1].next)
                    if (pool[e - 1].t.x == sv[k].x) {
                        if (o < cap) pairs[o] = make_uint2(pool[e - 1].t.y, sv[k].y);
                        o++;
                    }
            }
            __syncthreads();  // obase_s and wtot are reused by the next round
        }
    }
    matches = wave_reduce_u64(matches);
    cks = wave_reduce_u64(cks);
    if (lane_id() == 0) {
        red_m[wave_id()] = matches;
        red_c[wave_id()] = cks;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = 0, cc = 0;
        for (int w = 0; w < kProbeBlock / TBJ_WAVE; w++) {
            m += red_m[w];
            cc += red_c[w];
        }
        if (m) atomicAdd(&counters[0], m);
        if (cc) atomicAdd(&counters[1], cc);
    }
}

// ---------------------------------------------------------------------------
// CHT

struct ChtParams {
    uint32_t bmask;   // bitmap size - 1 (bitmap size = nextpow2(|R|), >= 64)
    uint32_t pmask;   // partition size - 1 (partition = bitmap / 128 bits, >= 32)
    uint32_t ov_mask; // overflow hash table slots - 1
};

__device__ __forceinline__ uint32_t cht_next_in_part(uint32_t pos, uint32_t pmask) {
    return (pos & ~pmask) | ((pos + 1) & pmask);
}

// Phase A: set the home bit, or the next bit inside the partition
// (CHT::setBit).  Tuples that find both taken go to the overflow list.  The
// bitmap lives in words[w].x; words[w].y later receives the rank of the
// word's first bit, so a lookup reads one 8-byte entry.
__global__ __launch_bounds__(256) void k_cht_bits(const uint2* __restrict__ R, uint64_t nR,
                                                  ChtParams p, uint2* __restrict__ words,
                                                  uint32_t* __restrict__ spill,
                                                  ChtState* __restrict__ state) {
    uint32_t* bw = reinterpret_cast<uint32_t*>(words);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nR;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t h = R[i].x & p.bmask;
        const uint32_t m = 1u << (h & 31);
        if (!(atomicOr(&bw[2 * (h >> 5)], m) & m)) continue;
        const uint32_t h2 = cht_next_in_part(h, p.pmask);
        const uint32_t m2 = 1u << (h2 & 31);
        if (!(atomicOr(&bw[2 * (h2 >> 5)], m2) & m2)) continue;
        spill[atomicAdd(&state->spill_ctr, 1u)] = (uint32_t)i;
    }
}

__global__ void k_cht_popc(const uint2* __restrict__ words, uint32_t nwords,
                           uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nwords) cnt[i] = __popc(words[i].x);
    else if (i == nwords) cnt[i] = 0;
}

// words[w].y = rank of word w's first set bit (the scanned popcounts)
__global__ void k_cht_ranks(const uint32_t* __restrict__ cnt, uint32_t nwords,
                            uint2* __restrict__ words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nwords) words[i].y = cnt[i];
}

__device__ __forceinline__ uint32_t cht_pos(const uint2 w, uint32_t h) {
    return w.y + __popc(w.x & ((1u << (h & 31)) - 1u));
}

// Phase C: place each tuple at the dense slot of its home bit or the next one
// (CHT::setTuple); a slot is claimed through a taken-bit.  Tuples already in
// the overflow list from phase A are skipped (marked in `spilled`).
__global__ __launch_bounds__(256) void k_cht_place(const uint2* __restrict__ R, uint64_t nR,
                                                   ChtParams p, const uint2* __restrict__ words,
                                                   const uint32_t* __restrict__ spilled,
                                                   uint32_t* __restrict__ taken,
                                                   uint2* __restrict__ arr,
                                                   uint32_t* __restrict__ spill,
                                                   ChtState* __restrict__ state) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nR;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (spilled[i >> 5] & (1u << (i & 31))) continue;
        const uint2 t = R[i];
        const uint32_t h = t.x & p.bmask;
        const uint32_t pos = cht_pos(words[h >> 5], h);
        uint32_t m = 1u << (pos & 31);
        if (!(atomicOr(&taken[pos >> 5], m) & m)) {
            arr[pos] = t;
            continue;
        }
        const uint32_t pos2 = pos + 1;
        m = 1u << (pos2 & 31);
        if (!(atomicOr(&taken[pos2 >> 5], m) & m)) {
            arr[pos2] = t;
            continue;
        }
        spill[atomicAdd(&state->spill_ctr, 1u)] = (uint32_t)i;
    }
}

// After phase C: the dense array is exactly full when no tuple spilled in
// phase C and none went past the last set bit's slot.  nset tuples then fill
// slots [0, nset) and nothing lies beyond, so a probe can tell a taken slot
// by its index alone.
__global__ void k_cht_dense(const uint32_t* __restrict__ nset_src,
                            const uint32_t* __restrict__ taken, ChtState* __restrict__ state) {
    if (threadIdx.x != 0) return;
    const uint32_t nset = *nset_src;
    const bool past = (taken[nset >> 5] >> (nset & 31)) & 1u;
    state->nset = nset;
    state->dense = (state->spill_ctr == state->spill_a && !past) ? 1u : 0u;
}

// Marks the phase-A spills in `spilled` and records their number.
__global__ void k_cht_mark_spilled(const uint32_t* __restrict__ spill,
                                   ChtState* __restrict__ state, uint32_t* __restrict__ spilled) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = state->spill_ctr;
    if (i == 0) state->spill_a = n;
    if (i < n) atomicOr(&spilled[spill[i] >> 5], 1u << (spill[i] & 31));
}

// ---------------------------------------------------------------------------
// CHT bulk build.  R is partitioned beforehand on the high pbits bits of its
// bitmap position h = key & bmask, so that partition p holds the tuples whose
// home lies in segment q (the p -> q map of seg_of) = positions [q << sbits,
// (q + 1) << sbits); sbits <= kChtSegBits and no segment straddles a CHT
// partition (sbits <= log2 of its size).  Phase A, one workgroup per segment,
// in LDS: set the home bit, or the next one in the CHT partition when that
// one lies in the segment too, else spill; write the segment's bitmap words
// and its number of set bits.  A scan of those numbers gives each segment
// the rank of its first set bit.  Phase C, again per segment: the words'
// ranks, and each tuple at the dense slot of its home bit's rank or the next
// one inside the segment's dense range (staged in LDS, written whole); the
// rest spill.  The result is the structure the per-tuple kernels build -- the
// same bitmap words, ranks and dense slots, read by the same probe -- except
// that a second choice leaving the segment spills to the overflow table.
constexpr uint32_t kChtSegBits = 12;
constexpr int kChtSegBlock = 256;

// Segment of partition p (p = (position bits [sbits, sbits + b1)) * F2 +
// (the bits above)).
__device__ __forceinline__ uint32_t cht_seg_of(uint32_t p, uint32_t b1, uint32_t F2) {
    return (p / F2) | ((p % F2) << b1);
}

__global__ __launch_bounds__(kChtSegBlock) void k_cht_seg_bits(
    const uint2* __restrict__ Rb, const uint2* __restrict__ bounds, uint32_t b1, uint32_t F2,
    uint32_t sbits, ChtParams cp, uint2* __restrict__ words, uint32_t* __restrict__ segcnt,
    uint32_t* __restrict__ spill, uint32_t* __restrict__ spilled, ChtState* __restrict__ state) {
    __shared__ uint32_t bw[1u << (kChtSegBits - 5)];
    __shared__ uint32_t red[kChtSegBlock / TBJ_WAVE];
    const uint32_t p = blockIdx.x, q = cht_seg_of(p, b1, F2);
    const uint32_t nw = 1u << (sbits - 5), smask = (1u << sbits) - 1;
    const uint2 bd = bounds[p];
    for (uint32_t w = threadIdx.x; w < nw; w += kChtSegBlock) bw[w] = 0;
    __syncthreads();
    for (uint32_t j = bd.x + threadIdx.x; j < bd.y; j += kChtSegBlock) {
        const uint32_t h = Rb[j].x & cp.bmask;
        const uint32_t hl = h & smask, m = 1u << (hl & 31);
        if (!(atomicOr(&bw[hl >> 5], m) & m)) continue;
        const uint32_t h2 = cht_next_in_part(h, cp.pmask);
        if ((h2 >> sbits) == q) {
            const uint32_t hl2 = h2 & smask, m2 = 1u << (hl2 & 31);
            if (!(atomicOr(&bw[hl2 >> 5], m2) & m2)) continue;
        }
        spill[atomicAdd(&state->spill_ctr, 1u)] = j;
        atomicOr(&spilled[j >> 5], 1u << (j & 31));
    }
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t w = threadIdx.x; w < nw; w += kChtSegBlock) {
        words[((uint64_t)q << (sbits - 5)) + w].x = bw[w];
        c += __popc(bw[w]);
    }
    c = (uint32_t)wave_reduce_u64(c);
    if (lane_id() == 0) red[wave_id()] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kChtSegBlock / TBJ_WAVE; w++) t += red[w];
        segcnt[q] = t;
    }
}

// Phase A's spills are the overflow table's first entries.
__global__ void k_cht_spill_a(ChtState* __restrict__ state) {
    if (threadIdx.x == 0) state->spill_a = state->spill_ctr;
}

// segbase: the scanned segcnt (segment q's first rank; segbase[nseg] = all
// set bits).
__global__ __launch_bounds__(kChtSegBlock) void k_cht_seg_place(
    const uint2* __restrict__ Rb, const uint2* __restrict__ bounds, uint32_t b1, uint32_t F2,
    uint32_t sbits, ChtParams cp, uint2* __restrict__ words,
    const uint32_t* __restrict__ segbase, const uint32_t* __restrict__ spilled,
    uint32_t* __restrict__ taken, uint2* __restrict__ arr, uint32_t* __restrict__ spill,
    ChtState* __restrict__ state) {
    __shared__ uint32_t bw[1u << (kChtSegBits - 5)], rk[1u << (kChtSegBits - 5)];
    __shared__ uint32_t tk[1u << (kChtSegBits - 5)];
    __shared__ uint2 dense[1u << kChtSegBits];
    __shared__ uint32_t scratch[kChtSegBlock / TBJ_WAVE + 1];
    const uint32_t p = blockIdx.x, q = cht_seg_of(p, b1, F2);
    const uint32_t nw = 1u << (sbits - 5), smask = (1u << sbits) - 1;
    const uint64_t w0 = (uint64_t)q << (sbits - 5);
    const uint32_t base = segbase[q], n = segbase[q + 1] - base;
    const uint2 bd = bounds[p];
    for (uint32_t w = threadIdx.x; w < nw; w += kChtSegBlock) {
        const uint32_t x = words[w0 + w].x;
        bw[w] = x;
        rk[w] = __popc(x);
        tk[w] = 0;
    }
    for (uint32_t i = threadIdx.x; i < n; i += kChtSegBlock) dense[i] = make_uint2(0, 0);
    __syncthreads();
    block_scan_lds<kChtSegBlock>(rk, nw, scratch);  // local ranks; ends with a barrier
    for (uint32_t w = threadIdx.x; w < nw; w += kChtSegBlock)
        words[w0 + w] = make_uint2(bw[w], base + rk[w]);
    for (uint32_t j = bd.x + threadIdx.x; j < bd.y; j += kChtSegBlock) {
        if ((spilled[j >> 5] >> (j & 31)) & 1u) continue;
        const uint2 t = Rb[j];
        const uint32_t hl = t.x & cp.bmask & smask, w = hl >> 5;
        const uint32_t lp = rk[w] + __popc(bw[w] & ((1u << (hl & 31)) - 1u));
        const uint32_t m = 1u << (lp & 31);
        if (!(atomicOr(&tk[lp >> 5], m) & m)) {
            dense[lp] = t;
            continue;
        }
        const uint32_t lp2 = lp + 1;
        if (lp2 < n) {
            const uint32_t m2 = 1u << (lp2 & 31);
            if (!(atomicOr(&tk[lp2 >> 5], m2) & m2)) {
                dense[lp2] = t;
                continue;
            }
        }
        spill[atomicAdd(&state->spill_ctr, 1u)] = j;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kChtSegBlock) arr[base + i] = dense[i];
    // the taken bits of slots [base, base + n): local word lw covers global
    // bits base + 32 lw ..., which straddle two global words unless aligned
    for (uint32_t lw = threadIdx.x; lw * 32 < n; lw += kChtSegBlock) {
        const uint32_t v = tk[lw];
        if (!v) continue;
        const uint32_t g = base + 32 * lw, sh = g & 31;
        atomicOr(&taken[g >> 5], v << sh);
        if (sh && (v >> (32 - sh))) atomicOr(&taken[(g >> 5) + 1], v >> (32 - sh));
    }
}

// Overflow table: open addressing, linear probing, (key, payload) with an
// occupancy word per slot.
__global__ void k_cht_overflow(const uint2* __restrict__ R, const uint32_t* __restrict__ spill,
                               const uint32_t* __restrict__ spill_ctr, uint32_t ov_mask,
                               uint32_t* __restrict__ ov_used, uint2* __restrict__ ov) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *spill_ctr) return;
    const uint2 t = R[spill[i]];
    uint32_t h = (t.x * 2654435761u) & ov_mask;
    while (atomicCAS(&ov_used[h], 0u, 1u) != 0u) h = (h + 1) & ov_mask;
    ov[h] = t;
}

// DENSE: the dense array is exactly full (k_cht_dense), so slot p is taken
// iff p < nset and the taken bitmap is not read.
template <bool MAT, bool DENSE>
__global__ __launch_bounds__(kProbeBlock) void k_cht_probe(
    const uint2* __restrict__ S, uint64_t nS, ChtParams p, const uint2* __restrict__ words,
    const uint32_t* __restrict__ taken, const uint2* __restrict__ arr, uint32_t nslots,
    const uint32_t* __restrict__ ov_used, const uint2* __restrict__ ov,
    const ChtState* __restrict__ state, uint2* __restrict__ pairs, uint64_t cap,
    unsigned long long* __restrict__ counters) {
    __shared__ uint32_t wtot[kProbeItems * kProbeWaves];
    __shared__ unsigned long long obase_s;
    __shared__ unsigned long long red_m[kProbeBlock / TBJ_WAVE], red_c[kProbeBlock / TBJ_WAVE];
    const bool any_ov = state->spill_ctr != 0;
    const uint32_t nset = DENSE ? state->nset : 0;
    uint64_t matches = 0, cks = 0;
    const uint64_t per_round = (uint64_t)kProbeBlock * kProbeItems;
    for (uint64_t base = (uint64_t)blockIdx.x * per_round; base < nS;
         base += (uint64_t)gridDim.x * per_round) {
        // every level of the lookup is issued for all items before any is
        // used: S, then the bitmap word with its rank, then the two dense
        // slots (and their taken bits unless DENSE)
        uint2 sv[kProbeItems], bw[kProbeItems];
        uint32_t pos[kProbeItems];
        uint2 a0[kProbeItems], a1[kProbeItems];
        uint32_t t0[kProbeItems], t1[kProbeItems];
        uint32_t cq[kProbeItems];
#pragma unroll
        for (int k = 0; k < kProbeItems; k++) {
            const uint64_t j = base + k * kProbeBlock + threadIdx.x;
            sv[k] = j < nS ? S[j] : make_uint2(0, 0);
        }
#pragma unroll
        for (int k = 0; k < kProbeItems; k++) bw[k] = words[(sv[k].x & p.bmask) >> 5];
#pragma unroll
        for (int k = 0; k < kProbeItems; k++) {
            pos[k] = cht_pos(bw[k], sv[k].x & p.bmask);
            const uint32_t p0 = min(pos[k], nslots - 1), p1 = min(pos[k] + 1, nslots - 1);
            a0[k] = arr[p0];
            a1[k] = arr[p1];
            if (!DENSE) {
                t0[k] = taken[p0 >> 5];
                t1[k] = taken[p1 >> 5];
            }
        }
        uint32_t hit[kProbeItems];  // bit 0: slot pos matches, bit 1: slot pos + 1
#pragma unroll
        for (int k = 0; k < kProbeItems; k++) {
            const uint64_t j = base + k * kProbeBlock + threadIdx.x;
            const uint32_t p0 = pos[k], p1 = pos[k] + 1;
            const bool ok = j < nS;
            bool u0, u1;  // slot taken
            if (DENSE) {
                u0 = p0 < nset;
                u1 = p1 < nset;
            } else {
                u0 = p0 < nslots && ((t0[k] >> (p0 & 31)) & 1u);
                u1 = p1 < nslots && ((t1[k] >> (p1 & 31)) & 1u);
            }
            const bool m0 = ok && u0 && a0[k].x == sv[k].x;
            const bool m1 = ok && u1 && a1[k].x == sv[k].x;
            hit[k] = (m0 ? 1u : 0u) | (m1 ? 2u : 0u);
            if (m0) cks += (uint64_t)a0[k].y + sv[k].y;
            if (m1) cks += (uint64_t)a1[k].y + sv[k].y;
            cq[k] = (uint32_t)m0 + (uint32_t)m1;
            if (any_ov && ok) {
                for (uint32_t e = (sv[k].x * 2654435761u) & p.ov_mask; ov_used[e];
                     e = (e + 1) & p.ov_mask)
                    if (ov[e].x == sv[k].x) {
                        cq[k]++;
                        cks += (uint64_t)ov[e].y + sv[k].y;
                    }
            }
            matches += cq[k];
        }
        if (MAT) {
            uint32_t oq[kProbeItems];
            probe_offsets(cq, oq, wtot, &counters[2], &obase_s);
            const uint64_t obase = obase_s;
#pragma unroll
            for (int k = 0; k < kProbeItems; k++) {
                if (!cq[k]) continue;
                const uint64_t j = base + k * kProbeBlock + threadIdx.x;
                uint64_t o = obase + oq[k];
                if (hit[k] & 1u) {
                    if (o < cap) pairs[o] = make_uint2(a0[k].y, sv[k].y);
                    o++;
                }
                if (hit[k] & 2u) {
                    if (o < cap) pairs[o] = make_uint2(a1[k].y, sv[k].y);
                    o++;
                }
                if (any_ov && j < nS) {
                    for (uint32_t e = (sv[k].x * 2654435761u) & p.ov_mask; ov_used[e];
                         e = (e + 1) & p.ov_mask)
                        if (ov[e].x == sv[k].x) {
                            if (o < cap) pairs[o] = make_uint2(ov[e].y, sv[k].y);
                            o++;
                        }
                }
            }
            __syncthreads();  // obase_s and wtot are reused by the next round
        }
    }
    matches = wave_reduce_u64(matches);
    cks = wave_reduce_u64(cks);
    if (lane_id() == 0) {
        red_m[wave_id()] = matches;
        red_c[wave_id()] = cks;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = 0, cc = 0;
        for (int w = 0; w < kProbeBlock / TBJ_WAVE; w++) {
            m += red_m[w];
            cc += red_c[w];
        }
        if (m) atomicAdd(&counters[0], m);
        if (cc) atomicAdd(&counters[1], cc);
    }
}

// ---------------------------------------------------------------------------
// Host launchers

static uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap_blocks) {
    uint64_t g = (n + block - 1) / block;
    if (g > cap_blocks) g = cap_blocks;
    return g ? (uint32_t)g : 1;
}

void launch_fill_u4(hipStream_t st, void* p, uint64_t bytes, uint32_t v) {
    const uint64_t n = bytes / 16;
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u4, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, (uint4*)p, n, v);
}

size_t pht_bucket_bytes() { return sizeof(PhtBucket); }
size_t pht_node_bytes() { return sizeof(OvNode); }

void launch_pht_build(hipStream_t st, const uint2* R, uint64_t nR, void* tab, uint32_t mask,
                      void* pool, uint32_t* pool_ctr) {
    hipLaunchKernelGGL(k_pht_build, dim3((uint32_t)((nR + 256 * kBuildItems - 1) / (256 * kBuildItems))),
                       dim3(256), 0, st, R, nR, (PhtBucket*)tab, mask, (OvNode*)pool, pool_ctr);
}

void launch_pht_place(hipStream_t st, const uint2* Rb, const uint2* bounds, uint32_t b1,
                      uint32_t pbits, uint32_t hbits, void* tab, void* pool, uint32_t* pool_ctr) {
    const uint32_t F2 = 1u << (pbits - b1);
    hipLaunchKernelGGL(k_pht_place, dim3(1u << pbits), dim3(kPhtPlaceBlock), 0, st, Rb, bounds, b1,
                       F2, pbits, hbits, (PhtBucket*)tab, (OvNode*)pool, pool_ctr);
}

void launch_pht_probe(hipStream_t st, bool mat, const uint2* S, uint64_t nS, const void* tab,
                      uint32_t mask, uint32_t pbits, uint32_t hbits, const void* pool,
                      uint2* pairs, uint64_t cap, unsigned long long* counters) {
    const uint32_t g = grid_for(nS, kProbeBlock * kProbeItems, 16384);
    if (mat)
        hipLaunchKernelGGL(k_pht_probe<true>, dim3(g), dim3(kProbeBlock), 0, st, S, nS,
                           (const PhtBucket*)tab, mask, pbits, hbits, (const OvNode*)pool, pairs,
                           cap, counters);
    else
        hipLaunchKernelGGL(k_pht_probe<false>, dim3(g), dim3(kProbeBlock), 0, st, S, nS,
                           (const PhtBucket*)tab, mask, pbits, hbits, (const OvNode*)pool, pairs,
                           cap, counters);
}

void launch_cht_build(hipStream_t st, const uint2* R, uint64_t nR, uint32_t bmask,
                      uint32_t pmask, uint2* words, uint32_t* cnt, uint32_t* partial,
                      uint32_t* taken, uint2* arr, uint32_t* spill, uint32_t* spilled,
                      ChtState* state) {
    ChtParams p{bmask, pmask, 0};
    const uint32_t nwords = (bmask + 1) / 32;
    hipLaunchKernelGGL(k_cht_bits, dim3(grid_for(nR, 256, 16384)), dim3(256), 0, st, R, nR, p,
                       words, spill, state);
    hipLaunchKernelGGL(k_cht_popc, dim3((nwords + 1 + 255) / 256), dim3(256), 0, st, words,
                       nwords, cnt);
    launch_scan(st, cnt, (uint64_t)nwords + 1, partial, nullptr);
    hipLaunchKernelGGL(k_cht_ranks, dim3((nwords + 255) / 256), dim3(256), 0, st, cnt, nwords,
                       words);
    hipLaunchKernelGGL(k_cht_mark_spilled, dim3(grid_for(nR, 256, 1u << 30)), dim3(256), 0, st,
                       spill, state, spilled);
    hipLaunchKernelGGL(k_cht_place, dim3(grid_for(nR, 256, 16384)), dim3(256), 0, st, R, nR, p,
                       words, spilled, taken, arr, spill, state);
    hipLaunchKernelGGL(k_cht_dense, dim3(1), dim3(64), 0, st, cnt + nwords, taken, state);
}

void launch_cht_build_bulk(hipStream_t st, const uint2* Rb, const uint2* bounds, uint32_t b1,
                           uint32_t pbits, uint32_t sbits, uint32_t bmask, uint32_t pmask,
                           uint2* words, uint32_t* segcnt, uint32_t* partial, uint32_t* taken,
                           uint2* arr, uint32_t* spill, uint32_t* spilled, ChtState* state) {
    ChtParams p{bmask, pmask, 0};
    const uint32_t nseg = 1u << pbits, F2 = 1u << (pbits - b1);
    hipLaunchKernelGGL(k_cht_seg_bits, dim3(nseg), dim3(kChtSegBlock), 0, st, Rb, bounds, b1, F2,
                       sbits, p, words, segcnt, spill, spilled, state);
    hipLaunchKernelGGL(k_cht_spill_a, dim3(1), dim3(64), 0, st, state);
    launch_fill_u32(st, segcnt + nseg, 1, 0);
    launch_scan(st, segcnt, (uint64_t)nseg + 1, partial, &state->nset);
    hipLaunchKernelGGL(k_cht_seg_place, dim3(nseg), dim3(kChtSegBlock), 0, st, Rb, bounds, b1, F2,
                       sbits, p, words, segcnt, spilled, taken, arr, spill, state);
    hipLaunchKernelGGL(k_cht_dense, dim3(1), dim3(64), 0, st, segcnt + nseg, taken, state);
}

void launch_cht_overflow(hipStream_t st, const uint2* R, const uint32_t* spill,
                         const ChtState* state, uint32_t nspill, uint32_t ov_mask,
                         uint32_t* ov_used, uint2* ov) {
    if (!nspill) return;
    hipLaunchKernelGGL(k_cht_overflow, dim3((nspill + 255) / 256), dim3(256), 0, st, R, spill,
                       &state->spill_ctr, ov_mask, ov_used, ov);
}

void launch_cht_probe(hipStream_t st, bool mat, bool dense, const uint2* S, uint64_t nS,
                      uint32_t bmask, uint32_t pmask, uint32_t ov_mask, const uint2* words,
                      const uint32_t* taken, const uint2* arr, uint32_t nslots,
                      const uint32_t* ov_used, const uint2* ov, const ChtState* state,
                      uint2* pairs, uint64_t cap, unsigned long long* counters) {
    ChtParams p{bmask, pmask, ov_mask};
    const uint32_t g = grid_for(nS, kProbeBlock * kProbeItems, 16384);
#define TBJ_CHT_PROBE(M, D)                                                                      \
    hipLaunchKernelGGL((k_cht_probe<M, D>), dim3(g), dim3(kProbeBlock), 0, st, S, nS, p, words, \
                       taken, arr, nslots, ov_used, ov, state, pairs, cap, counters)
    if (mat && dense) TBJ_CHT_PROBE(true, true);
    else if (mat) TBJ_CHT_PROBE(true, false);
    else if (dense) TBJ_CHT_PROBE(false, true);
    else TBJ_CHT_PROBE(false, false);
#undef TBJ_CHT_PROBE
}

}  // namespace tbj
