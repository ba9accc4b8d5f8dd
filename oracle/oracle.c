/*
 * ORACLE — test infrastructure only (see oracle.h).  A CPU restatement of
 * the reference's equi-join path, used as the checker of the HIP product.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* glibc rand(): TYPE_3, degree 31, separation 3.  srand() seeds the table
 * with the Park-Miller step 16807 (Schrage's method), then discards 310
 * outputs; each output is (tbl[f] += tbl[r]) >> 1.  This is the published
 * glibc random_r algorithm that srand/rand in generator.cpp:76-82,20 and
 * genzipf.c:44,122 resolve to. */
void orc_srand(orc_rng* g, uint32_t seed) {
    if (seed == 0) seed = 1;
    g->tbl[0] = (int32_t)seed;
    int32_t word = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
        long hi = word / 127773;
        long lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        g->tbl[i] = word;
    }
    g->f = 3;
    g->r = 0;
    for (int i = 0; i < 310; i++) (void)orc_rand(g);
}

int32_t orc_rand(orc_rng* g) {
    uint32_t v = (uint32_t)g->tbl[g->f] + (uint32_t)g->tbl[g->r];
    g->tbl[g->f] = (int32_t)v;
    if (++g->f >= 31) {
        g->f = 0;
        ++g->r;
    } else if (++g->r >= 31) {
        g->r = 0;
    }
    return (int32_t)(v >> 1);
}

void orc_rand_stream(uint32_t seed, uint64_t n, int32_t* out) {
    orc_rng g;
    orc_srand(&g, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = orc_rand(&g);
}

#define ORC_RAND_MAX 2147483647

/* ------------------------------------------------------------------------ */
/* Generators */

/* knuth_shuffle, generator.cpp:101-111: for i = n-1 .. 1,
 * j = (int64)(rand() / (RAND_MAX + 1.0) * i), swap keys i and j. */
static void knuth_shuffle(orc_rng* g, uint32_t* k, uint64_t n) {
    if (n < 2) return;
    for (uint64_t i = n - 1; i > 0; i--) {
        int64_t j = (int64_t)((double)orc_rand(g) / ((double)ORC_RAND_MAX + 1) * (double)i);
        uint32_t t = k[i];
        k[i] = k[j];
        k[j] = t;
    }
}

/* random_unique_gen, generator.cpp:144-155: keys 1..n then shuffle. */
static void unique_gen(orc_rng* g, uint32_t* k, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) k[i] = (uint32_t)(i + 1);
    knuth_shuffle(g, k, n);
}

/* random_unique_gen_maxid, generator.cpp:157-171: jump is the integer
 * quotient maxid / n held in a double; id starts at 1 (0 when maxid == 0). */
static void unique_gen_maxid(orc_rng* g, uint32_t* k, uint64_t n, uint32_t maxid) {
    double jump = (double)(maxid / n);
    double id = maxid == 0 ? 0 : 1;
    for (uint64_t i = 0; i < n; i++) {
        k[i] = (uint32_t)id;
        id += jump;
    }
    knuth_shuffle(g, k, n);
}

/* create_relation_pk, generator.cpp:353-379 (seeded as native.cpp:82). */
void orc_gen_pk(uint32_t seed, uint64_t n, uint32_t* keys) {
    orc_rng g;
    orc_srand(&g, seed);
    unique_gen(&g, keys, n);
}

/* create_relation_fk, generator.cpp:475-514: n / maxid shuffled blocks of
 * 1..maxid, then a shuffled remainder block of 1..(n mod maxid). */
void orc_gen_fk(uint32_t seed, uint64_t n, int64_t maxid, uint32_t* keys) {
    orc_rng g;
    orc_srand(&g, seed);
    int64_t iters = (int64_t)n / maxid;
    for (int64_t i = 0; i < iters; i++) unique_gen(&g, keys + maxid * i, (uint64_t)maxid);
    int64_t rem = (int64_t)n % maxid;
    if (rem > 0) unique_gen(&g, keys + maxid * iters, (uint64_t)rem);
}

/* create_relation_fk_sel, generator.cpp:516-555. */
void orc_gen_fk_sel(uint32_t seed, uint64_t n, int64_t maxid, uint32_t* keys) {
    orc_rng g;
    orc_srand(&g, seed);
    int64_t iters = maxid != 0 ? (int64_t)n / maxid : 0;
    for (int64_t i = 0; i < iters; i++)
        unique_gen_maxid(&g, keys + maxid * i, (uint64_t)maxid, (uint32_t)maxid);
    int64_t rem = maxid != 0 ? (int64_t)n % maxid : (int64_t)n;
    if (rem > 0) unique_gen_maxid(&g, keys + maxid * iters, (uint64_t)rem, (uint32_t)maxid);
}

/* gen_zipf, genzipf.c:28-158 (via create_relation_zipf, generator.cpp:639-662).
 * The alphabet permutation multiplies i * rand() in unsigned 32-bit
 * arithmetic (genzipf.c:44), so k only takes the values 0, 1 or 2. */
int orc_gen_zipf(uint32_t seed, uint64_t n, uint32_t alphabet, double z, uint32_t* keys) {
    orc_rng g;
    orc_srand(&g, seed);
    uint32_t* alpha = (uint32_t*)malloc((size_t)alphabet * sizeof(uint32_t));
    double* lut = (double*)malloc((size_t)alphabet * sizeof(double));
    if (!alpha || !lut) {
        free(alpha);
        free(lut);
        return -1;
    }
    for (uint32_t i = 0; i < alphabet; i++) alpha[i] = i + 1;
    for (uint32_t i = alphabet - 1; i > 0; i--) {
        uint32_t k = (uint32_t)(i * (uint32_t)orc_rand(&g)) / (uint32_t)ORC_RAND_MAX;
        uint32_t t = alpha[i];
        alpha[i] = alpha[k];
        alpha[k] = t;
    }
    double scale = 0.0;
    for (uint32_t i = 1; i <= alphabet; i++) scale += 1.0 / pow(i, z);
    double sum = 0.0;
    for (uint32_t i = 1; i <= alphabet; i++) {
        sum += 1.0 / pow(i, z);
        lut[i - 1] = sum / scale;
    }
    for (uint64_t i = 0; i < n; i++) {
        double r = ((double)orc_rand(&g)) / ORC_RAND_MAX;
        uint32_t left = 0, right = alphabet - 1, pos;
        if (lut[0] >= r) {
            pos = 0;
        } else {
            while (right - left > 1) {
                uint32_t m = (left + right) / 2;
                if (lut[m] < r) left = m;
                else right = m;
            }
            pos = right;
        }
        keys[i] = alpha[pos];
    }
    free(alpha);
    free(lut);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Radix join restatement (radix_join.c) */

#define ORC_RADIX_BITS 14          /* prj_params.h:18 */
#define ORC_PASS1_BITS 7           /* NUM_RADIX_BITS / NUM_PASSES */
#define ORC_PASS2_BITS 7
#define ORC_SMALL_PAD 24           /* prj_params.h: 3*64/8 */
#define ORC_PAD (ORC_SMALL_PAD * ((1 << ORC_PASS2_BITS) + 1))  /* 3096 */

typedef struct { uint32_t key, pay; } orc_tuple;

static uint32_t next_pow2(uint32_t v) {
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    v++;
    return v;
}

/* radix_cluster, radix_join.c:524-562: histogram of ((key & M) >> R),
 * offsets with `pad` tuples between clusters, stable scatter. hist[] gets
 * the per-cluster counts, start[] the first slot of each cluster. */
static void radix_cluster(orc_tuple* out, const orc_tuple* in, uint64_t n, int R, int D,
                          uint32_t pad, uint64_t* hist, uint64_t* start) {
    uint32_t fan = 1u << D, M = (fan - 1) << R;
    memset(hist, 0, fan * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) hist[(in[i].key & M) >> R]++;
    uint64_t off = 0;
    uint64_t* dst = (uint64_t*)malloc(fan * sizeof(uint64_t));
    for (uint32_t i = 0; i < fan; i++) {
        dst[i] = off + (uint64_t)i * pad;
        start[i] = dst[i];
        off += hist[i];
    }
    for (uint64_t i = 0; i < n; i++) out[dst[(in[i].key & M) >> R]++] = in[i];
    free(dst);
}

typedef struct {
    uint32_t* orp;
    uint32_t* osp;
    uint64_t cap;
    uint64_t n;
} orc_out;

static void emit(orc_out* o, uint32_t rp, uint32_t sp) {
    if (o->orp && o->n < o->cap) {
        o->orp[o->n] = rp;
        o->osp[o->n] = sp;
    }
    o->n++;
}

/* bucket_chaining_join, radix_join.c:186-244: N = nextpow2(|R_p|), bucket
 * index (key & ((N-1) << 14)) >> 14, chains through next[] (1-based). */
static void bucket_chaining(const orc_tuple* R, uint64_t nr, const orc_tuple* S, uint64_t ns,
                            orc_out* o) {
    uint32_t N = next_pow2((uint32_t)nr);
    uint32_t MASK = (N - 1) << ORC_RADIX_BITS;
    int32_t* next = (int32_t*)malloc(nr * sizeof(int32_t));
    int32_t* bucket = (int32_t*)calloc(N, sizeof(int32_t));
    for (uint32_t i = 0; i < nr;) {
        uint32_t idx = (R[i].key & MASK) >> ORC_RADIX_BITS;
        next[i] = bucket[idx];
        bucket[idx] = (int32_t)++i;
    }
    for (uint64_t i = 0; i < ns; i++) {
        uint32_t idx = (S[i].key & MASK) >> ORC_RADIX_BITS;
        for (int32_t hit = bucket[idx]; hit > 0; hit = next[hit - 1])
            if (S[i].key == R[hit - 1].key) emit(o, R[hit - 1].pay, S[i].pay);
    }
    free(next);
    free(bucket);
}

/* histogram_join (RHT), radix_join.c:353-440: Nhist = max(nextpow2(|R_p|)/4, 4)
 * buckets on bits 14.., R_p reordered by bucket, S probes bucket ranges. */
static void histogram_join(const orc_tuple* R, uint64_t nr, const orc_tuple* S, uint64_t ns,
                           orc_out* o) {
    uint32_t nh = next_pow2((uint32_t)nr) >> 2;
    if (nh < 4) nh = 4;
    uint32_t MASK = (nh - 1) << ORC_RADIX_BITS;
    int32_t* hist = (int32_t*)calloc(nh + 2, sizeof(int32_t));
    orc_tuple* tmp = (orc_tuple*)malloc(nr * sizeof(orc_tuple));
    for (uint64_t i = 0; i < nr; i++) hist[((R[i].key & MASK) >> ORC_RADIX_BITS) + 2]++;
    for (uint32_t i = 2, sum = 0; i <= nh + 1; i++) {
        sum += hist[i];
        hist[i] = (int32_t)sum;
    }
    for (uint64_t i = 0; i < nr; i++) {
        uint32_t idx = ((R[i].key & MASK) >> ORC_RADIX_BITS) + 1;
        tmp[hist[idx]++] = R[i];
    }
    for (uint64_t i = 0; i < ns; i++) {
        uint32_t idx = (S[i].key & MASK) >> ORC_RADIX_BITS;
        for (int32_t j = hist[idx]; j < hist[idx + 1]; j++)
            if (S[i].key == tmp[j].key) emit(o, tmp[j].pay, S[i].pay);
    }
    free(hist);
    free(tmp);
}

typedef void (*orc_joinfn)(const orc_tuple*, uint64_t, const orc_tuple*, uint64_t, orc_out*);

/* join_init_run + prj_thread with one thread (radix_join.c:839-1396):
 * pass 1 on bits 0..6 with PADDING_TUPLES gaps, pass 2 on bits 7..13 with
 * SMALL_PADDING_TUPLES gaps for partitions where both sides are non-empty,
 * then the per-partition join for sub-partitions non-empty on both sides. */
static int64_t radix_join(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                          const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                          uint32_t* orp, uint32_t* osp, uint64_t cap, orc_joinfn jf) {
    const uint32_t F1 = 1u << ORC_PASS1_BITS, F2 = 1u << ORC_PASS2_BITS;
    uint64_t padn = (uint64_t)ORC_PAD * F1;
    orc_tuple* R = (orc_tuple*)calloc(nr + padn, sizeof(orc_tuple));
    orc_tuple* S = (orc_tuple*)calloc(ns + padn, sizeof(orc_tuple));
    orc_tuple* tR = (orc_tuple*)calloc(nr + padn, sizeof(orc_tuple));
    orc_tuple* tS = (orc_tuple*)calloc(ns + padn, sizeof(orc_tuple));
    for (uint64_t i = 0; i < nr; i++) {
        R[i].key = rk[i];
        R[i].pay = rp ? rp[i] : 0;
    }
    for (uint64_t i = 0; i < ns; i++) {
        S[i].key = sk[i];
        S[i].pay = sp ? sp[i] : 0;
    }
    uint64_t hR1[128], hS1[128], stR1[128], stS1[128];
    radix_cluster(tR, R, nr, 0, ORC_PASS1_BITS, ORC_PAD, hR1, stR1);
    radix_cluster(tS, S, ns, 0, ORC_PASS1_BITS, ORC_PAD, hS1, stS1);
    orc_out o = {orp, osp, cap, 0};
    for (uint32_t p = 0; p < F1; p++) {
        if (!(hR1[p] > 0 && hS1[p] > 0)) continue;
        uint64_t hR2[128], hS2[128], stR2[128], stS2[128];
        /* pass 2 scatters back into the original buffers (radix_join.c:596-606) */
        radix_cluster(R + stR1[p], tR + stR1[p], hR1[p], ORC_PASS1_BITS, ORC_PASS2_BITS,
                      ORC_SMALL_PAD, hR2, stR2);
        radix_cluster(S + stS1[p], tS + stS1[p], hS1[p], ORC_PASS1_BITS, ORC_PASS2_BITS,
                      ORC_SMALL_PAD, hS2, stS2);
        for (uint32_t q = 0; q < F2; q++) {
            if (hR2[q] > 0 && hS2[q] > 0)
                jf(R + stR1[p] + stR2[q], hR2[q], S + stS1[p] + stS2[q], hS2[q], &o);
        }
    }
    free(R);
    free(S);
    free(tR);
    free(tS);
    return (int64_t)o.n;
}

int64_t orc_rho(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                uint32_t* out_rpay, uint32_t* out_spay, uint64_t cap) {
    return radix_join(rk, rp, nr, sk, sp, ns, out_rpay, out_spay, cap, bucket_chaining);
}

int64_t orc_rht(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                uint32_t* out_rpay, uint32_t* out_spay, uint64_t cap) {
    return radix_join(rk, rp, nr, sk, sp, ns, out_rpay, out_spay, cap, histogram_join);
}

/* ------------------------------------------------------------------------ */
/* PHT restatement, no_partitioning_join.h:119-356 (single thread): buckets
 * = nextpow2(|R| / 2) (allocate_hashtable), h = key & (buckets - 1), two
 * tuples per bucket, overflow buckets inserted right after the head. */
typedef struct orc_bucket {
    uint32_t count;
    uint32_t keys[2];
    struct orc_bucket* next;
} orc_bucket;

int64_t orc_pht(const uint32_t* rk, uint64_t nr, const uint32_t* sk, uint64_t ns) {
    uint32_t nb = next_pow2((uint32_t)(nr / 2));
    if (nb == 0) nb = 1;
    uint32_t mask = nb - 1;
    orc_bucket* b = (orc_bucket*)calloc(nb, sizeof(orc_bucket));
    orc_bucket* pool = (orc_bucket*)calloc(nr + 1, sizeof(orc_bucket));
    uint64_t npool = 0;
    for (uint64_t i = 0; i < nr; i++) {
        orc_bucket* cur = b + (rk[i] & mask);
        orc_bucket* nxt = cur->next;
        if (cur->count == 2) {
            if (!nxt || nxt->count == 2) {
                orc_bucket* nb2 = pool + npool++;
                cur->next = nb2;
                nb2->next = nxt;
                nb2->count = 1;
                nb2->keys[0] = rk[i];
            } else {
                nxt->keys[nxt->count++] = rk[i];
            }
        } else {
            cur->keys[cur->count++] = rk[i];
        }
    }
    int64_t m = 0;
    for (uint64_t i = 0; i < ns; i++) {
        for (orc_bucket* c = b + (sk[i] & mask); c; c = c->next)
            for (uint32_t j = 0; j < c->count; j++) m += (sk[i] == c->keys[j]);
    }
    free(b);
    free(pool);
    return m;
}

/* ------------------------------------------------------------------------ */
/* CHT restatement (CHT.hpp:26-254, CHTJoin.hpp:46-322) with 128 partitions,
 * identity hash, bitmap words of 32 bits + running count, one linear-probe
 * step inside a partition, two-slot probe and payload checksum.  Where the
 * reference would drop a tuple (a third colliding key, key 0) or abort
 * (tail miss, CHT.hpp:194-206) this restatement counts it in *unresolved
 * instead; on dense keys (every reference generator) unresolved is 0. */
int64_t orc_cht(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                uint64_t* checksum, int64_t* unresolved) {
    const uint32_t npart = 128;
    uint64_t bmsize = next_pow2((uint32_t)nr);
    if (bmsize < 64) bmsize = 64;
    uint64_t psize = bmsize / npart;
    if (psize < 32) psize = 32; /* the reference asserts bmsize % 32 == 0 */
    uint32_t nwords = (uint32_t)(bmsize / 32);
    uint32_t* bits = (uint32_t*)calloc(nwords, sizeof(uint32_t));
    uint32_t* cnt = (uint32_t*)calloc(nwords, sizeof(uint32_t));
    orc_tuple* arr = (orc_tuple*)calloc(nr + 2048, sizeof(orc_tuple));
    int64_t unres = 0;
    uint64_t mask = bmsize - 1;
#define NEXT_IN_PART(p) (((p) & ~(psize - 1)) | (((p) + 1) & (psize - 1)))
    /* partition by high bits = range partition; build order does not change
     * the bitmap, only which colliding key wins a slot */
    int shift = __builtin_ctz((uint32_t)bmsize) - 7;
    if (shift < 0) shift = 0;
    uint32_t* order = (uint32_t*)malloc(nr * sizeof(uint32_t));
    uint64_t hist[129] = {0};
    for (uint64_t i = 0; i < nr; i++) hist[((rk[i] >> shift) & (npart - 1)) + 1]++;
    for (uint32_t i = 1; i <= npart; i++) hist[i] += hist[i - 1];
    uint64_t* dst = (uint64_t*)malloc(npart * sizeof(uint64_t));
    for (uint32_t i = 0; i < npart; i++) dst[i] = hist[i];
    for (uint64_t i = 0; i < nr; i++) order[dst[(rk[i] >> shift) & (npart - 1)]++] = (uint32_t)i;
    for (uint32_t p = 0; p < npart; p++) {
        for (uint64_t t = hist[p]; t < hist[p + 1]; t++) {
            uint64_t h = rk[order[t]] & mask;
            if (!(bits[h >> 5] & (1u << (h & 31)))) bits[h >> 5] |= 1u << (h & 31);
            else {
                uint64_t h2 = NEXT_IN_PART(h);
                if (!(bits[h2 >> 5] & (1u << (h2 & 31)))) bits[h2 >> 5] |= 1u << (h2 & 31);
            }
        }
        /* computePartPopCount with startCount = first tuple of partition
         * in the partition buffer (CHTJoin.hpp:160) */
        uint64_t wpp = psize / 32, w0 = wpp * p;
        uint32_t c = (uint32_t)hist[p];
        if (p * psize < bmsize)
            for (uint64_t w = w0; w < w0 + wpp && w < nwords; w++) {
                cnt[w] = c;
                c += __builtin_popcount(bits[w]);
            }
        for (uint64_t t = hist[p]; t < hist[p + 1]; t++) {
            uint32_t k = rk[order[t]];
            uint64_t h = k & mask;
            uint64_t pos = cnt[h >> 5] + __builtin_popcount(bits[h >> 5] & ~((~0u) << (h & 31)));
            orc_tuple* ins = arr + pos;
            if (ins->key != 0) ins = arr + NEXT_IN_PART(pos);
            if (ins->key == 0 && k != 0) {
                ins->key = k;
                ins->pay = rp ? rp[order[t]] : 0;
            } else {
                unres++;
            }
        }
    }
    int64_t m = 0;
    uint64_t cs = 0;
    for (uint64_t i = 0; i < ns; i++) {
        uint64_t h = sk[i] & mask;
        uint64_t pos = cnt[h >> 5] + __builtin_popcount(bits[h >> 5] & ~((~0u) << (h & 31)));
        uint32_t spay = sp ? sp[i] : 0;
        if (arr[pos].key == sk[i]) {
            m++;
            cs += (uint64_t)arr[pos].pay + spay;
        } else if (arr[pos + 1].key == sk[i]) {
            m++;
            cs += (uint64_t)arr[pos + 1].pay + spay;
        }
    }
#undef NEXT_IN_PART
    free(bits);
    free(cnt);
    free(arr);
    free(order);
    free(dst);
    if (checksum) *checksum = cs;
    if (unresolved) *unresolved = unres;
    return m;
}

/* ------------------------------------------------------------------------ */
/* Sorting helpers: LSD radix sort on 64-bit words. */
static void radix_sort_u64(uint64_t* a, uint64_t n) {
    if (n < 2) return;
    uint64_t* tmp = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* src = a;
    uint64_t* dst = tmp;
    for (int pass = 0; pass < 4; pass++) {
        int sh = pass * 16;
        uint64_t* cnt = (uint64_t*)calloc(65537, sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) cnt[((src[i] >> sh) & 0xFFFF) + 1]++;
        for (int i = 1; i <= 65536; i++) cnt[i] += cnt[i - 1];
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 0xFFFF]++] = src[i];
        free(cnt);
        uint64_t* t = src;
        src = dst;
        dst = t;
    }
    free(tmp);
}

void orc_sort_pairs32(uint32_t* a, uint32_t* b, uint64_t n) {
    uint64_t* w = (uint64_t*)malloc(n * sizeof(uint64_t) + 8);
    for (uint64_t i = 0; i < n; i++) w[i] = ((uint64_t)a[i] << 32) | b[i];
    radix_sort_u64(w, n);
    for (uint64_t i = 0; i < n; i++) {
        a[i] = (uint32_t)(w[i] >> 32);
        b[i] = (uint32_t)w[i];
    }
    free(w);
}

/* ------------------------------------------------------------------------ */
/* Sort-merge restatement (PSM parallel_sortmerge_join.h:18-111 and RSM
 * radix_sortmerge_join.c:40-105): sort both relations by key, then merge.
 * The merge is bounded on both sides (the reference's inner loops are not,
 * which lets them read past the array end); bounded, it is the true join. */
int64_t orc_sortmerge(const uint32_t* rk, uint64_t nr, const uint32_t* sk, uint64_t ns) {
    uint64_t* r = (uint64_t*)malloc(nr * sizeof(uint64_t) + 8);
    uint64_t* s = (uint64_t*)malloc(ns * sizeof(uint64_t) + 8);
    for (uint64_t i = 0; i < nr; i++) r[i] = rk[i];
    for (uint64_t i = 0; i < ns; i++) s[i] = sk[i];
    radix_sort_u64(r, nr);
    radix_sort_u64(s, ns);
    int64_t m = 0;
    uint64_t i = 0, j = 0;
    while (i < nr && j < ns) {
        if (r[i] < s[j]) i++;
        else if (r[i] > s[j]) j++;
        else {
            uint64_t k = r[i], i2 = i, j2 = j;
            while (i2 < nr && r[i2] == k) i2++;
            while (j2 < ns && s[j2] == k) j2++;
            m += (int64_t)((i2 - i) * (j2 - j));
            i = i2;
            j = j2;
        }
    }
    free(r);
    free(s);
    return m;
}

/* MWAY's merge join (mway/joincommon.c:241-316) after MWAY has sorted both
 * relations (sortmergejoin_multiway.c): while both cursors are in range,
 * advance the smaller key; on equal keys, for every R tuple of the key walk
 * the S run of that key (bounded by |S|), then continue after the run. */
int64_t orc_mway(const uint32_t* rk, uint64_t nr, const uint32_t* sk, uint64_t ns) {
    uint64_t* r = (uint64_t*)malloc(nr * sizeof(uint64_t) + 8);
    uint64_t* s = (uint64_t*)malloc(ns * sizeof(uint64_t) + 8);
    for (uint64_t i = 0; i < nr; i++) r[i] = rk[i];
    for (uint64_t i = 0; i < ns; i++) s[i] = sk[i];
    radix_sort_u64(r, nr);
    radix_sort_u64(s, ns);
    int64_t m = 0;
    uint64_t i = 0, j = 0;
    while (i < nr && j < ns) {
        if (r[i] < s[j]) {
            i++;
        } else if (r[i] > s[j]) {
            j++;
        } else {
            uint64_t jj;
            do {
                jj = j;
                do {
                    m++;
                    jj++;
                } while (jj < ns && r[i] == s[jj]);
                i++;
            } while (i < nr && r[i] == s[j]);
            j = jj;
        }
    }
    free(r);
    free(s);
    return m;
}

/* The join written as a relation: for every match one (key, R.payload)
 * tuple, the fields to_relation copies out of each output_list_t node
 * (Enclave/TrustedJoins/secure_joins.cpp:277-301, nodes as insert_output
 * fills them, radix_join.c:231-233).  Tuples come out sorted by (key,
 * payload).  Return = matches; at most cap tuples are written. */
static int cmp_u64(const void* x, const void* y) {
    const uint64_t a = *(const uint64_t*)x, b = *(const uint64_t*)y;
    return a < b ? -1 : a > b;
}

int64_t orc_join_relation(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                          const uint32_t* sk, uint64_t ns, uint32_t* out_key,
                          uint32_t* out_pay, uint64_t cap) {
    uint64_t* r = (uint64_t*)malloc(nr * sizeof(uint64_t) + 8);
    uint64_t* s = (uint64_t*)malloc(ns * sizeof(uint64_t) + 8);
    for (uint64_t i = 0; i < nr; i++) r[i] = ((uint64_t)rk[i] << 32) | (rp ? rp[i] : 0);
    for (uint64_t i = 0; i < ns; i++) s[i] = sk[i];
    radix_sort_u64(r, nr);
    radix_sort_u64(s, ns);
    int64_t m = 0;
    uint64_t i = 0, j = 0;
    uint64_t* w = (uint64_t*)malloc(cap * sizeof(uint64_t) + 8);
    while (i < nr && j < ns) {
        const uint64_t k = r[i] >> 32;
        if (k < s[j]) i++;
        else if (k > s[j]) j++;
        else {
            uint64_t i2 = i, j2 = j;
            while (i2 < nr && (r[i2] >> 32) == k) i2++;
            while (j2 < ns && s[j2] == k) j2++;
            for (uint64_t a = i; a < i2; a++)
                for (uint64_t b = j; b < j2; b++) {
                    if ((uint64_t)m < cap) w[m] = r[a];
                    m++;
                }
            i = i2;
            j = j2;
        }
    }
    const uint64_t n = (uint64_t)m < cap ? (uint64_t)m : cap;
    qsort(w, n, sizeof(uint64_t), cmp_u64);
    for (uint64_t q = 0; q < n; q++) {
        out_key[q] = (uint32_t)(w[q] >> 32);
        out_pay[q] = (uint32_t)w[q];
    }
    free(w);
    free(r);
    free(s);
    return m;
}

/* ------------------------------------------------------------------------ */
/* Textbook equi-join: sort (key, payload) words of both sides, walk equal
 * key runs, emit the cross product of payloads, then sort the pairs. */
int64_t orc_textbook_count(const uint32_t* rk, uint64_t nr, const uint32_t* sk, uint64_t ns) {
    return orc_sortmerge(rk, nr, sk, ns);
}

int64_t orc_textbook_pairs(const uint32_t* rk, const uint32_t* rp, uint64_t nr,
                           const uint32_t* sk, const uint32_t* sp, uint64_t ns,
                           uint32_t* out_rpay, uint32_t* out_spay, uint64_t cap) {
    uint64_t* r = (uint64_t*)malloc(nr * sizeof(uint64_t) + 8);
    uint64_t* s = (uint64_t*)malloc(ns * sizeof(uint64_t) + 8);
    for (uint64_t i = 0; i < nr; i++) r[i] = ((uint64_t)rk[i] << 32) | (rp ? rp[i] : 0);
    for (uint64_t i = 0; i < ns; i++) s[i] = ((uint64_t)sk[i] << 32) | (sp ? sp[i] : 0);
    radix_sort_u64(r, nr);
    radix_sort_u64(s, ns);
    uint64_t n = 0, i = 0, j = 0;
    while (i < nr && j < ns) {
        uint32_t a = (uint32_t)(r[i] >> 32), b = (uint32_t)(s[j] >> 32);
        if (a < b) i++;
        else if (a > b) j++;
        else {
            uint64_t i2 = i, j2 = j;
            while (i2 < nr && (uint32_t)(r[i2] >> 32) == a) i2++;
            while (j2 < ns && (uint32_t)(s[j2] >> 32) == a) j2++;
            for (uint64_t x = i; x < i2; x++)
                for (uint64_t y = j; y < j2; y++) {
                    if (out_rpay && n < cap) {
                        out_rpay[n] = (uint32_t)r[x];
                        out_spay[n] = (uint32_t)s[y];
                    }
                    n++;
                }
            i = i2;
            j = j2;
        }
    }
    free(r);
    free(s);
    if (out_rpay) orc_sort_pairs32(out_rpay, out_spay, n < cap ? n : cap);
    return (int64_t)n;
}

typedef struct { uint64_t k, p; } orc_kp64;
static int cmp_kp64(const void* x, const void* y) {
    const orc_kp64* p = (const orc_kp64*)x;
    const orc_kp64* q = (const orc_kp64*)y;
    if (p->k != q->k) return p->k < q->k ? -1 : 1;
    if (p->p != q->p) return p->p < q->p ? -1 : 1;
    return 0;
}

/* Pairs come out ordered by (key, R.payload, S.payload): the order of a
 * range-partitioned sort-merge join's output. */
int64_t orc_textbook_pairs64(const uint64_t* rk, const uint64_t* rp, uint64_t nr,
                             const uint64_t* sk, const uint64_t* sp, uint64_t ns,
                             uint64_t* out_rpay, uint64_t* out_spay, uint64_t cap) {
    orc_kp64* r = (orc_kp64*)malloc(nr * sizeof(orc_kp64) + 16);
    orc_kp64* s = (orc_kp64*)malloc(ns * sizeof(orc_kp64) + 16);
    for (uint64_t i = 0; i < nr; i++) {
        r[i].k = rk[i];
        r[i].p = rp ? rp[i] : 0;
    }
    for (uint64_t i = 0; i < ns; i++) {
        s[i].k = sk[i];
        s[i].p = sp ? sp[i] : 0;
    }
    qsort(r, nr, sizeof(orc_kp64), cmp_kp64);
    qsort(s, ns, sizeof(orc_kp64), cmp_kp64);
    uint64_t n = 0, i = 0, j = 0;
    while (i < nr && j < ns) {
        if (r[i].k < s[j].k) i++;
        else if (r[i].k > s[j].k) j++;
        else {
            uint64_t i2 = i, j2 = j;
            while (i2 < nr && r[i2].k == r[i].k) i2++;
            while (j2 < ns && s[j2].k == r[i].k) j2++;
            for (uint64_t x = i; x < i2; x++)
                for (uint64_t y = j; y < j2; y++) {
                    if (out_rpay && n < cap) {
                        out_rpay[n] = r[x].p;
                        out_spay[n] = s[y].p;
                    }
                    n++;
                }
            i = i2;
            j = j2;
        }
    }
    free(r);
    free(s);
    return (int64_t)n;
}
