// Relation generators (include/tbj_gen.h), bit-exact with the reference's
// generator.cpp / genzipf.c on the same seeds, multi-threaded where the
// reference's rand() stream allows it: the stream is stepped once
// sequentially to snapshot the generator state at each block or chunk
// boundary, then independent blocks are shuffled (or Zipf ranks looked up)
// on separate threads from their snapshot.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/tbj_gen.h"

namespace {

constexpr int32_t kRandMax = 2147483647;

// glibc's rand(): additive feedback generator of degree 31, separation 3,
// seeded by the 16807 Park-Miller step (Schrage's method) and warmed up by
// 310 outputs; each output is (t[f] += t[r]) >> 1.
struct Rand3 {
    int32_t t[31];
    int f, r;

    explicit Rand3(uint32_t seed) {
        if (seed == 0) seed = 1;
        t[0] = (int32_t)seed;
        int32_t w = (int32_t)seed;
        for (int i = 1; i < 31; i++) {
            long hi = w / 127773, lo = w % 127773;
            w = (int32_t)(16807 * lo - 2836 * hi);
            if (w < 0) w += 2147483647;
            t[i] = w;
        }
        f = 3;
        r = 0;
        for (int i = 0; i < 310; i++) next();
    }
    inline int32_t next() {
        uint32_t v = (uint32_t)t[f] + (uint32_t)t[r];
        t[f] = (int32_t)v;
        if (++f >= 31) {
            f = 0;
            ++r;
        } else if (++r >= 31) {
            r = 0;
        }
        return (int32_t)(v >> 1);
    }
    void skip(uint64_t k) {
        for (uint64_t i = 0; i < k; i++) next();
    }
};

int clamp_threads(int n) {
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 64));
}

template <class F>
void parallel_for(size_t n, int nthreads, F fn) {
    nthreads = clamp_threads(nthreads);
    if (n == 0) return;
    if (nthreads == 1 || n == 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    for (int t = 0; t < std::min<size_t>(nthreads, n); t++)
        th.emplace_back([&]() {
            for (size_t i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto& x : th) x.join();
}

// knuth_shuffle of generator.cpp:101-111 over keys only; n - 1 rand() calls.
void shuffle_keys(Rand3& g, tbj_tuple* k, uint64_t n) {
    if (n < 2) return;
    for (uint64_t i = n - 1; i > 0; i--) {
        int64_t j = (int64_t)((double)g.next() / ((double)kRandMax + 1) * (double)i);
        uint32_t tmp = k[i].key;
        k[i].key = k[j].key;
        k[j].key = tmp;
    }
}

void set_payloads(tbj_tuple* out, uint64_t n, int mode, int nthreads) {
    const uint64_t chunk = 1 << 22;
    parallel_for((n + chunk - 1) / chunk, nthreads, [&](size_t c) {
        uint64_t a = c * chunk, b = std::min(n, a + chunk);
        for (uint64_t i = a; i < b; i++) out[i].payload = mode == TBJ_PAYLOAD_ROWID ? (uint32_t)i : 0;
    });
}

// Blocks [i*bs, i*bs + len_i) filled by `fill` then shuffled; each block
// consumes len_i - 1 rand() calls in order.
template <class Fill>
void shuffled_blocks(uint32_t seed, uint64_t n, uint64_t bs, tbj_tuple* out, int nthreads,
                     Fill fill) {
    const uint64_t nblocks = bs ? (n + bs - 1) / bs : 0;
    std::vector<Rand3> snaps;
    snaps.reserve(nblocks);
    Rand3 g(seed);
    for (uint64_t b = 0; b < nblocks; b++) {
        snaps.push_back(g);
        const uint64_t len = std::min(bs, n - b * bs);
        if (b + 1 < nblocks && len > 1) g.skip(len - 1);
    }
    parallel_for(nblocks, nthreads, [&](size_t b) {
        const uint64_t len = std::min(bs, n - b * bs);
        tbj_tuple* blk = out + b * bs;
        fill(blk, len);
        Rand3 gb = snaps[b];
        shuffle_keys(gb, blk, len);
    });
}

}  // namespace

extern "C" {

void tbj_rand_stream(uint32_t seed, uint64_t n, int32_t* out) {
    Rand3 g(seed);
    for (uint64_t i = 0; i < n; i++) out[i] = g.next();
}

int tbj_gen_pk(uint32_t seed, uint64_t n, tbj_tuple* out, int payload_mode, int nthreads) {
    if (!out && n) return TBJ_ERR_INVALID;
    set_payloads(out, n, payload_mode, nthreads);
    shuffled_blocks(seed, n, n, out, 1, [](tbj_tuple* b, uint64_t len) {
        for (uint64_t i = 0; i < len; i++) b[i].key = (uint32_t)(i + 1);
    });
    return TBJ_OK;
}

int tbj_gen_fk(uint32_t seed, uint64_t n, int64_t maxid, tbj_tuple* out, int payload_mode,
               int nthreads) {
    if ((!out && n) || maxid <= 0) return TBJ_ERR_INVALID;
    set_payloads(out, n, payload_mode, nthreads);
    shuffled_blocks(seed, n, (uint64_t)maxid, out, nthreads, [](tbj_tuple* b, uint64_t len) {
        for (uint64_t i = 0; i < len; i++) b[i].key = (uint32_t)(i + 1);
    });
    return TBJ_OK;
}

int tbj_gen_fk_sel(uint32_t seed, uint64_t n, int64_t maxid, tbj_tuple* out, int payload_mode,
                   int nthreads) {
    if ((!out && n) || maxid < 0) return TBJ_ERR_INVALID;
    set_payloads(out, n, payload_mode, nthreads);
    // random_unique_gen_maxid (generator.cpp:157-171): jump is the integer
    // quotient maxid / len held in a double, ids start at 1 (0 if maxid == 0)
    auto fill = [maxid](tbj_tuple* b, uint64_t len) {
        double jump = (double)((uint32_t)maxid / len);
        double id = maxid == 0 ? 0 : 1;
        for (uint64_t i = 0; i < len; i++) {
            b[i].key = (uint32_t)id;
            id += jump;
        }
    };
    if (maxid == 0) {
        shuffled_blocks(seed, n, n, out, 1, fill);
    } else {
        shuffled_blocks(seed, n, (uint64_t)maxid, out, nthreads, fill);
    }
    return TBJ_OK;
}

int tbj_gen_zipf(uint32_t seed, uint64_t n, uint32_t alphabet, double z, tbj_tuple* out,
                 int payload_mode, int nthreads) {
    if ((!out && n) || alphabet == 0) return TBJ_ERR_INVALID;
    set_payloads(out, n, payload_mode, nthreads);
    Rand3 g(seed);
    // gen_alphabet (genzipf.c:28-53): the unsigned 32-bit product i * rand()
    // wraps, so the swap partner k is 0, 1 or 2
    std::vector<uint32_t> alpha(alphabet);
    for (uint32_t i = 0; i < alphabet; i++) alpha[i] = i + 1;
    for (uint32_t i = alphabet - 1; i > 0; i--) {
        uint32_t k = (uint32_t)(i * (uint32_t)g.next()) / (uint32_t)kRandMax;
        std::swap(alpha[i], alpha[k]);
    }
    // gen_zipf_lut (genzipf.c:60-91): the terms 1/pow(i, z) are pure, so they
    // are computed in parallel; both running sums stay sequential
    std::vector<double> lut(alphabet);
    const uint64_t chunk = 1 << 20;
    parallel_for((alphabet + chunk - 1) / chunk, nthreads, [&](size_t c) {
        uint64_t a = c * chunk, b = std::min<uint64_t>(alphabet, a + chunk);
        for (uint64_t i = a; i < b; i++) lut[i] = 1.0 / pow((double)(i + 1), z);
    });
    double scale = 0.0;
    for (uint32_t i = 0; i < alphabet; i++) scale += lut[i];
    double sum = 0.0;
    for (uint32_t i = 0; i < alphabet; i++) {
        sum += lut[i];
        lut[i] = sum / scale;
    }
    // gen_zipf (genzipf.c:119-150): pos = first index with lut[pos] >= r,
    // at most alphabet - 1.  A guide table over r narrows the search.
    const uint32_t G = 1u << 20;
    std::vector<uint32_t> guide(G + 1);
    {
        uint32_t p = 0;
        for (uint32_t j = 0; j <= G; j++) {
            const double rj = (double)j / G;
            while (p < alphabet - 1 && lut[p] < rj) p++;
            guide[j] = p;  // first index with lut >= j/G (or alphabet-1)
        }
    }
    const int T = clamp_threads(nthreads);
    const uint64_t per = (n + T - 1) / T;
    std::vector<Rand3> snaps;
    for (int t = 0; t < T; t++) {
        snaps.push_back(g);
        if (t + 1 < T) g.skip(std::min<uint64_t>(per, n - std::min<uint64_t>(n, t * per)));
    }
    parallel_for(T, T, [&](size_t t) {
        Rand3 gt = snaps[t];
        const uint64_t a = std::min<uint64_t>(n, t * per), b = std::min<uint64_t>(n, a + per);
        for (uint64_t i = a; i < b; i++) {
            const double r = ((double)gt.next()) / kRandMax;
            uint32_t pos;
            if (lut[0] >= r) {
                pos = 0;
            } else {
                uint32_t j = (uint32_t)(r * G);
                if (j > G) j = G;
                uint32_t lo = j ? guide[j - 1] : 0;         // lut[x] < r for x < lo
                uint32_t hi = j < G ? guide[j + 1] : alphabet - 1;
                if (lo > 0) lo--;
                // first index in [lo, hi] with lut >= r, capped at alphabet - 1
                uint32_t L = lo, H = hi;
                while (L < H) {
                    uint32_t m = L + (H - L) / 2;
                    if (lut[m] < r) L = m + 1;
                    else H = m;
                }
                pos = L;
                if (lut[pos] < r) pos = alphabet - 1;
            }
            out[i].key = alpha[pos];
        }
    });
    return TBJ_OK;
}

int tbj_gen_from_keys(uint32_t seed, const uint32_t* keys, uint64_t n, tbj_tuple* out,
                      int payload_mode) {
    if (n && (!keys || !out)) return TBJ_ERR_INVALID;
    for (uint64_t i = 0; i < n; i++) out[i].key = keys[i];
    set_payloads(out, n, payload_mode, 1);
    Rand3 g(seed);
    shuffle_keys(g, out, n);
    return TBJ_OK;
}

int64_t tbj_gen_from_file(uint32_t seed, const char* path, tbj_tuple* out, uint64_t cap,
                          int payload_mode) {
    if (!path) return TBJ_ERR_INVALID;
    FILE* f = fopen(path, "r");
    if (!f) return TBJ_ERR_INVALID;
    // the tuple count is the number of newlines (generator.cpp:308-314)
    uint64_t lines = 0;
    for (int c = getc(f); c != EOF; c = getc(f))
        if (c == '\n') lines++;
    if (!out || lines > cap) {
        fclose(f);
        return (int64_t)lines;
    }
    rewind(f);
    std::vector<uint32_t> keys;
    keys.reserve(lines);
    char buf[256];
    while (keys.size() < lines && fgets(buf, sizeof(buf), f)) {
        keys.push_back((uint32_t)atoi(buf));
        // a line longer than the buffer: skip its rest
        while (!strchr(buf, '\n') && fgets(buf, sizeof(buf), f)) {
        }
    }
    fclose(f);
    keys.resize(lines, 0);
    tbj_gen_from_keys(seed, keys.data(), lines, out, payload_mode);
    return (int64_t)lines;
}

void tbj_remap_odd(tbj_tuple* rel, uint64_t n, uint32_t offset) {
    for (uint64_t i = 1; i < n; i += 2) rel[i].key += offset;
}

static inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// splitmix64's output is a bijection of its counter, so keys drawn from
// distinct counters are distinct.  Tuple i takes counter seed + (i+1)*gamma;
// the (never observed) zero output is replaced by the output of a counter
// past the last tuple's, which keeps every key distinct and nonzero.
int tbj_gen_unique64(uint64_t seed, uint64_t n, tbj_tuple64* out, int nthreads) {
    if (!out && n) return TBJ_ERR_INVALID;
    const uint64_t gamma = 0x9E3779B97F4A7C15ull, chunk = 1 << 20;
    std::atomic<uint64_t> extra{0};
    parallel_for((n + chunk - 1) / chunk, nthreads, [&](size_t c) {
        const uint64_t a = c * chunk, b = std::min<uint64_t>(n, a + chunk);
        for (uint64_t i = a; i < b; i++) {
            uint64_t x = seed + i * gamma;
            uint64_t k = splitmix64(x);
            while (k == 0) {
                x = seed + (n + extra.fetch_add(1)) * gamma;
                k = splitmix64(x);
            }
            out[i].key = k;
            out[i].payload = i;
        }
    });
    return TBJ_OK;
}

// S keys are R's keys at Zipf(z) ranks (rank 1 = R[0]): inverse-CDF lookup
// of a uniform double.  Blocks of 2^16 tuples own independent splitmix64
// streams, so the output does not depend on the thread count.
int tbj_gen_zipf_fk64(uint64_t seed, uint64_t n, const tbj_tuple64* R, uint64_t nR, double z,
                      tbj_tuple64* out, int nthreads) {
    if ((!out && n) || !R || nR == 0) return TBJ_ERR_INVALID;
    std::vector<double> cdf(nR);
    const uint64_t chunk = 1 << 20;
    parallel_for((nR + chunk - 1) / chunk, nthreads, [&](size_t c) {
        uint64_t a = c * chunk, b = std::min<uint64_t>(nR, a + chunk);
        for (uint64_t i = a; i < b; i++) cdf[i] = 1.0 / pow((double)(i + 1), z);
    });
    double s = 0;
    for (uint64_t i = 0; i < nR; i++) {
        s += cdf[i];
        cdf[i] = s;
    }
    const double inv = 1.0 / s;
    parallel_for((nR + chunk - 1) / chunk, nthreads, [&](size_t c) {
        uint64_t a = c * chunk, b = std::min<uint64_t>(nR, a + chunk);
        for (uint64_t i = a; i < b; i++) cdf[i] *= inv;
    });
    const uint64_t block = 1 << 16;
    parallel_for((n + block - 1) / block, nthreads, [&](size_t blk) {
        uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (blk + 1));
        const uint64_t a = blk * block, b = std::min<uint64_t>(n, a + block);
        for (uint64_t i = a; i < b; i++) {
            const double u = (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0);
            uint64_t pos = (uint64_t)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
            if (pos >= nR) pos = nR - 1;
            out[i].key = R[pos].key;
            out[i].payload = i;
        }
    });
    return TBJ_OK;
}

}  // extern "C"
