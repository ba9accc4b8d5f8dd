"""GPU parity of the 64-bit-key sort-merge join (tbj_join64, 16-byte tuples).

The reference has no 16-byte tuple path (Include/data-types.h:32-47), so
these tests hold the GPU to the textbook sorted-pair oracle
(orc_textbook_pairs64: every (R, S) pair with equal keys, ordered by (key,
R.payload, S.payload)) and to closed forms: with unique R keys and every S
key drawn from R, each S tuple matches exactly once.  The GPU's order inside
one key's block of pairs is unspecified, so the pairs are compared after
ordering each key block by (R.payload, S.payload); the key sequence itself
is checked as produced (non-decreasing for the learned range partitioner,
one contiguous block per key for radix).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U64 = (1 << 64) - 1


def dev(rel):
    import torch
    rel = np.ascontiguousarray(rel, np.uint64)
    return torch.from_numpy(rel.view(np.int64)).to("cuda:0")


def run64(tbj, algo, R, S, materialize=True, cap=None, **opts):
    import torch
    print(f"[case] join64 {algo} |R|={len(R)} |S|={len(S)} mat={materialize} {opts}",
          file=sys.stderr, flush=True)
    dR, dS = dev(R), dev(S)
    pairs = None
    if materialize:
        cap = len(S) if cap is None else cap
        pairs = torch.empty((max(cap, 1), 2), dtype=torch.int64, device="cuda:0")
    res = tbj.join64(algo, dR, dS, pairs, **opts)
    torch.cuda.synchronize()
    got = None
    if materialize:
        got = pairs[: res.pairs_written].cpu().numpy().view(np.uint64)
    return res, got


def rel64(keys, payload=None):
    keys = np.asarray(keys, np.uint64)
    p = np.arange(len(keys), dtype=np.uint64) if payload is None else payload
    return np.ascontiguousarray(np.stack([keys, np.asarray(p, np.uint64)], axis=1))


def checksum(a, b):
    return int((a.sum(dtype=np.uint64) + b.sum(dtype=np.uint64)) & np.uint64(U64))


def check_pairs(oracle, R, S, got, res, ordered):
    """got: GPU (R.payload, S.payload) pairs; R/S payload = row index."""
    n, oa, ob = oracle.pairs64(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    assert res.matches == n and len(got) == n
    assert res.checksum == checksum(oa, ob)
    keys = R[got[:, 0].astype(np.int64), 0]
    assert np.array_equal(keys, S[got[:, 1].astype(np.int64), 0])  # every pair joins
    if ordered:
        assert np.all(keys[1:] >= keys[:-1])
    else:  # each key's pairs form one contiguous block
        changes = np.count_nonzero(keys[1:] != keys[:-1]) + (1 if len(keys) else 0)
        assert changes == len(np.unique(keys))
    o = np.lexsort((got[:, 1], got[:, 0], keys))
    assert np.array_equal(got[o, 0], oa) and np.array_equal(got[o, 1], ob)


@pytest.mark.parametrize("algo", ["PSM", "RSM"])
def test_unique_r_zipf_s(tbj, oracle, algo):
    R = tbj.gen_unique64(5, 1 << 18)
    S = tbj.gen_zipf_fk64(6, 1 << 18, R, 1.5)
    res, got = run64(tbj, algo, R, S)
    assert res.matches == len(S)
    assert res.partitioner_used == (tbj.PART_LEARNED if algo == "PSM" else tbj.PART_RADIX)
    check_pairs(oracle, R, S, got, res, ordered=algo == "PSM")


@pytest.mark.parametrize("algo", ["PSM", "RSM"])
def test_count_only_matches_checksum(tbj, oracle, algo):
    R = tbj.gen_unique64(7, 100000)
    S = tbj.gen_zipf_fk64(8, 300000, R, 1.0)
    res, _ = run64(tbj, algo, R, S, materialize=False)
    n, oa, ob = oracle.pairs64(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    assert res.matches == n == len(S) and res.checksum == checksum(oa, ob)


@pytest.mark.parametrize("algo", ["PSM", "RSM"])
def test_duplicates_and_oversized_partitions(tbj, oracle, algo):
    # a few keys with thousands of R copies: partitions above one 2048-tuple
    # chunk are sorted by merge passes (2 and 3 passes, rounded up to even)
    rng = np.random.default_rng(11)
    hot = rng.integers(1, 1 << 62, 3, dtype=np.uint64)
    rk = np.concatenate([np.repeat(hot[0], 5000), np.repeat(hot[1], 9000),
                         np.repeat(hot[2], 700),
                         rng.integers(1, 1 << 62, 60000, dtype=np.uint64)])
    rk = rk[rng.permutation(len(rk))]
    sk = np.concatenate([np.repeat(hot, [3, 1, 40]), rng.choice(rk, 20000),
                         rng.integers(1, 1 << 62, 5000, dtype=np.uint64)])
    sk = sk[rng.permutation(len(sk))]
    R, S = rel64(rk), rel64(sk)
    n = oracle.pairs64(R[:, 0], R[:, 1], S[:, 0], S[:, 1])[0]
    res, got = run64(tbj, algo, R, S, cap=n)
    assert res.max_r_partition >= 9000
    check_pairs(oracle, R, S, got, res, ordered=algo == "PSM")


@pytest.mark.parametrize("algo", ["PSM", "RSM"])
def test_small_single_tasks_with_duplicates(tbj, oracle, algo):
    # partitions of at most one chunk, each probed by a few dozen S tuples:
    # the unsorted path (R chained in LDS, the task's S tuples ranked by key),
    # with keys of 1, 2 and 3 R copies, repeated S keys and S keys that miss
    rng = np.random.default_rng(41)
    base = rng.choice(1 << 62, 40000, replace=False).astype(np.uint64) + np.uint64(1)
    rk = np.concatenate([base, base[:5000], base[:1000]])
    rk = rk[rng.permutation(len(rk))]
    sk = np.concatenate([rng.choice(base, 3000), rng.choice(base[:1000], 500),
                         rng.integers(1, 1 << 62, 800, dtype=np.uint64)])
    sk = sk[rng.permutation(len(sk))]
    R, S = rel64(rk), rel64(sk)
    n = oracle.pairs64(R[:, 0], R[:, 1], S[:, 0], S[:, 1])[0]
    res, got = run64(tbj, algo, R, S, cap=n)
    assert res.max_r_partition <= 2048
    check_pairs(oracle, R, S, got, res, ordered=algo == "PSM")


@pytest.mark.parametrize("algo", ["PSM", "RSM"])
def test_small_tasks_staged_and_recounted(tbj, oracle, algo):
    # small single-task chunks whose pairs are at most their S tuples are laid
    # out by the count kernel and copied by the emit kernel; the chunk holding
    # a key with 40 R copies, hit by 24 S tuples, has more pairs than S tuples
    # and is placed from its R chunk instead -- both in one join
    rng = np.random.default_rng(43)
    base = rng.choice(1 << 62, 20000, replace=False).astype(np.uint64) + np.uint64(1)
    rk = np.concatenate([base, np.repeat(base[7], 40)])
    rk = rk[rng.permutation(len(rk))]
    sk = np.concatenate([rng.choice(base, 2000), np.repeat(base[7], 24),
                         rng.integers(1, 1 << 62, 300, dtype=np.uint64)])
    sk = sk[rng.permutation(len(sk))]
    R, S = rel64(rk), rel64(sk)
    n = oracle.pairs64(R[:, 0], R[:, 1], S[:, 0], S[:, 1])[0]
    res, got = run64(tbj, algo, R, S, cap=n)
    assert res.max_r_partition <= 2048
    check_pairs(oracle, R, S, got, res, ordered=algo == "PSM")


def sampled_positions(n):
    """Tuples a sample-sized first pass reads: exactly one 16-tuple group in
    every block of 256 tuples, at a hashed position."""
    i = np.arange(n, dtype=np.uint64)
    block = i >> np.uint64(8)
    pick = ((block * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) >> np.uint64(28)
    return ((i >> np.uint64(4)) & np.uint64(15)) == pick


@pytest.mark.parametrize("algo", ["PSM", "RSM"])
@pytest.mark.parametrize("z", [0.5, 1.5])
def test_sample_sized_first_pass(tbj, oracle, algo, z):
    # the first pass sized from a 1/16 sample (regions per tile row and bin,
    # runs claimed), the second pass exact over the filled regions
    R = tbj.gen_unique64(21, 1 << 18)
    S = tbj.gen_zipf_fk64(22, 1 << 19, R, z)
    res, got = run64(tbj, algo, R, S, sampled_regions=True, radix_bits=16, passes=2)
    assert res.passes == 2 and res.redone_exact == 0
    check_pairs(oracle, R, S, got, res, ordered=algo == "PSM")


def test_sample_sized_overflow_redoes_exact(tbj, oracle):
    # the tuples the sample sees carry one key and the others another, whose
    # regions the sample believes empty
    rng = np.random.default_rng(23)
    rk = rng.choice(1 << 62, 1 << 14, replace=False).astype(np.uint64) + np.uint64(1)
    nS = 1 << 17
    sk = np.where(sampled_positions(nS), rk[0], rk[1]).astype(np.uint64)
    R, S = rel64(rk), rel64(sk)
    res, got = run64(tbj, "RSM", R, S, sampled_regions=True, radix_bits=12, passes=2)
    assert res.redone_exact == 1 and res.matches == nS
    check_pairs(oracle, R, S, got, res, ordered=False)


def test_sample_sized_learned_drops_outside_range(tbj, oracle):
    # R's range measured by R's claiming scatter; S keys outside it dropped
    rng = np.random.default_rng(24)
    rk = rng.integers(1 << 40, 1 << 41, 200000, dtype=np.uint64)
    rk = np.unique(rk)
    sk = np.concatenate([rng.choice(rk, 300000), rng.integers(0, 1 << 40, 40000, dtype=np.uint64),
                         rng.integers(1 << 41, 1 << 63, 60000, dtype=np.uint64)])
    R, S = rel64(rk), rel64(sk[rng.permutation(len(sk))])
    res, got = run64(tbj, "PSM", R, S, sampled_regions=True, radix_bits=14, passes=2)
    assert res.passes == 2 and res.redone_exact == 0 and res.s_dropped == 100000
    check_pairs(oracle, R, S, got, res, ordered=True)


def test_sample_sized_learned_keeps_rare_outside_keys(tbj, oracle):
    # 20 of 400020 S keys outside R's range: the sample sees too few of them
    # for R's range to be measured, so nothing is dropped and those keys
    # simply find no match
    rng = np.random.default_rng(25)
    rk = np.unique(rng.integers(1 << 40, 1 << 41, 200000, dtype=np.uint64))
    sk = np.concatenate([rng.choice(rk, 400000), rng.integers(0, 1 << 40, 10, dtype=np.uint64),
                         rng.integers(1 << 41, 1 << 63, 10, dtype=np.uint64)])
    R, S = rel64(rk), rel64(sk[rng.permutation(len(sk))])
    res, got = run64(tbj, "PSM", R, S, sampled_regions=True, radix_bits=14, passes=2)
    assert res.passes == 2 and res.redone_exact == 0 and res.s_dropped == 0
    assert res.matches == 400000
    check_pairs(oracle, R, S, got, res, ordered=True)


def test_learned_oversized_unique_partitions(tbj, oracle):
    # 6 bits for 2^18 unique keys: every range partition spans two chunks
    R = tbj.gen_unique64(9, 1 << 18)
    S = tbj.gen_zipf_fk64(10, 1 << 17, R, 0.5)
    res, got = run64(tbj, "PSM", R, S, radix_bits=6)
    assert res.radix_bits == 6 and res.max_r_partition > 2048
    check_pairs(oracle, R, S, got, res, ordered=True)


def test_out_of_range_s_is_dropped(tbj, oracle):
    rng = np.random.default_rng(12)
    rk = rng.integers(1 << 40, 1 << 41, 50000, dtype=np.uint64)
    sk = np.concatenate([rng.choice(rk, 30000), rng.integers(0, 1 << 40, 7000, dtype=np.uint64),
                         rng.integers(1 << 41, 1 << 63, 9000, dtype=np.uint64)])
    R, S = rel64(rk), rel64(sk[rng.permutation(len(sk))])
    res, got = run64(tbj, "PSM", R, S)
    assert res.s_dropped == 16000
    check_pairs(oracle, R, S, got, res, ordered=True)


# 512 and 1024 bins in a pass exercise the scatter's byte-wide staged bins
# (bin >> 1, bin >> 2 and the low bits from the bin starts); learned with 20
# bits puts the model and 1024 bins into one workgroup's LDS
@pytest.mark.parametrize("opts", [dict(radix_bits=9, passes=1), dict(radix_bits=16, passes=2),
                                  dict(radix_bits=4), dict(partitioner="learned", radix_bits=12),
                                  dict(partitioner="learned", radix_bits=17),
                                  dict(radix_bits=10, passes=1), dict(radix_bits=20, passes=2),
                                  dict(partitioner="learned", radix_bits=20)])
def test_forced_plans(tbj, oracle, opts):
    R = tbj.gen_unique64(13, 200000)
    S = tbj.gen_zipf_fk64(14, 200000, R, 1.2)
    algo = "PSM" if opts.get("partitioner") == "learned" else "RSM"
    res, got = run64(tbj, algo, R, S, **opts)
    assert res.radix_bits == opts["radix_bits"]
    check_pairs(oracle, R, S, got, res, ordered=algo == "PSM")


@pytest.mark.parametrize("algo", ["PSM", "RSM"])
def test_clustered_keys_in_one_chunk(tbj, oracle, algo):
    # dense small keys next to a few huge ones: the equal-width key buckets
    # of a chunk overflow and the chunk falls back to a bitonic sort
    rng = np.random.default_rng(17)
    rk = np.concatenate([np.arange(1, 2001, dtype=np.uint64),
                         rng.integers(1 << 60, 1 << 63, 1000, dtype=np.uint64)])
    rk = rk[rng.permutation(len(rk))]
    sk = np.concatenate([rng.choice(rk, 4000), np.arange(1500, 2500, dtype=np.uint64)])
    R, S = rel64(rk), rel64(sk[rng.permutation(len(sk))])
    res, got = run64(tbj, algo, R, S)
    assert res.max_r_partition <= 2048
    check_pairs(oracle, R, S, got, res, ordered=algo == "PSM")


@pytest.mark.parametrize("nR,nS", [(0, 10), (10, 0), (1, 1), (1, 5), (700, 900), (1025, 3000)])
def test_tiny_and_empty(tbj, oracle, nR, nS):
    rng = np.random.default_rng(nR * 7 + nS)
    rk = rng.integers(1, 300, nR, dtype=np.uint64)
    sk = rng.integers(1, 300, nS, dtype=np.uint64)
    R, S = rel64(rk), rel64(sk)
    n = oracle.pairs64(rk, R[:, 1], sk, S[:, 1])[0]
    for algo in ("PSM", "RSM"):
        res, got = run64(tbj, algo, R, S, cap=max(n, 1))
        if nR and nS:
            check_pairs(oracle, R, S, got, res, ordered=algo == "PSM" or res.passes == 0)
        else:
            assert res.matches == 0 and len(got) == 0


def test_pairs_cap_truncates(tbj, oracle):
    R = tbj.gen_unique64(15, 50000)
    S = tbj.gen_zipf_fk64(16, 80000, R, 1.5)
    res, got = run64(tbj, "PSM", R, S, cap=1000)
    assert res.matches == len(S) and res.pairs_written == 1000 and len(got) == 1000
    keys = R[got[:, 0].astype(np.int64), 0]
    assert np.array_equal(keys, S[got[:, 1].astype(np.int64), 0])
    assert np.all(keys[1:] >= keys[:-1])


def test_rejects_other_algorithms(tbj):
    import torch
    R = dev(rel64([1, 2, 3]))
    with pytest.raises(tbj.TbjError):
        tbj.join64("RHO", R, R)
    torch.cuda.synchronize()


def test_large_unique_r_zipf_s_properties(tbj):
    # 2^24 x 2^24 at z = 1.5: every S tuple matches once, pairs in key order
    import torch
    n = 1 << 24
    R = tbj.gen_unique64(11111, n)
    S = tbj.gen_zipf_fk64(22222, n, R, 1.5)
    dR, dS = dev(R), dev(S)
    pairs = torch.empty((n, 2), dtype=torch.int64, device="cuda:0")
    print("[case] join64 PSM 2^24 x 2^24 z=1.5", file=sys.stderr, flush=True)
    res = tbj.join64("PSM", dR, dS, pairs)
    torch.cuda.synchronize()
    assert res.matches == n and res.pairs_written == n
    rk = dR[:, 0][pairs[:, 0]]
    sk = dS[:, 0][pairs[:, 1]]
    assert bool((rk == sk).all())
    # keys as unsigned: compare with the sign bit flipped
    ku = rk ^ (-(1 << 63))
    assert bool((ku[1:] >= ku[:-1]).all())
    sp = torch.sort(pairs[:, 1]).values
    assert bool((sp == torch.arange(n, device="cuda:0")).all())


@pytest.mark.parametrize("case", ["top", "narrow", "bottom"])
def test_learned_key_spans(tbj, oracle, case):
    # The learned model's cell table spans the sample's key range: keys at
    # the top of the 64-bit range (cell starts past 2^64 saturate), a span
    # narrower than the cell count (cells one key wide) and keys at zero.
    rng = np.random.default_rng({"top": 31, "narrow": 32, "bottom": 33}[case])
    n = 120000
    if case == "top":
        rk = np.uint64(2**64 - 1) - rng.choice(1 << 40, n, replace=False).astype(np.uint64)
    elif case == "narrow":
        rk = (np.uint64(5000) + rng.permutation(1500).astype(np.uint64))
    else:
        rk = rng.choice(1 << 30, n, replace=False).astype(np.uint64)
    sk = rk[rng.integers(0, len(rk), 3 * len(rk))]
    R, S = rel64(rk), rel64(sk)
    res, got = run64(tbj, "PSM", R, S, partitioner="learned")
    check_pairs(oracle, R, S, got, res, ordered=True)
    if case != "narrow":  # uniform inside their span: the linear model
        assert res.learned_model == 2


def test_learned_piecewise_model64_on_skewed_keys(tbj, oracle):
    # half of R packed into the bottom 2^20 keys, half spread up to 2^60:
    # far from linear, so the knots-and-cells model; output in key order
    rng = np.random.default_rng(34)
    n = 1 << 17
    rk = np.concatenate([rng.choice(1 << 20, n // 2, replace=False).astype(np.uint64),
                         (np.uint64(1 << 20) +
                          rng.choice(1 << 40, n // 2, replace=False).astype(np.uint64)
                          * np.uint64(1 << 20))])
    rk = rk[rng.permutation(len(rk))]
    sk = rk[rng.integers(0, len(rk), 3 * len(rk))]
    R, S = rel64(rk), rel64(sk)
    res, got = run64(tbj, "PSM", R, S, partitioner="learned")
    assert res.learned_model == 1
    check_pairs(oracle, R, S, got, res, ordered=True)
