"""GPU parity of the no-partitioning joins (PHT, NPO_st, CHT) against the
oracle: exact counts, CHT checksum and sorted (R.payload, S.payload) pairs."""
import numpy as np
import pytest

import gpu_util as gu
from test_gpu_radix import assert_join, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("algo", ["PHT", "NPO_st", "CHT"])
def test_reference_generators(tbj, oracle, algo):
    nR, nS = 1 << 16, 1 << 18
    R = tbj.gen_pk(11111, nR)
    for S in (tbj.gen_fk(22222, nS, nR), tbj.gen_fk_sel(22222, nS, 2 * nR),
              tbj.gen_zipf(22222, nS, nR, 1.0)):
        assert_join(tbj, oracle, algo, R, S)


@pytest.mark.parametrize("algo", ["PHT", "CHT"])
def test_random_keys_with_duplicates(tbj, oracle, algo):
    # non-dense keys: PHT overflow chains, CHT overflow table (the reference
    # CHT drops or aborts on such inputs)
    rng = np.random.default_rng(7)
    R = rel(np.concatenate([rng.integers(0, 1 << 20, 50000), [0, 0, 0xFFFFFFFF]]))
    S = rel(np.concatenate([rng.integers(0, 1 << 20, 200000), [0, 0xFFFFFFFF]]))
    n = oracle.count(R[:, 0], S[:, 0])
    res, got = gu.run(tbj, algo, R, S, cap=n)
    _, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert res.matches == n and np.array_equal(ga, ra) and np.array_equal(gb, sa)
    assert res.checksum == gu.checksum(ra, sa)


def test_cht_checksum_matches_reference_structure(tbj, oracle):
    # dense keys: the restated reference CHT resolves every tuple; GPU CHT
    # must report the same count and checksum
    nR, nS = 100000, 400000
    R = tbj.gen_pk(11111, nR)
    S = tbj.gen_zipf(22222, nS, nR, 0.9)
    m, cks, unresolved = oracle.cht(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    assert unresolved == 0
    res, _ = gu.run(tbj, "CHT", R, S, materialize=False)
    assert (res.matches, res.checksum) == (m, cks)


@pytest.mark.parametrize("algo", ["PHT", "CHT"])
def test_empty_and_tiny(tbj, oracle, algo):
    res, got = gu.run(tbj, algo, rel([9]), rel([9, 9, 1]))
    assert res.matches == 2 and sorted(got[:, 1].tolist()) == [0, 1]
    res, _ = gu.run(tbj, algo, rel([]), rel([1]), materialize=False)
    assert res.matches == 0


def test_cht_slot_past_last_bit(tbj, oracle):
    # Two R tuples of key 31 with a 64-bit bitmap of 32-bit partitions: the
    # second sets bit 0 (the partition wraps), so both place from rank 1 and
    # one lands past the last set bit's slot, leaving slot 0 empty.  A prior
    # run leaves key 0 in slot 0: the probe must not take the stale slot.
    R0 = rel(np.arange(64))
    res, _ = gu.run(tbj, "CHT", R0, rel([0, 5]))
    assert res.matches == 2
    R = rel([31, 31])
    S = rel([0, 31, 1, 31, 32])
    res, got = gu.run(tbj, "CHT", R, S)
    _, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert res.matches == 4 and np.array_equal(ga, ra) and np.array_equal(gb, sa)
    assert res.checksum == gu.checksum(ra, sa)


def join_all_pairs(tbj, oracle, algo, R, S, **opts):
    """Every pair written (output sized to the oracle's count) and compared."""
    n = oracle.count(R[:, 0], S[:, 0])
    res, got = gu.run(tbj, algo, R, S, cap=max(n, 1), **opts)
    _, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert res.matches == n and res.pairs_written == n
    assert np.array_equal(ga, ra) and np.array_equal(gb, sa)
    assert res.checksum == gu.checksum(ra, sa)
    return res


def pht_layout_bits(nR):
    """Partition bits of PHT's bulk build: the table has nextpow2(|R| / 2)
    buckets, blocks of at most 2^10 of them."""
    lg = max(0, int(max(1, nR // 2) - 1).bit_length())
    return max(0, lg - 10)


@pytest.mark.parametrize("bulk", [1, -1])
@pytest.mark.parametrize("nR", [1, 3, 1000, 5000, 70001, 1 << 20])
def test_pht_bulk_and_atomic_builds(tbj, oracle, bulk, nR):
    # blocks of buckets filled from R partitioned on the buckets' low bits (no
    # partitioning, one pass and two passes) against per-tuple inserts; keys
    # with a narrow range repeat, so buckets overflow into chains
    rng = np.random.default_rng(nR)
    rk = rng.integers(0, max(2, nR // 2), nR)
    rk[: min(nR, 3)] = [0xFFFFFFFF, 0, 7][: min(nR, 3)]
    sk = np.concatenate([rng.integers(0, max(2, nR // 2) + 50, 3 * nR + 5), [0xFFFFFFFF, 0]])
    res = join_all_pairs(tbj, oracle, "PHT", rel(rk), rel(sk), bulk_build=bulk)
    assert res.radix_bits == (pht_layout_bits(nR) if bulk == 1 else 0)


@pytest.mark.parametrize("algo", ["PHT", "NPO_st"])
def test_pht_bulk_build_long_chains(tbj, oracle, algo):
    # one key 10000 times (a bucket whose chain spans many pool nodes, placed
    # by one workgroup) next to distinct keys, through two partition passes
    rk = np.concatenate([np.full(10000, 42), np.arange(1, 1 << 20)])
    sk = np.concatenate([np.full(7, 42), np.arange(0, 1 << 20, 3)])
    res = join_all_pairs(tbj, oracle, algo, rel(rk), rel(sk), bulk_build=1)
    assert res.radix_bits == pht_layout_bits(len(rk)) and res.passes == 2


@pytest.mark.parametrize("bulk", [1, -1])
def test_cht_bulk_and_atomic_builds(tbj, oracle, bulk):
    # duplicates and misses (spills in both phases) through the bitmap built
    # per segment in LDS and through the per-tuple atomics
    rng = np.random.default_rng(31)
    rk = np.concatenate([rng.integers(0, 1 << 19, 300000), [0, 0, 0xFFFFFFFF]])
    sk = np.concatenate([rng.integers(0, 1 << 19, 500000), [0, 0xFFFFFFFF]])
    res = join_all_pairs(tbj, oracle, "CHT", rel(rk), rel(sk), bulk_build=bulk)
    assert (res.radix_bits > 0) == (bulk == 1)


def test_cht_bulk_second_choice_leaves_segment(tbj, oracle):
    # |R| > 2^19: a bitmap of 2^20 bits, CHT partitions of 8192 bits, two
    # 4096-bit segments each.  Key 4095 + 2^20 shares home 4095 (the last of
    # segment 0) with key 4095: its next bit, 4096, is free but lies in
    # segment 1, so it spills.  Key 8191 + 2^20 shares home 8191, the last of
    # partition 0, whose next bit wraps to 0 in segment 0: it spills too.
    bm = 1 << 20
    base = np.arange(1, 600000)
    base = base[base != 4096]
    rk = np.concatenate([base, [4095 + bm, 8191 + bm]])
    sk = np.concatenate([np.arange(0, 610000, 7), [4095, 4095 + bm, 4096, 8191, 8191 + bm,
                                                   4095 + 2 * bm]])
    res = join_all_pairs(tbj, oracle, "CHT", rel(rk), rel(sk), bulk_build=1)
    assert res.radix_bits == 8 and res.max_r_partition >= 2  # tuples in the overflow table


def test_cht_bulk_checksum_matches_reference_structure(tbj, oracle):
    nR, nS = 1 << 20, 1 << 21
    R = tbj.gen_pk(11111, nR)
    S = tbj.gen_zipf(22222, nS, nR, 0.9)
    m, cks, unresolved = oracle.cht(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    assert unresolved == 0
    res, _ = gu.run(tbj, "CHT", R, S, materialize=False, bulk_build=1)
    assert (res.matches, res.checksum) == (m, cks) and res.radix_bits == 8
