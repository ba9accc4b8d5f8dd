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
