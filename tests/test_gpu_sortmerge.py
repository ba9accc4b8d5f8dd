"""GPU parity of the sort-merge joins (RSM, PSM) against the oracle.

The reference's own merges are unbounded and miscount (PSM counts one extra
match, Joins/psm/parallel_sortmerge_join.h:26-46; RSM miscounts under skew);
the oracle's restatement uses a bounded merge, which is the true equi-join
these tests hold the GPU to."""
import numpy as np
import pytest

import gpu_util as gu
from test_gpu_radix import assert_join, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("algo", ["RSM", "PSM"])
def test_reference_generators(tbj, oracle, algo):
    nR, nS = 1 << 16, 1 << 18
    R = tbj.gen_pk(11111, nR)
    for S in (tbj.gen_fk(22222, nS, nR), tbj.gen_fk_sel(22222, nS, 2 * nR),
              tbj.gen_zipf(22222, nS, nR, 1.5)):
        res = assert_join(tbj, oracle, algo, R, S)
        assert res.matches == oracle.sortmerge(R[:, 0], S[:, 0])


@pytest.mark.parametrize("algo", ["RSM", "PSM"])
def test_duplicates_both_sides(tbj, oracle, algo):
    rng = np.random.default_rng(3)
    R = rel(rng.integers(0, 3000, 40000))
    S = rel(rng.integers(0, 4000, 60000))
    n = oracle.count(R[:, 0], S[:, 0])
    res, got = gu.run(tbj, algo, R, S, cap=n)
    _, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert res.matches == n and np.array_equal(ga, ra) and np.array_equal(gb, sa)


def test_psm_uses_range_partitions(tbj, oracle):
    nR, nS = 1 << 18, 1 << 20
    R = tbj.gen_pk(1, nR)
    S = tbj.gen_zipf(2, nS, nR, 1.0)
    res = assert_join(tbj, oracle, "PSM", R, S)
    assert res.partitioner_used == tbj.PART_LEARNED
