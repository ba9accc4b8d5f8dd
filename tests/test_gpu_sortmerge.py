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


# The sort-merge join claims its output in chunks per workgroup (one atomic
# per chunk instead of one per group of S tuples) and afterwards moves the
# pairs above the total into the unused tails of the workgroups' last chunks;
# slots past the caller's capacity go to a side buffer meanwhile.  These cases
# hold the result to the oracle's pairs when the output is sized exactly to
# the matches, when it is truncated, and when one group's pairs span several
# chunks.

def assert_pairs_equal(oracle, R, S, got, res, n):
    _, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert res.matches == n and res.pairs_written == n
    assert np.array_equal(ga, ra) and np.array_equal(gb, sa)


@pytest.mark.parametrize("algo", ["RSM", "PSM", "MWAY"])
def test_chunked_output_sized_to_the_matches(tbj, oracle, algo):
    nR, nS = 1 << 18, 1 << 21  # hundreds of workgroups claim chunks of 512
    R = tbj.gen_pk(5, nR)
    S = tbj.gen_fk(6, nS, nR)
    res, got = gu.run(tbj, algo, R, S, cap=nS)
    assert_pairs_equal(oracle, R, S, got, res, nS)


@pytest.mark.parametrize("algo", ["RSM", "PSM"])
def test_chunked_output_truncated(tbj, oracle, algo):
    nR, nS = 1 << 17, 1 << 20
    R = tbj.gen_pk(7, nR)
    S = tbj.gen_fk(8, nS, nR)
    cap = nS // 3 + 17
    res, got = gu.run(tbj, algo, R, S, cap=cap)
    assert res.matches == nS and res.pairs_written == cap
    assert np.array_equal(R[got[:, 0], 0], S[got[:, 1], 0])  # every pair joins
    # no pair twice: S payloads are row indexes and each S row matches once
    assert len(np.unique(got[:, 1])) == cap


@pytest.mark.parametrize("algo", ["RSM", "PSM"])
def test_chunked_output_hot_key_spans_chunks(tbj, oracle, algo):
    # one R key with 3000 copies probed by 400 S tuples: 1.2 M pairs from a
    # few groups, each far above one chunk, beside ordinary unique keys
    rng = np.random.default_rng(9)
    base = rng.choice(1 << 31, 1 << 17, replace=False).astype(np.uint64) + np.uint64(1)
    rk = np.concatenate([base, np.repeat(base[5], 3000)])
    rk = rk[rng.permutation(len(rk))]
    sk = np.concatenate([rng.choice(base, 1 << 19), np.repeat(base[5], 400)])
    sk = sk[rng.permutation(len(sk))]
    R, S = rel(rk), rel(sk)
    n = oracle.count(R[:, 0], S[:, 0])
    res, got = gu.run(tbj, algo, R, S, cap=n)
    assert_pairs_equal(oracle, R, S, got, res, n)


def test_sort_join_under_workspace_caps(tbj, oracle):
    # the side buffer of the chunked claims is the join's last request, so
    # at the smallest workspace cap that still runs the join it is refused
    # and the join claims once per group: every run that succeeds, down to
    # that cap, must give the oracle's pairs
    import torch
    nR, nS = 1 << 16, 1 << 19
    R = tbj.gen_pk(13, nR)
    S = tbj.gen_fk(14, nS, nR)
    stream = torch.cuda.current_stream().cuda_stream

    def attempt(cap):
        tbj.release_stream(stream)
        try:
            res, got = gu.run(tbj, "RSM", R, S, cap=nS, workspace_cap=cap)
        except tbj.TbjError as e:
            assert e.status == tbj.ERR_OOM
            return False
        assert_pairs_equal(oracle, R, S, got, res, nS)
        return True

    lo, hi = 1 << 20, 1 << 30  # MiB .. GiB: fails at lo, runs at hi
    assert attempt(hi) and not attempt(lo)
    while hi - lo > (1 << 16):
        mid = (lo + hi) // 2
        if attempt(mid):
            hi = mid
        else:
            lo = mid
    assert attempt(hi)
    tbj.release_stream(stream)
