"""GPU test of tbj_join_noalloc, the entry point that allocates no device
memory (its workspace is a fixed device array the library holds per GPU).
Called through the C ABI and checked against tbj_join's output on the same
inputs (and against the oracle): bit-exact sorted pairs, count and checksum,
on two calls in a row, the second with a smaller input."""
import ctypes
import sys

import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu


def noalloc(tbj, algo, dR, nR, dS, nS, pairs, cap, **opts):
    o = tbj.make_options(materialize=pairs is not None, **opts)
    res = tbj.Result()
    rc = tbj.lib().tbj_join_noalloc(tbj.ALGOS[algo], dR.data_ptr(), nR, dS.data_ptr(), nS,
                                    pairs.data_ptr() if pairs is not None else None, cap,
                                    ctypes.byref(o), ctypes.byref(res), None)
    return rc, res


@pytest.mark.parametrize("algo", ["RHO", "RSM", "PSM"])
def test_noalloc_matches_join_on_two_calls(tbj, oracle, algo):
    import torch
    # the second call is smaller: stale scratch contents from the first must
    # not leak into it
    for nR, nS, z in ((1 << 20, (1 << 22) - (1 << 20), 1.0), (50000, 200003, 0.5)):
        R = tbj.gen_pk(11111 + nR, nR)
        S = tbj.gen_zipf(22222 + nS, nS, nR, z)
        dR, dS = gu.to_dev(R), gu.to_dev(S)
        p1 = torch.empty(nS, dtype=torch.int64, device="cuda:0")
        p2 = torch.full((nS,), -1, dtype=torch.int64, device="cuda:0")
        print(f"[case] noalloc {algo} |R|={nR} |S|={nS}", file=sys.stderr, flush=True)
        ref = tbj.join(algo, dR, dS, p1)
        rc, res = noalloc(tbj, algo, dR, nR, dS, nS, p2, nS)
        torch.cuda.synchronize()
        assert rc == tbj.TBJ_OK, tbj.lib().tbj_status_str(rc)
        assert (res.matches, res.checksum, res.pairs_written) == \
            (ref.matches, ref.checksum, ref.pairs_written)
        a = p1[: ref.pairs_written].cpu().numpy().view(np.uint32).reshape(-1, 2)
        b = p2[: res.pairs_written].cpu().numpy().view(np.uint32).reshape(-1, 2)
        aa, ab = oracle.sort_pairs(a[:, 0], a[:, 1])
        ba, bb = oracle.sort_pairs(b[:, 0], b[:, 1])
        assert np.array_equal(aa, ba) and np.array_equal(ab, bb)
        n, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
        assert res.matches == n and np.array_equal(ba, ra) and np.array_equal(bb, sa)


def test_noalloc_count_only_and_limits(tbj):
    import torch
    nR, nS = 1 << 16, 1 << 18
    R = tbj.gen_pk(1, nR)
    S = tbj.gen_fk(2, nS, nR)
    dR, dS = gu.to_dev(R), gu.to_dev(S)
    rc, res = noalloc(tbj, "RHO", dR, nR, dS, nS, None, 0)
    assert rc == tbj.TBJ_OK and res.matches == nS
    # above the maximum size: a clear error, nothing launched
    big = torch.empty(1, dtype=torch.int64, device="cuda:0")
    rc, _ = noalloc(tbj, "RHO", big, 1 << 21, big, (1 << 21) + 1, None, 0)
    assert rc == tbj.ERR_TOO_LARGE
    assert b"TBJ_NOALLOC_MAX_TUPLES" in tbj.lib().tbj_status_str(rc)
    rc, _ = noalloc(tbj, "PHT", dR, nR, dS, nS, None, 0)
    assert rc == tbj.ERR_UNSUPPORTED
    # a forced plan whose workspace exceeds the array is refused, not allocated
    rc, _ = noalloc(tbj, "RHO", dR, nR, dS, nS, None, 0, radix_bits=22, passes=2)
    assert rc == tbj.ERR_TOO_LARGE
