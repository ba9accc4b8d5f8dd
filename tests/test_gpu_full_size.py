"""GPU checks at the benchmark workloads' full sizes.

A CPU join of 2^27 x 2^29 tuples is too slow for a test, so these pin the
output through properties instead.  R's keys are unique in every workload
here, so an S tuple has at most one match, and the pair set is fixed by:
  - every pair joins: the R key at the pair's R payload equals the S key at
    its S payload (payloads are row indices);
  - the S payloads of the pairs are exactly the S rows whose key is in R
    (each once).
Both partitioners must produce that same set.
"""
import sys

import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu

NT = 16  # generator threads


def _cols32(t):
    """int64 [n] device tuples or pairs -> int32 [n, 2] view (low word first)."""
    import torch
    return t.view(torch.int32).view(-1, 2)


def _assert_pairs_join32(dR, dS, pairs):
    import torch
    pv = _cols32(pairs)
    rk = torch.index_select(_cols32(dR)[:, 0], 0, pv[:, 0].contiguous())
    sk = torch.index_select(_cols32(dS)[:, 0], 0, pv[:, 1].contiguous())
    assert torch.equal(rk, sk)


def _r_row_of_s_row(pairs, nS):
    """Map S row -> matched R row (-1: no match); asserts no S row twice."""
    import torch
    pv = _cols32(pairs)
    sp = pv[:, 1].long()
    m = torch.full((nS,), -1, dtype=torch.int32, device="cuda:0")
    m[sp] = pv[:, 0]
    assert int(torch.bincount(sp, minlength=nS).max().item()) <= 1
    return m


@pytest.fixture(scope="module")
def c2_rels(tbj):
    """|R| = 2^27 unique shuffled keys, |S| = 2^29 foreign keys over R, on the device."""
    nR, nS = 1 << 27, 1 << 29
    R = tbj.gen_pk(11111, nR, nthreads=NT)
    S = tbj.gen_fk(22222, nS, nR, nthreads=NT)
    return gu.to_dev(R), gu.to_dev(S)


@pytest.mark.parametrize("algo", ["RHO", "RHT", "RJ", "PHT", "NPO_st", "CHT", "RSM", "PSM"])
def test_c2_full_size_pairs(tbj, c2_rels, algo):
    import torch
    dR, dS = c2_rels
    nS = dS.numel()
    pairs = torch.empty(nS, dtype=torch.int64, device="cuda:0")
    print(f"[case] {algo} 2^27 x 2^29 foreign keys, pairs", file=sys.stderr, flush=True)
    res = tbj.join(algo, dR, dS, pairs)
    torch.cuda.synchronize()
    assert res.matches == nS and res.pairs_written == nS
    _assert_pairs_join32(dR, dS, pairs)
    # every S row exactly once
    sp = torch.sort(_cols32(pairs)[:, 1]).values
    assert torch.equal(sp, torch.arange(nS, dtype=torch.int32, device="cuda:0"))


def test_c3_learned_and_radix_same_pairs_full_size(tbj):
    # |R| = 2^27 unique, |S| = 2^29 Zipf(1.0) over R's keys with every
    # odd-indexed S key moved past R's largest key: the even rows match
    import torch
    nR, nS = 1 << 27, 1 << 29
    R = tbj.gen_pk(11111, nR, nthreads=NT)
    S = tbj.gen_zipf(22222, nS, nR, 1.0, nthreads=NT)
    tbj.remap_odd(S, nR)
    dR, dS = gu.to_dev(R), gu.to_dev(S)
    del R, S
    half = nS // 2
    maps = []
    for part in ("learned", "radix"):
        pairs = torch.empty(half, dtype=torch.int64, device="cuda:0")
        print(f"[case] RHO 2^27 x 2^29 Zipf 1.0, {part}, pairs", file=sys.stderr, flush=True)
        res = tbj.join("RHO", dR, dS, pairs, partitioner=part)
        torch.cuda.synchronize()
        assert res.partitioner_used == (2 if part == "learned" else 1)
        assert res.matches == half and res.pairs_written == half
        _assert_pairs_join32(dR, dS, pairs)
        maps.append(_r_row_of_s_row(pairs, nS))
        del pairs
    assert torch.equal(maps[0], maps[1])
    assert bool((maps[0][0::2] >= 0).all()) and bool((maps[0][1::2] == -1).all())


def test_c5_psm_and_rsm_same_pairs_full_size(tbj):
    # |R| = |S| = 2^28 16-byte tuples, unique 64-bit R keys, S Zipf(1.5)
    # over R's keys: every S row matches once; PSM (learned range
    # partitions) writes the pairs in key order
    import torch
    n = 1 << 28
    R = tbj.gen_unique64(11111, n, nthreads=NT)
    S = tbj.gen_zipf_fk64(22222, n, R, 1.5, nthreads=NT)
    dR = torch.from_numpy(R.view(np.int64)).to("cuda:0")
    dS = torch.from_numpy(S.view(np.int64)).to("cuda:0")
    del R, S
    maps = []
    for algo, part in (("PSM", "learned"), ("RSM", "radix")):
        pairs = torch.empty((n, 2), dtype=torch.int64, device="cuda:0")
        print(f"[case] {algo} 2^28 x 2^28 64-bit, {part}, sorted pairs", file=sys.stderr,
              flush=True)
        res = tbj.join64(algo, dR, dS, pairs, partitioner=part)
        torch.cuda.synchronize()
        assert res.matches == n and res.pairs_written == n
        rk = torch.index_select(dR[:, 0], 0, pairs[:, 0].contiguous())
        sk = torch.index_select(dS[:, 0], 0, pairs[:, 1].contiguous())
        assert torch.equal(rk, sk)
        del sk
        if part == "learned":  # key order over the whole output (keys unsigned)
            ku = rk ^ (-(1 << 63))
            assert bool((ku[1:] >= ku[:-1]).all())
            del ku
        del rk
        m = torch.full((n,), -1, dtype=torch.int64, device="cuda:0")
        m[pairs[:, 1]] = pairs[:, 0]
        assert bool((m >= 0).all())  # n pairs and every S row present: each once
        maps.append(m)
        del pairs
    assert torch.equal(maps[0], maps[1])


def test_c4_full_batch_closed_form(tbj, oracle):
    # The benchmark batch: 4096 generated joins (|R_i| log-uniform in
    # [2^10, 2^22], |S_i| = k |R_i|, Zipf z ~ U[0, 1.5) over R_i's keys
    # 1..|R_i|, R_i unique, each join RHO or PHT), so count_i = |S_i|; eager
    # run, then the captured graph twice
    import torch
    torch.cuda.empty_cache()  # the arena alone is 170 GB
    specs, total = tbj.gen_batch_specs(0xC4C4000000000000, 4096)
    arena = torch.empty(total, dtype=torch.int64, device="cuda:0")
    print(f"[case] gen_batch_device 4096 joins, {total} tuples", file=sys.stderr, flush=True)
    tbj.gen_batch_device(arena, specs)
    items = tbj.batch_items(list(specs))
    counts = torch.zeros(len(items), dtype=torch.int64, device="cuda:0")
    want = [s.nS for s in specs]
    b = tbj.Batch(arena, items, counts, None)
    print("[case] batch 4096 joins eager", file=sys.stderr, flush=True)
    res = b.run()
    torch.cuda.synchronize()
    assert counts.cpu().tolist() == want and res.matches == sum(want)
    b.capture()
    for rep in range(2):
        counts.fill_(0)
        print(f"[case] batch 4096 joins graph replay {rep}", file=sys.stderr, flush=True)
        res = b.run()
        torch.cuda.synchronize()
        assert counts.cpu().tolist() == want and res.matches == sum(want)
    b.close()
    # a few joins against the oracle on their generated contents
    for i in (0, 1, 777, 2048, 4095):
        s = specs[i]
        R = arena[s.r_off:s.r_off + s.nR].cpu().numpy().view(np.uint32).reshape(-1, 2)
        S = arena[s.s_off:s.s_off + s.nS].cpu().numpy().view(np.uint32).reshape(-1, 2)
        assert oracle.count(R[:, 0], S[:, 0]) == s.nS
