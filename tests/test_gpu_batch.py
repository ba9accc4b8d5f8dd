"""GPU parity of batched joins (tbj_batch_*) and of the device batch generator.

Every join of a batch is checked against the oracle's count on the same
tuples (copied back from the arena) or, for generated batches, against the
closed form: every S key is drawn from R's key set, so join i has exactly
|S_i| matches.  The batch path has no reference counterpart; its per-join
semantics are RHO's and PHT's (true equi-join counts, checked against the
oracle's count and, for PHT items, the oracle's restatement of the
reference's PHT table).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def arena_of(rels, algos=None):
    """Packs (R, S) pairs of (n, 2) uint32 arrays into one arena; returns
    (device arena tensor, BatchItem array).  algos: per-join algo names."""
    import torch
    import tbjoin as tb
    parts, items, off = [], [], 0
    for i, (R, S) in enumerate(rels):
        parts += [R, S]
        items.append((off, len(R), off + len(R), len(S), algos[i] if algos else "RHO"))
        off += len(R) + len(S)
    host = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros((1, 2), np.uint32))
    arena = torch.from_numpy(host.view(np.int64).reshape(-1)).to("cuda:0")
    return arena, tb.batch_items(items)


def run_batch(tb, arena, items, capture=False, algo="per_item", **tuning):
    import torch
    counts = torch.zeros(len(items), dtype=torch.int64, device="cuda:0")
    t = tb.BatchTuning(**tuning) if tuning else None
    print(f"[case] batch n={len(items)} capture={capture} algo={algo} {tuning}",
          file=sys.stderr, flush=True)
    b = tb.Batch(arena, items, counts, t, algo=algo)
    if capture:
        b.capture()
    res = b.run()
    torch.cuda.synchronize()
    return b, res, counts.cpu().numpy()


def rel(keys):
    keys = np.asarray(keys, np.uint32)
    return np.ascontiguousarray(np.stack([keys, np.arange(len(keys), dtype=np.uint32)], 1))


def mixed_rels(tb, rng):
    """Joins of every plan group: tiny, one-pass and two-pass sizes, reference
    generators, duplicate R keys, keys outside R's range, empty S."""
    rels = []
    for nR, nS, kind in [(1, 1, "fk"), (100, 5000, "fk"), (2048, 2048, "zipf"),
                         (3000, 40000, "fk"), (50000, 200000, "zipf"), (1 << 18, 1 << 20, "fk"),
                         (1 << 20, 1 << 21, "zipf"), (300000, 0, "fk")]:
        R = tb.gen_pk(int(rng.integers(1, 1 << 30)), nR)
        if nS == 0:
            S = np.zeros((0, 2), np.uint32)
        elif kind == "fk":
            S = tb.gen_fk(int(rng.integers(1, 1 << 30)), nS, nR)
        else:
            S = tb.gen_zipf(int(rng.integers(1, 1 << 30)), nS, nR, 1.2)
        rels.append((R, S))
    R = rel(rng.integers(0, 5000, 60000))  # duplicates on both sides, misses
    S = rel(rng.integers(0, 9000, 90000))
    rels.append((R, S))
    return rels


def test_mixed_batch_matches_oracle(tbj, oracle):
    rels = mixed_rels(tbj, np.random.default_rng(1))
    arena, items = arena_of(rels)
    _, res, counts = run_batch(tbj, arena, items)
    want = [oracle.count(R[:, 0], S[:, 0]) if len(S) else 0 for R, S in rels]
    assert counts.tolist() == want
    assert res.matches == sum(want)
    assert res.groups >= 4


def test_pht_batch_matches_oracle(tbj, oracle):
    # every join PHT (the batch's algo argument): one table per join, built
    # and probed for all of them in one launch each; the counts are the
    # oracle's restatement of the reference PHT (no_partitioning_join.h)
    rels = mixed_rels(tbj, np.random.default_rng(5))
    arena, items = arena_of(rels)
    _, res, counts = run_batch(tbj, arena, items, algo="PHT")
    want = [oracle.pht(R[:, 0], S[:, 0]) if len(S) else 0 for R, S in rels]
    assert want == [oracle.count(R[:, 0], S[:, 0]) if len(S) else 0 for R, S in rels]
    assert counts.tolist() == want and res.matches == sum(want)
    assert res.groups == 1  # one wave, one PHT group


@pytest.mark.parametrize("seed", [6, 7])
def test_mixed_rho_pht_batch_matches_oracle(tbj, oracle, seed):
    # per-item algos: RHO and PHT joins interleaved in one arena
    rng = np.random.default_rng(seed)
    rels = mixed_rels(tbj, rng)
    algos = [("RHO", "PHT")[int(b)] for b in rng.integers(0, 2, len(rels))]
    algos[0], algos[-1] = "PHT", "RHO"
    arena, items = arena_of(rels, algos)
    _, res, counts = run_batch(tbj, arena, items)
    want = [oracle.count(R[:, 0], S[:, 0]) if len(S) else 0 for R, S in rels]
    assert counts.tolist() == want and res.matches == sum(want)


@pytest.mark.parametrize("tuning", [dict(wave_tuples=1 << 20), dict(one_pass_bits=4),
                                    dict(target_r=1024, s_chunk=4096),
                                    dict(one_pass_bits=11, target_r=4096),
                                    dict(tile=2048), dict(tile=8192),
                                    dict(tile=8192, one_pass_bits=4, wave_tuples=1 << 20)])
def test_tunings_and_waves_agree(tbj, oracle, tuning):
    rels = mixed_rels(tbj, np.random.default_rng(2))
    # half of the joins PHT, so that waves hold both kinds of group
    arena, items = arena_of(rels, ["RHO", "PHT"] * (len(rels) // 2) + ["RHO"] * (len(rels) % 2))
    _, res, counts = run_batch(tbj, arena, items, **tuning)
    want = [oracle.count(R[:, 0], S[:, 0]) if len(S) else 0 for R, S in rels]
    assert counts.tolist() == want
    if "wave_tuples" in tuning:
        assert res.waves > 1


def test_graph_replay_sees_new_arena_contents(tbj, oracle):
    import torch
    rels = mixed_rels(tbj, np.random.default_rng(3))
    arena, items = arena_of(rels)
    b, _, first = run_batch(tbj, arena, items, capture=True)
    # replays agree with the eager run, with eager work between them
    for _ in range(3):
        res = b.run()
        torch.cuda.synchronize()
        assert b.counts.cpu().numpy().tolist() == first.tolist()
        _ = (arena[:1024] + 1).sum().item()
    # change one join's S keys in place: the replay must see them
    k = 4
    R, S = rels[k]
    S2 = S.copy()
    S2[::2, 0] = S2[::2, 0] + np.uint32(1 << 28)  # half of S now misses
    off = items[k].s_off
    arena[off:off + len(S2)] = torch.from_numpy(S2.view(np.int64).reshape(-1)).to("cuda:0")
    res = b.run()
    torch.cuda.synchronize()
    got = b.counts.cpu().numpy()
    assert got[k] == oracle.count(R[:, 0], S2[:, 0])
    assert all(got[i] == first[i] for i in range(len(rels)) if i != k)
    b.close()


def small_specs(tb, n, seed, max_r=1 << 16):
    """Generated-batch specs with sizes scaled down to at most max_r R tuples."""
    specs, _ = tb.gen_batch_specs(seed, n)
    off = 0
    for s in specs:
        s.nR = max(1, min(int(s.nR), max_r) >> (int(s.seed) % 4))
        s.nS = s.nR * (1 + int(s.seed >> 8) % 16)
        s.r_off, s.s_off = off, off + s.nR
        off += s.nR + s.nS
    return specs, off


def test_generated_batch_closed_form_and_samples(tbj, oracle):
    import torch
    specs, total = small_specs(tbj, 300, 0xC4C4C4C4C4C4)
    arena = torch.empty(total, dtype=torch.int64, device="cuda:0")
    print("[case] gen_batch_device 300 joins", file=sys.stderr, flush=True)
    tbj.gen_batch_device(arena, specs)
    items = tbj.batch_items(list(specs))  # the specs' own RHO / PHT choice
    assert {it.algo for it in items} == {tbj.ALGOS["RHO"], tbj.ALGOS["PHT"]}
    _, res, counts = run_batch(tbj, arena, items)
    assert counts.tolist() == [s.nS for s in specs]
    assert res.matches == sum(s.nS for s in specs)
    host = arena.cpu().numpy().view(np.uint32).reshape(-1, 2)
    for i in list(range(0, 300, 37)) + [299]:
        s = specs[i]
        R = host[s.r_off:s.r_off + s.nR]
        S = host[s.s_off:s.s_off + s.nS]
        assert np.array_equal(np.sort(R[:, 0]), np.arange(1, s.nR + 1, dtype=np.uint32))
        assert np.array_equal(R[:, 1], np.arange(s.nR, dtype=np.uint32))
        assert S[:, 0].min() >= 1 and S[:, 0].max() <= s.nR
        assert oracle.count(R[:, 0], S[:, 0]) == s.nS


def test_generated_zipf_frequencies(tbj):
    import torch
    specs, _ = tbj.gen_batch_specs(5, 2)
    for s, z in zip(specs, (1.0, 0.0)):
        s.nR, s.nS, s.z = 1000, 400000, z
    specs[0].r_off, specs[0].s_off = 0, 1000
    specs[1].r_off, specs[1].s_off = 401000, 402000
    arena = torch.empty(802000, dtype=torch.int64, device="cuda:0")
    tbj.gen_batch_device(arena, specs)
    host = arena.cpu().numpy().view(np.uint32).reshape(-1, 2)
    k = host[1000:401000, 0].astype(np.int64)
    freq = np.bincount(k, minlength=1001)[1:] / len(k)
    p = 1.0 / np.arange(1, 1001)
    p /= p.sum()
    # rank frequencies within 5 standard errors of Zipf(1.0) on 1000 ranks
    se = np.sqrt(p * (1 - p) / len(k))
    assert np.all(np.abs(freq[:50] - p[:50]) < 5 * se[:50] + 1e-4)
    u = host[402000:802000, 0].astype(np.int64)
    assert u.min() >= 1 and u.max() <= 1000
    assert abs(np.bincount(u, minlength=1001)[1:].std() / 400 - 0.05) < 0.02


def test_autotune_returns_a_working_tuning(tbj, oracle):
    import torch
    rels = mixed_rels(tbj, np.random.default_rng(4))[:6]
    arena, items = arena_of(rels)
    counts = torch.zeros(len(items), dtype=torch.int64, device="cuda:0")
    print("[case] batch autotune", file=sys.stderr, flush=True)
    best, ms = tbj.batch_autotune(arena, items, counts, reps=1)
    assert ms > 0 and best.target_r in (1024, 2048) and best.one_pass_bits in (6, 8, 10)
    assert best.tile in (2048, 4096, 8192)
    _, _, got = run_batch(tbj, arena, items, **best.as_dict())
    assert got.tolist() == [oracle.count(R[:, 0], S[:, 0]) if len(S) else 0 for R, S in rels]


def test_rejects_bad_arguments(tbj):
    import torch
    arena, items = arena_of([(rel([1, 2]), rel([2]))])
    counts = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    # joins other than RHO and PHT, for the batch or one item
    for algo in ("CHT", "RSM"):
        with pytest.raises(tbj.TbjError) as e:
            tbj.Batch(arena, items, counts, algo=algo)
        assert e.value.status == tbj.ERR_UNSUPPORTED
    _, bad = arena_of([(rel([1, 2]), rel([2]))], ["CHT"])
    with pytest.raises(tbj.TbjError) as e:
        tbj.Batch(arena, bad, counts)
    assert e.value.status == tbj.ERR_UNSUPPORTED
    # tile shapes other than 2048, 4096, 8192
    with pytest.raises(tbj.TbjError) as e:
        tbj.Batch(arena, items, counts, tbj.BatchTuning(tile=1000))
    assert e.value.status == tbj.ERR_INVALID
    big = tbj.batch_items([(0, 1 << 33, 0, 1)])
    with pytest.raises(tbj.TbjError):
        tbj.Batch(arena, big, counts)
