"""GPU tests of join output chaining, MWAY, file relations and concurrent calls.

* Relation output (tbj_options.output = TBJ_OUT_RELATION): each match written
  as the (key, R.payload) tuple that the reference's to_relation copies out
  of an output_list_t node (Enclave/TrustedJoins/secure_joins.cpp:277-301),
  so the output is a relation that feeds the next join on the device; the
  three-way join (R join S) join T of secure_joins.cpp:348-399 chains two
  such joins with no host round trip.  Checked against the oracle's
  join_relation / three_way and against the golden three-way outputs built
  from the reference RHO's own materialised result.
* MWAY (mway/sortmergejoin_multiway.c:327): counts against the reference
  build's MWAY counts (golden) and the oracle's restatement of its merge join.
* File relations (-t/-u): the app's counts on the committed key files against
  the reference's counts on the same files.
* Concurrent calls on two streams from two host threads.
"""
import json
import os
import re
import subprocess
import sys
import threading

import hashlib
import numpy as np
import pytest

import gpu_util as gu

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
GOLDEN = json.load(open(os.path.join(GOLDEN_DIR, "golden.json")))
APP = os.path.join(ROOT, "tee-bench_amd", "app")


def rel(keys, payload=None):
    keys = np.asarray(keys, np.uint32)
    out = np.empty((len(keys), 2), np.uint32)
    out[:, 0] = keys
    out[:, 1] = np.arange(len(keys), dtype=np.uint32) if payload is None else payload
    return out


def sorted_sha(k, p):
    w = (k.astype(np.uint64) << np.uint64(32)) | p.astype(np.uint64)
    return hashlib.sha256(np.sort(w).astype("<u8").tobytes()).hexdigest()


def relation_join(tbj, algo, dR, dS, cap, **opts):
    """Joins device relations with relation output; returns (res, out tensor)."""
    import torch
    out = torch.empty(max(cap, 1), dtype=torch.int64, device="cuda:0")
    res = tbj.join(algo, dR, dS, out, output=tbj.OUT_RELATION, **opts)
    return res, out


@pytest.mark.parametrize("algo", ["RHO", "RHT", "RJ", "RSM", "PSM", "MWAY"])
def test_relation_output_matches_oracle(tbj, oracle, algo):
    rng = np.random.default_rng(5)
    R = rel(rng.integers(1, 3000, 5000))  # duplicate R keys: several matches per S tuple
    S = rel(rng.integers(1, 3500, 20000))
    n, k, p = oracle.join_relation(R[:, 0], R[:, 1], S[:, 0])
    print(f"[case] relation output {algo} n={n}", file=sys.stderr, flush=True)
    res, out = relation_join(tbj, algo, gu.to_dev(R), gu.to_dev(S), n)
    assert res.matches == n and res.pairs_written == n
    got = out[:n].cpu().numpy().view(np.uint32).reshape(-1, 2)
    order = np.lexsort((got[:, 1], got[:, 0]))
    assert np.array_equal(got[order, 0], k) and np.array_equal(got[order, 1], p)


def test_relation_output_rejected_where_unsupported(tbj):
    import torch
    R, S = gu.to_dev(rel([1, 2])), gu.to_dev(rel([1, 2]))
    out = torch.empty(4, dtype=torch.int64, device="cuda:0")
    for algo in ("PHT", "CHT"):
        with pytest.raises(tbj.TbjError) as e:
            tbj.join(algo, R, S, out, output=tbj.OUT_RELATION)
        assert e.value.status == tbj.ERR_UNSUPPORTED


@pytest.mark.parametrize("case", GOLDEN["three_way"], ids=lambda c: c["name"])
def test_three_way_join_matches_reference(tbj, oracle, case):
    nR, nS, nT = case["nR"], case["nS"], case["nT"]
    R = tbj.gen_pk(11111, nR)
    S = tbj.gen_fk(22222, nS, nR)
    T = (tbj.gen_fk(case["t_seed"], nT, nR) if case["T"] == "fk"
         else tbj.gen_zipf(case["t_seed"], nT, nR, 1.0))
    print(f"[case] three-way {case['name']}", file=sys.stderr, flush=True)
    # join 1: out1 = (key, R.payload); join 2 probes T with out1 as its S side
    res1, out1 = relation_join(tbj, "RHO", gu.to_dev(R), gu.to_dev(S), nS)
    assert res1.matches == case["join1"] == res1.pairs_written
    res2, out2 = relation_join(tbj, "RHO", gu.to_dev(T), out1[: res1.pairs_written],
                               4 * case["matches"])
    assert res2.matches == case["matches"] == res2.pairs_written
    got = out2[: res2.pairs_written].cpu().numpy().view(np.uint32).reshape(-1, 2)
    assert sorted_sha(got[:, 0], got[:, 1]) == case["sorted_out2_sha256"]
    n1, k2, p2 = oracle.three_way(R[:, 0], R[:, 1], S[:, 0], T[:, 0], T[:, 1])
    assert n1 == case["join1"] and len(k2) == case["matches"]
    assert sorted_sha(k2, p2) == case["sorted_out2_sha256"]


def test_three_way_entry_point(tbj, oracle):
    import torch
    rng = np.random.default_rng(9)
    R = rel(rng.integers(1, 2000, 3000))
    S = rel(rng.integers(1, 2500, 6000))
    T = rel(rng.integers(1, 2200, 4000))
    _, k2, p2 = oracle.three_way(R[:, 0], R[:, 1], S[:, 0], T[:, 0], T[:, 1])
    out = torch.empty(len(k2) + 16, dtype=torch.int64, device="cuda:0")
    res = tbj.join3(gu.to_dev(R), gu.to_dev(S), gu.to_dev(T), out)
    assert res.matches == len(k2) == res.pairs_written
    got = out[: res.pairs_written].cpu().numpy().view(np.uint32).reshape(-1, 2)
    order = np.lexsort((got[:, 1], got[:, 0]))
    assert np.array_equal(got[order, 0], k2) and np.array_equal(got[order, 1], p2)


@pytest.mark.parametrize("case", GOLDEN["joins"], ids=lambda c: c["name"])
def test_mway_counts_match_reference(tbj, oracle, case):
    R = tbj.gen_pk(11111, case["nR"])
    kind, nS, p, z = case["S"], case["nS"], case["p"], case["z"]
    S = (tbj.gen_fk(22222, nS, p) if kind == "fk" else tbj.gen_fk_sel(22222, nS, p)
         if kind == "fk_sel" else tbj.gen_zipf(22222, nS, p, z))
    res, _ = gu.run(tbj, "MWAY", R, S, materialize=False)
    assert res.matches == case["counts"]["MWAY"] == oracle.mway(R[:, 0], S[:, 0])


def test_mway_duplicates_both_sides(tbj, oracle):
    g = GOLDEN["mway_dups"]
    R = tbj.gen_fk(11111, g["nR"], g["pR"])
    S = tbj.gen_fk(22222, g["nS"], g["pS"])
    n = g["counts"]["MWAY"]
    res, got = gu.run(tbj, "MWAY", R, S, cap=n)
    assert res.matches == n == oracle.mway(R[:, 0], S[:, 0])
    _, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
    assert np.array_equal(ga, ra) and np.array_equal(gb, sa)


@pytest.mark.parametrize("algo", ["RHO", "PHT", "RSM", "MWAY"])
def test_app_relations_from_files(algo):
    f = GOLDEN["files"][0]
    args = [APP, "-a", algo, "-t", os.path.join(GOLDEN_DIR, f["r_file"]),
            "-u", os.path.join(GOLDEN_DIR, f["s_file"])]
    print(f"[case] app {' '.join(args[1:])}", file=sys.stderr, flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    m = re.search(r"Matches = (\d+)", p.stdout)
    assert m and int(m.group(1)) == f["counts"][algo]


def test_concurrent_calls_on_two_streams(tbj, oracle):
    import torch
    nR, nS = 1 << 17, 1 << 19
    cases = []
    for seed in (1, 2):
        R = tbj.gen_pk(seed, nR)
        S = tbj.gen_zipf(seed + 10, nS, nR, 0.9)
        cases.append((R, S, gu.to_dev(R), gu.to_dev(S),
                      torch.empty(nS, dtype=torch.int64, device="cuda:0")))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    results, errors = [None, None], []

    def worker(i):
        try:
            _, _, dR, dS, pairs = cases[i]
            with torch.cuda.stream(streams[i]):
                rs = []
                for _ in range(3):
                    rs.append(tbj.join("RHO", dR, dS, pairs))
                streams[i].synchronize()
            results[i] = (rs, pairs.cpu().numpy())
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    print("[case] two threads, two streams, 3 RHO joins each", file=sys.stderr, flush=True)
    th = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i, (R, S, _, _, _) in enumerate(cases):
        rs, pairs = results[i]
        n, ra, sa = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
        assert all(r.matches == n for r in rs)
        got = pairs[: rs[-1].pairs_written].view(np.uint32).reshape(-1, 2)
        ga, gb = oracle.sort_pairs(got[:, 0], got[:, 1])
        assert np.array_equal(ga, ra) and np.array_equal(gb, sa)
    for s in streams:
        tbj.release_stream(s.cuda_stream)
