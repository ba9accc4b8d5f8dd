"""GPU tests of the reference-facing boundary: the `app` harness (the plain-CPU
harness's command line, native.cpp:48-130 / commons.c:13-198) and the C++
operator API of libteebench.so (`result_t* RHO(relation_t*, relation_t*, int)`,
radix_join.h:17-24), on the reference's generators and seeds.

Expected counts are the true equi-join counts recorded from the reference
build in tests/golden/golden.json (RHO's column); materialised output is
compared with the oracle's pairs as sorted (key, Rpayload, Spayload) triples.
"""
import ctypes
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tee-bench_amd")
APP = os.path.join(PKG, "app")
GOLDEN = {c["name"]: c for c in json.load(open(os.path.join(ROOT, "tests", "golden",
                                                            "golden.json")))["joins"]}
ALGOS = ["RHO", "RHT", "RJ", "RHO_atomic", "PHT", "NPO_st", "CHT", "RSM", "PSM"]


def run_app(*args):
    print(f"[case] app {' '.join(args)}", file=sys.stderr, flush=True)
    p = subprocess.run([APP, *args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    m = re.search(r"Matches = (\d+)", p.stdout)
    assert m, p.stdout
    return int(m.group(1)), p.stdout


@pytest.mark.parametrize("algo", ALGOS)
def test_app_default_workload(algo):
    # -r/-s default to 2^21 (commons.c): the C1 workload
    n, out = run_app("-a", algo)
    assert n == GOLDEN["C1 default"]["counts"]["RHO"]
    assert "Running algorithm" in out


@pytest.mark.parametrize("algo", ["RHO", "PHT", "CHT", "PSM"])
@pytest.mark.parametrize("case,args", [("sel 50", ["-r", "1048576", "-s", "2097152", "-l", "50"]),
                                       ("zipf 1.0", ["-r", "1048576", "-s", "4194304", "-z", "1.0"]),
                                       ("zipf 1.5", ["-r", "262144", "-s", "1048576", "-z", "1.5"])])
def test_app_reference_workloads(algo, case, args):
    n, _ = run_app("-a", algo, *args)
    assert n == GOLDEN[case]["counts"]["RHO"]


def test_app_learned_partitioner_same_count():
    n, _ = run_app("-a", "RHO", "--learned", "-r", "1048576", "-s", "4194304", "-z", "1.0")
    assert n == GOLDEN["zipf 1.0"]["counts"]["RHO"]


# --- libteebench.so through its C++ symbols --------------------------------

class Row(ctypes.Structure):
    _fields_ = [("key", ctypes.c_uint32), ("payload", ctypes.c_uint32)]


class Table(ctypes.Structure):
    _fields_ = [("tuples", ctypes.POINTER(Row)), ("num_tuples", ctypes.c_uint64),
                ("ratio_holes", ctypes.c_int), ("sorted", ctypes.c_int)]


class OutputList(ctypes.Structure):
    pass


OutputList._fields_ = [("key", ctypes.c_uint32), ("Rpayload", ctypes.c_uint32),
                       ("Spayload", ctypes.c_uint32), ("next", ctypes.POINTER(OutputList))]


class ThreadResult(ctypes.Structure):
    _fields_ = [("nresults", ctypes.c_int64), ("results", ctypes.POINTER(OutputList)),
                ("threadid", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = [("totalresults", ctypes.c_int64), ("resultlist", ctypes.POINTER(ThreadResult)),
                ("nthreads", ctypes.c_int)]


@pytest.fixture(scope="module")
def teebench():
    import torch  # noqa: F401  (load the HIP runtime first, as tbjoin does)
    L = ctypes.CDLL(os.path.join(PKG, "libteebench.so"))
    L.teebench_set_materialize.argtypes = [ctypes.c_int]
    L.teebench_set_partitioner.argtypes = [ctypes.c_int]
    L.teebench_free_result.argtypes = [ctypes.POINTER(Result)]
    return L


def call(L, name, R, S):
    fn = getattr(L, f"_Z{len(name)}{name}P7table_tS0_i")
    fn.restype = ctypes.POINTER(Result)
    fn.argtypes = [ctypes.POINTER(Table), ctypes.POINTER(Table), ctypes.c_int]
    tR = Table(R.ctypes.data_as(ctypes.POINTER(Row)), len(R), 0, 0)
    tS = Table(S.ctypes.data_as(ctypes.POINTER(Row)), len(S), 0, 0)
    print(f"[case] libteebench {name} |R|={len(R)} |S|={len(S)}", file=sys.stderr, flush=True)
    return fn(ctypes.byref(tR), ctypes.byref(tS), 4)


@pytest.mark.parametrize("algo", ["RHO", "RHT", "PHT", "CHT", "RSM", "PSM"])
def test_operator_api_materialised_output(teebench, tbj, oracle, algo):
    rng = np.random.default_rng(21)
    R = tbj.gen_pk(11111, 100000)
    # distinct scrambled payloads (an odd multiplier is a bijection mod 2^32)
    R[:, 1] = (np.arange(len(R), dtype=np.uint64) * 2654435761 + 12345).astype(np.uint32)
    S = tbj.gen_zipf(22222, 300000, 100000, 1.0)
    S[:, 1] = rng.integers(0, 1 << 32, len(S), dtype=np.uint64).astype(np.uint32)
    R0, S0 = R.copy(), S.copy()
    teebench.teebench_set_materialize(1)
    try:
        res = call(teebench, algo, R, S)
    finally:
        teebench.teebench_set_materialize(0)
    r = res.contents
    n, oa, ob = oracle.pairs(R[:, 0], R[:, 1], S[:, 0], S[:, 1])
    assert r.totalresults == n
    assert np.array_equal(R, R0) and np.array_equal(S, S0)  # inputs left unmodified
    triples = []
    for t in range(r.nthreads):
        tr = r.resultlist[t]
        node = tr.results
        for _ in range(tr.nresults):
            triples.append((node.contents.key, node.contents.Rpayload, node.contents.Spayload))
            node = node.contents.next
    assert len(triples) == n
    got = np.array(triples, dtype=np.uint64).reshape(-1, 3)
    # every triple joins: its key is the key of both payloads' tuples
    rkey = dict(zip(R[:, 1].tolist(), R[:, 0].tolist()))
    assert all(rkey[int(rp)] == int(k) for k, rp, _ in got[:1000])
    ga, gb = oracle.sort_pairs(got[:, 1].astype(np.uint32), got[:, 2].astype(np.uint32))
    assert np.array_equal(ga, oa) and np.array_equal(gb, ob)
    teebench.teebench_free_result(res)


def test_operator_api_count_only(teebench, tbj):
    R = tbj.gen_pk(11111, 1 << 21)
    S = tbj.gen_fk(22222, 1 << 21, 1 << 21)
    for algo in ALGOS:
        res = call(teebench, algo, R, S)
        assert res.contents.totalresults == GOLDEN["C1 default"]["counts"]["RHO"], algo
        teebench.teebench_free_result(res)
