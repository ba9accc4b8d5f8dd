"""Writes tests/golden/golden.json from the reference's own code.

Runs in the build container only: loads oracle/_ref/libref_count.so and
libref_mat.so, which oracle/Makefile (`make -C oracle ref`) compiles from the
reference's generator and join sources under /root/reference, and records
the reference's outputs on its own seeds:
  * glibc rand() streams after srand(seed)             (generator.cpp:76-82)
  * generator keys: sha256 + head                      (generator.cpp, genzipf.c)
  * match counts of every -a selector, quirks included (native.cpp:34-46)
  * CHT's payload checksum with payload = row index    (CHTJoin.hpp:287-302)
  * RHO's materialised pairs with payload = row index  (radix_join.c:231-233)
  * MWAY counts, duplicate keys on both sides too      (mway/joincommon.c:241-316)
  * create_relation_from_file on the committed key files (generator.cpp:290-351)
  * the three-way join (R join S) join T of secure_joins.cpp:348-399 built from
    the reference RHO's materialised output and to_relation's tuple copy
    (secure_joins.cpp:277-301)

Usage: python3 tests/golden/make_golden.py
"""
import ctypes
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
vp = ctypes.c_void_p


def load(name):
    L = ctypes.CDLL(os.path.join(REF, name))
    L.ref_rand_stream.argtypes = [ctypes.c_uint32, ctypes.c_uint64, i32p]
    L.ref_gen_pk.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p, ctypes.c_int]
    for f in ("ref_gen_fk", "ref_gen_fk_sel"):
        getattr(L, f).argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, u32p,
                                  ctypes.c_int]
    L.ref_gen_zipf.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_double,
                               u32p, ctypes.c_int]
    L.ref_gen_from_file.restype = ctypes.c_int64
    L.ref_gen_from_file.argtypes = [ctypes.c_uint32, ctypes.c_char_p, u32p, ctypes.c_uint64,
                                    ctypes.c_int]
    L.ref_join_count.restype = ctypes.c_int64
    L.ref_join_count.argtypes = [ctypes.c_char_p, u32p, vp, ctypes.c_uint64, u32p, vp,
                                 ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_uint64), vp]
    if hasattr(L, "ref_join_pairs"):
        L.ref_join_pairs.restype = ctypes.c_int64
        L.ref_join_pairs.argtypes = [ctypes.c_char_p, u32p, u32p, ctypes.c_uint64, u32p, u32p,
                                     ctypes.c_uint64, ctypes.c_int, ctypes.c_int, u32p, u32p,
                                     u32p, ctypes.c_uint64]
    return L


def in_child(fn):
    """Runs fn() in a forked child; returns its result, or "crash" when the
    reference code dies (its unbounded merges can run off the arrays)."""
    r, w = os.pipe()
    pid = os.fork()
    if pid == 0:
        os.close(r)
        try:
            out = json.dumps(fn()).encode()
        except Exception as e:  # noqa: BLE001
            out = json.dumps(f"error: {e}").encode()
        os.write(w, out)
        os._exit(0)
    os.close(w)
    data = b""
    while True:
        chunk = os.read(r, 65536)
        if not chunk:
            break
        data += chunk
    os.close(r)
    _, status = os.waitpid(pid, 0)
    if not data or status != 0:
        return "crash"
    return json.loads(data)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.uint32).tobytes()).hexdigest()


def gen(L, kind, seed, n, p=None, z=None):
    out = np.zeros(n, np.uint32)
    if kind == "pk":
        L.ref_gen_pk(seed, n, out, 1)
    elif kind == "fk":
        L.ref_gen_fk(seed, n, p, out, 1)
    elif kind == "fk_sel":
        L.ref_gen_fk_sel(seed, n, p, out, 1)
    elif kind == "zipf":
        L.ref_gen_zipf(seed, n, p, z, out, 1)
    return out


def sorted_pairs_sha(a, b):
    w = (a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64)
    w.sort()
    return hashlib.sha256(w.astype("<u8").tobytes()).hexdigest()


def main():
    C = load("libref_count.so")
    M = load("libref_mat.so")
    g = {"source": "reference native build (oracle/Makefile `ref`), seeds as native.cpp:63-64"}

    g["rand"] = {}
    for seed in (0, 1, 11111, 22222, 4294967295):
        out = np.zeros(64, np.int32)
        C.ref_rand_stream(seed, 64, out)
        g["rand"][str(seed)] = out.tolist()

    gens = [("pk", 11111, 1000, None, None), ("pk", 11111, 1 << 21, None, None),
            ("fk", 22222, 1 << 21, 1 << 21, None), ("pk", 7, 100003, None, None),
            ("fk", 22222, 4000037, 1000003, None), ("fk_sel", 22222, 4000037, 2000006, None),
            ("fk_sel", 22222, 1 << 20, 0, None),
            ("zipf", 22222, 1 << 20, 1 << 18, 0.5), ("zipf", 22222, 1 << 20, 1 << 18, 1.0),
            ("zipf", 22222, 1 << 20, 1 << 18, 1.5), ("zipf", 22222, 4000037, 1000003, 1.0)]
    g["generators"] = []
    for kind, seed, n, p, z in gens:
        k = gen(C, kind, seed, n, p, z)
        g["generators"].append({"kind": kind, "seed": seed, "n": n, "p": p, "z": z,
                                "sha256": sha(k), "head": k[:16].tolist(),
                                "sum": int(k.astype(np.uint64).sum())})

    # join counts of every selector on reference-generated inputs
    cases = [
        ("C1 default", 1 << 21, "fk", 1 << 21, 1 << 21, None),
        ("sel 50", 1 << 20, "fk_sel", 1 << 21, 2 << 20, None),
        ("zipf 1.0", 1 << 20, "zipf", 1 << 22, 1 << 20, 1.0),
        ("zipf 1.5", 1 << 18, "zipf", 1 << 20, 1 << 18, 1.5),
        ("odd sel 50", 1000003, "fk_sel", 4000037, 2000006, None),
    ]
    g["joins"] = []
    for name, nR, skind, nS, p, z in cases:
        rk = gen(C, "pk", 11111, nR)
        sk = gen(C, skind, 22222, nS, p, z)
        rp = np.arange(nR, dtype=np.uint32)
        sp = np.arange(nS, dtype=np.uint32)
        counts = {}
        for algo in ("RHO", "RHT", "RJ", "PHT", "NPO_st", "PSM", "RSM", "CHT", "MWAY"):
            # CHT aborts on a probe miss in its unbatched tail (CHT.hpp:194-206)
            if algo == "CHT" and nS % 16 and skind == "fk_sel":
                counts[algo] = None
                continue
            def one(algo=algo):
                cs = ctypes.c_uint64(0)
                m = C.ref_join_count(algo.encode(), rk, rp.ctypes.data, nR, sk, sp.ctypes.data,
                                     nS, 4, 1, ctypes.byref(cs), None)
                return [m, cs.value]
            got = in_child(one)
            counts[algo] = got if got == "crash" else got[0]
            if algo == "CHT" and got != "crash":
                counts["CHT_checksum"] = got[1]
        g["joins"].append({"name": name, "nR": nR, "S": skind, "nS": nS, "p": p, "z": z,
                           "r_sha256": sha(rk), "s_sha256": sha(sk), "counts": counts})

    # RHO materialised pairs (payload = row index)
    g["pairs"] = []
    for name, nR, skind, nS, p, z, explicit in [
            ("small fk", 1000, "fk", 3000, 1000, None, True),
            ("zipf 1.0", 1 << 16, "zipf", 1 << 18, 1 << 16, 1.0, False),
            ("sel 50", 1 << 16, "fk_sel", 1 << 18, 1 << 17, None, False)]:
        rk = gen(C, "pk", 11111, nR)
        sk = gen(C, skind, 22222, nS, p, z)
        rp = np.arange(nR, dtype=np.uint32)
        sp = np.arange(nS, dtype=np.uint32)
        cap = nS * 2
        a, b, k = (np.zeros(cap, np.uint32) for _ in range(3))
        n = M.ref_join_pairs(b"RHO", rk, rp, nR, sk, sp, nS, 4, 1, a, b, k, cap)
        a, b, k = a[:n], b[:n], k[:n]
        assert np.all(rk[a] == sk[b]) and np.all(k == sk[b])
        entry = {"name": name, "nR": nR, "S": skind, "nS": nS, "p": p, "z": z, "matches": int(n),
                 "sorted_pairs_sha256": sorted_pairs_sha(a, b)}
        if explicit:
            order = np.lexsort((b, a))
            entry["pairs"] = [[int(x), int(y)] for x, y in zip(a[order], b[order])]
        g["pairs"].append(entry)

    # MWAY with duplicate keys on both sides: R = fk over 2^14 (4 copies of
    # each key), S = fk over 2^14 (16 copies); true count 2^14 * 4 * 16
    rk = gen(C, "fk", 11111, 1 << 16, 1 << 14)
    sk = gen(C, "fk", 22222, 1 << 18, 1 << 14)
    rp = np.arange(1 << 16, dtype=np.uint32)
    sp = np.arange(1 << 18, dtype=np.uint32)
    g["mway_dups"] = {"nR": 1 << 16, "R": "fk", "pR": 1 << 14, "nS": 1 << 18, "S": "fk",
                      "pS": 1 << 14, "r_sha256": sha(rk), "s_sha256": sha(sk), "counts": {}}
    for algo in ("MWAY", "RHO", "PHT"):
        g["mway_dups"]["counts"][algo] = in_child(
            lambda algo=algo: C.ref_join_count(algo.encode(), rk, rp.ctypes.data, len(rk), sk,
                                               sp.ctypes.data, len(sk), 4, 1, None, None))

    # relations from files (-t / -u): keys as the reference loads and
    # shuffles them, and the join counts on them
    g["files"] = []
    for rfile, sfile in (("r_keys.txt", "s_keys.txt"),):
        ent = {"r_file": rfile, "s_file": sfile, "r_seed": 11111, "s_seed": 22222}
        keys = {}
        for side, fname, seed in (("r", rfile, 11111), ("s", sfile, 22222)):
            path = os.path.join(HERE, fname)
            n = sum(1 for _ in open(path))
            out = np.zeros(n, np.uint32)
            got = in_child(lambda: [C.ref_gen_from_file(seed, path.encode(), out, n, 1),
                                    out.tolist()])
            assert got != "crash" and got[0] == n, got
            keys[side] = np.array(got[1], np.uint32)
            ent[side + "_n"] = n
            ent[side + "_sha256"] = sha(keys[side])
            ent[side + "_head"] = keys[side][:16].tolist()
        ent["counts"] = {}
        zr = np.zeros(len(keys["r"]), np.uint32)
        zs = np.zeros(len(keys["s"]), np.uint32)
        for algo in ("RHO", "PHT", "RSM", "MWAY"):
            ent["counts"][algo] = in_child(
                lambda algo=algo: C.ref_join_count(algo.encode(), keys["r"], zr.ctypes.data,
                                                   len(keys["r"]), keys["s"], zs.ctypes.data,
                                                   len(keys["s"]), 2, 1, None, None))
        g["files"].append(ent)

    # three-way join: out1 = to_relation(RHO(R, S)) = (key, R.payload) per
    # match; out2 = to_relation(RHO(T, out1)) = (key, T.payload) per match
    g["three_way"] = []
    for name, nR, nS, nT, tkind in (("fk T", 1 << 14, 1 << 16, 1 << 15, "fk"),
                                    ("zipf T", 1 << 14, 1 << 16, 1 << 15, "zipf")):
        rk = gen(C, "pk", 11111, nR)
        sk = gen(C, "fk", 22222, nS, nR)
        tk = gen(C, "fk", 33333, nT, nR) if tkind == "fk" else gen(C, "zipf", 33333, nT, nR, 1.0)
        rp = np.arange(nR, dtype=np.uint32)
        sp = np.arange(nS, dtype=np.uint32)
        tp = np.arange(nT, dtype=np.uint32)
        cap = 4 * (nS + nT) + (1 << 16)
        a, b, k = (np.zeros(cap, np.uint32) for _ in range(3))
        n1 = M.ref_join_pairs(b"RHO", rk, rp, nR, sk, sp, nS, 4, 1, a, b, k, cap)
        o1k, o1p = k[:n1].copy(), a[:n1].copy()  # (key, R.payload)
        a2, b2, k2 = (np.zeros(cap, np.uint32) for _ in range(3))
        n2 = M.ref_join_pairs(b"RHO", tk, tp, nT, o1k, o1p, n1, 4, 1, a2, b2, k2, cap)
        assert n2 >= 0
        o2k, o2p = k2[:n2], a2[:n2]  # (key, T.payload)
        g["three_way"].append({"name": name, "nR": nR, "nS": nS, "nT": nT, "T": tkind,
                               "t_seed": 33333, "join1": int(n1), "matches": int(n2),
                               "sorted_out2_sha256": sorted_pairs_sha(o2k, o2p)})

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(g, f, indent=1)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
