"""ctypes binding of the oracle (oracle/liboracle.so): the CPU checker.

Test infrastructure only: tests compare the HIP product against it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_u32 = np.ctypeslib.ndpointer(np.uint32, flags="C")
_u64 = np.ctypeslib.ndpointer(np.uint64, flags="C")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C")
_vp = ctypes.c_void_p
_U64 = ctypes.c_uint64


def _ptr(a):
    return None if a is None else a.ctypes.data


class Oracle:
    def __init__(self):
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(path)
        L.orc_rand_stream.argtypes = [ctypes.c_uint32, _U64, _i32]
        L.orc_gen_pk.argtypes = [ctypes.c_uint32, _U64, _u32]
        L.orc_gen_fk.argtypes = [ctypes.c_uint32, _U64, ctypes.c_int64, _u32]
        L.orc_gen_fk_sel.argtypes = [ctypes.c_uint32, _U64, ctypes.c_int64, _u32]
        L.orc_gen_zipf.argtypes = [ctypes.c_uint32, _U64, ctypes.c_uint32, ctypes.c_double, _u32]
        L.orc_gen_zipf.restype = ctypes.c_int
        for f in ("orc_rho", "orc_rht"):
            getattr(L, f).argtypes = [_u32, _vp, _U64, _u32, _vp, _U64, _vp, _vp, _U64]
            getattr(L, f).restype = ctypes.c_int64
        L.orc_pht.argtypes = [_u32, _U64, _u32, _U64]
        L.orc_pht.restype = ctypes.c_int64
        L.orc_cht.argtypes = [_u32, _vp, _U64, _u32, _vp, _U64,
                              ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int64)]
        L.orc_cht.restype = ctypes.c_int64
        L.orc_sortmerge.argtypes = [_u32, _U64, _u32, _U64]
        L.orc_sortmerge.restype = ctypes.c_int64
        L.orc_textbook_count.argtypes = [_u32, _U64, _u32, _U64]
        L.orc_textbook_count.restype = ctypes.c_int64
        L.orc_textbook_pairs.argtypes = [_u32, _vp, _U64, _u32, _vp, _U64, _vp, _vp, _U64]
        L.orc_textbook_pairs.restype = ctypes.c_int64
        L.orc_textbook_pairs64.argtypes = [_u64, _vp, _U64, _u64, _vp, _U64, _vp, _vp, _U64]
        L.orc_textbook_pairs64.restype = ctypes.c_int64
        L.orc_sort_pairs32.argtypes = [_u32, _u32, _U64]
        L.orc_mway.argtypes = [_u32, _U64, _u32, _U64]
        L.orc_mway.restype = ctypes.c_int64
        L.orc_join_relation.argtypes = [_u32, _vp, _U64, _u32, _U64, _vp, _vp, _U64]
        L.orc_join_relation.restype = ctypes.c_int64
        self.L = L

    # -- generators (keys only, like the reference)
    def rand_stream(self, seed, n):
        out = np.zeros(n, np.int32)
        self.L.orc_rand_stream(seed, n, out)
        return out

    def gen_pk(self, seed, n):
        out = np.zeros(n, np.uint32)
        self.L.orc_gen_pk(seed, n, out)
        return out

    def gen_fk(self, seed, n, maxid):
        out = np.zeros(n, np.uint32)
        self.L.orc_gen_fk(seed, n, maxid, out)
        return out

    def gen_fk_sel(self, seed, n, maxid):
        out = np.zeros(n, np.uint32)
        self.L.orc_gen_fk_sel(seed, n, maxid, out)
        return out

    def gen_zipf(self, seed, n, alphabet, z):
        out = np.zeros(n, np.uint32)
        assert self.L.orc_gen_zipf(seed, n, alphabet, z, out) == 0
        return out

    # -- joins
    def rho(self, rk, rp, sk, sp, pairs=False):
        return self._radix("orc_rho", rk, rp, sk, sp, pairs)

    def rht(self, rk, rp, sk, sp, pairs=False):
        return self._radix("orc_rht", rk, rp, sk, sp, pairs)

    def _radix(self, fn, rk, rp, sk, sp, pairs):
        rk, sk = np.ascontiguousarray(rk, np.uint32), np.ascontiguousarray(sk, np.uint32)
        rp = None if rp is None else np.ascontiguousarray(rp, np.uint32)
        sp = None if sp is None else np.ascontiguousarray(sp, np.uint32)
        f = getattr(self.L, fn)
        if not pairs:
            return f(rk, _ptr(rp), len(rk), sk, _ptr(sp), len(sk), None, None, 0)
        n = f(rk, _ptr(rp), len(rk), sk, _ptr(sp), len(sk), None, None, 0)
        a, b = np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1), np.uint32)
        f(rk, _ptr(rp), len(rk), sk, _ptr(sp), len(sk), a.ctypes.data, b.ctypes.data, n)
        self.L.orc_sort_pairs32(a, b, n)
        return n, a[:n], b[:n]

    def pht(self, rk, sk):
        return self.L.orc_pht(np.ascontiguousarray(rk, np.uint32), len(rk),
                              np.ascontiguousarray(sk, np.uint32), len(sk))

    def cht(self, rk, rp, sk, sp):
        cs, un = ctypes.c_uint64(), ctypes.c_int64()
        rk, sk = np.ascontiguousarray(rk, np.uint32), np.ascontiguousarray(sk, np.uint32)
        rp = None if rp is None else np.ascontiguousarray(rp, np.uint32)
        sp = None if sp is None else np.ascontiguousarray(sp, np.uint32)
        m = self.L.orc_cht(rk, _ptr(rp), len(rk), sk, _ptr(sp), len(sk),
                           ctypes.byref(cs), ctypes.byref(un))
        return m, cs.value, un.value

    def sortmerge(self, rk, sk):
        return self.L.orc_sortmerge(np.ascontiguousarray(rk, np.uint32), len(rk),
                                    np.ascontiguousarray(sk, np.uint32), len(sk))

    def mway(self, rk, sk):
        return self.L.orc_mway(np.ascontiguousarray(rk, np.uint32), len(rk),
                               np.ascontiguousarray(sk, np.uint32), len(sk))

    def join_relation(self, rk, rp, sk):
        """The join as sorted (key, R.payload) tuples (to_relation's format)."""
        rk, sk = np.ascontiguousarray(rk, np.uint32), np.ascontiguousarray(sk, np.uint32)
        rp = np.ascontiguousarray(rp, np.uint32)
        n = self.L.orc_join_relation(rk, rp.ctypes.data, len(rk), sk, len(sk), None, None, 0)
        k, p = np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1), np.uint32)
        self.L.orc_join_relation(rk, rp.ctypes.data, len(rk), sk, len(sk), k.ctypes.data,
                                 p.ctypes.data, n)
        return n, k[:n], p[:n]

    def three_way(self, rk, rp, sk, tk, tp):
        """(R join S) join T as secure_joins.cpp:348-399 runs it: out1 = the
        (key, R.payload) relation of R join S, out2 = the (key, T.payload)
        relation of T join out1.  Returns (|out1|, sorted out2 keys, payloads)."""
        n1, k1, _ = self.join_relation(rk, rp, sk)
        n2, k2, p2 = self.join_relation(tk, tp, k1)
        return n1, k2, p2

    def count(self, rk, sk):
        return self.L.orc_textbook_count(np.ascontiguousarray(rk, np.uint32), len(rk),
                                         np.ascontiguousarray(sk, np.uint32), len(sk))

    def pairs(self, rk, rp, sk, sp):
        """Sorted (R.payload, S.payload) pairs of the textbook equi-join."""
        rk, sk = np.ascontiguousarray(rk, np.uint32), np.ascontiguousarray(sk, np.uint32)
        rp = np.ascontiguousarray(rp, np.uint32)
        sp = np.ascontiguousarray(sp, np.uint32)
        n = self.L.orc_textbook_pairs(rk, rp.ctypes.data, len(rk), sk, sp.ctypes.data, len(sk),
                                      None, None, 0)
        a, b = np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1), np.uint32)
        self.L.orc_textbook_pairs(rk, rp.ctypes.data, len(rk), sk, sp.ctypes.data, len(sk),
                                  a.ctypes.data, b.ctypes.data, n)
        return n, a[:n], b[:n]

    def pairs64(self, rk, rp, sk, sp):
        """(R.payload, S.payload) pairs ordered by (key, R.payload, S.payload)."""
        rk, sk = np.ascontiguousarray(rk, np.uint64), np.ascontiguousarray(sk, np.uint64)
        rp, sp = np.ascontiguousarray(rp, np.uint64), np.ascontiguousarray(sp, np.uint64)
        n = self.L.orc_textbook_pairs64(rk, rp.ctypes.data, len(rk), sk, sp.ctypes.data,
                                        len(sk), None, None, 0)
        a, b = np.zeros(max(n, 1), np.uint64), np.zeros(max(n, 1), np.uint64)
        self.L.orc_textbook_pairs64(rk, rp.ctypes.data, len(rk), sk, sp.ctypes.data, len(sk),
                                    a.ctypes.data, b.ctypes.data, n)
        return n, a[:n], b[:n]

    def sort_pairs(self, a, b):
        a, b = np.ascontiguousarray(a, np.uint32).copy(), np.ascontiguousarray(b, np.uint32).copy()
        self.L.orc_sort_pairs32(a, b, len(a))
        return a, b
