"""Helpers for GPU tests: upload relations, run a join, fetch sorted pairs."""
import sys

import numpy as np


def to_dev(rel):
    """(n, 2) uint32 tuples -> contiguous int64 device tensor (one word per tuple)."""
    import torch
    rel = np.ascontiguousarray(rel, np.uint32)
    return torch.from_numpy(rel.view(np.int64).reshape(-1)).to("cuda:0")


def run(tbj, algo, R, S, materialize=True, cap=None, dR=None, dS=None, **opts):
    import torch
    dR = to_dev(R) if dR is None else dR
    dS = to_dev(S) if dS is None else dS
    pairs = None
    if materialize:
        cap = S.shape[0] if cap is None else cap
        pairs = torch.empty(max(cap, 1), dtype=torch.int64, device="cuda:0")
    print(f"[case] {algo} |R|={R.shape[0]} |S|={S.shape[0]} mat={materialize} {opts}",
          file=sys.stderr, flush=True)
    res = tbj.join(algo, dR, dS, pairs, **opts) if materialize else tbj.join(algo, dR, dS, **opts)
    torch.cuda.synchronize()
    got = None
    if materialize:
        got = pairs[: res.pairs_written].cpu().numpy().view(np.uint32).reshape(-1, 2)
    return res, got


def checksum(rp, sp):
    """CHT's checksum: sum of (R.payload + S.payload) mod 2^64."""
    return int((rp.astype(np.uint64).sum() + sp.astype(np.uint64).sum()) & 0xFFFFFFFFFFFFFFFF)
