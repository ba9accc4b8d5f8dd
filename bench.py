"""Benchmark of the equi-join hot path on one MI355X.

Default workload: RHO radix hash join, |R| = 2^27 unique shuffled keys
(create_relation_pk, seed 11111), |S| = 2^29 foreign keys (create_relation_fk
over R's keys, seed 22222), 8-byte tuples (uint32 key, uint32 payload =
row index), materialised (R.payload, S.payload) pairs.  A step is one full
join call (partition passes, build/probe, pair output) with R and S already
resident in HBM.

Prints one JSON line.  `value` is (|R|+|S|) tuples per second over the timed
steps.  `roofline` rates the longest single kernel of the step against the
HBM peak from its algorithmic bytes and its average duration (HIP events on
the library's stream, over the timed steps); `pipeline_roofline` rates the
whole step with the partitioned-join byte floor 8 B x (|R| + |S|) +
16 B x (|R| + |S_in|) + 8 B x matches, S_in being the S tuples inside R's key
range.  The CPU baseline is the reference's own RHO (built from its sources
into oracle/_ref) timed on this host's cores on the same inputs.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tee-bench_amd"))

METRIC = "(|R|+|S|) tuples/s + % HBM roofline; learned vs radix partitioning, 1 GPU"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3"])
    ap.add_argument("--algo", default="RHO")
    ap.add_argument("--partitioner", default=None, choices=["radix", "learned"])
    ap.add_argument("--count-only", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=3)
    return ap.parse_args()


def workload(tb, name, nthreads):
    """Generates (R, S, description) with the reference's generators."""
    import numpy as np
    if name == "c1":
        nR, nS = 1 << 21, 1 << 21
        R = tb.gen_pk(11111, nR, nthreads=nthreads)
        S = tb.gen_fk(22222, nS, nR, nthreads=nthreads)
        desc = "RHO |R|=|S|=2^21, S = fk over R (100% selectivity)"
    elif name == "c2":
        nR, nS = 1 << 27, 1 << 29
        R = tb.gen_pk(11111, nR, nthreads=nthreads)
        S = tb.gen_fk(22222, nS, nR, nthreads=nthreads)
        desc = "RHO |R|=2^27 unique shuffled, |S|=2^29 uniform foreign keys over R"
    else:
        nR, nS = 1 << 27, 1 << 29
        R = tb.gen_pk(11111, nR, nthreads=nthreads)
        S = tb.gen_zipf(22222, nS, nR, 1.0, nthreads=nthreads)
        tb.remap_odd(S, nR)  # odd-indexed S keys moved outside R's key range
        desc = ("|R|=2^27 unique, |S|=2^29 Zipf(1.0) over R's keys with every odd-indexed "
                "S key shifted by 2^27 out of R's range (50% selectivity)")
    assert R.dtype == np.uint32
    return R, S, desc


def s_in_range(R, S):
    import numpy as np
    lo, hi = int(R[:, 0].min()), int(R[:, 0].max())
    k = S[:, 0]
    return int(np.count_nonzero((k >= lo) & (k <= hi)))


def cpu_baseline(R, S, reps):
    """Times the reference's RHO (oracle/_ref, built from its sources) here."""
    import ctypes
    import numpy as np
    path = os.path.join(ROOT, "oracle", "_ref", "libref_count.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
    L.ref_join_count.restype = ctypes.c_int64
    L.ref_join_count.argtypes = [ctypes.c_char_p, u32p, ctypes.c_void_p, ctypes.c_uint64, u32p,
                                 ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    cores = min(16, os.cpu_count() or 1)
    rk, rp = np.ascontiguousarray(R[:, 0]), np.ascontiguousarray(R[:, 1])
    sk, sp = np.ascontiguousarray(S[:, 0]), np.ascontiguousarray(S[:, 1])
    times, m = [], None
    for _ in range(reps):
        us = ctypes.c_double()
        m = L.ref_join_count(b"RHO", rk, rp.ctypes.data, len(rk), sk, sp.ctypes.data, len(sk),
                             cores, 1, None, ctypes.byref(us))
        times.append(us.value * 1e-6)
    t = statistics.median(times)
    return {"value": (len(rk) + len(sk)) / t, "unit": "tuples/s", "cores": cores,
            "kind": "reference",
            "sample": (f"reference RHO (native build, count only, its default) on the full "
                       f"workload, {cores} threads, median of {reps}; matches={m}; "
                       f"seconds={t:.3f}")}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import tbjoin as tb

    tb.init(0)
    nthreads = min(16, os.cpu_count() or 1)
    t0 = time.time()
    R, S, desc = workload(tb, args.workload, nthreads)
    nR, nS = R.shape[0], S.shape[0]
    gen_s = time.time() - t0
    part = args.partitioner or ("learned" if args.workload == "c3" else "radix")
    mat = not args.count_only

    dev = torch.device("cuda:0")
    dR = torch.from_numpy(R.view(np.int64).reshape(-1)).to(dev)
    dS = torch.from_numpy(S.view(np.int64).reshape(-1)).to(dev)
    pairs = torch.empty(nS if mat else 1, dtype=torch.int64, device=dev)
    opts = dict(partitioner=part, profile=True)

    def step():
        return tb.join(args.algo, dR, dS, pairs if mat else None, **opts)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    results = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        results.append(step())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start

    matches = results[-1].matches
    expect = nS if args.workload in ("c1", "c2") else None
    for r in results:
        if r.matches != matches or (expect is not None and r.matches != expect):
            raise SystemExit(f"wrong match count {r.matches}")
    if mat:
        # spot-check materialised pairs: R[rp].key == S[sp].key
        got = pairs[: min(1 << 20, results[-1].pairs_written)].cpu().numpy()
        got = got.view(np.uint32).reshape(-1, 2)
        if not np.array_equal(R[got[:, 0], 0], S[got[:, 1], 0]):
            raise SystemExit("materialised pairs do not join")

    ms_per_step = elapsed / args.steps * 1e3
    value = (nR + nS) * args.steps / elapsed
    phases = {k: statistics.mean(r.as_dict()["phase_ms"][k] for r in results)
              for k in tb.Result.PHASES}

    # dominant single-launch kernel and its algorithmic bytes per launch
    n = nR + nS
    s_in = s_in_range(R, S) if part == "learned" else nS
    pair_b = 8 * matches if mat else 0
    kernels = {
        # k_join: reads every partitioned tuple once, writes the pairs
        "join": (phases["join"], 8 * (nR + s_in) + pair_b, "k_join"),
        # second-pass scatter: reads and writes every surviving tuple once
        "scatter2": (phases["scatter2"], 16 * (nR + s_in), "k_scatter (pass 2)"),
        # second-pass sizing: the per-tile histogram read
        "size2": (phases["size2"], 8 * (nR + s_in), "k_hist (pass 2) + scan"),
    }
    kname = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes, ksym = kernels[kname]
    achieved = kbytes / (kms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{ksym.split()[0]}_{args.workload}.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": ksym, "kernel_ms": round(kms, 4), "algorithmic_bytes": kbytes}
    # whole-step floor: every input tuple read once, the surviving ones
    # written once to their partition and read once for build/probe, plus pairs
    b_alg = 8 * n + 16 * (nR + s_in) + pair_b
    pipe = b_alg / (ms_per_step * 1e-3) / 1e9

    line = {
        "metric": METRIC, "value": value, "unit": "tuples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic: reference generators and seeds, payload = row index",
        "config": {"workload": desc, "algo": args.algo, "partitioner": part,
                   "partitioner_used": ["auto", "radix", "learned"][results[-1].partitioner_used],
                   "materialized_pairs": mat, "R": nR, "S": nS, "matches": matches,
                   "radix_bits": results[-1].radix_bits, "passes": results[-1].passes,
                   "phase_ms": {k: round(v, 4) for k, v in phases.items()},
                   "gen_seconds": round(gen_s, 1)},
        "roofline": roofline,
        "pipeline_roofline": {"achieved": round(pipe, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(pipe / HBM_PEAK_GBS, 4), "algorithmic_bytes": b_alg},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        del dR, dS, pairs
        torch.cuda.empty_cache()
        line["cpu_baseline"] = cpu_baseline(R, S, args.cpu_reps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
