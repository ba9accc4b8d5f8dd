"""Benchmark of the equi-join hot path on one MI355X.

Default workload (c2): RHO radix hash join, |R| = 2^27 unique shuffled keys
(create_relation_pk, seed 11111), |S| = 2^29 foreign keys (create_relation_fk
over R's keys, seed 22222), 8-byte tuples (uint32 key, uint32 payload = row
index), materialised (R.payload, S.payload) pairs.  A step is one full join
call (partition passes, build/probe, pair output) with R and S already
resident in HBM.

Other workloads: c3 = |R| = 2^27, |S| = 2^29 Zipf(1.0) over R's keys with
every odd-indexed S key moved out of R's range, learned partitioner, radix
timed beside it on the same inputs; c4 = a batch of 4096 generated joins
(count only; one step = the whole batch, replayed as one HIP graph, tuning
from the autotuner); c5 = 64-bit keys, |R| = |S| = 2^28 16-byte tuples, S
Zipf(1.5) over R's keys, PSM with the learned range partitioner, sorted
materialised output; c1 = the reference's CPU-sized case (2^21 x 2^21).

Prints one JSON line.  `value` is (|R|+|S|) tuples per second over the timed
steps.  `roofline` rates the longest single kernel of the step against the
HBM peak from its algorithmic bytes and its average duration (HIP events on
the library's stream, over the timed steps); `pipeline_roofline` rates the
whole step with the partitioned-join byte floor T x (|R| + |S|) +
2T x (|R| + |S_in|) + P x matches (T, P = tuple and pair bytes, S_in = S
tuples inside R's key range); for the no-partitioning joins (--algo PHT,
NPO_st, CHT) the rated kernel is the probe (T x |S| + (T + P) x matches) and
the floor 2T x |R| + T x |S| + (T + P) x matches.  The CPU baseline is the reference's own RHO
(built from its sources into oracle/_ref) timed on this host's cores on the
same inputs.
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
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--algo", default=None)
    ap.add_argument("--partitioner", default=None, choices=["radix", "learned"])
    ap.add_argument("--count-only", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--joins", type=int, default=4096, help="c4: joins in the batch")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, run one plain 1 GiB device copy (a "
                         "known-size kernel for scaling HBM byte counters)")
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


def measured_traffic(kernel, workload_name):
    """HBM bytes per launch from the committed counter summary, if any."""
    # a phase of several kernels ("a+b+c") is filed as a_b_c
    path = os.path.join(ROOT, "profiles",
                        f"traffic_{kernel.replace('+', '_')}_{workload_name}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch"):
            return int(line.split()[1])
    return None


def roofline(kbytes, kms, ksym, wl, **extra):
    achieved = kbytes / (kms * 1e-3) / 1e9
    d = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(ksym, wl),
         "kernel": ksym, "kernel_ms": round(kms, 4), "algorithmic_bytes": kbytes}
    d.update(extra)
    return d


def calibrate_copy():
    """One plain device copy of 1 GiB (read 1 GiB, write 1 GiB)."""
    import torch
    x = torch.empty(1 << 28, dtype=torch.int32, device="cuda:0")
    x.fill_(1)
    torch.cuda.synchronize()
    y = x.clone()
    torch.cuda.synchronize()
    del x, y


def timed(step, steps, warmup):
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    results = []
    t0 = time.perf_counter()
    for _ in range(steps):
        results.append(step())
    torch.cuda.synchronize()
    return results, time.perf_counter() - t0


def bench_join32(args, tb):
    """c1, c2, c3: 8-byte tuples through tbj_join."""
    import numpy as np
    import torch
    nthreads = min(16, os.cpu_count() or 1)
    t0 = time.time()
    R, S, desc = workload(tb, args.workload, nthreads)
    nR, nS = R.shape[0], S.shape[0]
    gen_s = time.time() - t0
    algo = args.algo or "RHO"
    part = args.partitioner or ("learned" if args.workload == "c3" else "radix")
    mat = not args.count_only
    dev = torch.device("cuda:0")
    dR = torch.from_numpy(R.view(np.int64).reshape(-1)).to(dev)
    dS = torch.from_numpy(S.view(np.int64).reshape(-1)).to(dev)
    pairs = torch.empty(nS if mat else 1, dtype=torch.int64, device=dev)

    def make_step(p):
        return lambda: tb.join(algo, dR, dS, pairs if mat else None, partitioner=p, profile=True)

    results, elapsed = timed(make_step(part), args.steps, args.warmup)
    matches = results[-1].matches
    expect = {"c1": nS, "c2": nS}.get(args.workload)
    for r in results:
        if r.matches != matches or (expect is not None and r.matches != expect):
            raise SystemExit(f"wrong match count {r.matches}")
    if mat:  # spot-check materialised pairs: R[rp].key == S[sp].key
        got = pairs[: min(1 << 20, results[-1].pairs_written)].cpu().numpy()
        got = got.view(np.uint32).reshape(-1, 2)
        if not np.array_equal(R[got[:, 0], 0], S[got[:, 1], 0]):
            raise SystemExit("materialised pairs do not join")
    comparison = None
    if args.workload == "c3":  # the other partitioner on the same inputs
        other = "radix" if part == "learned" else "learned"
        res2, el2 = timed(make_step(other), max(3, args.steps // 2), 1)
        if res2[-1].matches != matches:
            raise SystemExit("learned and radix partitioners disagree")
        comparison = {"partitioner": other,
                      "value": (nR + nS) * len(res2) / el2,
                      "ms_per_step": round(el2 / len(res2) * 1e3, 4),
                      "partitioner_used": ["auto", "radix", "learned"][res2[-1].partitioner_used]}
    if args.calibrate_copy:
        calibrate_copy()

    ms_per_step = elapsed / args.steps * 1e3
    value = (nR + nS) * args.steps / elapsed
    phases = {k: statistics.mean(r.as_dict()["phase_ms"][k] for r in results)
              for k in tb.Result.PHASES}
    s_in = s_in_range(R, S) if args.workload == "c3" else nS
    pair_b = 8 * matches if mat else 0
    nopart = algo in ("PHT", "NPO_st", "CHT")
    if nopart:
        # one shared table, no partitioning: the probe kernel is the step's
        # longest (its time is the library's probe interval, HIP events)
        probe_ms = statistics.mean(r.ms_join for r in results)
        psym = "k_cht_probe" if algo == "CHT" else "k_pht_probe"
        # S read once, one R tuple read per match, pairs written
        kernels = {"probe": (probe_ms, 8 * nS + 8 * matches + pair_b, psym)}
        # + R read once and written once into the table
        b_alg = 16 * nR + 8 * nS + 8 * matches + pair_b
    else:
        kernels = None
    if algo in ("RSM", "PSM"):
        jsym = "k_join_sort"
    else:
        jsym = "k_join_mat" if mat else "k_join"
    kernels = kernels or {
        # k_join / k_join_mat / k_join_sort: reads every partitioned tuple
        # once, writes the pairs
        "join": (phases["join"], 8 * (nR + s_in) + pair_b, jsym),
        # second-pass scatter: reads and writes every surviving tuple once
        "scatter2": (phases["scatter2"], 16 * (nR + s_in), "k_scatter"),
        # second-pass sizing: the per-tile histogram read
        "size2": (phases["size2"], 8 * (nR + s_in), "k_hist"),
    }
    kname = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes, ksym = kernels[kname]
    if not nopart:
        b_alg = 8 * (nR + nS) + 16 * (nR + s_in) + pair_b
    pipe = b_alg / (ms_per_step * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": value, "unit": "tuples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic: reference generators and seeds, payload = row index",
        "config": {"workload": desc, "algo": algo, "partitioner": part,
                   "partitioner_used": ["auto", "radix", "learned"][results[-1].partitioner_used],
                   "materialized_pairs": mat, "R": nR, "S": nS, "matches": matches,
                   "radix_bits": results[-1].radix_bits, "passes": results[-1].passes,
                   "max_r_partition": results[-1].max_r_partition,
                   "s_dropped": results[-1].s_dropped,
                   "phase_ms": {k: round(v, 4) for k, v in phases.items()},
                   "gen_seconds": round(gen_s, 1)},
        "roofline": roofline(kbytes, kms, ksym, args.workload),
        "pipeline_roofline": {"achieved": round(pipe, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(pipe / HBM_PEAK_GBS, 4), "algorithmic_bytes": b_alg},
        "cpu_baseline": None,
    }
    if comparison:
        line["config"]["comparison"] = comparison
    if not args.no_cpu_baseline:
        del dR, dS, pairs
        torch.cuda.empty_cache()
        line["cpu_baseline"] = cpu_baseline(R, S, args.cpu_reps)
    return line


def bench_batch(args, tb):
    """c4: a generated batch of count-only joins, one step = the whole batch."""
    import numpy as np
    import torch
    specs, total = tb.gen_batch_specs(0xC4C4000000000000, args.joins)
    arena = torch.empty(total, dtype=torch.int64, device="cuda:0")
    t0 = time.time()
    tb.gen_batch_device(arena, specs)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    items = tb.batch_items(list(specs))
    counts = torch.zeros(args.joins, dtype=torch.int64, device="cuda:0")
    t0 = time.time()
    tuning, tune_ms = tb.batch_autotune(arena, items, counts, reps=1)
    tune_s = time.time() - t0
    b = tb.Batch(arena, items, counts, tuning)
    b.capture()
    results, elapsed = timed(b.run, args.steps, args.warmup)
    want = np.array([s.nS for s in specs], dtype=np.int64)
    if not np.array_equal(counts.cpu().numpy(), want):
        raise SystemExit("batch counts differ from |S_i|")
    if args.calibrate_copy:
        calibrate_copy()
    ms_per_step = elapsed / args.steps * 1e3
    dev_ms = statistics.mean(r.ms for r in results)
    b_alg = 24 * total
    nR = sum(s.nR for s in specs)
    line = {
        "metric": METRIC, "value": total * args.steps / elapsed, "unit": "tuples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32",
        "data": "synthetic: device-generated batch (Feistel-permuted pk, Zipf/uniform fk)",
        "config": {"workload": (f"{args.joins} RHO joins, |R_i| log-uniform in [2^10, 2^22), "
                                "|S_i| = k|R_i|, k in 1..16, z in {0, 0.1, ..., 1.5}; count only"),
                   "joins": args.joins, "tuples": total, "R": nR, "S": total - nR,
                   "matches": results[-1].matches, "waves": results[-1].waves,
                   "groups": results[-1].groups, "tuning": tuning.as_dict(),
                   "autotune_best_ms": round(tune_ms, 3), "autotune_seconds": round(tune_s, 1),
                   "graph_replay": True, "device_ms": round(dev_ms, 3),
                   "gen_seconds": round(gen_s, 1)},
        "roofline": roofline(b_alg, dev_ms, "batch", "c4",
                             note="whole batch (all kernels of the graph), not one kernel"),
        "pipeline_roofline": {"achieved": round(b_alg / (ms_per_step * 1e-3) / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(b_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                              "algorithmic_bytes": b_alg},
        "cpu_baseline": None,
    }
    b.close()
    return line


def bench_join64(args, tb):
    """c5: 64-bit keys, sorted materialised output."""
    import numpy as np
    import torch
    n = 1 << 28
    t0 = time.time()
    R = tb.gen_unique64(11111, n, nthreads=16)
    S = tb.gen_zipf_fk64(22222, n, R, 1.5, nthreads=16)
    gen_s = time.time() - t0
    algo = args.algo or "PSM"
    part = args.partitioner or ("learned" if algo == "PSM" else "radix")
    mat = not args.count_only
    dR = torch.from_numpy(R.view(np.int64)).to("cuda:0")
    dS = torch.from_numpy(S.view(np.int64)).to("cuda:0")
    pairs = torch.empty((n, 2), dtype=torch.int64, device="cuda:0") if mat else None

    def make_step(p):
        return lambda: tb.join64(algo, dR, dS, pairs, partitioner=p, profile=True)

    results, elapsed = timed(make_step(part), args.steps, args.warmup)
    for r in results:
        if r.matches != n:
            raise SystemExit(f"wrong match count {r.matches}")
    if mat:  # every pair joins and keys are non-decreasing (spot check)
        got = pairs[: 1 << 20].cpu().numpy().view(np.uint64)
        keys = R[got[:, 0].astype(np.int64), 0]
        if not np.array_equal(keys, S[got[:, 1].astype(np.int64), 0]):
            raise SystemExit("materialised pairs do not join")
        if part == "learned" and not np.all(keys[1:] >= keys[:-1]):
            raise SystemExit("pairs are not in key order")
    other = "radix" if part == "learned" else "learned"
    # range partitions are key ranges in key order; radix partitions are not
    order = {"learned": "key order over the whole output",
             "radix": "key order inside each radix partition"}
    res2, el2 = timed(make_step(other), max(3, args.steps // 2), 1)
    if args.calibrate_copy:
        calibrate_copy()
    ms_per_step = elapsed / args.steps * 1e3
    phases = {k: statistics.mean(r.as_dict()["phase_ms"][k] for r in results)
              for k in tb.Result.PHASES}
    b_alg = 16 * 2 * n + 32 * 2 * n + (16 * n if mat else 0)
    kernels = {
        "scatter1": (phases["scatter1"] / 2, 32 * n, "k_scatter64"),  # two launches (R, S)
        "scatter2": (phases["scatter2"], 32 * 2 * n, "k_scatter64"),
        "join": (phases["join"], 16 * 2 * n + (16 * n if mat else 0), "count+offsets+emit"),
    }
    kname = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes, ksym = kernels[kname]
    return {
        "metric": METRIC, "value": 2 * n * args.steps / elapsed, "unit": "tuples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: unique splitmix64 R keys, Zipf(1.5) S over R's keys, payload = row",
        "config": {"workload": "|R|=|S|=2^28 16-byte tuples, S Zipf(1.5) over R's keys",
                   "algo": algo, "partitioner": part, "materialized_pairs": mat,
                   "sorted_output": part == "learned", "matches": results[-1].matches,
                   "radix_bits": results[-1].radix_bits,
                   "max_r_partition": results[-1].max_r_partition,
                   "phase_ms": {k: round(v, 4) for k, v in phases.items()},
                   "comparison": {"partitioner": other,
                                  "value": 2 * n * len(res2) / el2,
                                  "ms_per_step": round(el2 / len(res2) * 1e3, 4),
                                  "output_order": order[other]},
                   "output_order": order[part],
                   "gen_seconds": round(gen_s, 1)},
        "roofline": roofline(kbytes, kms, ksym, "c5"),
        "pipeline_roofline": {"achieved": round(b_alg / (ms_per_step * 1e-3) / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(b_alg / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                              "algorithmic_bytes": b_alg},
        "cpu_baseline": None,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import tbjoin as tb
    tb.init(0)
    if args.workload == "c4":
        line = bench_batch(args, tb)
    elif args.workload == "c5":
        line = bench_join64(args, tb)
    else:
        line = bench_join32(args, tb)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
