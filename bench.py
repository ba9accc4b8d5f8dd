"""Benchmark of the equi-join hot path on one MI355X.

Headline workload (c2): RHO radix hash join, |R| = 2^27 unique shuffled keys
(create_relation_pk, seed 11111), |S| = 2^29 foreign keys (create_relation_fk
over R's keys, seed 22222), 8-byte tuples (uint32 key, uint32 payload = row
index), materialised (R.payload, S.payload) pairs.  A step is one full join
call (partition passes, build/probe, pair output) with R and S already
resident in HBM.  `value` is (|R|+|S|) tuples per second over the timed steps.

The default run (--workload c2) then measures the other GPU workloads on the
same invocation and reports each under its own key:
  c3: |R| = 2^27, |S| = 2^29 Zipf(1.0) over R's keys with every odd-indexed
      S key moved out of R's range (50% selectivity), materialised; the
      learned partitioner and plain radix on the same inputs;
  c5: 64-bit keys, |R| = |S| = 2^28 16-byte tuples, S Zipf(1.5) over R's
      keys, materialised; PSM with the learned range partitioner (output in
      key order) and RSM with radix partitions on the same inputs;
  c4: a generated batch of 4096 count-only RHO/PHT joins (|R_i| log-uniform
      in [2^10, 2^22), |S_i| = k |R_i|, k in 1..16, z ~ U[0, 1.5)), tuned by
      the autotuner, one step = the whole batch replayed as one HIP graph.
Each of them can also be run alone (--workload c3|c4|c5; c1 = the
reference's CPU-sized case, 2^21 x 2^21).

`roofline` rates the step's dominant kernel launch (the longest single
launch: the second-pass scatter or the join) against the HBM peak, from its
algorithmic bytes per launch and its average duration (HIP events on the
library's stream, over the timed steps); `traffic` is that launch's HBM bytes
from the committed counter summary.  `step_roofline` rates the whole step:
the partitioned-join byte floor T x (|R| + |S|) + 2T x (|R| + |S_in|) + P x
matches (T, P = tuple and pair bytes, S_in = S tuples inside R's key range)
over the step time, with the counter-measured HBM bytes of one whole step
(`measured_bytes`) beside it.  The CPU baseline is the reference's own RHO
(built from its sources into oracle/_ref) timed on this host on the same
inputs: materialising (one output-list node per match, like-for-like with
the GPU line) and count-only.
"""
import argparse
import json
import math
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
    ap.add_argument("--no-extras", action="store_true",
                    help="c2: skip the c3, c5 and c4 measurements")
    ap.add_argument("--extra-steps", type=int, default=8, help="timed steps of c3/c5 per side")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--joins", type=int, default=4096, help="c4: joins in the batch")
    ap.add_argument("--ceilings", action="store_true",
                    help="instead of a workload: time the chip's ceilings for the path's "
                         "access patterns (copy, scatter runs, random reads and atomics)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, run one plain 1 GiB device copy (a "
                         "known-size kernel for scaling HBM byte counters)")
    return ap.parse_args()


def host_cores():
    """CPUs this process may run on: its affinity set, capped by a cgroup CPU
    quota when one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, math.ceil(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


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
    """The reference's RHO (oracle/_ref, built from its sources) on this host:
    materialising once (the value), count-only `reps` times (median)."""
    import ctypes
    import numpy as np
    count_path = os.path.join(ROOT, "oracle", "_ref", "libref_count.so")
    mat_path = os.path.join(ROOT, "oracle", "_ref", "libref_mat.so")
    if not os.path.exists(count_path) or not os.path.exists(mat_path):
        return None
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
    args = [ctypes.c_char_p, u32p, ctypes.c_void_p, ctypes.c_uint64, u32p, ctypes.c_void_p,
            ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    Lc = ctypes.CDLL(count_path)
    Lc.ref_join_count.restype = ctypes.c_int64
    Lc.ref_join_count.argtypes = args + [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    Lm = ctypes.CDLL(mat_path)
    Lm.ref_join_mat_timed.restype = ctypes.c_int64
    Lm.ref_join_mat_timed.argtypes = args + [ctypes.POINTER(ctypes.c_double)]
    cores = host_cores()
    rk, rp = np.ascontiguousarray(R[:, 0]), np.ascontiguousarray(R[:, 1])
    sk, sp = np.ascontiguousarray(S[:, 0]), np.ascontiguousarray(S[:, 1])
    n = len(rk) + len(sk)
    us = ctypes.c_double()
    mm = Lm.ref_join_mat_timed(b"RHO", rk, rp.ctypes.data, len(rk), sk, sp.ctypes.data, len(sk),
                               cores, 1, ctypes.byref(us))
    t_mat = us.value * 1e-6
    times, mc = [], None
    for _ in range(reps):
        mc = Lc.ref_join_count(b"RHO", rk, rp.ctypes.data, len(rk), sk, sp.ctypes.data,
                               len(sk), cores, 1, None, ctypes.byref(us))
        times.append(us.value * 1e-6)
    t_cnt = statistics.median(times)
    return {"value": n / t_mat, "unit": "tuples/s", "cores": cores, "kind": "reference",
            "sample": (f"reference RHO (native build, JOIN_MATERIALIZE: one output-list node "
                       f"per match) on the full workload, {cores} threads on {cpu_model()}, "
                       f"one run; matches={mm}; seconds={t_mat:.3f}"),
            "count_only": {"value": n / t_cnt, "seconds": round(t_cnt, 4), "matches": mc,
                           "sample": f"reference RHO count only, median of {reps}"}}


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


def measured_step_bytes(workload_name, algo="RHO"):
    """HBM bytes of one whole step (all kernels) from the committed counter
    summary of that algorithm (RHO's is traffic_step_<workload>.txt)."""
    tag = workload_name if algo == "RHO" else f"{algo.lower()}_{workload_name}"
    path = os.path.join(ROOT, "profiles", f"traffic_step_{tag}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_step"):
            return int(line.split()[1])
    return None


def pattern_ceiling(ksym, bins=256):
    """The committed calibration's rate (GB/s) for the access pattern of a
    kernel (profiles/ceilings_mi355x.md, from bench.py --ceilings): a
    scatter's runs of 4096 / bins tuples into `bins` regions per XCD, or a
    plain copy for a kernel that streams its input and output.  None if not
    calibrated."""
    path = os.path.join(ROOT, "profiles", "ceilings_mi355x.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    if ksym == "k_scatter":
        for line in text.splitlines():
            cells = [c.strip() for c in line.strip("|").split("|")]
            run = 4096 // bins
            if len(cells) == 5 and cells[:3] == [str(bins), str(run), str(8 * run)]:
                return {"GB/s": float(cells[4]) * 1e3,
                        "pattern": f"runs of {8 * run} B, {bins} bins/XCD"}
    if ksym in ("k_join_mat", "k_join", "k_join_sort"):
        # the faster of torch's copy and one 32-KB run per tile (a streamed copy)
        best = None
        for line in text.splitlines():
            if line.startswith("Plain device copy"):
                best = float(line.split(":")[1].split()[0]) * 1e3
            cells = [c.strip() for c in line.strip("|").split("|")]
            if len(cells) == 5 and cells[:2] == ["1", "4096"]:
                best = max(best or 0.0, float(cells[4]) * 1e3)
        return {"GB/s": best, "pattern": "streamed copy"} if best else None
    return None


def random_read_ceiling(table_bytes):
    """Calibrated random 32-byte reads per second (profiles/ceilings_mi355x.md)
    for the smallest calibrated table at least table_bytes large (the largest
    calibrated one beyond it)."""
    path = os.path.join(ROOT, "profiles", "ceilings_mi355x.md")
    if not os.path.exists(path):
        return None
    rows = []
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0].endswith(" MB"):
            rows.append((int(cells[0].split()[0]) << 20, float(cells[2]) * 1e9))
    if not rows:
        return None
    fit = [r for r in sorted(rows) if r[0] >= table_bytes]
    return (fit[0] if fit else sorted(rows)[-1])


def roofline(kbytes, kms, ksym, wl, bins=256, **extra):
    achieved = kbytes / (kms * 1e-3) / 1e9
    d = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(ksym, wl),
         "kernel": ksym, "kernel_ms": round(kms, 4), "algorithmic_bytes": kbytes}
    pc = pattern_ceiling(ksym, bins)
    if pc:
        # the same launch against the measured rate of its access pattern
        d["pattern_ceiling"] = dict(pc, frac=round(achieved / pc["GB/s"], 4))
    d.update(extra)
    return d


def step_roofline(b_alg, ms_per_step, wl, algo):
    a = b_alg / (ms_per_step * 1e-3) / 1e9
    return {"achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(a / HBM_PEAK_GBS, 4), "algorithmic_bytes": b_alg,
            "measured_bytes": measured_step_bytes(wl, algo)}


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


def release_device_memory(tb):
    """Frees the library's workspace on the current stream and torch's cache."""
    import torch
    torch.cuda.synchronize()
    tb.release_stream(torch.cuda.current_stream().cuda_stream)
    torch.cuda.empty_cache()


def phases_of(tb, results):
    return {k: statistics.mean(r.as_dict()["phase_ms"][k] for r in results)
            for k in tb.Result.PHASES}


def side(tb, results, elapsed, n, part=None):
    """One side of a comparison: time, throughput and what the join reported."""
    r = results[-1]
    d = {"ms_per_step": round(elapsed / len(results) * 1e3, 4),
         "value": n * len(results) / elapsed, "matches": r.matches, "checksum": r.checksum,
         "phase_ms": {k: round(v, 4) for k, v in phases_of(tb, results).items()}}
    if part:
        d["partitioner"] = part
        d["partitioner_used"] = ["auto", "radix", "learned"][r.partitioner_used]
        d["fallback"] = r.fallback
        d["s_dropped"] = r.s_dropped
        d["max_r_partition"] = r.max_r_partition
        d["learned_model"] = ["none", "knots+cells", "linear"][r.learned_model]
    return d


def bench_join32(args, tb):
    """c1, c2, c3: 8-byte tuples through tbj_join."""
    import numpy as np
    import torch
    nthreads = min(16, host_cores())
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
        if res2[-1].matches != matches or res2[-1].checksum != results[-1].checksum:
            raise SystemExit("learned and radix partitioners disagree")
        comparison = side(tb, res2, el2, nR + nS, other)
    if args.calibrate_copy:
        calibrate_copy()

    ms_per_step = elapsed / args.steps * 1e3
    value = (nR + nS) * args.steps / elapsed
    phases = phases_of(tb, results)
    s_in = s_in_range(R, S) if args.workload == "c3" else nS
    pair_b = 8 * matches if mat else 0
    nopart = algo in ("PHT", "NPO_st", "CHT")
    probe_pattern = None
    if nopart:
        # one shared table, no partitioning: the probe kernel is the step's
        # longest (its time is the library's probe interval, HIP events)
        probe_ms = statistics.mean(r.ms_join for r in results)
        psym = "k_cht_probe" if algo == "CHT" else "k_pht_probe"
        # S read once, one R tuple read per match, pairs written
        kernels = {"probe": (probe_ms, 8 * nS + 8 * matches + pair_b, psym)}
        # + R read once and written once into the table
        b_alg = 16 * nR + 8 * nS + 8 * matches + pair_b
        # the probe's own bound: random reads into a table far larger than L2
        # (PHT: one 32-byte bucket per S tuple in nextpow2(|R|/2) x 32 B; CHT:
        # a bitmap entry and a dense slot per S tuple, the dense array |R| x 8 B)
        nb = 1 << max(0, (max(1, nR // 2) - 1).bit_length())
        table = nb * 32 if algo != "CHT" else nR * 8
        per_probe = 2 if algo == "CHT" else 1
        rc = random_read_ceiling(table)
        probe_pattern = None
        if rc:
            rate = per_probe * nS / (probe_ms * 1e-3)
            probe_pattern = {"unit": "random reads/s", "achieved": round(rate / 1e9, 2),
                             "ceiling": round(rc[1] / 1e9, 2), "table_MB": rc[0] >> 20,
                             "frac": round(rate / rc[1], 4),
                             "pattern": f"{per_probe} random read(s) per S tuple"}
    else:
        jsym = "k_join_sort" if algo in ("RSM", "PSM", "MWAY") else (
            "k_join_mat" if mat else "k_join")
        kernels = {
            # one launch: reads every partitioned tuple once, writes the pairs
            "join": (phases["join"], 8 * (nR + s_in) + pair_b, jsym),
            # second-pass scatter, one launch over R and S: reads and writes
            # every surviving tuple once
            "scatter2": (phases["scatter2"], 16 * (nR + s_in), "k_scatter"),
        }
        if phases["size2"] > phases["scatter2"]:  # exact plans: the per-tile histogram read
            kernels["size2"] = (phases["size2"], 8 * (nR + s_in), "k_hist")
        b_alg = 8 * (nR + nS) + 16 * (nR + s_in) + pair_b
    kname = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes, ksym = kernels[kname]
    line = {
        "metric": METRIC, "value": value, "unit": "tuples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic: reference generators and seeds, payload = row index",
        "config": {"workload": desc, "algo": algo, "partitioner": part,
                   "partitioner_used": ["auto", "radix", "learned"][results[-1].partitioner_used],
                   "materialized_pairs": mat, "R": nR, "S": nS, "matches": matches,
                   "checksum": results[-1].checksum,
                   "radix_bits": results[-1].radix_bits, "passes": results[-1].passes,
                   "max_r_partition": results[-1].max_r_partition,
                   "s_dropped": results[-1].s_dropped,
                   "phase_ms": {k: round(v, 4) for k, v in phases.items()},
                   "gen_seconds": round(gen_s, 1)},
        # the pass-2 scatter's bins: the plan's bits less the first pass's
        "roofline": roofline(kbytes, kms, ksym, args.workload,
                             bins=1 << (results[-1].radix_bits - (results[-1].radix_bits + 1) // 2)
                             if results[-1].passes == 2 else 256),
        "step_roofline": step_roofline(b_alg, ms_per_step, args.workload, algo),
        "cpu_baseline": None,
    }
    if nopart and probe_pattern:
        line["roofline"]["pattern_ceiling"] = probe_pattern
    if comparison:
        line["config"]["comparison"] = comparison
    del dR, dS, pairs
    release_device_memory(tb)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(R, S, args.cpu_reps)
    return line


def extra_c3(args, tb):
    """Learned vs radix partitioning on the c3 inputs: both sides' times, the
    same matches and checksum, the S tuples the learned pass 1 dropped."""
    import numpy as np
    import torch
    t0 = time.time()
    R, S, desc = workload(tb, "c3", min(16, host_cores()))
    gen_s = time.time() - t0
    n = R.shape[0] + S.shape[0]
    dR = torch.from_numpy(R.view(np.int64).reshape(-1)).to("cuda:0")
    dS = torch.from_numpy(S.view(np.int64).reshape(-1)).to("cuda:0")
    pairs = torch.empty(S.shape[0], dtype=torch.int64, device="cuda:0")
    del R, S
    out = {"workload": desc + "; RHO, materialised pairs", "gen_seconds": round(gen_s, 1)}
    for p in ("learned", "radix"):
        res, el = timed(lambda: tb.join("RHO", dR, dS, pairs, partitioner=p, profile=True),
                        args.extra_steps, 2)
        out[p] = side(tb, res, el, n, p)
    L, X = out["learned"], out["radix"]
    out["matches_equal"] = L["matches"] == X["matches"]
    out["checksum_equal"] = L["checksum"] == X["checksum"]
    if not (out["matches_equal"] and out["checksum_equal"]):
        raise SystemExit("c3: learned and radix partitioners disagree")
    out["learned_speedup"] = round(X["ms_per_step"] / L["ms_per_step"], 4)
    del dR, dS, pairs
    release_device_memory(tb)
    return out


def extra_c5(args, tb):
    """PSM with the learned range partitioner vs RSM with radix partitions on
    the c5 inputs (64-bit keys, sorted materialised output)."""
    import numpy as np
    import torch
    n = 1 << 28
    t0 = time.time()
    R = tb.gen_unique64(11111, n, nthreads=min(16, host_cores()))
    S = tb.gen_zipf_fk64(22222, n, R, 1.5, nthreads=min(16, host_cores()))
    gen_s = time.time() - t0
    dR = torch.from_numpy(R.view(np.int64)).to("cuda:0")
    dS = torch.from_numpy(S.view(np.int64)).to("cuda:0")
    pairs = torch.empty((n, 2), dtype=torch.int64, device="cuda:0")
    out = {"workload": ("|R|=|S|=2^28 16-byte tuples (uint64 key, uint64 payload), unique "
                        "random 64-bit R keys, S Zipf(1.5) over R's keys; materialised"),
           "gen_seconds": round(gen_s, 1)}
    for name, algo, p in (("psm_learned", "PSM", "learned"), ("rsm_radix", "RSM", "radix")):
        res, el = timed(lambda: tb.join64(algo, dR, dS, pairs, partitioner=p, profile=True),
                        args.extra_steps, 2)
        if any(r.matches != n for r in res):
            raise SystemExit(f"c5: {algo} wrong match count")
        out[name] = side(tb, res, el, 2 * n, p)
        if name == "psm_learned":  # key order over the whole output (spot check)
            got = pairs[: 1 << 20].cpu().numpy().view(np.uint64)
            keys = R[got[:, 0].astype(np.int64), 0]
            if not (np.array_equal(keys, S[got[:, 1].astype(np.int64), 0])
                    and np.all(keys[1:] >= keys[:-1])):
                raise SystemExit("c5: PSM pairs do not join in key order")
    out["psm_learned"]["output_order"] = "key order over the whole output"
    out["rsm_radix"]["output_order"] = "key order inside each radix partition"
    out["matches_equal"] = out["psm_learned"]["matches"] == out["rsm_radix"]["matches"]
    out["checksum_equal"] = out["psm_learned"]["checksum"] == out["rsm_radix"]["checksum"]
    out["learned_speedup"] = round(out["rsm_radix"]["ms_per_step"] /
                                   out["psm_learned"]["ms_per_step"], 4)
    del dR, dS, pairs, R, S
    release_device_memory(tb)
    return out


def run_batch(args, tb, steps, warmup):
    """The generated c4 batch: autotuned, captured, replayed."""
    import numpy as np
    import torch
    specs, total = tb.gen_batch_specs(0xC4C4000000000000, args.joins)
    arena = torch.empty(total, dtype=torch.int64, device="cuda:0")
    t0 = time.time()
    tb.gen_batch_device(arena, specs)
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    items = tb.batch_items(list(specs))  # each join's own RHO / PHT choice
    counts = torch.zeros(args.joins, dtype=torch.int64, device="cuda:0")
    t0 = time.time()
    tuning, tune_ms = tb.batch_autotune(arena, items, counts, reps=1)
    tune_s = time.time() - t0
    b = tb.Batch(arena, items, counts, tuning)
    b.capture()
    results, elapsed = timed(b.run, steps, warmup)
    want = np.array([s.nS for s in specs], dtype=np.int64)
    if not np.array_equal(counts.cpu().numpy(), want):
        raise SystemExit("batch counts differ from |S_i|")
    if args.calibrate_copy:
        calibrate_copy()
    b.close()
    del arena, counts
    torch.cuda.empty_cache()
    nR = sum(s.nR for s in specs)
    npht = sum(1 for s in specs if s.algo == tb.ALGOS["PHT"])
    return {"specs": specs, "total": total, "nR": nR, "npht": npht, "results": results,
            "elapsed": elapsed, "tuning": tuning, "tune_ms": tune_ms, "tune_s": tune_s,
            "gen_s": gen_s}


C4_DESC = ("{joins} joins ({nrho} RHO, {npht} PHT), |R_i| log-uniform in [2^10, 2^22), "
           "|S_i| = k|R_i|, k in 1..16, Zipf z ~ U[0, 1.5) over R_i's keys; count only")


def bench_batch(args, tb):
    """c4 alone: one step = the whole batch."""
    B = run_batch(args, tb, args.steps, args.warmup)
    total, results, elapsed = B["total"], B["results"], B["elapsed"]
    ms_per_step = elapsed / args.steps * 1e3
    dev_ms = statistics.mean(r.ms for r in results)
    b_alg = 24 * total
    return {
        "metric": METRIC, "value": total * args.steps / elapsed, "unit": "tuples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32",
        "data": "synthetic: device-generated batch (Feistel-permuted pk, Zipf fk)",
        "config": {"workload": C4_DESC.format(joins=args.joins, nrho=args.joins - B["npht"],
                                              npht=B["npht"]),
                   "joins": args.joins, "tuples": total, "R": B["nR"], "S": total - B["nR"],
                   "matches": results[-1].matches, "waves": results[-1].waves,
                   "groups": results[-1].groups, "tuning": B["tuning"].as_dict(),
                   "autotune_best_ms": round(B["tune_ms"], 3),
                   "autotune_seconds": round(B["tune_s"], 1),
                   "graph_replay": True, "device_ms": round(dev_ms, 3),
                   "gen_seconds": round(B["gen_s"], 1)},
        "roofline": roofline(b_alg, dev_ms, "batch", "c4",
                             note="whole batch (all kernels of the graph), not one kernel"),
        "step_roofline": step_roofline(b_alg, ms_per_step, "c4", "batch"),
        "cpu_baseline": None,
    }


def extra_c4(args, tb):
    B = run_batch(args, tb, 3, 1)
    total, results, elapsed = B["total"], B["results"], B["elapsed"]
    return {"workload": C4_DESC.format(joins=args.joins, nrho=args.joins - B["npht"],
                                       npht=B["npht"]),
            "ms_per_step": round(elapsed / len(results) * 1e3, 4),
            "value": total * len(results) / elapsed, "unit": "tuples/s",
            "device_ms": round(statistics.mean(r.ms for r in results), 3),
            "tuples": total, "matches": results[-1].matches,
            "matches_closed_form": True, "waves": results[-1].waves,
            "groups": results[-1].groups, "tuning": B["tuning"].as_dict(),
            "autotune_seconds": round(B["tune_s"], 1), "graph_replay": True}


def bench_join64(args, tb):
    """c5 alone: 64-bit keys, sorted materialised output."""
    import numpy as np
    import torch
    n = 1 << 28
    t0 = time.time()
    R = tb.gen_unique64(11111, n, nthreads=min(16, host_cores()))
    S = tb.gen_zipf_fk64(22222, n, R, 1.5, nthreads=min(16, host_cores()))
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
    res2, el2 = timed(make_step(other), max(3, args.steps // 2), 1)
    if args.calibrate_copy:
        calibrate_copy()
    ms_per_step = elapsed / args.steps * 1e3
    phases = phases_of(tb, results)
    b_alg = 16 * 2 * n + 32 * 2 * n + (16 * n if mat else 0)
    kernels = {
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
                   "comparison": side(tb, res2, el2, 2 * n, other),
                   "gen_seconds": round(gen_s, 1)},
        "roofline": roofline(kbytes, kms, ksym, "c5"),
        "step_roofline": step_roofline(b_alg, ms_per_step, "c5", algo),
        "cpu_baseline": None,
    }


def ceilings(tb):
    """Rates of the access patterns that bound the path's kernels, on this
    device: a plain copy; a partition pass's writes as runs of `run` 8-byte
    tuples into `bins` regions per XCD (tiles of bins x run tuples); random
    8- and 32-byte reads and u32 atomics into tables of growing size."""
    import torch
    n = 1 << 29
    src = torch.ones(n, dtype=torch.int64, device="cuda:0")
    dst = torch.empty_like(src)

    def best(f, reps=3):
        f()
        return min(f() for _ in range(reps))

    def copy_ms():
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1)

    out = {"copy": {"bytes": 16 * n, "ms": round(best(copy_ms), 4)}}
    out["copy"]["TB/s"] = round(16 * n / out["copy"]["ms"] / 1e9, 3)
    runs = []
    for bins, run in ((1, 4096), (64, 64), (128, 32), (256, 16), (256, 32), (512, 8), (512, 16),
                      (2048, 2), (2048, 4), (2048, 8), (8192, 1), (8192, 2), (16384, 1)):
        ms = best(lambda: tb.calibrate_runs(src, dst, bins, run))
        runs.append({"bins_per_xcd": bins, "run_tuples": run, "run_bytes": 8 * run,
                     "ms": round(ms, 4), "TB/s": round(16 * n / ms / 1e9, 3)})
    out["scatter_runs"] = runs
    del dst
    table = torch.full((1 << 28,), 0x01010101, dtype=torch.int64, device="cuda:0")
    sink = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    nprobe = 1 << 28
    rnd = []
    for mb in (2, 8, 32, 256, 1024, 2048):
        row = {"table_MB": mb}
        for kind, name in ((tb.CAL_READ8, "read8"), (tb.CAL_READ32, "read32"),
                           (tb.CAL_ATOMIC, "atomic_u32")):
            ms = best(lambda: tb.calibrate_random(table, mb << 20, nprobe, kind, sink))
            row[name + "_G/s"] = round(nprobe / ms / 1e6, 1)
        rnd.append(row)
    out["random"] = rnd
    return {"ceilings": out, "device": torch.cuda.get_device_name(0),
            "note": ("scatter_runs: TB/s counts 8 B read + 8 B written per tuple; random: "
                     "accesses per second (2^28 per launch, 8 in flight per thread)")}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import tbjoin as tb
    tb.init(0)
    if args.ceilings:
        print(json.dumps(ceilings(tb)), flush=True)
        return
    if args.workload == "c4":
        line = bench_batch(args, tb)
    elif args.workload == "c5":
        line = bench_join64(args, tb)
    else:
        line = bench_join32(args, tb)
        if args.workload == "c2" and not args.no_extras:
            for key, fn in (("c3", extra_c3), ("c5", extra_c5), ("c4", extra_c4)):
                t0 = time.time()
                print(f"[bench] {key}", file=sys.stderr, flush=True)
                line[key] = fn(args, tb)
                line[key]["wall_seconds"] = round(time.time() - t0, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
