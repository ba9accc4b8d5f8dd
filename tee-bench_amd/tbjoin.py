"""ctypes binding of libtbj.so, the MI355X equi-join library (include/tbj.h).

Mirrors the reference's join operator API: ``join("RHO", R, S)`` is the
device-side counterpart of ``result_t* RHO(relation_t*, relation_t*, int)``
(Joins/radix_join.c:1557-1561), selected by the same names as the reference's
``-a`` flag (native.cpp:34-46).  Relations are device tensors of 8-byte tuples
viewed as int64/uint32 pairs; PyTorch only provides the HBM buffers.

The library is required: every entry point raises when libtbj.so is missing
or no HIP device is visible.  There is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtbj.so")

ALGOS = {"RHO": 0, "RHT": 1, "RJ": 2, "PHT": 3, "CHT": 4, "RSM": 5, "PSM": 6,
         "RHO_atomic": 7, "NPO_st": 8, "MWAY": 9}
OUT_PAIRS, OUT_RELATION = 0, 1
PART_AUTO, PART_RADIX, PART_LEARNED = 0, 1, 2
PARTITIONERS = {"auto": PART_AUTO, "radix": PART_RADIX, "learned": PART_LEARNED}

TBJ_OK = 0
ERR_INVALID, ERR_NO_DEVICE, ERR_OOM, ERR_HIP, ERR_UNSUPPORTED, ERR_TOO_LARGE = -1, -2, -3, -4, -5, -6


class Options(ctypes.Structure):
    _fields_ = [("partitioner", ctypes.c_int32), ("radix_bits", ctypes.c_int32),
                ("passes", ctypes.c_int32), ("materialize", ctypes.c_int32),
                ("workspace_cap", ctypes.c_uint64), ("learned_bound", ctypes.c_uint32),
                ("learned_samples", ctypes.c_int32), ("s_chunk", ctypes.c_int32),
                ("sampled_regions", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("output", ctypes.c_int32), ("bulk_build", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 1)]


class Result(ctypes.Structure):
    _fields_ = [("matches", ctypes.c_uint64), ("checksum", ctypes.c_uint64),
                ("pairs_written", ctypes.c_uint64), ("partitioner_used", ctypes.c_int32),
                ("fallback", ctypes.c_int32), ("radix_bits", ctypes.c_int32),
                ("passes", ctypes.c_int32), ("num_partitions", ctypes.c_uint32),
                ("max_r_partition", ctypes.c_uint32), ("s_dropped", ctypes.c_uint64),
                ("ms_partition", ctypes.c_float), ("ms_join", ctypes.c_float),
                ("ms_total", ctypes.c_float), ("redone_exact", ctypes.c_int32),
                ("learned_model", ctypes.c_int32),
                ("phase_ms", ctypes.c_float * 8)]

    PHASES = ("hist1", "scan1", "scatter1", "size2", "scatter2", "tasks", "join", "other")

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "phase_ms"}
        d["phase_ms"] = dict(zip(self.PHASES, list(self.phase_ms)))
        return d


class TbjError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"tbj status {status}: {msg}")
        self.status = status


_lib = None


def lib():
    """Loads libtbj.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # torch ships its own HIP runtime: load it first so that libtbj.so
        # binds to that same runtime instead of a second copy from /opt/rocm
        # (two runtimes in one process hide the GPU from one of them)
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        vp, u64 = ctypes.c_void_p, ctypes.c_uint64
        L.tbj_init.argtypes = [ctypes.c_int]
        L.tbj_init.restype = ctypes.c_int
        L.tbj_status_str.argtypes = [ctypes.c_int]
        L.tbj_status_str.restype = ctypes.c_char_p
        L.tbj_algo_name.argtypes = [ctypes.c_int]
        L.tbj_algo_name.restype = ctypes.c_char_p
        L.tbj_algo_from_name.argtypes = [ctypes.c_char_p]
        L.tbj_algo_from_name.restype = ctypes.c_int
        L.tbj_default_options.argtypes = [ctypes.POINTER(Options)]
        L.tbj_join.argtypes = [ctypes.c_int, vp, u64, vp, u64, vp, u64,
                               ctypes.POINTER(Options), ctypes.POINTER(Result), vp]
        L.tbj_join.restype = ctypes.c_int
        L.tbj_join_host.argtypes = [ctypes.c_int, vp, u64, vp, u64, vp, u64,
                                    ctypes.POINTER(Options), ctypes.POINTER(Result)]
        L.tbj_join_host.restype = ctypes.c_int
        L.tbj_join64.argtypes = [ctypes.c_int, vp, u64, vp, u64, vp, u64,
                                 ctypes.POINTER(Options), ctypes.POINTER(Result), vp]
        L.tbj_join64.restype = ctypes.c_int
        L.tbj_join3.argtypes = [ctypes.c_int, vp, u64, vp, u64, vp, u64, vp, u64,
                                ctypes.POINTER(Options), ctypes.POINTER(Result), vp]
        L.tbj_join3.restype = ctypes.c_int
        L.tbj_join_noalloc.argtypes = L.tbj_join.argtypes
        L.tbj_join_noalloc.restype = ctypes.c_int
        L.tbj_release_stream.argtypes = [vp]
        L.tbj_release_stream.restype = ctypes.c_int
        L.tbj_calibrate_runs.argtypes = [vp, vp, u64, ctypes.c_uint32, ctypes.c_uint32, vp,
                                         ctypes.POINTER(ctypes.c_float)]
        L.tbj_calibrate_runs.restype = ctypes.c_int
        L.tbj_calibrate_random.argtypes = [vp, u64, u64, ctypes.c_int, vp, vp,
                                           ctypes.POINTER(ctypes.c_float)]
        L.tbj_calibrate_random.restype = ctypes.c_int
        _lib = L
    return _lib


def check(status):
    if status != TBJ_OK:
        raise TbjError(status, lib().tbj_status_str(status).decode())


def init(device=0):
    check(lib().tbj_init(device))


def make_options(partitioner="auto", radix_bits=0, passes=0, materialize=False,
                 workspace_cap=0, learned_bound=0, learned_samples=0, s_chunk=0,
                 sampled_regions=False, profile=False, output=OUT_PAIRS, bulk_build=0):
    o = Options()
    lib().tbj_default_options(ctypes.byref(o))
    o.partitioner = PARTITIONERS[partitioner] if isinstance(partitioner, str) else int(partitioner)
    o.radix_bits, o.passes = radix_bits, passes
    o.materialize = 1 if materialize else 0
    o.workspace_cap, o.learned_bound = workspace_cap, learned_bound
    o.learned_samples, o.s_chunk = learned_samples, s_chunk
    # True = always, False = auto; -1 = never
    o.sampled_regions = int(sampled_regions)
    o.profile = 1 if profile else 0
    o.output = int(output)
    # PHT / NPO_st table build: 0 = auto, 1 = always bulk, -1 = atomic inserts
    o.bulk_build = int(bulk_build)
    return o


def _algo(name):
    if isinstance(name, int):
        return name
    if name not in ALGOS:
        raise ValueError(f"unknown join algorithm {name!r}")
    return ALGOS[name]


def join_ptr(algo, r_ptr, nR, s_ptr, nS, pairs_ptr=0, pairs_cap=0, options=None, stream=0):
    """Joins device relations given as raw HBM pointers; returns a Result."""
    o = options if options is not None else make_options()
    res = Result()
    check(lib().tbj_join(_algo(algo), r_ptr, nR, s_ptr, nS, pairs_ptr or None, pairs_cap,
                         ctypes.byref(o), ctypes.byref(res), stream or None))
    return res


def join(algo, R, S, pairs=None, **opts):
    """Joins device tensors R, S (int64 [n] or uint32/int32 [n, 2], 8-byte tuples).

    pairs: optional device tensor with room for the (R.payload, S.payload)
    matches (int64 [cap] or int32 [cap, 2]); materialises when given.
    """
    for t in (R, S):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("relations must be contiguous device tensors")
    nR = R.numel() * R.element_size() // 8
    nS = S.numel() * S.element_size() // 8
    cap = 0
    if pairs is not None:
        cap = pairs.numel() * pairs.element_size() // 8
        opts.setdefault("materialize", True)
    o = make_options(**opts)
    import torch
    stream = torch.cuda.current_stream(R.device).cuda_stream
    return join_ptr(algo, R.data_ptr(), nR, S.data_ptr(), nS,
                    pairs.data_ptr() if pairs is not None else 0, cap, o, stream)


def join3(R, S, T, out, algo="RHO", **opts):
    """Three-way join (R join S) join T on the device: out receives the
    (key, T.payload) tuples of T joined with out1 = the (key, R.payload)
    tuples of R join S (int64 [cap] or uint32 [cap, 2])."""
    for t in (R, S, T, out):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("relations must be contiguous device tensors")
    n = [x.numel() * x.element_size() // 8 for x in (R, S, T, out)]
    o = make_options(**opts)
    import torch
    stream = torch.cuda.current_stream(R.device).cuda_stream
    res = Result()
    check(lib().tbj_join3(_algo(algo), R.data_ptr(), n[0], S.data_ptr(), n[1], T.data_ptr(),
                          n[2], out.data_ptr(), n[3], ctypes.byref(o), ctypes.byref(res),
                          stream or None))
    return res


def release_stream(stream):
    """Frees the library's session (workspace) of a raw hipStream_t."""
    check(lib().tbj_release_stream(stream or None))


def join64_ptr(algo, r_ptr, nR, s_ptr, nS, pairs_ptr=0, pairs_cap=0, options=None, stream=0):
    """64-bit-key sort-merge join (PSM / RSM) of 16-byte tuples at raw HBM pointers."""
    o = options if options is not None else make_options()
    res = Result()
    check(lib().tbj_join64(_algo(algo), r_ptr, nR, s_ptr, nS, pairs_ptr or None, pairs_cap,
                           ctypes.byref(o), ctypes.byref(res), stream or None))
    return res


def join64(algo, R, S, pairs=None, **opts):
    """Joins device tensors of 16-byte tuples (int64/uint64 [n, 2]: key, payload).

    pairs: optional device tensor (int64 [cap, 2]) for the (R.payload,
    S.payload) matches; materialises when given.  PSM output is in key order.
    """
    for t in (R, S):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("relations must be contiguous device tensors")
    nR = R.numel() * R.element_size() // 16
    nS = S.numel() * S.element_size() // 16
    cap = 0
    if pairs is not None:
        cap = pairs.numel() * pairs.element_size() // 16
        opts.setdefault("materialize", True)
    o = make_options(**opts)
    import torch
    stream = torch.cuda.current_stream(R.device).cuda_stream
    return join64_ptr(algo, R.data_ptr(), nR, S.data_ptr(), nS,
                      pairs.data_ptr() if pairs is not None else 0, cap, o, stream)


# ---------------------------------------------------------------------------
# Relation generators (host, include/tbj_gen.h): numpy arrays of tuples,
# shape (n, 2), columns (key, payload).

PAYLOAD_ZERO, PAYLOAD_ROWID = 0, 1
_gen_ready = False


def _gen_lib():
    global _gen_ready
    L = lib()
    if not _gen_ready:
        vp, u64 = ctypes.c_void_p, ctypes.c_uint64
        L.tbj_rand_stream.argtypes = [ctypes.c_uint32, u64, vp]
        L.tbj_gen_pk.argtypes = [ctypes.c_uint32, u64, vp, ctypes.c_int, ctypes.c_int]
        L.tbj_gen_fk.argtypes = [ctypes.c_uint32, u64, ctypes.c_int64, vp, ctypes.c_int,
                                 ctypes.c_int]
        L.tbj_gen_fk_sel.argtypes = L.tbj_gen_fk.argtypes
        L.tbj_gen_zipf.argtypes = [ctypes.c_uint32, u64, ctypes.c_uint32, ctypes.c_double, vp,
                                   ctypes.c_int, ctypes.c_int]
        L.tbj_remap_odd.argtypes = [vp, u64, ctypes.c_uint32]
        L.tbj_gen_unique64.argtypes = [u64, u64, vp, ctypes.c_int]
        L.tbj_gen_zipf_fk64.argtypes = [u64, u64, vp, u64, ctypes.c_double, vp, ctypes.c_int]
        for f in ("tbj_gen_pk", "tbj_gen_fk", "tbj_gen_fk_sel", "tbj_gen_zipf",
                  "tbj_gen_unique64", "tbj_gen_zipf_fk64"):
            getattr(L, f).restype = ctypes.c_int
        _gen_ready = True
    return L


def _np():
    import numpy
    return numpy


def _alloc(n, out, dtype_bits=32):
    np = _np()
    dt = np.uint32 if dtype_bits == 32 else np.uint64
    if out is None:
        out = np.empty((n, 2), dt)
    assert out.dtype == dt and out.shape == (n, 2) and out.flags.c_contiguous
    return out


def rand_stream(seed, n):
    np = _np()
    out = np.empty(n, np.int32)
    _gen_lib().tbj_rand_stream(seed, n, out.ctypes.data)
    return out


def gen_pk(seed, n, payload=PAYLOAD_ROWID, nthreads=0, out=None):
    out = _alloc(n, out)
    check(_gen_lib().tbj_gen_pk(seed, n, out.ctypes.data, payload, nthreads))
    return out


def gen_fk(seed, n, maxid, payload=PAYLOAD_ROWID, nthreads=0, out=None):
    out = _alloc(n, out)
    check(_gen_lib().tbj_gen_fk(seed, n, maxid, out.ctypes.data, payload, nthreads))
    return out


def gen_fk_sel(seed, n, maxid, payload=PAYLOAD_ROWID, nthreads=0, out=None):
    out = _alloc(n, out)
    check(_gen_lib().tbj_gen_fk_sel(seed, n, maxid, out.ctypes.data, payload, nthreads))
    return out


def gen_zipf(seed, n, alphabet, z, payload=PAYLOAD_ROWID, nthreads=0, out=None):
    out = _alloc(n, out)
    check(_gen_lib().tbj_gen_zipf(seed, n, alphabet, z, out.ctypes.data, payload, nthreads))
    return out


def remap_odd(rel, offset):
    _gen_lib().tbj_remap_odd(rel.ctypes.data, rel.shape[0], offset)
    return rel


def gen_unique64(seed, n, nthreads=0):
    out = _alloc(n, None, 64)
    check(_gen_lib().tbj_gen_unique64(seed, n, out.ctypes.data, nthreads))
    return out


def gen_zipf_fk64(seed, n, R, z, nthreads=0):
    out = _alloc(n, None, 64)
    check(_gen_lib().tbj_gen_zipf_fk64(seed, n, R.ctypes.data, R.shape[0], z, out.ctypes.data,
                                       nthreads))
    return out


# ---------------------------------------------------------------------------
# Batches of joins (include/tbj.h: tbj_batch_*; include/tbj_gen.h:
# tbj_gen_batch_*).  All relations of a batch live in one device arena.

BATCH_PER_ITEM = -1  # tbj_batch_create: each item's own algo


class BatchItem(ctypes.Structure):
    _fields_ = [("r_off", ctypes.c_uint64), ("nR", ctypes.c_uint64),
                ("s_off", ctypes.c_uint64), ("nS", ctypes.c_uint64),
                ("algo", ctypes.c_int32), ("reserved", ctypes.c_uint32)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("target_r", ctypes.c_uint32), ("one_pass_bits", ctypes.c_uint32),
                ("s_chunk", ctypes.c_uint32), ("wave_tuples", ctypes.c_uint32),
                ("tile", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchResult(ctypes.Structure):
    _fields_ = [("matches", ctypes.c_uint64), ("tuples", ctypes.c_uint64),
                ("waves", ctypes.c_uint32), ("groups", ctypes.c_uint32), ("ms", ctypes.c_float)]


class BatchSpec(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("nR", ctypes.c_uint64), ("nS", ctypes.c_uint64),
                ("z", ctypes.c_double), ("r_off", ctypes.c_uint64), ("s_off", ctypes.c_uint64),
                ("algo", ctypes.c_int32), ("reserved", ctypes.c_uint32)]


_batch_ready = False


def _batch_lib():
    global _batch_ready
    L = lib()
    if not _batch_ready:
        vp, u32, ci = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
        L.tbj_batch_create.argtypes = [ci, vp, ctypes.POINTER(BatchItem), u32, vp,
                                       ctypes.POINTER(BatchTuning), ctypes.POINTER(vp)]
        L.tbj_batch_run.argtypes = [vp, vp, ctypes.POINTER(BatchResult)]
        L.tbj_batch_capture.argtypes = [vp, vp]
        L.tbj_batch_destroy.argtypes = [vp]
        L.tbj_batch_destroy.restype = None
        L.tbj_batch_autotune.argtypes = [ci, vp, ctypes.POINTER(BatchItem), u32, vp, ci,
                                         ctypes.POINTER(BatchTuning),
                                         ctypes.POINTER(ctypes.c_float)]
        L.tbj_gen_batch_specs.argtypes = [ctypes.c_uint64, u32, ctypes.POINTER(BatchSpec)]
        L.tbj_gen_batch_specs.restype = ctypes.c_uint64
        L.tbj_gen_batch_device.argtypes = [vp, ctypes.POINTER(BatchSpec), u32, vp]
        for f in ("tbj_batch_create", "tbj_batch_run", "tbj_batch_capture",
                  "tbj_batch_autotune", "tbj_gen_batch_device"):
            getattr(L, f).restype = ctypes.c_int
        _batch_ready = True
    return L


def batch_items(specs_or_tuples):
    """BatchItem array from BatchSpecs or (r_off, nR, s_off, nS[, algo])
    tuples (algo: a name or enum value, default RHO)."""
    arr = (BatchItem * len(specs_or_tuples))()
    for i, x in enumerate(specs_or_tuples):
        if isinstance(x, BatchSpec):
            arr[i] = BatchItem(x.r_off, x.nR, x.s_off, x.nS, x.algo, 0)
        else:
            algo = _algo(x[4]) if len(x) > 4 else ALGOS["RHO"]
            arr[i] = BatchItem(*x[:4], algo, 0)
    return arr


def _batch_algo(algo):
    return BATCH_PER_ITEM if algo in (None, "per_item", BATCH_PER_ITEM) else _algo(algo)


class Batch:
    """A planned batch of count-only RHO/PHT joins over one device arena.

    arena: device tensor holding the tuples (8 bytes each); items: BatchItem
    array; counts: device int64 tensor with one entry per join; algo: "RHO"
    or "PHT" for every item, or "per_item" (each item's algo field).
    """

    def __init__(self, arena, items, counts, tuning=None, algo="per_item"):
        L = _batch_lib()
        self.arena, self.items, self.counts = arena, items, counts
        self._plan = ctypes.c_void_p()
        t = ctypes.byref(tuning) if tuning is not None else None
        check(L.tbj_batch_create(_batch_algo(algo), arena.data_ptr(), items, len(items),
                                 counts.data_ptr(), t, ctypes.byref(self._plan)))

    def _stream(self):
        import torch
        return torch.cuda.current_stream(self.arena.device).cuda_stream

    def capture(self):
        check(_batch_lib().tbj_batch_capture(self._plan, self._stream()))

    def run(self):
        res = BatchResult()
        check(_batch_lib().tbj_batch_run(self._plan, self._stream(), ctypes.byref(res)))
        return res

    def close(self):
        if self._plan:
            _batch_lib().tbj_batch_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_autotune(arena, items, counts, reps=2, wave_tuples=0, algo="per_item"):
    """Times candidate tunings on the batch; returns (BatchTuning, ms)."""
    best = BatchTuning(0, 0, 0, wave_tuples, 0)
    ms = ctypes.c_float()
    check(_batch_lib().tbj_batch_autotune(_batch_algo(algo), arena.data_ptr(), items, len(items),
                                          counts.data_ptr(), reps, ctypes.byref(best),
                                          ctypes.byref(ms)))
    return best, ms.value


def gen_batch_specs(base_seed, n):
    """Sizes, seeds and arena offsets of a generated batch: (specs, total tuples)."""
    specs = (BatchSpec * n)()
    total = _batch_lib().tbj_gen_batch_specs(base_seed, n, specs)
    return specs, total


def gen_batch_device(arena, specs):
    """Writes the batch's relations into the device arena tensor."""
    import torch
    stream = torch.cuda.current_stream(arena.device).cuda_stream
    check(_batch_lib().tbj_gen_batch_device(arena.data_ptr(), specs, len(specs), stream))


def calibrate_runs(src, dst, bins, run):
    """Device ms of a partition-pass write pattern over the int64 device
    tensors src -> dst (same length n): tiles of bins * run tuples, each
    written as one run of `run` tuples into each of `bins` regions of its XCD
    (include/tbj_calib.h)."""
    import torch
    ms = ctypes.c_float()
    check(lib().tbj_calibrate_runs(src.data_ptr(), dst.data_ptr(), src.numel(), bins, run,
                                   torch.cuda.current_stream(src.device).cuda_stream,
                                   ctypes.byref(ms)))
    return ms.value


CAL_READ8, CAL_READ32, CAL_ATOMIC = 0, 1, 2


def calibrate_random(table, nbytes, nprobe, kind, sink):
    """Device ms of nprobe random accesses (kind CAL_*) into the first nbytes
    (a power of two) of the device tensor `table`; sink: a device tensor of
    at least 8 bytes."""
    import torch
    assert table.numel() * table.element_size() >= nbytes
    ms = ctypes.c_float()
    check(lib().tbj_calibrate_random(table.data_ptr(), nbytes, nprobe, kind, sink.data_ptr(),
                                     torch.cuda.current_stream(table.device).cuda_stream,
                                     ctypes.byref(ms)))
    return ms.value
