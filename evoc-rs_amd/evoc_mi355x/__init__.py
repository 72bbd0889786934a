"""Python mirror of the reference's kNN-graph API over the MI355X C ABI.

    knn_graph(data, options, rng) -> (indices, distances)      src/knn.rs:345-359
    KnnGraphOptions                                            src/knn.rs:43-76
    NumpyRandomState / check_random_state                      src/numpy_rng.rs:12-226
    KnnError(EmptyData | InvalidNeighbors | Backend(...))      src/knn.rs:33-41

Everything runs in ``libevoc_mi355x.so`` (gfx950 HIP kernels behind the C ABI
of include/evoc_mi355x.h).  There is no CPU fallback: if the library or the GPU
is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# EVOC_MI355X_LIB names another build of the library (same ABI) to load instead
LIB_PATH = os.environ.get("EVOC_MI355X_LIB") or os.path.join(_HERE, "libevoc_mi355x.so")

DTYPE_F32, DTYPE_I8, DTYPE_U8 = 0, 1, 2
MEM_HOST, MEM_DEVICE = 0, 1
MODE_EXACT, MODE_NNDESCENT = 0, 1
PREC_FP32, PREC_BF16_RESCORE = 0, 1

OK = 0
ERR_EMPTY_DATA = 1
ERR_INVALID_NEIGHBORS = 2
ERR_INVALID_ARGUMENT = 3
ERR_UNSUPPORTED = 4
ERR_OUT_OF_BUDGET = 5
ERR_HIP = 6
ERR_CANCELLED = 7
ERR_NOT_BITWISE = 8
ERR_NO_DEVICE = 9
ERR_PENDING = 10

_DTYPES = {np.dtype(np.float32): DTYPE_F32, np.dtype(np.int8): DTYPE_I8, np.dtype(np.uint8): DTYPE_U8}

INT32_MIN_P1 = -(2**31) + 1  # src/knn.rs:11
INT32_MAX_M1 = 2**31 - 2     # src/knn.rs:12


class KnnError(RuntimeError):
    """src/knn.rs:33-41 plus the backend's own failures (status code in .code)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class EmptyData(KnnError):
    pass


class InvalidNeighbors(KnnError):
    pass


class BackendError(KnnError):
    pass


class _Options(C.Structure):
    _fields_ = [
        ("n_neighbors", C.c_int64), ("n_trees", C.c_int64), ("leaf_size", C.c_int64),
        ("max_candidates", C.c_int64), ("max_rptree_depth", C.c_int64), ("n_iters", C.c_int64),
        ("delta", C.c_float), ("delta_improv", C.c_float), ("use_sorted_updates", C.c_int32),
        ("deterministic", C.c_int32), ("brute_force_max_n", C.c_int64), ("mode", C.c_int32),
        ("precision", C.c_int32), ("margin", C.c_int32), ("strict_precision", C.c_int32),
        ("rayon_threads", C.c_int32), ("reserved0", C.c_int32),
    ]


class _Input(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("location", C.c_int32), ("data", C.c_void_p),
                ("n", C.c_int64), ("d", C.c_int64), ("ld", C.c_int64)]


class _Output(C.Structure):
    _fields_ = [("location", C.c_int32), ("reserved0", C.c_int32), ("indices", C.c_void_p),
                ("distances", C.c_void_p), ("capacity_k", C.c_int64), ("k_out", C.c_int64),
                ("seconds", C.c_double), ("device_seconds", C.c_double), ("score_seconds", C.c_double),
                ("iterations", C.c_int64), ("rescore_rounds", C.c_int64), ("offers_dropped", C.c_int64),
                ("queue_seconds", C.c_double),
                ("score_pairs", C.c_int64), ("slabs", C.c_int64), ("join_bytes", C.c_int64),
                ("join_seconds", C.c_double), ("join_launches", C.c_int64)]


class _ExactPlan(C.Structure):
    _fields_ = [("symmetric", C.c_int32), ("n_splits", C.c_int32), ("caps", C.c_int32), ("ccap", C.c_int32),
                ("stride", C.c_int32), ("n_sample", C.c_int32), ("segments", C.c_int32),
                ("tile_shape", C.c_int32), ("candidates", C.c_int32), ("bound_rank", C.c_int32),
                ("work_items", C.c_int64), ("scratch_bytes", C.c_uint64)]


class _FuzzyInput(C.Structure):
    _fields_ = [("location", C.c_int32), ("symmetrize", C.c_int32), ("knn_indices", C.c_void_p),
                ("knn_dists", C.c_void_p), ("n", C.c_int64), ("k", C.c_int64), ("n_neighbors", C.c_float),
                ("reserved0", C.c_int32)]


class _FuzzyOutput(C.Structure):
    _fields_ = [("location", C.c_int32), ("reserved0", C.c_int32), ("indptr", C.c_void_p),
                ("indices", C.c_void_p), ("data", C.c_void_p), ("capacity", C.c_int64),
                ("sigmas", C.c_void_p), ("rhos", C.c_void_p), ("nnz", C.c_int64),
                ("seconds", C.c_double), ("device_seconds", C.c_double)]


_lib = None

# Every function include/evoc_mi355x.h declares, with its ctypes signature.
_P = C.c_void_p
SIGNATURES = {
    "evoc_mi355x_version": ([], C.c_char_p),
    "evoc_mi355x_device_count": ([C.POINTER(C.c_int32)], C.c_int),
    "evoc_mi355x_create": ([C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
    "evoc_mi355x_destroy": ([_P], None),
    "evoc_mi355x_last_error": ([_P], C.c_char_p),
    "evoc_mi355x_trim": ([_P], None),
    "evoc_knn_options_default": ([C.POINTER(_Options)], None),
    "evoc_knn_effective_k": ([C.c_int64, C.c_int64], C.c_int64),
    "evoc_knn_rng_draws": ([C.c_int32, C.c_int64, C.POINTER(_Options)], C.c_int64),
    "evoc_mi355x_knn_graph": ([_P, C.POINTER(_Input), C.POINTER(_Options), _P, C.POINTER(_Output)],
                              C.c_int),
    "evoc_mi355x_knn_graph_seeded": ([_P, C.POINTER(_Input), C.POINTER(_Options), _P, _P, C.c_int64,
                                      C.POINTER(_Output)], C.c_int),
    "evoc_mi355x_knn_graph_noalloc": ([_P, C.POINTER(_Input), C.POINTER(_Options), _P, _P, C.c_int64,
                                       C.POINTER(_Output)], C.c_int),
    "evoc_mi355x_request_bytes": ([_P, C.POINTER(_Input), C.POINTER(_Options), C.c_int32,
                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
    "evoc_mi355x_exact_plan": ([C.POINTER(_Input), C.POINTER(_Options), C.c_int32, C.c_int32, _P], C.c_int),
    "evoc_mi355x_make_forest": ([_P, C.POINTER(_Input), C.POINTER(_Options), _P, C.c_int64, _P, C.c_int64,
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64), _P, _P], C.c_int),
    "evoc_mi355x_autotune": ([_P, C.POINTER(_Input), C.POINTER(_Options), C.c_int64, C.POINTER(C.c_int32),
                              C.POINTER(C.c_double)], C.c_int),
    "evoc_mi355x_server_create": ([C.c_int32, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
    "evoc_mi355x_server_destroy": ([_P], None),
    "evoc_mi355x_submit": ([_P, C.POINTER(_Input), C.POINTER(_Options), _P, _P, C.c_int64,
                            C.POINTER(_Output), C.POINTER(C.c_uint64)], C.c_int),
    "evoc_mi355x_poll": ([_P, C.c_uint64], C.c_int),
    "evoc_mi355x_wait": ([_P, C.c_uint64, C.c_double, C.c_char_p, C.c_int64], C.c_int),
    "evoc_mi355x_server_request_bytes": ([_P, C.POINTER(_Input), C.POINTER(_Options), C.c_int32,
                                          C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
    "evoc_mi355x_server_set_batching": ([_P, C.c_int64], C.c_int),
    "evoc_mi355x_server_batch_stats": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64)], None),
    "evoc_mi355x_server_stats": ([_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)], None),
    "evoc_mi355x_fuzzy_graph": ([_P, C.POINTER(_FuzzyInput), C.POINTER(_FuzzyOutput)], C.c_int),
    "evoc_mi355x_find_duplicates": ([_P, C.c_int32, _P, _P, C.c_int64, C.c_int64, C.c_int32, _P, C.c_int64,
                                     C.POINTER(C.c_int64)], C.c_int),
    "evoc_npy_read": ([C.c_char_p, _P, C.c_char_p, C.c_int64], C.c_int),
    "evoc_npz_read": ([C.c_char_p, C.c_char_p, _P, C.c_char_p, C.c_int64], C.c_int),
    "evoc_npy_write": ([C.c_char_p, _P, C.c_char_p, C.c_int64], C.c_int),
    "evoc_npy_free": ([_P], None),
    "evoc_numpy_rng_new": ([C.c_uint32], _P),
    "evoc_numpy_rng_free": ([_P], None),
    "evoc_numpy_rng_randint": ([_P, C.c_int64, C.c_int64], C.c_int64),
    "evoc_numpy_rng_normal": ([_P], C.c_double),
    "evoc_numpy_rng_random": ([_P], C.c_double),
    "evoc_numpy_rng_get_state": ([_P, _P, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_double)], None),
    "evoc_numpy_rng_set_state": ([_P, _P, C.c_int32, C.c_int32, C.c_double], None),
}


def lib():
    """Load the in-tree HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it first (python __graft_entry__.py build "
                              "or make -C evoc-rs_amd/csrc)")
        # One HIP runtime per process: when PyTorch is installed, load its bundled
        # libamdhip64.so.7 first so this library binds to the same runtime (same
        # SONAME) instead of a second copy that cannot see the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def version() -> str:
    return lib().evoc_mi355x_version().decode()


def device_count() -> int:
    c = C.c_int32(0)
    lib().evoc_mi355x_device_count(C.byref(c))
    return int(c.value)


class NumpyRandomState:
    """``numpy.random.RandomState`` as restated by src/numpy_rng.rs (MT19937)."""

    def __init__(self, seed: int = 0):
        self._h = lib().evoc_numpy_rng_new(C.c_uint32(seed & 0xFFFFFFFF))
        if not self._h:
            raise MemoryError("evoc_numpy_rng_new")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.evoc_numpy_rng_free(self._h)
            self._h = None

    def randint(self, low: int, high: int) -> int:
        return int(lib().evoc_numpy_rng_randint(self._h, low, high))

    def normal(self) -> float:
        return float(lib().evoc_numpy_rng_normal(self._h))

    def random(self) -> float:
        return float(lib().evoc_numpy_rng_random(self._h))

    def randint3_for_tau(self) -> np.ndarray:
        return np.array([self.randint(INT32_MIN_P1, INT32_MAX_M1) for _ in range(3)], np.int64)

    def get_state(self):
        key = np.zeros(624, np.uint32)
        pos, hg, g = C.c_int32(), C.c_int32(), C.c_double()
        lib().evoc_numpy_rng_get_state(self._h, key.ctypes.data, C.byref(pos), C.byref(hg), C.byref(g))
        return key, int(pos.value), int(hg.value), float(g.value)

    def set_state(self, key: np.ndarray, pos: int, has_gauss: int = 0, gauss: float = 0.0):
        k = np.ascontiguousarray(key, np.uint32)
        lib().evoc_numpy_rng_set_state(self._h, k.ctypes.data, pos, has_gauss, gauss)


def check_random_state(seed: Optional[int]) -> NumpyRandomState:
    """src/numpy_rng.rs:216-226."""
    if seed is None:
        import time

        seed = time.time_ns() & 0xFFFFFFFF
    return NumpyRandomState(seed)


@dataclass
class KnnGraphOptions:
    """src/knn.rs:43-76 (None = derived default, src/knn.rs:402-414) + mi355x extensions."""

    n_neighbors: int = 30
    n_trees: Optional[int] = None
    leaf_size: Optional[int] = None
    max_candidates: Optional[int] = None
    max_rptree_depth: int = 200
    n_iters: Optional[int] = None
    delta: float = 0.001
    delta_improv: float = 0.001
    use_sorted_updates: bool = True
    brute_force_max_n: int = 0
    deterministic: bool = False
    # mi355x extensions
    mode: int = MODE_EXACT
    precision: int = PREC_BF16_RESCORE
    margin: int = 16
    strict_precision: bool = False
    # rayon::current_num_threads() of the caller, for the n_trees default (src/knn.rs:404-407);
    # 0: the count rayon's global pool has in this process (RAYON_NUM_THREADS, else the CPUs
    # this process may run on), as for a caller that runs in that pool
    rayon_threads: int = 0

    def _c(self) -> _Options:
        o = _Options()
        lib().evoc_knn_options_default(C.byref(o))
        o.n_neighbors = self.n_neighbors
        o.n_trees = self.n_trees or 0
        o.leaf_size = self.leaf_size or 0
        o.max_candidates = self.max_candidates or 0
        o.max_rptree_depth = self.max_rptree_depth
        o.n_iters = self.n_iters or 0
        o.delta = self.delta
        o.delta_improv = self.delta_improv
        o.use_sorted_updates = int(self.use_sorted_updates)
        o.brute_force_max_n = self.brute_force_max_n
        o.deterministic = int(self.deterministic)
        o.mode = self.mode
        o.precision = self.precision
        o.margin = self.margin
        o.strict_precision = int(self.strict_precision)
        o.rayon_threads = self.rayon_threads
        return o


def effective_k(n_neighbors: int, n: int) -> int:
    return int(lib().evoc_knn_effective_k(n_neighbors, n))


def rng_draws(dtype_code: int, n: int, options: KnnGraphOptions) -> int:
    o = options._c()
    return int(lib().evoc_knn_rng_draws(dtype_code, n, C.byref(o)))


def exact_plan(dtype_code: int, n: int, d: int, options: KnnGraphOptions, n_cu: int = 256,
               serving: bool = False) -> dict:
    """The plan an exact search of an n x d matrix follows on a device with n_cu compute units
    (evoc_mi355x_exact_plan; host only, no device touched)."""
    i = _Input(dtype_code, MEM_DEVICE, None, n, d, d)
    o = options._c()
    p = _ExactPlan()
    rc = lib().evoc_mi355x_exact_plan(C.byref(i), C.byref(o), n_cu, int(serving), C.byref(p))
    if rc != OK:
        raise BackendError(rc, "evoc_mi355x_exact_plan")
    return {name: getattr(p, name) for name, _ in _ExactPlan._fields_}


def _raise(code: int, ctx) -> None:
    msg = lib().evoc_mi355x_last_error(ctx).decode() if ctx else ""
    if code == ERR_EMPTY_DATA:
        raise EmptyData(code, msg or "empty embedding matrix")
    if code == ERR_INVALID_NEIGHBORS:
        raise InvalidNeighbors(code, msg or "n_neighbors must be positive")
    raise BackendError(code, msg or "mi355x backend error")


class Context:
    """One GPU, one stream, one HBM budget (evoc_mi355x_ctx)."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        h = C.c_void_p()
        rc = lib().evoc_mi355x_create(device, hbm_budget_bytes, C.byref(h))
        if rc != OK:
            raise BackendError(rc, f"evoc_mi355x_create(device={device}) failed with status {rc} "
                                   "(no gfx950 device visible?)")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().evoc_mi355x_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def last_error(self) -> str:
        return lib().evoc_mi355x_last_error(self._h).decode()

    def trim(self) -> None:
        """Return the device scratch kept from the last call (evoc_mi355x_trim)."""
        lib().evoc_mi355x_trim(self._h)


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class CallStats:
    seconds: float = 0.0
    device_seconds: float = 0.0
    score_seconds: float = 0.0
    iterations: int = 0
    rescore_rounds: int = 0
    offers_dropped: int = 0
    queue_seconds: float = 0.0
    score_pairs: int = 0
    slabs: int = 0
    join_bytes: int = 0
    join_seconds: float = 0.0
    join_launches: int = 0
    extra: dict = field(default_factory=dict)


def _sync_torch(*tensors) -> None:
    """The library runs on its own streams: finish the caller's pending torch
    work on any device tensor it is handed."""
    for t in tensors:
        if t is not None and type(t).__module__.startswith("torch") and t.is_cuda:
            import torch

            torch.cuda.current_stream(t.device).synchronize()
            return


def _fill_stats(stats, res) -> None:
    if stats is None:
        return
    stats.seconds = res.seconds
    stats.device_seconds = res.device_seconds
    stats.score_seconds = res.score_seconds
    stats.iterations = res.iterations
    stats.rescore_rounds = res.rescore_rounds
    stats.offers_dropped = res.offers_dropped
    stats.queue_seconds = res.queue_seconds
    stats.score_pairs = res.score_pairs
    stats.slabs = res.slabs
    stats.join_bytes = res.join_bytes
    stats.join_seconds = res.join_seconds
    stats.join_launches = res.join_launches


def _input_from(data, dtype_code=None):
    """numpy array (host) or torch tensor (device) -> (_Input, keepalive)."""
    try:
        import torch  # noqa: F401

        is_torch = type(data).__module__.startswith("torch")
    except ImportError:
        is_torch = False
    if is_torch:
        t = data
        tmap = {"torch.float32": DTYPE_F32, "torch.int8": DTYPE_I8, "torch.uint8": DTYPE_U8}
        code = tmap.get(str(t.dtype))
        if code is None or t.dim() != 2 or t.stride(1) != 1:
            raise BackendError(ERR_INVALID_ARGUMENT, "expected a 2-D row-major float32/int8/uint8 tensor")
        loc = MEM_DEVICE if t.is_cuda else MEM_HOST
        inp = _Input(code, loc, C.c_void_p(t.data_ptr()), t.shape[0], t.shape[1], t.stride(0))
        return inp, t
    a = np.asarray(data)
    if a.ndim != 2:
        raise BackendError(ERR_INVALID_ARGUMENT, "expected a 2-D array")
    if a.dtype not in _DTYPES:
        raise BackendError(ERR_INVALID_ARGUMENT, f"unsupported dtype {a.dtype} (float32, int8, uint8)")
    a = np.ascontiguousarray(a)
    inp = _Input(_DTYPES[a.dtype], MEM_HOST, C.c_void_p(a.ctypes.data if a.size else 1), a.shape[0],
                 a.shape[1], a.shape[1])
    return inp, a


def knn_graph(data, options: KnnGraphOptions, rng: NumpyRandomState,
              ctx: Optional[Context] = None, stats: Optional[CallStats] = None,
              out=None):
    """src/knn.rs:345-359 on the MI355X.  Advances `rng` exactly like the reference.

    `data`: (n, d) float32 / int8 / uint8 numpy array, or a torch tensor on the
    GPU (the graph is then built from HBM-resident rows).  Returns numpy
    (indices int32 [n, k], distances float32 [n, k]) unless `out` = (idx, dist)
    device tensors are given, which are filled in place.
    """
    ctx = ctx or default_context()
    inp, keep = _input_from(data)
    o = options._c()
    n = inp.n
    if options.n_neighbors <= 0:
        _raise(ERR_INVALID_NEIGHBORS, None)
    if n <= 0:
        _raise(ERR_EMPTY_DATA, None)
    k = effective_k(options.n_neighbors, n)
    res = _Output()
    if out is not None:
        idx_t, dist_t = out
        res.location = MEM_DEVICE if idx_t.is_cuda else MEM_HOST
        res.indices = C.c_void_p(idx_t.data_ptr())
        res.distances = C.c_void_p(dist_t.data_ptr())
        idx = dist = None
    else:
        idx = np.empty((n, k), np.int32)
        dist = np.empty((n, k), np.float32)
        res.location = MEM_HOST
        res.indices = C.c_void_p(idx.ctypes.data)
        res.distances = C.c_void_p(dist.ctypes.data)
    res.capacity_k = k
    _sync_torch(keep, *(out or ()))
    rc = lib().evoc_mi355x_knn_graph(ctx.handle, C.byref(inp), C.byref(o), rng._h, C.byref(res))
    del keep
    if rc != OK:
        _raise(rc, ctx.handle)
    _fill_stats(stats, res)
    if out is not None:
        return out
    return idx, dist


NOALLOC_SCRATCH_BYTES = 256 << 20  # EVOC_MI355X_NOALLOC_SCRATCH_BYTES


def knn_graph_seeded(data, options: KnnGraphOptions, tau_state, tree_states=None,
                     ctx: Optional[Context] = None, stats: Optional[CallStats] = None, out=None,
                     noalloc: bool = False):
    """evoc_mi355x_knn_graph_seeded: the caller made the RNG draws itself.  With
    noalloc, evoc_mi355x_knn_graph_noalloc: the same result, every device buffer of
    the call carved from the library's fixed per-GPU scratch (no device allocation)."""
    ctx = ctx or default_context()
    inp, keep = _input_from(data)
    o = options._c()
    n = inp.n
    k = effective_k(options.n_neighbors, n) if n > 0 else 1
    tau = np.ascontiguousarray(tau_state, np.int64)
    trees = None if tree_states is None else np.ascontiguousarray(tree_states, np.int64)
    res = _Output()
    if out is not None:
        idx_t, dist_t = out
        res.location = MEM_DEVICE if idx_t.is_cuda else MEM_HOST
        res.indices = C.c_void_p(idx_t.data_ptr())
        res.distances = C.c_void_p(dist_t.data_ptr())
        idx = dist = None
    else:
        idx = np.empty((max(n, 0), k), np.int32)
        dist = np.empty((max(n, 0), k), np.float32)
        res.location = MEM_HOST
        res.indices = C.c_void_p(idx.ctypes.data)
        res.distances = C.c_void_p(dist.ctypes.data)
    res.capacity_k = k
    _sync_torch(keep, *(out or ()))
    fn = lib().evoc_mi355x_knn_graph_noalloc if noalloc else lib().evoc_mi355x_knn_graph_seeded
    rc = fn(
        ctx.handle, C.byref(inp), C.byref(o), C.c_void_p(tau.ctypes.data),
        C.c_void_p(trees.ctypes.data) if trees is not None else None,
        0 if trees is None else trees.shape[0], C.byref(res))
    del keep
    if rc != OK:
        _raise(rc, ctx.handle)
    _fill_stats(stats, res)
    if out is not None:
        return out
    return idx, dist


def make_forest(data, options: KnnGraphOptions, tree_states, ctx: Optional[Context] = None):
    """The device random-projection forest (evoc_mi355x_make_forest; make_forest,
    src/knn.rs:115-154): (leaf_array [leaves, cols] int32 padded with -1,
    leaves per tree [n_trees] int64, depth per tree [n_trees] int32)."""
    ctx = ctx or default_context()
    inp, keep = _input_from(data)
    o = options._c()
    ts = np.ascontiguousarray(tree_states, np.int64).reshape(-1, 3)
    nt = ts.shape[0]
    rows, cols = C.c_int64(), C.c_int64()
    per_tree = np.zeros(max(nt, 1), np.int64)
    depths = np.zeros(max(nt, 1), np.int32)
    _sync_torch(keep)
    rc = lib().evoc_mi355x_make_forest(ctx.handle, C.byref(inp), C.byref(o), C.c_void_p(ts.ctypes.data), nt, None, 0,
                                       C.byref(rows), C.byref(cols), None, None)
    if rc != OK:
        _raise(rc, ctx.handle)
    leaf = np.empty((rows.value, cols.value), np.int32)
    rc = lib().evoc_mi355x_make_forest(ctx.handle, C.byref(inp), C.byref(o), C.c_void_p(ts.ctypes.data), nt,
                                       C.c_void_p(leaf.ctypes.data), rows.value, C.byref(rows), C.byref(cols),
                                       C.c_void_p(per_tree.ctypes.data), C.c_void_p(depths.ctypes.data))
    del keep
    if rc != OK:
        _raise(rc, ctx.handle)
    return leaf, per_tree[:nt], depths[:nt]


def request_bytes(data, options: KnnGraphOptions, out_on_device: bool = False,
                  ctx: Optional[Context] = None):
    """(peak, minimum) device bytes of a knn_graph call (evoc_mi355x_request_bytes)."""
    ctx = ctx or default_context()
    inp, keep = _input_from(data)
    o = options._c()
    full, low = C.c_uint64(), C.c_uint64()
    rc = lib().evoc_mi355x_request_bytes(ctx.handle, C.byref(inp), C.byref(o),
                                         MEM_DEVICE if out_on_device else MEM_HOST,
                                         C.byref(full), C.byref(low))
    del keep
    if rc != OK:
        raise BackendError(rc, "invalid request")
    return int(full.value), int(low.value)


def autotune(data, options: KnnGraphOptions, max_rows: int = 65536, ctx: Optional[Context] = None):
    """Pick the fastest MFMA tile shape for searches like knn_graph(data, options)
    (evoc_mi355x_autotune).  Returns {"shape": s, "ms": [t0, t1, t2, t3]}."""
    ctx = ctx or default_context()
    inp, keep = _input_from(data)
    _sync_torch(keep)
    o = options._c()
    best = C.c_int32(-1)
    ms = (C.c_double * 4)()
    rc = lib().evoc_mi355x_autotune(ctx.handle, C.byref(inp), C.byref(o), max_rows, C.byref(best), ms)
    del keep
    if rc != OK:
        _raise(rc, ctx.handle)
    return {"shape": best.value, "ms": [float(v) for v in ms]}


class Ticket:
    """A submitted request: keeps its buffers alive until it is waited for."""

    def __init__(self, server, ticket, res, keep, result):
        self.server, self.id, self._res, self._keep, self._result = server, ticket, res, keep, result
        self.status = None
        self.stats = CallStats()
        self.submitted = None

    def poll(self) -> bool:
        """True once the request has finished (successfully or not)."""
        if self.status is not None:
            return True
        return lib().evoc_mi355x_poll(self.server._h, self.id) != ERR_PENDING

    def wait(self, timeout: float = -1.0):
        """Block until done; returns the result (numpy pair or the caller's out tensors)."""
        if self.status is None:
            buf = C.create_string_buffer(1024)
            rc = lib().evoc_mi355x_wait(self.server._h, self.id, timeout, buf, 1024)
            if rc == ERR_PENDING:
                raise TimeoutError(f"request {self.id} still running after {timeout} s")
            self.status = rc
            self._msg = buf.value.decode()
            _fill_stats(self.stats, self._res)
            self._keep = None
        if self.status == ERR_EMPTY_DATA:
            raise EmptyData(self.status, self._msg or "empty embedding matrix")
        if self.status == ERR_INVALID_NEIGHBORS:
            raise InvalidNeighbors(self.status, self._msg or "n_neighbors must be positive")
        if self.status != OK:
            raise BackendError(self.status, self._msg or "mi355x serving error")
        return self._result


class Server:
    """Continuous-batching request loop on one GPU (evoc_mi355x_server_*).

    `slots` requests run concurrently on their own HIP streams; further
    requests queue and are admitted as slots and HBM budget free up.
    """

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget_bytes: int = 0):
        h = C.c_void_p()
        rc = lib().evoc_mi355x_server_create(device, slots, hbm_budget_bytes, C.byref(h))
        if rc != OK:
            raise BackendError(rc, f"evoc_mi355x_server_create(device={device}) failed with status {rc}")
        self._h = h

    def submit(self, data, options: KnnGraphOptions, rng: Optional[NumpyRandomState] = None,
               tau_state=None, tree_states=None, out=None) -> Ticket:
        """Queue knn_graph(data, options, rng).  With `rng`, its draws are made
        here, in submission order, exactly as knn_graph would make them."""
        inp, keep = _input_from(data)
        n = inp.n
        if options.n_neighbors <= 0:
            raise InvalidNeighbors(ERR_INVALID_NEIGHBORS, "n_neighbors must be positive")
        if n <= 0:
            raise EmptyData(ERR_EMPTY_DATA, "empty embedding matrix")
        if tau_state is None:
            if rng is None:
                raise ValueError("submit needs rng or tau_state")
            draws = rng_draws(inp.dtype, n, options)
            vals = [rng.randint(-2147483647, 2147483646) for _ in range(draws)]
            tau_state, tree_states = vals[:3], (np.array(vals[3:], np.int64).reshape(-1, 3)
                                                if draws > 3 else None)
        tau = np.ascontiguousarray(tau_state, np.int64)
        trees = None if tree_states is None else np.ascontiguousarray(tree_states, np.int64)
        k = effective_k(options.n_neighbors, n)
        res = _Output()
        if out is not None:
            idx_t, dist_t = out
            res.location = MEM_DEVICE if idx_t.is_cuda else MEM_HOST
            res.indices = C.c_void_p(idx_t.data_ptr())
            res.distances = C.c_void_p(dist_t.data_ptr())
            result = out
        else:
            idx = np.empty((n, k), np.int32)
            dist = np.empty((n, k), np.float32)
            res.location = MEM_HOST
            res.indices = C.c_void_p(idx.ctypes.data)
            res.distances = C.c_void_p(dist.ctypes.data)
            result = (idx, dist)
        res.capacity_k = k
        o = options._c()
        _sync_torch(keep, *(out or ()))
        t = C.c_uint64()
        rc = lib().evoc_mi355x_submit(self._h, C.byref(inp), C.byref(o), C.c_void_p(tau.ctypes.data),
                                      C.c_void_p(trees.ctypes.data) if trees is not None else None,
                                      0 if trees is None else trees.shape[0], C.byref(res), C.byref(t))
        if rc != OK:
            raise BackendError(rc, f"submit failed with status {rc}")
        return Ticket(self, t.value, res, (keep, res, tau, trees, o), result)

    def request_bytes(self, data, options: KnnGraphOptions, out_on_device: bool = False):
        """(full, minimum) device bytes this server's admission reserves for such a request
        (evoc_mi355x_server_request_bytes)."""
        inp, keep = _input_from(data)
        o = options._c()
        full, low = C.c_uint64(), C.c_uint64()
        rc = lib().evoc_mi355x_server_request_bytes(self._h, C.byref(inp), C.byref(o),
                                                    MEM_DEVICE if out_on_device else MEM_HOST,
                                                    C.byref(full), C.byref(low))
        del keep
        if rc != OK:
            raise BackendError(rc, "invalid request")
        return int(full.value), int(low.value)

    def set_batching(self, max_rows: int) -> None:
        """Batch requests of at most max_rows rows (0: off); see evoc_mi355x_server_set_batching."""
        rc = lib().evoc_mi355x_server_set_batching(self._h, max_rows)
        if rc != OK:
            raise BackendError(rc, "evoc_mi355x_server_set_batching")

    def batch_stats(self) -> dict:
        b, r, m, g = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        lib().evoc_mi355x_server_batch_stats(self._h, C.byref(b), C.byref(r), C.byref(m), C.byref(g))
        return {"batches": b.value, "requests": r.value, "largest": m.value, "graph_replays": g.value}

    def stats(self):
        q, r, g, b = C.c_int64(), C.c_int64(), C.c_uint64(), C.c_uint64()
        lib().evoc_mi355x_server_stats(self._h, C.byref(q), C.byref(r), C.byref(g), C.byref(b))
        return {"queued": q.value, "running": r.value, "granted_bytes": g.value, "budget_bytes": b.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().evoc_mi355x_server_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()


# ---------------------------------------------------------------------------
# The steps after knn_graph that consume its output
# ---------------------------------------------------------------------------
def _knn_arrays(knn_indices, knn_dists):
    """(indices, distances) as numpy arrays or GPU tensors -> (location, ptrs, n, k, keepalive)."""
    is_torch = type(knn_indices).__module__.startswith("torch")
    if is_torch:
        if str(knn_indices.dtype) != "torch.int32" or str(knn_dists.dtype) != "torch.float32":
            raise BackendError(ERR_INVALID_ARGUMENT, "expected int32 indices and float32 distances")
        ind, dst = knn_indices.contiguous(), knn_dists.contiguous()
        loc = MEM_DEVICE if ind.is_cuda else MEM_HOST
        return loc, ind.data_ptr(), dst.data_ptr(), ind.shape[0], ind.shape[1], (ind, dst)
    ind = np.ascontiguousarray(knn_indices, dtype=np.int32)
    dst = np.ascontiguousarray(knn_dists, dtype=np.float32)
    if ind.ndim != 2 or ind.shape != dst.shape:
        raise BackendError(ERR_INVALID_ARGUMENT, "knn_indices and knn_dists must be n x k")
    return MEM_HOST, ind.ctypes.data, dst.ctypes.data, ind.shape[0], ind.shape[1], (ind, dst)


def neighbor_graph_matrix(n_neighbors: float, knn_indices, knn_dists, symmetrize: bool = True,
                          ctx: Optional[Context] = None, stats: Optional[CallStats] = None,
                          return_sigmas: bool = False, out=None):
    """src/graph_construction.rs:162-173 (neighbor_graph_matrix) on the MI355X: the
    fuzzy graph of a kNN graph as CSR (indptr int64 [n + 1], indices int32,
    data float32), each row's columns ascending.  With return_sigmas also the
    per-row (sigma, rho) of smooth_knn_dist.  Inputs: numpy arrays or torch
    tensors (host or GPU).  Outputs are numpy arrays, or with `out` = (indptr,
    indices, data) GPU tensors (indices / data of capacity 2 n k) those are
    filled in place and (indptr, indices[:nnz], data[:nnz]) views returned."""
    ctx = ctx or default_context()
    loc, pi, pd, n, k, keep = _knn_arrays(knn_indices, knn_dists)
    if n <= 0:
        _raise(ERR_EMPTY_DATA, None)
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    if out is not None:
        indptr, indices, data = out
        cap = min(indices.numel(), data.numel())
        ptrs = (indptr.data_ptr(), indices.data_ptr(), data.data_ptr())
        oloc = MEM_DEVICE if indptr.is_cuda else MEM_HOST
        sig_p = rho_p = None  # (host sigma / rho would need a host copy per call)
    else:
        cap = max(1, 2 * n * k)
        indptr = np.empty(n + 1, np.int64)
        indices = np.empty(cap, np.int32)
        data = np.empty(cap, np.float32)
        ptrs = (indptr.ctypes.data, indices.ctypes.data, data.ctypes.data)
        oloc = MEM_HOST
        sig_p, rho_p = sig.ctypes.data, rho.ctypes.data
    fin = _FuzzyInput(loc, 1 if symmetrize else 0, C.c_void_p(pi), C.c_void_p(pd), n, k, float(n_neighbors), 0)
    fo = _FuzzyOutput(oloc, 0, C.c_void_p(ptrs[0]), C.c_void_p(ptrs[1]), C.c_void_p(ptrs[2]), cap,
                      C.c_void_p(sig_p), C.c_void_p(rho_p), 0, 0.0, 0.0)
    _sync_torch(*keep, *(out or ()))
    rc = lib().evoc_mi355x_fuzzy_graph(ctx.handle, C.byref(fin), C.byref(fo))
    del keep
    if rc != OK:
        _raise(rc, ctx.handle)
    if stats is not None:
        stats.seconds = fo.seconds
        stats.device_seconds = fo.device_seconds
    nnz = fo.nnz
    if out is not None:
        res = (indptr, indices[:nnz], data[:nnz])
    else:
        res = (indptr, indices[:nnz].copy(), data[:nnz].copy())
    return res + ((sig, rho),) if return_sigmas else res


def smooth_knn_dist(distances, k: float, ctx: Optional[Context] = None):
    """src/graph_construction.rs:15-81 on the MI355X: per-row (sigma, rho)."""
    d = np.ascontiguousarray(distances, dtype=np.float32)
    ind = np.full(d.shape, -1, np.int32)  # no edges: only the per-row bisection runs
    _, _, _, (sig, rho) = neighbor_graph_matrix(k, ind, d, True, ctx=ctx, return_sigmas=True)
    return sig, rho


def find_duplicates(knn_inds, knn_dists, ctx: Optional[Context] = None):
    """src/cluster_util.rs:430-449 on the MI355X: the set of pairs (i, j), i < j,
    listed at a distance <= the largest first-column distance."""
    ctx = ctx or default_context()
    loc, pi, pd, n, k, keep = _knn_arrays(knn_inds, knn_dists)
    if n <= 0:
        _raise(ERR_EMPTY_DATA, None)
    cap = max(1, n * k)
    pairs = np.empty(cap, np.uint64)
    cnt = C.c_int64(0)
    _sync_torch(*keep)
    rc = lib().evoc_mi355x_find_duplicates(ctx.handle, loc, C.c_void_p(pi), C.c_void_p(pd), n, k, MEM_HOST,
                                           C.c_void_p(pairs.ctypes.data), cap, C.byref(cnt))
    del keep
    if rc != OK:
        _raise(rc, ctx.handle)
    p = pairs[: cnt.value]
    return {(int(a), int(b)) for a, b in zip(p >> np.uint64(32), p & np.uint64(0xFFFFFFFF))}


# ---------------------------------------------------------------------------
# .npy / .npz arrays through the library's own reader and writer
# (include/evoc_mi355x.h: evoc_npy_read / evoc_npz_read / evoc_npy_write)
# ---------------------------------------------------------------------------
class _NpyArray(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("ndim", C.c_int32), ("shape", C.c_int64 * 8),
                ("data", C.c_void_p), ("nbytes", C.c_uint64)]


_NPY_TYPES = [np.bool_, np.int8, np.uint8, np.int32, np.uint32, np.int64, np.uint64, np.float32,
              np.float64]


def _npy_take(a: "_NpyArray") -> np.ndarray:
    try:
        shape = tuple(a.shape[i] for i in range(a.ndim))
        out = np.empty(shape, _NPY_TYPES[a.dtype])
        if a.nbytes:
            C.memmove(out.ctypes.data, a.data, a.nbytes)
        return out
    finally:
        lib().evoc_npy_free(C.byref(a))


def npy_read(path: str) -> np.ndarray:
    """ndarray-npy's ReadNpyExt::read_npy (tests/parity.rs:22-30) in the library."""
    a, err = _NpyArray(), C.create_string_buffer(512)
    rc = lib().evoc_npy_read(os.fsencode(path), C.byref(a), err, len(err))
    if rc != OK:
        raise BackendError(rc, err.value.decode())
    return _npy_take(a)


def npz_read(path: str, member: str) -> np.ndarray:
    """ndarray-npy's NpzReader::by_name (src/graph_construction.rs:117-140)."""
    a, err = _NpyArray(), C.create_string_buffer(512)
    rc = lib().evoc_npz_read(os.fsencode(path), member.encode(), C.byref(a), err, len(err))
    if rc != OK:
        raise BackendError(rc, err.value.decode())
    return _npy_take(a)


def npy_write(path: str, arr) -> None:
    """Writes a C-order little-endian array as .npy (format 1.0)."""
    x = np.asarray(arr, order="C")  # (ascontiguousarray would make a 0-d array 1-d)
    codes = {np.dtype(t): i for i, t in enumerate(_NPY_TYPES)}
    if x.dtype.newbyteorder("=") not in codes or x.dtype.byteorder == ">":
        raise BackendError(ERR_UNSUPPORTED, f"dtype {x.dtype} is not supported")
    if x.ndim > 8:
        raise BackendError(ERR_UNSUPPORTED, "more than 8 dimensions")
    a = _NpyArray()
    a.dtype = codes[x.dtype.newbyteorder("=")]
    a.ndim = x.ndim
    for i, d in enumerate(x.shape):
        a.shape[i] = d
    a.data = x.ctypes.data if x.nbytes else None
    a.nbytes = x.nbytes
    err = C.create_string_buffer(512)
    rc = lib().evoc_npy_write(os.fsencode(path), C.byref(a), err, len(err))
    if rc != OK:
        raise BackendError(rc, err.value.decode())
