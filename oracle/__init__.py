"""ORACLE — test infrastructure only.

ctypes bindings for ``liboracle.so``, the plain-C restatement of the reference's
CPU kNN-graph path (``evoc_oracle.c``; every C function cites the reference
file:line it follows).  Only the tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker / the timed CPU baseline.  The HIP library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_fast_cosine.so")

F32, I8, U8 = 0, 1, 2
_DTYPES = {np.dtype(np.float32): F32, np.dtype(np.int8): I8, np.dtype(np.uint8): U8}

INT32_MIN_P1 = -(2**31) + 1  # src/knn.rs:11
INT32_MAX_M1 = 2**31 - 2     # src/knn.rs:12

_lib = None


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree is present)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class _NpRng(C.Structure):
    _fields_ = [("key", C.c_uint32 * 624), ("pos", C.c_int32), ("has_gauss", C.c_int32),
                ("gauss", C.c_double)]


class _KnnOpts(C.Structure):
    _fields_ = [("n_neighbors", C.c_int64), ("n_trees", C.c_int64), ("leaf_size", C.c_int64),
                ("max_candidates", C.c_int64), ("max_rptree_depth", C.c_int64),
                ("n_iters", C.c_int64), ("delta", C.c_float), ("delta_improv", C.c_float),
                ("use_sorted_updates", C.c_int32), ("deterministic", C.c_int32),
                ("brute_force_max_n", C.c_int64)]


_P = C.c_void_p
_I64 = C.c_int64


def _declare(L):
    L.orc_nprng_seed.argtypes = [C.POINTER(_NpRng), C.c_uint32]
    L.orc_nprng_set_state.argtypes = [C.POINTER(_NpRng), _P, C.c_int32, C.c_int32, C.c_double]
    L.orc_nprng_randint.argtypes = [C.POINTER(_NpRng), _I64, _I64]
    L.orc_nprng_randint.restype = _I64
    L.orc_nprng_normal.argtypes = [C.POINTER(_NpRng)]
    L.orc_nprng_normal.restype = C.c_double
    L.orc_nprng_random.argtypes = [C.POINTER(_NpRng)]
    L.orc_nprng_random.restype = C.c_double
    L.orc_nprng_next_u32.argtypes = [C.POINTER(_NpRng)]
    L.orc_nprng_next_u32.restype = C.c_uint32
    L.orc_tau_rand_int.argtypes = [_P]
    L.orc_tau_rand_int.restype = C.c_int32
    L.orc_tau_rand.argtypes = [_P]
    L.orc_tau_rand.restype = C.c_float
    L.orc_fast_cosine_c.argtypes = [_P, _P, C.c_int]
    L.orc_fast_cosine_c.restype = C.c_float
    L.orc_dot_serial.argtypes = [_P, _P, C.c_int]
    L.orc_dot_serial.restype = C.c_float
    L.orc_distance.argtypes = [C.c_int, _P, _I64, C.c_int, _I64, _I64]
    L.orc_distance.restype = C.c_float
    L.orc_set_strict_cosine.argtypes = [_P]
    L.orc_make_forest.argtypes = [C.c_int, _P, _I64, _I64, _I64, _I64, _I64, C.POINTER(_NpRng),
                                  _I64, C.POINTER(_I64), C.POINTER(_I64), C.c_int]
    L.orc_make_forest.restype = C.POINTER(C.c_int32)
    L.orc_free.argtypes = [_P]
    L.orc_rp_split.argtypes = [C.c_int, _P, _I64, _I64, _P, _I64, _P, _P, _P]
    L.orc_rp_split.restype = _I64
    L.orc_nn_descent.argtypes = [C.c_int, _P, _I64, _I64, C.c_int, _I64, _P, _I64, _I64,
                                 C.c_float, C.c_float, C.c_int, _P, _I64, _I64, C.c_int, _I64,
                                 C.c_int, _P, _P, C.POINTER(_I64)]
    L.orc_nn_descent.restype = C.c_int
    L.orc_brute_force_knn.argtypes = [C.c_int, _P, _I64, _I64, _I64, _P, _P, C.c_int]
    L.orc_exact_knn.argtypes = [C.c_int, _P, _I64, _I64, _I64, C.c_int, C.c_int, _P, _P, C.c_int]
    L.orc_transform.argtypes = [C.c_int, _P, _I64]
    L.orc_prepare_float.argtypes = [_P, _I64, _I64]
    L.orc_deheap_sort.argtypes = [_P, _P, _I64, _I64, C.c_int]
    L.orc_knn_graph.argtypes = [C.c_int, _P, _I64, _I64, C.POINTER(_KnnOpts), C.POINTER(_NpRng),
                                _I64, C.c_int, _P, _P, C.POINTER(_I64), C.POINTER(_I64)]
    L.orc_knn_graph.restype = C.c_int


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def use_reference_fast_cosine(enable: bool = True) -> bool:
    """Route the deterministic f32 distance through oracle/_ref (the reference's own
    native/fast_cosine.c).  Returns False when that build is absent."""
    L = lib()
    if not enable:
        L.orc_set_strict_cosine(None)
        return True
    if not os.path.exists(_REF_PATH):
        return False
    ref = C.CDLL(_REF_PATH)
    fn = ref.fast_cosine_numba
    fn.restype = C.c_float
    addr = C.cast(fn, C.c_void_p).value
    L.orc_set_strict_cosine(C.c_void_p(addr))
    use_reference_fast_cosine._keep = ref  # keep the handle alive
    return True


class NumpyRandomState:
    """MT19937 ``RandomState`` restated from src/numpy_rng.rs."""

    def __init__(self, seed: int = 0):
        self._s = _NpRng()
        lib().orc_nprng_seed(C.byref(self._s), C.c_uint32(seed & 0xFFFFFFFF))

    @classmethod
    def from_state(cls, key: np.ndarray, pos: int, has_gauss: int = 0, gauss: float = 0.0):
        r = cls(0)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        lib().orc_nprng_set_state(C.byref(r._s), _ptr(k), pos, has_gauss, gauss)
        return r

    def randint(self, low: int, high: int) -> int:
        return int(lib().orc_nprng_randint(C.byref(self._s), low, high))

    def normal(self) -> float:
        return float(lib().orc_nprng_normal(C.byref(self._s)))

    def random(self) -> float:
        return float(lib().orc_nprng_random(C.byref(self._s)))

    def randint3_for_tau(self) -> np.ndarray:
        return np.array([self.randint(INT32_MIN_P1, INT32_MAX_M1) for _ in range(3)], np.int64)

    def key_state(self) -> np.ndarray:
        return np.frombuffer(bytes(self._s.key), dtype=np.uint32).copy()

    @property
    def pos(self) -> int:
        return int(self._s.pos)


def tau_rand_int(state: np.ndarray) -> int:
    assert state.dtype == np.int64 and state.shape == (3,)
    return int(lib().orc_tau_rand_int(_ptr(state)))


def tau_rand(state: np.ndarray) -> float:
    assert state.dtype == np.int64 and state.shape == (3,)
    return float(lib().orc_tau_rand(_ptr(state)))


def fast_cosine(x: np.ndarray, y: np.ndarray) -> float:
    """native/fast_cosine.c restated (fast-math translation unit)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    return float(lib().orc_fast_cosine_c(_ptr(x), _ptr(y), x.shape[0]))


def distance(data: np.ndarray, a: int, b: int, strict: bool = True) -> float:
    data = np.ascontiguousarray(data)
    return float(lib().orc_distance(_DTYPES[data.dtype], _ptr(data), data.shape[1], int(strict), a, b))


def make_forest(data: np.ndarray, n_neighbors: int, n_trees: int, leaf_size: int,
                rng: NumpyRandomState, max_depth: int = 200, workers: int = 1) -> Optional[np.ndarray]:
    data = np.ascontiguousarray(data)
    rows, cols = _I64(0), _I64(0)
    p = lib().orc_make_forest(_DTYPES[data.dtype], _ptr(data), data.shape[0], data.shape[1],
                              n_neighbors, n_trees, leaf_size, C.byref(rng._s), max_depth,
                              C.byref(rows), C.byref(cols), workers)
    if not p:
        return None
    out = np.ctypeslib.as_array(p, shape=(rows.value * cols.value,)).copy().reshape(rows.value, cols.value)
    lib().orc_free(C.cast(p, C.c_void_p))
    return out


def rp_split(data: np.ndarray, ind: np.ndarray, state: np.ndarray):
    """One random-projection split of rows `ind` (float.rs:13-95 / int8.rs:24-123 /
    uint8.rs:38-111) with tau state `state` (3 int64, advanced in place).
    Returns (left rows, right rows, (left pivot, right pivot))."""
    data = np.ascontiguousarray(data)
    ind = np.ascontiguousarray(ind, np.int32)
    assert state.dtype == np.int64 and state.shape == (3,) and state.flags.c_contiguous
    out = np.empty(ind.shape[0], np.int32)
    piv = np.zeros(2, np.int64)
    nl = lib().orc_rp_split(_DTYPES[data.dtype], _ptr(data), data.shape[0], data.shape[1], _ptr(ind),
                            ind.shape[0], _ptr(state), _ptr(out), _ptr(piv))
    if nl < 0:
        raise MemoryError("orc_rp_split")
    return out[:nl].copy(), out[nl:].copy(), (int(piv[0]), int(piv[1]))


def nn_descent(data: np.ndarray, k: int, rng_state: np.ndarray, max_candidates: int, n_iters: int,
               delta: float = 0.001, delta_improv: Optional[float] = 0.001,
               leaf_array: Optional[np.ndarray] = None, n_threads: int = 1, workers: int = 1,
               strict: bool = True, use_sorted: bool = True):
    """Raw (untransformed) NN-descent output, rows ascending; returns (idx, dist, iters)."""
    data = np.ascontiguousarray(data)
    n = data.shape[0]
    st = np.ascontiguousarray(rng_state, np.int64).copy()
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float32)
    if leaf_array is not None:
        leaf = np.ascontiguousarray(leaf_array, np.int32)
        lp, lr, lc = _ptr(leaf), leaf.shape[0], leaf.shape[1]
    else:
        lp, lr, lc = None, 0, 0
    it = _I64(0)
    rc = lib().orc_nn_descent(_DTYPES[data.dtype], _ptr(data), n, data.shape[1], int(strict), k,
                              _ptr(st), max_candidates, n_iters, delta,
                              0.0 if delta_improv is None else delta_improv,
                              int(delta_improv is not None), lp, lr, lc, int(use_sorted),
                              n_threads, workers, _ptr(idx), _ptr(dist), C.byref(it))
    if rc:
        raise MemoryError("oracle nn_descent failed")
    return idx, dist, int(it.value)


def brute_force_knn(data: np.ndarray, n_neighbors: int, workers: int = 1):
    """src/knn.rs:189-281: self skipped, unsorted heap order, raw distances."""
    data = np.ascontiguousarray(data)
    n = data.shape[0]
    k = max(min(n_neighbors, n - 1), 1)
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float32)
    lib().orc_brute_force_knn(_DTYPES[data.dtype], _ptr(data), n, data.shape[1], n_neighbors,
                              _ptr(idx), _ptr(dist), workers)
    return idx, dist


ACC_FASTMATH, ACC_SERIAL, ACC_F64 = 0, 1, 2


def exact_knn(data: np.ndarray, k: int, acc: int = ACC_FASTMATH, include_self: bool = True,
              workers: int = 1):
    """Exact k smallest raw distances with the NN-descent layout (self seeded at -1.0),
    ordered by (distance, self first, index).  Raw distances returned as float64."""
    data = np.ascontiguousarray(data)
    n = data.shape[0]
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float64)
    lib().orc_exact_knn(_DTYPES[data.dtype], _ptr(data), n, data.shape[1], k, acc,
                        int(include_self), _ptr(idx), _ptr(dist), workers)
    return idx, dist


def transform(dtype_code: int, raw: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(raw, np.float32).copy()
    lib().orc_transform(dtype_code, _ptr(out), out.size)
    return out


def prepare_float(x: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(x, np.float32).copy()
    lib().orc_prepare_float(_ptr(out), out.shape[0], out.shape[1])
    return out


@dataclass
class KnnGraphOptions:
    """src/knn.rs:43-76 (None = derived default)."""
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


def knn_graph(data: np.ndarray, options: KnnGraphOptions, rng: NumpyRandomState,
              n_threads: int = 1, workers: int = 1):
    """src/knn.rs:345-454.  Returns (indices, transformed distances, iterations run)."""
    if options.n_neighbors <= 0:
        raise ValueError("n_neighbors must be positive")
    x = np.ascontiguousarray(data).copy()
    n = x.shape[0]
    if n == 0:
        raise ValueError("empty embedding matrix")
    k = max(min(options.n_neighbors, n - 1), 1)
    o = _KnnOpts(options.n_neighbors, options.n_trees or 0, options.leaf_size or 0,
                 options.max_candidates or 0, options.max_rptree_depth, options.n_iters or 0,
                 options.delta, options.delta_improv, int(options.use_sorted_updates),
                 int(options.deterministic), options.brute_force_max_n)
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float32)
    kout, it = _I64(0), _I64(0)
    rc = lib().orc_knn_graph(_DTYPES[x.dtype], _ptr(x), n, x.shape[1], C.byref(o), C.byref(rng._s),
                             n_threads, workers, _ptr(idx), _ptr(dist), C.byref(kout), C.byref(it))
    if rc:
        raise RuntimeError(f"oracle knn_graph failed ({rc})")
    return idx, dist, int(it.value)


def _declare_fuzzy(L):
    L.orc_smooth_knn_dist.argtypes = [_P, _I64, _I64, C.c_float, _P, _P]
    L.orc_fuzzy_graph.argtypes = [_P, _P, _I64, _I64, C.c_float, C.c_int, _P, _P, _P, _P, _P, _I64]
    L.orc_fuzzy_graph.restype = _I64


def smooth_knn_dist(distances: np.ndarray, k: float):
    """sigma, rho per row (src/graph_construction.rs:15-81)."""
    d = np.ascontiguousarray(distances, dtype=np.float32)
    n, kk = d.shape
    L = lib()
    _declare_fuzzy(L)
    sigma = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    L.orc_smooth_knn_dist(_ptr(d), n, kk, float(k), _ptr(sigma), _ptr(rho))
    return sigma, rho


def neighbor_graph_matrix(n_neighbors: float, knn_indices: np.ndarray, knn_dists: np.ndarray,
                          symmetrize: bool = True):
    """Fuzzy graph as CSR (indptr int64, indices int32, data float32), rows sorted
    by column, plus (sigma, rho): src/graph_construction.rs:162-226."""
    ind = np.ascontiguousarray(knn_indices, dtype=np.int32)
    dst = np.ascontiguousarray(knn_dists, dtype=np.float32)
    n, k = ind.shape
    L = lib()
    _declare_fuzzy(L)
    cap = max(1, 2 * n * k)
    sigma = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    indptr = np.empty(n + 1, np.int64)
    indices = np.empty(cap, np.int32)
    data = np.empty(cap, np.float32)
    nnz = L.orc_fuzzy_graph(_ptr(ind), _ptr(dst), n, k, float(n_neighbors), int(bool(symmetrize)), _ptr(sigma),
                            _ptr(rho), _ptr(indptr), _ptr(indices), _ptr(data), cap)
    if nnz < 0:
        raise MemoryError("orc_fuzzy_graph failed")
    return indptr, indices[:nnz].copy(), data[:nnz].copy(), sigma, rho


def find_duplicates(knn_inds: np.ndarray, knn_dists: np.ndarray) -> set:
    """src/cluster_util.rs:430-449: threshold = max of the first distance column
    (NaN ignored, as f32::max folds from NaN); every listed pair (i, j), i != j,
    at a distance <= threshold, as (min, max).  -1 entries are skipped (the
    reference would cast them to usize::MAX)."""
    ind = np.asarray(knn_inds)
    dst = np.asarray(knn_dists, dtype=np.float32)
    col = dst[:, 0]
    col = col[~np.isnan(col)]
    if col.size == 0:
        return set()
    thr = col.max()
    out = set()
    for i in range(ind.shape[0]):
        for j in range(ind.shape[1]):
            kk = int(ind[i, j])
            if kk >= 0 and dst[i, j] <= thr and kk != i:
                out.add((min(i, kk), max(i, kk)))
    return out
