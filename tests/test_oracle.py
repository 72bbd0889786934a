"""Pins the CPU oracle (oracle/evoc_oracle.c) before it is trusted as the checker.

Every check here is against something other than the oracle's own formula:
the reference's golden fixtures and known-answer tests, the reference's own
C source compiled in oracle/_ref, numpy brute force, or a mathematical identity.
"""
import numpy as np
import pytest

import oracle as O
from parity_util import (FIXTURES, assert_knn_parity, blobs, boundary_mismatch_rows,
                         load_fixture, recall)

pytestmark = pytest.mark.filterwarnings("ignore::DeprecationWarning")


# --- known-answer tests --------------------------------------------------------------

def test_numpy_rng_kat():
    """src/numpy_rng.rs:232-252."""
    r = O.NumpyRandomState(42)
    lo, hi = O.INT32_MIN_P1, O.INT32_MAX_M1
    assert [r.randint(lo, hi) for _ in range(3)] == [-538846105, 1273642420, 1935803229]
    assert [r.randint(lo, hi) for _ in range(5)] == [-1359637233, 996406379, 1201263688,
                                                      423734973, 415968277]
    nrm = [r.normal() for _ in range(3)]
    assert abs(nrm[0] - (-0.23415337472333597)) < 1e-12
    assert abs(nrm[1] - (-0.23413695694918055)) < 1e-12
    assert abs(nrm[2] - 1.5792128155073915) < 1e-12


def test_numpy_rng_uniform_matches_numpy():
    """The legacy MT19937 stream equals numpy's own RandomState (random_sample)."""
    ref = np.random.RandomState(1234)
    r = O.NumpyRandomState(1234)
    got = np.array([r.random() for _ in range(100)])
    np.testing.assert_array_equal(got, ref.random_sample(100))


def test_fast_cosine_kat():
    """src/fast_cosine.rs:70-85."""
    a = np.array([1.0, 0.0], np.float32)
    assert abs(O.fast_cosine(a, np.array([0.5, 0.8660254], np.float32)) + 0.5) < 1e-6
    assert O.fast_cosine(a, np.array([-1.0, 0.0], np.float32)) == np.float32(1.17549435e-38)


def test_fast_cosine_bitwise_vs_reference_build():
    """oracle_fastmath.c against the reference's own native/fast_cosine.c (oracle/_ref)."""
    import ctypes as C
    import os

    path = os.path.join(os.path.dirname(O.__file__), "_ref", "libref_fast_cosine.so")
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    ref = C.CDLL(path)
    ref.fast_cosine_numba.restype = C.c_float
    ref.fast_cosine_numba.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    g = np.random.default_rng(0)
    for _ in range(2000):
        d = int(g.integers(1, 1025))
        a = g.standard_normal(d).astype(np.float32)
        b = g.standard_normal(d).astype(np.float32)
        v_ref = np.float32(ref.fast_cosine_numba(a.ctypes.data, b.ctypes.data, d))
        assert np.float32(O.fast_cosine(a, b)).tobytes() == v_ref.tobytes()


def test_tau_rand_range_and_offsets():
    """tau_rand maps |int32| onto [0,1]; offset streams differ (src/rng.rs:17-35)."""
    st = np.array([-538846105, 1273642420, 1935803229], np.int64)
    vals = [O.tau_rand(st) for _ in range(2000)]
    assert min(vals) >= 0.0 and max(vals) <= 1.0
    assert 0.45 < float(np.mean(vals)) < 0.55


# --- reference goldens ---------------------------------------------------------------

@pytest.mark.parametrize("name,seed", FIXTURES)
def test_rng_after_knn_matches_golden(name, seed):
    """tests/parity_intermediates.rs:656-687: knn_graph draws 3 + 3*4 values."""
    x, _, _, _ = load_fixture(name)
    rng = O.NumpyRandomState(seed)
    O.knn_graph(x, O.KnnGraphOptions(n_neighbors=15, deterministic=True), rng)
    key = np.load(f"{O.__file__.rsplit('/', 2)[0]}/tests/golden/{name}/rng_after_knn_key.npy")
    meta = np.load(f"{O.__file__.rsplit('/', 2)[0]}/tests/golden/{name}/rng_after_knn_meta.npz")
    assert np.array_equal(rng.key_state(), key)
    assert rng.pos == int(meta["pos"])


def test_leaf_array_golden():
    """medium_800/leaf_array.npy: 4 RP trees from a fresh RandomState(42)."""
    x, _, _, _ = load_fixture("medium_800")
    gl = np.load(f"{O.__file__.rsplit('/', 2)[0]}/tests/golden/medium_800/leaf_array.npy")
    la = O.make_forest(O.prepare_float(x), 15, 4, 15, O.NumpyRandomState(42))
    assert la.shape == gl.shape
    assert np.array_equal(la, gl)


@pytest.mark.parametrize("name,seed", FIXTURES)
def test_knn_graph_deterministic_goldens(name, seed):
    """tests/parity.rs:135-189 (assert_knn_parity, deterministic=true)."""
    x, ei, ed, _ = load_fixture(name)
    idx, dist, _ = O.knn_graph(x, O.KnnGraphOptions(n_neighbors=15, deterministic=True),
                               O.NumpyRandomState(seed))
    assert_knn_parity(idx, dist, ei, ed)


def test_knn_graph_deterministic_with_reference_cosine():
    """Same golden check with the reference's own compiled fast_cosine in the loop."""
    if not O.use_reference_fast_cosine(True):
        pytest.skip("oracle/_ref not built")
    try:
        for name, seed in FIXTURES:
            x, ei, ed, _ = load_fixture(name)
            idx, dist, _ = O.knn_graph(x, O.KnnGraphOptions(n_neighbors=15, deterministic=True),
                                       O.NumpyRandomState(seed))
            assert_knn_parity(idx, dist, ei, ed)
    finally:
        O.use_reference_fast_cosine(False)


def test_knn_only_small_200_nondeterministic():
    """tests/parity.rs:235-260: default (non-deterministic) path, exact index equality."""
    x, ei, _, _ = load_fixture("small_200")
    idx, _, _ = O.knn_graph(x, O.KnnGraphOptions(n_neighbors=15), O.NumpyRandomState(42))
    assert np.array_equal(idx, ei)


def test_knn_graph_float_smoke_shapes():
    """tests/smoke.rs:32-47: 500x32, k=10."""
    r = O.NumpyRandomState(1)
    x = np.array([[r.normal() for _ in range(32)] for _ in range(500)], np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    idx, dist, _ = O.knn_graph(x, O.KnnGraphOptions(n_neighbors=10), O.NumpyRandomState(1))
    assert idx.shape == (500, 10) and dist.shape == (500, 10)


def test_raw_heap_distances_nonpositive():
    """src/knn.rs:464-519: raw NN-descent distances on medium_800 are <= 0."""
    x, _, _, _ = load_fixture("medium_800")
    x = O.prepare_float(x)
    rng = O.NumpyRandomState(42)
    st = rng.randint3_for_tau()
    leaf = O.make_forest(x, 15, 4, 15, rng)
    _, raw, _ = O.nn_descent(x, 15, st, 22, 10, leaf_array=leaf)
    assert raw.max() <= 0.0


# --- mathematics ---------------------------------------------------------------------

def test_exact_knn_f64_matches_numpy():
    x = blobs(300, 24, 6, 0.3, seed=3)
    idx, dist = O.exact_knn(x, 12, acc=O.ACC_F64)
    s = x.astype(np.float64) @ x.astype(np.float64).T
    raw = np.where(s > 0, -s, np.float64(np.finfo(np.float32).tiny))
    np.fill_diagonal(raw, -1.0)
    order = np.lexsort((np.arange(300)[None, :].repeat(300, 0),
                        ~np.eye(300, dtype=bool), raw), axis=1)[:, :12]
    np.testing.assert_array_equal(idx, order)
    np.testing.assert_allclose(dist, np.take_along_axis(raw, order, 1), rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", ["i8", "u8"])
def test_brute_force_equals_exact_without_self(dtype):
    """knn.rs:189-281 heap output, once sorted, equals the exact k smallest (self skipped)."""
    g = np.random.default_rng(5)
    if dtype == "i8":
        x = g.integers(-127, 128, size=(200, 40)).astype(np.int8)
    else:
        x = g.integers(0, 256, size=(200, 16)).astype(np.uint8)
    bi, bd = O.brute_force_knn(x, 10)
    ei, ed = O.exact_knn(x, 10, include_self=False)
    for r in range(200):
        o = np.lexsort((bi[r], bd[r]))
        np.testing.assert_array_equal(bd[r][o], ed[r].astype(np.float32))
        # members strictly inside the k-th distance are fixed; at the boundary the heap
        # layout (not the index) decides which tied candidate survives
        inner = ed[r] < ed[r][-1]
        assert set(bi[r][o][inner].tolist()) == set(ei[r][inner].tolist())


def test_int8_distance_closed_form():
    g = np.random.default_rng(6)
    x = g.integers(-127, 128, size=(50, 512)).astype(np.int8)
    for a, b in [(0, 1), (3, 3), (10, 40)]:
        exp = -float(np.dot(x[a].astype(np.int64), x[b].astype(np.int64)))
        assert O.distance(x, a, b) == np.float32(exp)


def test_jaccard_closed_form():
    g = np.random.default_rng(7)
    x = g.integers(0, 256, size=(50, 128)).astype(np.uint8)
    x[5] = 0
    bits = np.unpackbits(x, axis=1).astype(bool)
    for a, b in [(0, 1), (2, 2), (10, 40), (5, 5), (5, 6)]:
        inter = int((bits[a] & bits[b]).sum())
        union = int((bits[a] | bits[b]).sum())
        exp = -(np.float32(inter) / np.float32(union)) if union else np.float32(0.0)
        assert O.distance(x, a, b) == np.float32(exp)


def test_transform_closed_forms():
    raw = np.array([-1.0, -0.5, -0.25, 1.17549435e-38, -1.0000001, 3.4e38], np.float32)
    t = O.transform(O.F32, raw)
    np.testing.assert_allclose(t, [0.0, 1.0, 2.0, 0.0, 0.0, 0.0], atol=1e-7)
    t8 = O.transform(O.I8, np.array([-4.0, -0.5, 0.0, 7.0], np.float32))
    np.testing.assert_array_equal(t8, np.array([0.25, 2.0, 0.0, 0.0], np.float32))
    tu = O.transform(O.U8, np.array([-1.0, -0.5, -0.125, 0.0], np.float32))
    assert tu[0] == 0.0 and np.signbit(tu[0])  # -log2(1) = -0.0
    np.testing.assert_array_equal(tu[1:], np.array([1.0, 3.0, 0.0], np.float32))


def test_int8_forest_tracks_float_forest():
    """The i8 split normalises both pivot rows (int8.rs:43-76); on an int8 matrix and its
    L2-normalised f32 copy the hyperplanes agree (a margin's sign does not change when the
    tested row is scaled), so the forests agree up to rounding.  A third of the rows are
    scaled by 0.3 and a third by 0.08 before quantising: with such different pivot norms a
    split without the per-pivot normalisation picks other hyperplanes."""
    g = np.random.default_rng(8)
    c = g.standard_normal((8, 64))
    x = c[g.integers(0, 8, 600)] + 0.3 * g.standard_normal((600, 64))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    x[200:400] *= 0.3
    x[400:] *= 0.08
    xi = np.clip(np.round(x * 127), -127, 127).astype(np.int8)
    xf = xi.astype(np.float32)
    xf /= np.linalg.norm(xf, axis=1, keepdims=True)
    li = O.make_forest(xi, 10, 4, 10, O.NumpyRandomState(9))
    lf = O.make_forest(xf, 10, 4, 10, O.NumpyRandomState(9))
    si = set(frozenset(r[r >= 0].tolist()) for r in li)
    sf = set(frozenset(r[r >= 0].tolist()) for r in lf)
    assert len(si & sf) >= 0.9 * len(sf)


@pytest.mark.parametrize("dtype", ["f32", "i8", "u8"])
def test_forest_is_a_partition_per_tree(dtype):
    g = np.random.default_rng(10)
    n = 700
    if dtype == "f32":
        x = blobs(n, 32, 5, 0.3, seed=10)
    elif dtype == "i8":
        x = g.integers(-127, 128, size=(n, 32)).astype(np.int8)
    else:
        x = g.integers(0, 256, size=(n, 16)).astype(np.uint8)
    leaf = O.make_forest(x, 15, 5, 15, O.NumpyRandomState(11))
    vals = leaf[leaf >= 0]
    assert np.array_equal(np.bincount(vals, minlength=n), np.full(n, 5))
    assert ((leaf >= 0).sum(1) <= 15).all()


@pytest.mark.parametrize("dtype", ["f32", "i8", "u8"])
def test_nn_descent_reaches_exact_on_small_data(dtype):
    g = np.random.default_rng(12)
    n, k = 400, 10
    if dtype == "f32":
        x = blobs(n, 16, 8, 0.3, seed=12)
    elif dtype == "i8":
        xf = blobs(n, 32, 8, 0.3, seed=13)
        x = np.clip(np.round(xf * 127), -127, 127).astype(np.int8)
    else:
        protos = g.random((8, 256)) < 0.3
        bits = protos[g.integers(0, 8, n)] ^ (g.random((n, 256)) < 0.08)
        x = np.packbits(bits, axis=1)
    rng = O.NumpyRandomState(13)
    idx, raw, _ = O.knn_graph(x, O.KnnGraphOptions(n_neighbors=k, n_iters=30, delta=0.0,
                                                    delta_improv=0.0), rng)
    if dtype == "i8":
        # int8 evicts the -1.0 self seed, but a new x old pair with p == q re-inserts self
        # at -|x|^2 (int8.rs:420-572 + heap.rs:143-200): score the non-self entries only
        ei, _ = O.exact_knn(x, k, acc=O.ACC_SERIAL, include_self=False)
        hit = tot = 0
        for r in range(n):
            got = set(idx[r].tolist()) - {r}
            hit += len(got & set(ei[r].tolist()))
            tot += len(got)
        assert hit / tot >= 0.995
    elif dtype == "u8":
        # Jaccard values are small rationals: many exact ties at the k-th boundary
        ei, ed = O.exact_knn(x, k, acc=O.ACC_SERIAL)
        exp_t = O.transform(O.U8, ed.astype(np.float32))
        assert boundary_mismatch_rows(idx, raw, ei, exp_t, 1e-6) <= n // 100
    else:
        ei, _ = O.exact_knn(x, k, acc=O.ACC_SERIAL)
        assert recall(idx, ei) >= 0.995
    # rows ascending, no duplicate neighbours
    assert (np.diff(raw, axis=1) >= 0).all()
    for r in idx:
        assert len(set(r.tolist())) == k
    if dtype in ("f32", "u8"):
        assert (idx[:, 0] == np.arange(n)).mean() > 0.99  # self kept at column 0


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_nn_descent_worker_count_does_not_change_result(threads):
    """Running the T simulated workers on several OS threads is deterministic."""
    x = blobs(1500, 32, 10, 0.3, seed=14)
    rng = O.NumpyRandomState(3)
    st = rng.randint3_for_tau()
    leaf = O.make_forest(x, 15, 4, 15, rng)
    a = O.nn_descent(x, 15, st, 22, 6, leaf_array=leaf, n_threads=8, workers=1)
    b = O.nn_descent(x, 15, st, 22, 6, leaf_array=leaf, n_threads=8, workers=threads)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_unsorted_path_matches_goldens():
    """float.rs:672-772 (use_sorted_updates=false) also reproduces the goldens."""
    for name, seed in FIXTURES:
        x, ei, ed, _ = load_fixture(name)
        idx, dist, _ = O.knn_graph(x, O.KnnGraphOptions(n_neighbors=15, deterministic=True,
                                                        use_sorted_updates=False),
                                   O.NumpyRandomState(seed))
        assert_knn_parity(idx, dist, ei, ed)


def test_brute_force_branch_rng_advance():
    """knn.rs:418-422: brute force for int8 when n <= brute_force_max_n; only 3 draws."""
    g = np.random.default_rng(15)
    x = g.integers(-127, 128, size=(100, 16)).astype(np.int8)
    rng = O.NumpyRandomState(5)
    idx, dist, _ = O.knn_graph(x, O.KnnGraphOptions(n_neighbors=5, brute_force_max_n=1000), rng)
    ref = O.NumpyRandomState(5)
    for _ in range(3):
        ref.randint(O.INT32_MIN_P1, O.INT32_MAX_M1)
    assert np.array_equal(rng.key_state(), ref.key_state()) and rng.pos == ref.pos
    assert not (idx == np.arange(100)[:, None]).any()  # self skipped


def test_error_cases():
    x = np.zeros((0, 4), np.float32)
    with pytest.raises(ValueError):
        O.knn_graph(x, O.KnnGraphOptions(n_neighbors=5), O.NumpyRandomState(0))
    with pytest.raises(ValueError):
        O.knn_graph(np.ones((4, 4), np.float32), O.KnnGraphOptions(n_neighbors=0),
                    O.NumpyRandomState(0))


def _tau_mod(state, modulus):
    """tau_rand_int(state).rem_euclid(modulus) (src/rng.rs:22-26), state advanced in place."""
    return O.tau_rand_int(state) % modulus


def _split_setup(n, m, seed):
    g = np.random.default_rng(seed)
    ind = g.permutation(n)[:m].astype(np.int32)
    state = O.NumpyRandomState(seed).randint3_for_tau()
    return ind, state


def _check_split(ind, state0, left, right, piv, margin, tol):
    """The split a margin implies: the pivots the tau draws pick (rem_euclid, right index
    bumped past the left one, src/nndescent/uint8.rs:45-49), rows with margin > tol left and
    < -tol right, each side in input order.  Returns the rows the check decided."""
    st = state0.copy()
    m = ind.shape[0]
    li = _tau_mod(st, m)
    ri = _tau_mod(st, m)
    ri = (ri + (li == ri)) % m
    assert piv == (int(ind[li]), int(ind[ri]))
    pos = {int(p): t for t, p in enumerate(ind)}
    lset, rset = set(left.tolist()), set(right.tolist())
    assert len(lset) + len(rset) == m and not (lset & rset)
    assert sorted(pos[p] for p in left.tolist()) == [pos[p] for p in left.tolist()]
    assert sorted(pos[p] for p in right.tolist()) == [pos[p] for p in right.tolist()]
    sure = np.abs(margin) > tol
    exp_left = set(ind[sure & (margin > 0)].tolist())
    exp_right = set(ind[sure & (margin < 0)].tolist())
    assert exp_left and exp_right  # both sides non-empty: no random fallback split
    assert exp_left <= lset and exp_right <= rset
    return sure


def test_uint8_split_margin_closed_form():
    """uint8.rs:56-82: with positive = (a ^ b) & a and negative = (a ^ b) & b, the margin of
    row x is popc(x & a & ~b) - popc(x & b & ~a), an exact integer; positive rows go left,
    negative right, zero margins by a tau draw.  Counting only the positive term (a plausible
    slip) classifies some of these rows differently, and the test would catch it."""
    g = np.random.default_rng(30)
    n, nbytes = 500, 24
    x = np.packbits(g.random((n, nbytes * 8)) < g.uniform(0.1, 0.6, (n, 1)), axis=1)
    wrong_disagrees = 0
    for seed in range(31, 37):
        ind, st = _split_setup(n, 300, seed)
        st0 = st.copy()
        left, right, piv = O.rp_split(x, ind, st)
        bits = np.unpackbits(x, axis=1).astype(np.int64)
        a, b = bits[piv[0]], bits[piv[1]]
        xs = bits[ind]
        margin = (xs * (a * (1 - b))).sum(1) - (xs * (b * (1 - a))).sum(1)
        _check_split(ind, st0, left, right, piv, margin.astype(np.float64), 0.5)
        wrong = (xs * (a * (1 - b))).sum(1)
        in_left = np.isin(ind, left)
        wrong_disagrees += int(((wrong > 0) != in_left)[margin != 0].sum())
    assert wrong_disagrees > 0


def test_int8_split_normalises_each_pivot():
    """int8.rs:43-76: the hyperplane is a/|a| - b/|b| (then unit length), so the margin of
    row x has the sign of x . (a/|a| - b/|b|).  Rows of very different norms (some scaled by
    0.3 or 0.08 before quantising): a split on a - b alone (no per-pivot normalisation)
    classifies some rows differently, and the test would catch it."""
    g = np.random.default_rng(40)
    n, d = 600, 96
    c = g.standard_normal((6, d))
    xf = c[g.integers(0, 6, n)] + 0.5 * g.standard_normal((n, d))
    xf /= np.linalg.norm(xf, axis=1, keepdims=True)
    xf *= g.choice([1.0, 0.3, 0.08], size=(n, 1))
    x = np.clip(np.round(xf * 127), -127, 127).astype(np.int8)
    x[::97] = 0  # zero rows: a zero pivot contributes 0 (int8.rs:59-68)
    wrong_disagrees = 0
    for seed in range(41, 49):
        ind, st = _split_setup(n, 400, seed)
        st0 = st.copy()
        left, right, piv = O.rp_split(x, ind, st)
        X = x.astype(np.float64)
        a, b = X[piv[0]], X[piv[1]]
        na, nb = np.linalg.norm(a), np.linalg.norm(b)
        hv = (a / na if na > 0 else 0 * a) - (b / nb if nb > 0 else 0 * b)
        xs = X[ind]
        margin = xs @ hv
        tol = 1e-4 * (np.linalg.norm(xs, axis=1) * np.linalg.norm(hv) + 1e-30)
        sure = _check_split(ind, st0, left, right, piv, margin, tol)
        wrong = xs @ (a - b)
        in_left = np.isin(ind, left)
        wrong_disagrees += int(((wrong > 0) != in_left)[sure].sum())
    assert wrong_disagrees > 0


def test_float_split_margin_closed_form():
    """float.rs:13-95: the margin of row x has the sign of x . (a - b)."""
    x = blobs(500, 48, 6, 0.3, seed=50)
    for seed in range(51, 55):
        ind, st = _split_setup(500, 350, seed)
        st0 = st.copy()
        left, right, piv = O.rp_split(x, ind, st)
        X = x.astype(np.float64)
        hv = X[piv[0]] - X[piv[1]]
        xs = X[ind]
        margin = xs @ hv
        _check_split(ind, st0, left, right, piv, margin, 1e-4 * np.linalg.norm(hv))
