"""The fuzzy graph and find_duplicates on the MI355X (evoc_mi355x_fuzzy_graph,
evoc_mi355x_find_duplicates) against the reference's goldens and the oracle.

Bar: the reference's own rule for the fuzzy graph (tests/graph_parity.rs:29-90):
the same sparsity pattern and |delta| <= 1e-6 per entry, sigma / rho within
1e-6 (both sides compute exp of the same operands; f32 exp may differ in the
last bit).  find_duplicates is integer work: exact set equality."""
import os

import numpy as np
import pytest

import evoc_mi355x as E
import oracle as O
from evoc_mi355x import synth

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _csr_dict(indptr, indices, data):
    out = {}
    for i in range(len(indptr) - 1):
        for p in range(indptr[i], indptr[i + 1]):
            out[(i, int(indices[p]))] = float(data[p])
    return out


def _knn(name):
    return (np.load(os.path.join(GOLD, name, "nn_inds.npy")), np.load(os.path.join(GOLD, name, "nn_dists.npy")))


def assert_csr_close(g, o, tol=1e-6):
    """Same CSR layout, values within tol."""
    gi, gx, gd = g[:3]
    oi, ox, od = o[:3]
    assert np.array_equal(gi, oi), "row starts differ"
    assert np.array_equal(gx, ox), "columns differ"
    if gd.size:
        assert np.abs(gd - od).max() <= tol, np.abs(gd - od).max()


@pytest.fixture(scope="module")
def ctx():
    c = E.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", ["small_200", "medium_800", "large_2000"])
def test_fuzzy_graph_matches_golden(ctx, name):
    ind, dst = _knn(name)
    g = E.neighbor_graph_matrix(15.0, ind, dst, True, ctx=ctx)
    z = np.load(os.path.join(GOLD, name, "graph_coo.npz"))
    want = {(int(r), int(c)): float(v) for r, c, v in zip(z["rows"], z["cols"], z["data"])}
    got = _csr_dict(*g)
    assert set(got) == set(want)
    assert max(abs(got[k] - want[k]) for k in want) <= 1e-6
    assert_csr_close(g, O.neighbor_graph_matrix(15.0, ind, dst, True))


def test_sigma_rho_match_golden(ctx):
    _, dst = _knn("medium_800")
    sg, rh = E.smooth_knn_dist(dst, 15.0, ctx=ctx)
    assert np.abs(sg - np.load(os.path.join(GOLD, "medium_800", "sigmas.npy"))).max() < 1e-6
    assert np.abs(rh - np.load(os.path.join(GOLD, "medium_800", "rhos.npy"))).max() < 1e-6


def test_directed_graph_matches_golden(ctx):
    ind, dst = _knn("medium_800")
    g = E.neighbor_graph_matrix(15.0, ind, dst, False, ctx=ctx)
    z = np.load(os.path.join(GOLD, "medium_800", "directed_coo.npz"))
    want = {(int(r), int(c)): float(v) for r, c, v in zip(z["rows"], z["cols"], z["data"])}
    got = _csr_dict(*g)
    assert set(got) == set(want)
    assert max(abs(got[k] - want[k]) for k in want) <= 1e-6


@pytest.mark.parametrize("n,k,sym", [(20000, 30, True), (20000, 30, False), (5000, 15, True)])
def test_fuzzy_graph_vs_oracle_on_device_knn(ctx, n, k, sym):
    """The device kNN graph of Gaussian blobs, then the fuzzy graph from device
    tensors, against the oracle's fuzzy graph of the same kNN arrays."""
    import torch

    x = synth.blobs_f32(n, 128, 25, 0.3, seed=n + k)
    xd = torch.from_numpy(x).cuda()
    idx = torch.empty((n, k), dtype=torch.int32, device="cuda")
    dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
    E.knn_graph(xd, E.KnnGraphOptions(n_neighbors=k), E.NumpyRandomState(1), ctx=ctx, out=(idx, dist))
    torch.cuda.synchronize()
    g = E.neighbor_graph_matrix(float(k), idx, dist, sym, ctx=ctx, return_sigmas=True)
    o = O.neighbor_graph_matrix(float(k), idx.cpu().numpy(), dist.cpu().numpy(), sym)
    assert_csr_close(g, o)
    assert np.abs(g[3][0] - o[3]).max() <= 1e-6 and np.array_equal(g[3][1], o[4])
    if sym:  # the fuzzy union is symmetric, entry for entry
        d = _csr_dict(*g[:3])
        assert all(d[(j, i)] == v for (i, j), v in d.items())
    # device outputs: the same CSR, bit for bit
    cap = 2 * n * k
    dout = (torch.empty(n + 1, dtype=torch.int64, device="cuda"), torch.empty(cap, dtype=torch.int32, device="cuda"),
            torch.empty(cap, dtype=torch.float32, device="cuda"))
    gd = E.neighbor_graph_matrix(float(k), idx, dist, sym, ctx=ctx, out=dout)
    assert all(np.array_equal(a.cpu().numpy(), b) for a, b in zip(gd, g[:3]))


def test_missing_neighbours_and_errors(ctx):
    ind, dst = _knn("small_200")
    ind = ind.copy()
    ind[::7, 5] = -1  # holes are skipped
    assert_csr_close(E.neighbor_graph_matrix(15.0, ind, dst, True, ctx=ctx),
                     O.neighbor_graph_matrix(15.0, ind, dst, True))
    bad = ind.copy()
    bad[3, 2] = 200  # outside [0, n)
    with pytest.raises(E.BackendError):
        E.neighbor_graph_matrix(15.0, bad, dst, True, ctx=ctx)
    with pytest.raises(E.EmptyData):
        E.neighbor_graph_matrix(15.0, ind[:0], dst[:0], True, ctx=ctx)


def _planted_knn(groups, k, rs):
    """kNN arrays listing self, then the other members of its group at distance
    0, then other rows at distances in [0.1, 1)."""
    gid = np.repeat(np.arange(len(groups)), groups)
    n = len(gid)
    ind = np.empty((n, k), np.int32)
    dst = np.empty((n, k), np.float32)
    for i in range(n):
        mates = [j for j in np.nonzero(gid == gid[i])[0] if j != i][: k - 1]
        others = [j for j in rs.permutation(n) if gid[j] != gid[i]][: k - 1 - len(mates)]
        ind[i] = [i] + mates + others
        dst[i] = [0.0] * (1 + len(mates)) + sorted(rs.uniform(0.1, 1.0, len(others)).tolist())
    return gid, ind, dst


def test_find_duplicates(ctx):
    """Planted duplicate groups: exactly the pairs within groups; a device kNN
    graph and the goldens: the oracle's set."""
    rs = np.random.default_rng(3)
    gid, ind, dst = _planted_knn(rs.integers(1, 5, size=300), 8, rs)
    want = {(i, j) for i in range(len(gid)) for j in range(i + 1, len(gid)) if gid[i] == gid[j]}
    assert E.find_duplicates(ind, dst, ctx=ctx) == want == O.find_duplicates(ind, dst)
    x = synth.blobs_f32(3000, 64, 10, 0.3, seed=5)
    x[1::3] = x[0::3][: len(x[1::3])]  # repeated rows
    idx, dist = E.knn_graph(x, E.KnnGraphOptions(n_neighbors=10), E.NumpyRandomState(2), ctx=ctx)
    assert E.find_duplicates(idx, dist, ctx=ctx) == O.find_duplicates(idx, dist)
    for name in ["small_200", "medium_800"]:
        ind, dst = _knn(name)
        assert E.find_duplicates(ind, dst, ctx=ctx) == O.find_duplicates(ind, dst)
