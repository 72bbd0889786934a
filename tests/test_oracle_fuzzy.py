"""The oracle's fuzzy graph and find_duplicates (CPU), pinned against the
reference's goldens (tests/graph_parity.rs:29-90 rules: |delta| <= 1e-6 per
entry, sigma / rho within 1e-6) and closed forms."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _coo_dict(rows, cols, data):
    return {(int(r), int(c)): float(v) for r, c, v in zip(rows, cols, data)}


def _csr_dict(indptr, indices, data):
    out = {}
    for i in range(len(indptr) - 1):
        for p in range(indptr[i], indptr[i + 1]):
            out[(i, int(indices[p]))] = float(data[p])
    return out


def _knn(name):
    return (np.load(os.path.join(GOLD, name, "nn_inds.npy")), np.load(os.path.join(GOLD, name, "nn_dists.npy")))


def assert_graph_matches(got: dict, want: dict, tol=1e-6):
    assert set(got) == set(want), (len(set(got) - set(want)), len(set(want) - set(got)))
    worst = max(abs(got[k] - want[k]) for k in want)
    assert worst <= tol, worst


@pytest.mark.parametrize("name", ["small_200", "medium_800", "large_2000"])
def test_fuzzy_union_matches_golden_coo(name):
    ind, dst = _knn(name)
    indptr, indices, data, _, _ = O.neighbor_graph_matrix(15.0, ind, dst, True)
    z = np.load(os.path.join(GOLD, name, "graph_coo.npz"))
    assert_graph_matches(_csr_dict(indptr, indices, data), _coo_dict(z["rows"], z["cols"], z["data"]))
    # rows ascending by column, no explicit zeros
    for i in range(len(indptr) - 1):
        assert np.all(np.diff(indices[indptr[i]:indptr[i + 1]]) > 0)
    assert np.all(data > 0) and np.all(data <= 1)


def test_sigma_rho_match_golden():
    _, dst = _knn("medium_800")
    sg, rh = O.smooth_knn_dist(dst, 15.0)
    assert np.abs(sg - np.load(os.path.join(GOLD, "medium_800", "sigmas.npy"))).max() < 1e-6
    assert np.abs(rh - np.load(os.path.join(GOLD, "medium_800", "rhos.npy"))).max() < 1e-6


def test_directed_strengths_match_golden():
    ind, dst = _knn("medium_800")
    indptr, indices, data, _, _ = O.neighbor_graph_matrix(15.0, ind, dst, False)
    z = np.load(os.path.join(GOLD, "medium_800", "directed_coo.npz"))
    assert_graph_matches(_csr_dict(indptr, indices, data), _coo_dict(z["rows"], z["cols"], z["data"]))


def test_csr_layout_matches_golden_csr():
    ind, dst = _knn("medium_800")
    indptr, indices, data, _, _ = O.neighbor_graph_matrix(15.0, ind, dst, True)
    z = np.load(os.path.join(GOLD, "medium_800", "graph_csr.npz"))
    assert np.array_equal(indptr, z["indptr"]) and np.array_equal(indices, z["indices"])
    assert np.abs(data - z["data"]).max() <= 1e-6


def test_fuzzy_union_closed_form():
    """A + A^T - A o A^T on a tiny hand-made graph, each entry by its formula."""
    ind = np.array([[0, 1, 2], [1, 0, 3], [2, 3, 0], [3, 2, 1]], np.int32)
    dst = np.array([[0.0, 0.5, 1.5], [0.0, 0.5, 2.0], [0.0, 0.25, 1.5], [0.0, 0.25, 2.0]], np.float32)
    _, _, _, sg, rh = O.neighbor_graph_matrix(3.0, ind, dst, False)
    a = {}
    for i in range(4):
        for t in range(3):
            j = int(ind[i, t])
            if j == i:
                continue
            x = np.float32(dst[i, t]) - rh[i]
            a[(i, j)] = np.float32(1.0) if x <= 0 else np.float32(np.exp(np.float32(-(x / sg[i]))))
    want = {}
    for (i, j), v in a.items():
        b = a.get((j, i))
        want[(i, j)] = float(v if b is None else np.float32(np.float32(v + b) - np.float32(v * b)))
        if b is None:
            want[(j, i)] = float(v)
    indptr, indices, data, _, _ = O.neighbor_graph_matrix(3.0, ind, dst, True)
    assert_graph_matches(_csr_dict(indptr, indices, data), want)
    got = _csr_dict(indptr, indices, data)
    assert all(got[(i, j)] == got[(j, i)] for (i, j) in got)


def test_find_duplicates_planted_groups():
    """kNN arrays that list self, then the other members of the row's group at
    distance 0, then other rows further away: the pairs are exactly those within
    groups (the first-column threshold is 0)."""
    rs = np.random.default_rng(3)
    groups = rs.integers(1, 5, size=60)
    gid = np.repeat(np.arange(len(groups)), groups)
    n, k = len(gid), 6
    ind = np.empty((n, k), np.int32)
    dst = np.empty((n, k), np.float32)
    for i in range(n):
        mates = [j for j in np.nonzero(gid == gid[i])[0] if j != i]
        others = [j for j in rs.permutation(n) if gid[j] != gid[i]][: k - 1 - len(mates)]
        ind[i] = [i] + mates + others
        dst[i] = [0.0] * (1 + len(mates)) + sorted(rs.uniform(0.1, 1.0, len(others)).tolist())
    want = {(i, j) for i in range(n) for j in range(i + 1, n) if gid[i] == gid[j]}
    assert O.find_duplicates(ind, dst) == want
    # the threshold is the first column's largest value, not zero
    dst2 = dst.copy()
    dst2[0, 0] = 0.5
    extra = {(min(i, int(ind[i, t])), max(i, int(ind[i, t]))) for i in range(n) for t in range(1, k)
             if dst2[i, t] <= 0.5 and ind[i, t] != i}
    assert O.find_duplicates(ind, dst2) == want | extra
