"""Shared parity helpers for the tests."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = [("small_200", 42), ("medium_800", 42), ("large_2000", 7)]


def load_fixture(name):
    d = os.path.join(GOLDEN, name)
    with open(os.path.join(d, "meta.json")) as f:
        meta = json.load(f)
    return (np.load(os.path.join(d, "data.npy")), np.load(os.path.join(d, "nn_inds.npy")),
            np.load(os.path.join(d, "nn_dists.npy")), meta)


def knn_parity_counts(inds, dists, exp_inds, exp_dists, tol=1e-6):
    """tests/parity.rs:34-68: (#|dd|>tol, #index mismatches with different distance,
    #index mismatches at equal distance)."""
    inds = np.asarray(inds)
    dd = np.abs(np.asarray(dists, np.float64) - np.asarray(exp_dists, np.float64))
    dist_m = int((dd > tol).sum())
    bad = int(((inds != exp_inds) & (dd > tol)).sum())
    tie = int(((inds != exp_inds) & (dd <= tol)).sum())
    return dist_m, bad, tie


def assert_knn_parity(inds, dists, exp_inds, exp_dists, max_ties=8, tol=1e-6):
    dist_m, bad, tie = knn_parity_counts(inds, dists, exp_inds, exp_dists, tol)
    assert dist_m == 0, f"kNN distance mismatch count {dist_m}"
    assert bad == 0, f"kNN index mismatch with different distance {bad}"
    assert tie <= max_ties, f"excessive tied-neighbor reordering: {tie}"


def boundary_mismatch_rows(inds, dists, exp_inds, exp_dists, tol):
    """Rows whose neighbour set differs from the expected one in a member that is not
    tied (within `tol`) with the expected row's k-th distance."""
    bad = 0
    for r in range(inds.shape[0]):
        a = dict(zip(inds[r].tolist(), np.asarray(dists[r], np.float64).tolist()))
        b = dict(zip(exp_inds[r].tolist(), np.asarray(exp_dists[r], np.float64).tolist()))
        if a.keys() == b.keys():
            continue
        kth = max(b.values())
        for j in a.keys() ^ b.keys():
            dj = a[j] if j in a else b[j]
            if abs(dj - kth) > tol:
                bad += 1
                break
    return bad


def recall(inds, truth):
    """Mean |S ∩ S*| / k over rows."""
    k = truth.shape[1]
    hit = 0
    for a, b in zip(inds, truth):
        hit += len(set(a.tolist()) & set(b.tolist()))
    return hit / (k * truth.shape[0])


def blobs(n, d, centers, std, seed, normalize=True, zipf=None):
    """Gaussian blobs from a numpy Generator (test data, not a reference format)."""
    g = np.random.default_rng(seed)
    c = g.standard_normal((centers, d))
    if zipf is None:
        sizes = np.full(centers, n // centers)
        sizes[: n - sizes.sum()] += 1
    else:
        w = np.arange(1, centers + 1, dtype=np.float64) ** (-zipf)
        sizes = np.floor(w / w.sum() * n).astype(np.int64)
        sizes[: n - sizes.sum()] += 1
    lab = np.repeat(np.arange(centers), sizes)
    g.shuffle(lab)
    x = c[lab] + std * g.standard_normal((n, d))
    x = x.astype(np.float32)
    if normalize:
        x /= np.linalg.norm(x, axis=1, keepdims=True).astype(np.float32)
    return np.ascontiguousarray(x, np.float32)
