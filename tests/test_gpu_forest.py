"""The device random-projection forest (evoc_mi355x_make_forest), checked directly:
the reference's make_forest (src/knn.rs:115-154, float.rs:97-177, int8.rs:125-204,
uint8.rs:113-192) makes each tree a partition of the rows into leaves of at most
leaf_size rows, splitting at most max_rptree_depth levels, and its hyperplane splits
keep near neighbours together.  The device draws its random choices from a hash, not
the tau RNG, so the trees are compared with the oracle's by what they achieve (the
share of leaf-mate pairs that are true neighbours), not leaf for leaf."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

E = pytest.importorskip("evoc_mi355x")
from evoc_mi355x import synth  # noqa: E402

TREES = np.arange(1, 13, dtype=np.int64).reshape(4, 3) * 7919


def _data(kind, n):
    if kind == "f32":
        return synth.blobs_f32(n, 64, 40, 0.3, seed=61)
    if kind == "i8":
        return synth.blobs_i8(n, 128, 40, 0.3, seed=62)
    return synth.bits_u8(n, 512, 40, 0.3, 0.08, seed=63)


def _split_trees(leaf, per_tree):
    out, r = [], 0
    for c in per_tree:
        out.append(leaf[r:r + c])
        r += c
    assert r == leaf.shape[0]
    return out


def _check_partition(tree, n):
    vals = tree[tree >= 0]
    assert np.array_equal(np.sort(vals), np.arange(n))
    # -1 only as padding after a leaf's members
    for row in tree:
        m = (row >= 0).sum()
        assert (row[:m] >= 0).all() and (row[m:] == -1).all() and m >= 1


def _neighbour_share(trees, truth):
    """Share of leaf-mate pairs (p, q), p != q, with q among p's exact neighbours."""
    tset = [set(r.tolist()) for r in truth]
    hit = tot = 0
    for tree in trees:
        for row in tree:
            mem = row[row >= 0]
            for p in mem.tolist():
                for q in mem.tolist():
                    if p != q:
                        tot += 1
                        hit += q in tset[p]
    return hit / max(tot, 1)


@pytest.mark.parametrize("kind", ["f32", "i8", "u8"])
def test_device_forest_partitions_and_bounds(kind):
    n, k = 8000, 15
    print(f"[gpu] forest {kind} n={n}", flush=True)
    x = _data(kind, n)
    opts = E.KnnGraphOptions(n_neighbors=k)
    leaf, per_tree, depths = E.make_forest(x, opts, TREES)
    assert leaf.shape[1] >= 15 and len(per_tree) == 4
    trees = _split_trees(leaf, per_tree)
    for t, tree in enumerate(trees):
        _check_partition(tree, n)
        sizes = (tree >= 0).sum(1)
        assert sizes.max() <= 15  # leaf_size = max(10, k) (src/knn.rs:408-411)
        # a tree whose leaves hold at most 15 of n rows has at least n / 15 leaves
        assert int(np.ceil(np.log2(n / 15))) <= depths[t] <= 200
    # the same seeds give the same forest
    leaf2, per2, dep2 = E.make_forest(x, opts, TREES)
    assert np.array_equal(leaf, leaf2) and np.array_equal(per_tree, per2) and np.array_equal(depths, dep2)
    # the splits follow the data as the reference's do: leaf mates are true neighbours about
    # as often as in the oracle's forest (same options, its own tau-RNG draws), and far more
    # often than in a random partition into leaves of the same sizes
    if kind == "f32":
        truth, _ = O.exact_knn(x, k, acc=O.ACC_F64, workers=8)
    else:
        truth, _ = O.exact_knn(x, k, include_self=(kind == "u8"), workers=8)
    ref = O.make_forest(x, k, 4, 15, O.NumpyRandomState(3))
    ref_trees = [ref]  # the oracle returns the trees concatenated; the share is per pair
    dev = _neighbour_share(trees, truth)
    orc = _neighbour_share(ref_trees, truth)
    rng = np.random.default_rng(0)
    perm = rng.permutation(n)
    rand_tree = np.full((n // 15 + 1, 15), -1, np.int32)
    rand_tree.reshape(-1)[:n] = perm
    rnd = _neighbour_share([rand_tree], truth)
    print(f"[gpu] forest {kind}: neighbour share device={dev:.4f} oracle={orc:.4f} random={rnd:.4f} "
          f"depths={depths.tolist()} leaves={per_tree.tolist()}", flush=True)
    assert dev >= 0.85 * orc
    assert dev >= 10 * rnd


def test_device_forest_depth_limit():
    """max_rptree_depth = 3: three levels of splits, at most 8 leaves per tree, leaves
    larger than leaf_size, still a partition; the array is as wide as the widest leaf."""
    n = 5000
    x = _data("f32", n)
    leaf, per_tree, depths = E.make_forest(x, E.KnnGraphOptions(n_neighbors=15, max_rptree_depth=3), TREES)
    assert (depths == 3).all() and (per_tree <= 8).all()
    trees = _split_trees(leaf, per_tree)
    widest = 0
    for tree in trees:
        _check_partition(tree, n)
        widest = max(widest, int((tree >= 0).sum(1).max()))
    assert widest > 15 and leaf.shape[1] == widest


def test_device_forest_small_and_errors():
    """n <= leaf_size: one leaf per tree holding every row, depth 0 (float.rs:97-126);
    zero trees: an empty array; bad arguments are errors."""
    x = _data("f32", 10)  # leaf_size = max(10, k) = 10
    leaf, per_tree, depths = E.make_forest(x, E.KnnGraphOptions(n_neighbors=5), TREES)
    assert (per_tree == 1).all() and (depths == 0).all() and leaf.shape == (4, 10)
    for row in leaf:
        assert sorted(row.tolist()) == list(range(10))
    leaf0, per0, _ = E.make_forest(x, E.KnnGraphOptions(n_neighbors=5), np.zeros((0, 3), np.int64))
    assert leaf0.shape[0] == 0 and len(per0) == 0
    with pytest.raises(E.KnnError):
        E.make_forest(x, E.KnnGraphOptions(n_neighbors=0), TREES)
