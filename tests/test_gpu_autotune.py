"""Tile-shape autotuner (evoc_mi355x_autotune) on the MI355X.

The autotuner runs the exact search once per MFMA tile shape and fails if any
shape's graph differs bit for bit from the first one's, so a passing call is
also the check that tile shapes change speed only.
"""
import numpy as np
import pytest

import oracle as O
from parity_util import boundary_mismatch_rows

pytestmark = pytest.mark.gpu

E = pytest.importorskip("evoc_mi355x")
from evoc_mi355x import synth  # noqa: E402


@pytest.mark.parametrize("prec_name,prec", [("bf16", E.PREC_BF16_RESCORE), ("fp32", E.PREC_FP32)])
def test_autotune_f32_shapes_agree(prec_name, prec):
    x = synth.blobs_f32(20000, 384, 40, 0.3, seed=21)
    r = E.autotune(x, E.KnnGraphOptions(n_neighbors=15, precision=prec))
    print(f"[gpu] autotune f32 {prec_name}: {r}", flush=True)
    assert r["shape"] in (0, 1, 2, 3) and all(t > 0 for t in r["ms"])


def test_autotune_int8_shapes_agree_and_tuned_search_is_exact():
    x = synth.blobs_i8(9000, 256, 20, 0.25, seed=22)
    r = E.autotune(x, E.KnnGraphOptions(n_neighbors=15))
    print(f"[gpu] autotune int8: {r}", flush=True)
    assert r["shape"] in (0, 1, 2, 3)
    # later searches in the bucket use the tuned shape and stay bit-exact
    idx, dist = E.knn_graph(x, E.KnnGraphOptions(n_neighbors=15), E.NumpyRandomState(0))
    ei, ed = O.exact_knn(x, 15, workers=8)
    np.testing.assert_array_equal(idx, ei)
    np.testing.assert_array_equal(dist, O.transform(O.I8, ed.astype(np.float32)))


def test_autotune_binary_has_one_engine():
    x = synth.bits_u8(3000, 512, 20, 0.3, 0.08, seed=23)
    assert E.autotune(x, E.KnnGraphOptions(n_neighbors=10))["shape"] == -1
