"""The allocation-free entry point (evoc_mi355x_knn_graph_noalloc) on the MI355X,
through the C ABI: the same graph, bit for bit, as evoc_mi355x_knn_graph_seeded on
the same inputs, for two calls in a row on the shared per-GPU scratch (the second
smaller, so stale contents of the first are in the scratch), for each element kind
and both search modes; a call larger than the scratch runs in query slabs that fit
it; one that cannot fit at all is refused before any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E = pytest.importorskip("evoc_mi355x")
torch = pytest.importorskip("torch")
from evoc_mi355x import synth  # noqa: E402

TAU = [123456789, -987654321, 555555555]


def _pair(ctx, x, opts, trees=None):
    """(seeded, noalloc) outputs and stats of the same call."""
    a_st, b_st = E.CallStats(), E.CallStats()
    a = E.knn_graph_seeded(x, opts, TAU, trees, ctx=ctx, stats=a_st)
    b = E.knn_graph_seeded(x, opts, TAU, trees, ctx=ctx, stats=b_st, noalloc=True)
    return a, b, a_st, b_st


def _equal(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("kind", ["f32-bf16", "f32-fp32", "i8", "u8"])
def test_noalloc_equals_seeded_twice_in_a_row(kind):
    ctx = E.Context(0)
    for n in (6000, 2500):  # the second call smaller: the scratch holds the first's contents
        print(f"[gpu] noalloc {kind} n={n}", flush=True)
        if kind.startswith("f32"):
            x = synth.blobs_f32(n, 384, 20, 0.3, seed=n)
            prec = E.PREC_BF16_RESCORE if kind == "f32-bf16" else E.PREC_FP32
            opts = E.KnnGraphOptions(n_neighbors=15, precision=prec)
        elif kind == "i8":
            x = synth.blobs_i8(n, 256, 20, 0.25, seed=n)
            opts = E.KnnGraphOptions(n_neighbors=15)
        else:
            x = synth.bits_u8(n, 512, 20, 0.3, 0.08, seed=n)
            opts = E.KnnGraphOptions(n_neighbors=15)
        a, b, _, _ = _pair(ctx, x, opts)
        _equal(a, b)
    ctx.close()


def test_noalloc_nndescent_equals_seeded():
    ctx = E.Context(0)
    trees = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9], [10, 11, 12]], np.int64)
    for n in (5000, 3000):
        print(f"[gpu] noalloc nndescent n={n}", flush=True)
        x = synth.blobs_f32(n, 96, 20, 0.3, seed=n + 1)
        opts = E.KnnGraphOptions(n_neighbors=10, n_iters=4, delta=0.0, delta_improv=0.0,
                                 mode=E.MODE_NNDESCENT, precision=E.PREC_FP32)
        a, b, _, _ = _pair(ctx, x, opts, trees)
        _equal(a, b)
    ctx.close()


def test_noalloc_device_buffers_and_query_slabs():
    """30,000 x 768 f32 rows and outputs on the device: more than the scratch holds in
    one slab, so the search runs in several; still the seeded call's graph."""
    ctx = E.Context(0)
    n, d, k = 30000, 768, 30
    print(f"[gpu] noalloc slabs n={n}", flush=True)
    x = synth.blobs_f32_device(n, d, 30, 0.25, seed=3)
    opts = E.KnnGraphOptions(n_neighbors=k)
    outs, stats = [], []
    for noalloc in (False, True, True):
        o = (torch.empty((n, k), dtype=torch.int32, device="cuda"), torch.empty((n, k), dtype=torch.float32, device="cuda"))
        st = E.CallStats()
        E.knn_graph_seeded(x, opts, TAU, None, ctx=ctx, stats=st, out=o, noalloc=noalloc)
        outs.append(o)
        stats.append(st)
    print(f"[gpu] noalloc slabs: seeded {stats[0].slabs}, noalloc {stats[1].slabs}", flush=True)
    assert stats[1].slabs > 1
    for o in outs[1:]:
        assert torch.equal(outs[0][0], o[0]) and torch.equal(outs[0][1], o[1])
    ctx.close()


def test_noalloc_refuses_what_cannot_fit():
    ctx = E.Context(0)
    x = torch.zeros((300000, 768), dtype=torch.float32, device="cuda")  # 460 MB of BF16 rows alone
    full, minimum = E.request_bytes(x, E.KnnGraphOptions(n_neighbors=30), out_on_device=False, ctx=ctx)
    assert minimum > E.NOALLOC_SCRATCH_BYTES
    with pytest.raises(E.BackendError) as ei:
        E.knn_graph_seeded(x, E.KnnGraphOptions(n_neighbors=30), TAU, None, ctx=ctx, noalloc=True)
    assert ei.value.code == E.ERR_OUT_OF_BUDGET and "scratch" in str(ei.value)
    # the context still works afterwards
    y = synth.blobs_f32(1000, 64, 5, 0.3, seed=1)
    _equal(*_pair(ctx, y, E.KnnGraphOptions(n_neighbors=10))[:2])
    ctx.close()
