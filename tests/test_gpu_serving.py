"""Serving loop (evoc_mi355x_server_*) on the MI355X, through the C ABI.

The contract: every request returns exactly what evoc_mi355x_knn_graph_seeded
returns for it alone (exact mode is deterministic, so bit-identical), whatever
else runs beside it; requests are admitted by HBM budget; a small request is
not held behind a large one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E = pytest.importorskip("evoc_mi355x")
from evoc_mi355x import synth  # noqa: E402


def _requests():
    rs = np.random.default_rng(3)
    reqs = []
    for i in range(14):
        kind = ["f32", "f32", "i8", "u8"][i % 4]
        n = int(rs.integers(300, 9000))
        k = int(rs.choice([10, 15, 30]))
        if kind == "f32":
            d = int(rs.choice([64, 384, 768]))
            x = synth.blobs_f32(n, d, int(rs.integers(10, 60)), 0.3, seed=100 + i)
            prec = E.PREC_BF16_RESCORE if i % 2 == 0 else E.PREC_FP32
        elif kind == "i8":
            x = synth.blobs_i8(n, 256, 20, 0.25, seed=100 + i)
            prec = E.PREC_FP32
        else:
            x = synth.bits_u8(n, 512, 30, 0.3, 0.08, seed=100 + i)
            prec = E.PREC_FP32
        reqs.append((x, E.KnnGraphOptions(n_neighbors=k, precision=prec), 1000 + i))
    return reqs


def test_results_equal_alone_and_in_the_mix():
    reqs = _requests()
    print(f"[gpu] serving: {len(reqs)} requests alone", flush=True)
    ctx = E.Context(0)
    alone = [E.knn_graph(x, o, E.NumpyRandomState(seed), ctx=ctx) for x, o, seed in reqs]
    ctx.close()
    print("[gpu] serving: same requests in the mix", flush=True)
    srv = E.Server(0, slots=4)
    tickets = [srv.submit(x, o, E.NumpyRandomState(seed)) for x, o, seed in reqs]
    mixed = [t.wait(timeout=120.0) for t in tickets]
    srv.close()
    for (ia, da), (im, dm) in zip(alone, mixed):
        np.testing.assert_array_equal(ia, im)
        np.testing.assert_array_equal(da, dm)


def test_rng_draws_match_knn_graph():
    """submit(rng=...) takes the same draws from the caller's RNG as knn_graph."""
    x = synth.blobs_f32(1500, 32, 5, 0.3, seed=1)
    a, b = E.NumpyRandomState(7), E.NumpyRandomState(7)
    E.knn_graph(x, E.KnnGraphOptions(n_neighbors=10), a)
    srv = E.Server(0, slots=2)
    srv.submit(x, E.KnnGraphOptions(n_neighbors=10), b).wait(timeout=60.0)
    srv.close()
    assert np.array_equal(a.get_state()[0], b.get_state()[0]) and a.get_state()[1] == b.get_state()[1]


def test_admission_by_budget():
    """A budget that holds about one request at a time still completes every request,
    never grants more than the budget, and refuses a request that can never fit."""
    xs = [synth.blobs_f32(6000, 256, 20, 0.3, seed=s) for s in range(6)]
    o = E.KnnGraphOptions(n_neighbors=15)
    probe = E.Server(0, slots=1)
    full, low = probe.request_bytes(xs[0], o)
    probe.close()
    budget = int(full * 1.5)
    print(f"[gpu] admission: request {full / 2**20:.1f} MiB (min {low / 2**20:.1f}), "
          f"budget {budget / 2**20:.1f} MiB", flush=True)
    srv = E.Server(0, slots=4, hbm_budget_bytes=budget)
    tickets = [srv.submit(x, o, E.NumpyRandomState(s)) for s, x in enumerate(xs)]
    seen = []
    while not all(t.poll() for t in tickets):
        st = srv.stats()
        assert st["granted_bytes"] <= st["budget_bytes"]
        seen.append(st["running"])
    res = [t.wait(timeout=60.0) for t in tickets]
    ctx = E.Context(0)
    for s, (x, r) in enumerate(zip(xs, res)):
        ia, da = E.knn_graph(x, o, E.NumpyRandomState(s), ctx=ctx)
        np.testing.assert_array_equal(ia, r[0])
    ctx.close()
    big = synth.blobs_f32(200000, 256, 20, 0.3, seed=9)
    probe = E.Server(0, slots=1)
    _, big_low = probe.request_bytes(big, o)
    probe.close()
    assert big_low > budget  # the request can never fit, even in the smallest slabs
    with pytest.raises(E.BackendError) as ei:
        srv.submit(big, o, E.NumpyRandomState(0))
    assert ei.value.code == E.ERR_OUT_OF_BUDGET
    srv.close()


def test_small_request_not_held_behind_large():
    torch = pytest.importorskip("torch")
    big = torch.from_numpy(synth.blobs_f32(100000, 768, 50, 0.25, seed=42)).cuda()
    small = [synth.blobs_f32(2000, 128, 10, 0.3, seed=s) for s in range(4)]
    n, k = big.shape[0], 30
    bi = torch.empty((n, k), dtype=torch.int32, device="cuda")
    bd = torch.empty((n, k), dtype=torch.float32, device="cuda")
    srv = E.Server(0, slots=4)
    # warm every slot (first calls pay for pool growth and code-object loads)
    for t in [srv.submit(x, E.KnnGraphOptions(n_neighbors=15), E.NumpyRandomState(2)) for x in small * 3]:
        t.wait(timeout=60.0)
    tb = srv.submit(big, E.KnnGraphOptions(n_neighbors=k), E.NumpyRandomState(0), out=(bi, bd))
    ts = [srv.submit(x, E.KnnGraphOptions(n_neighbors=15), E.NumpyRandomState(1)) for x in small]
    for t in ts:
        t.wait(timeout=60.0)
    big_done_when_small_done = tb.poll()
    tb.wait(timeout=120.0)
    srv.close()
    print(f"[gpu] large request done before the small ones: {big_done_when_small_done}; "
          f"small latencies {[round(t.stats.seconds * 1e3, 2) for t in ts]} ms "
          f"(queued {[round(t.stats.queue_seconds * 1e3, 2) for t in ts]} ms), "
          f"large {tb.stats.seconds * 1e3:.1f} ms", flush=True)
    assert not big_done_when_small_done
    ctx = E.Context(0)
    ri, rd = E.knn_graph(big, E.KnnGraphOptions(n_neighbors=k), E.NumpyRandomState(0), ctx=ctx)
    ctx.close()
    assert torch.equal(bi.cpu(), torch.from_numpy(ri))


def test_submit_errors():
    srv = E.Server(0, slots=1)
    with pytest.raises(E.InvalidNeighbors):
        srv.submit(np.ones((10, 4), np.float32), E.KnnGraphOptions(n_neighbors=0), E.NumpyRandomState(0))
    with pytest.raises(E.EmptyData):
        srv.submit(np.ones((0, 4), np.float32), E.KnnGraphOptions(n_neighbors=3), E.NumpyRandomState(0))
    t = srv.submit(np.ones((10, 4), np.float32), E.KnnGraphOptions(n_neighbors=3, strict_precision=True),
                   E.NumpyRandomState(0))
    with pytest.raises(E.BackendError) as ei:
        t.wait(timeout=30.0)
    assert ei.value.code == E.ERR_NOT_BITWISE
    srv.close()


def test_c5_sized_requests_equal_alone_and_in_the_mix():
    """Requests at the serving workload's sizes (n up to 200,000, d in {384, 768, 1024},
    k in {10, 15, 30}, BF16 + FP32 re-score) submitted together to four slots: the large
    ones run in query slabs that shrink while others wait or run, and every result still
    equals the same request run alone, bit for bit."""
    torch = pytest.importorskip("torch")
    specs = [(200000, 1024, 30, 10), (150000, 768, 15, 32), (60000, 1024, 10, 100), (3000, 384, 30, 10),
             (20000, 1024, 15, 32), (1000, 768, 10, 100), (90000, 384, 30, 32), (7000, 1024, 30, 10)]
    xs = [synth.blobs_f32_device(n, d, nb, 0.25, seed=300 + i) for i, (n, d, k, nb) in enumerate(specs)]
    print(f"[gpu] serving c5: {len(specs)} requests alone", flush=True)
    ctx = E.Context(0)
    alone = []
    for (n, d, k, nb), x in zip(specs, xs):
        out = (torch.empty((n, k), dtype=torch.int32, device="cuda"), torch.empty((n, k), dtype=torch.float32, device="cuda"))
        E.knn_graph(x, E.KnnGraphOptions(n_neighbors=k), E.NumpyRandomState(5), ctx=ctx, out=out)
        alone.append(out)
    ctx.close()
    print("[gpu] serving c5: same requests in the mix", flush=True)
    srv = E.Server(0, slots=4)
    tickets = []
    for (n, d, k, nb), x in zip(specs, xs):
        out = (torch.empty((n, k), dtype=torch.int32, device="cuda"), torch.empty((n, k), dtype=torch.float32, device="cuda"))
        tickets.append((srv.submit(x, E.KnnGraphOptions(n_neighbors=k), E.NumpyRandomState(5), out=out), out))
    for t, _ in tickets:
        t.wait(timeout=300.0)
    srv.close()
    slabs = [t.stats.slabs for t, _ in tickets]
    print(f"[gpu] serving c5: slabs per request {slabs}", flush=True)
    # the 200k x 1024 request is cut into time-bounded slabs, smaller while others run
    assert slabs[0] >= 8
    for (a_i, a_d), (t, (m_i, m_d)) in zip(alone, tickets):
        assert torch.equal(a_i, m_i)
        assert torch.equal(a_d, m_d)
