"""Continuous batching of small requests in the serving loop (the batch lane:
evoc_mi355x_server_set_batching), through the C ABI.

Small exact-mode requests (f32 BF16 + re-score, f32 FP32, int8) queued together
run as one batch: each pass is one persistent launch over all of them, replayed
from a hipGraph captured once per element kind.  Contract: every request's graph
is the one it gets alone, bit for bit; replays of the captured sequence (with
other work between them) keep giving those graphs; a request whose candidate
buffers overflow in the batch is handed back and still comes out exact."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E = pytest.importorskip("evoc_mi355x")
torch = pytest.importorskip("torch")
from evoc_mi355x import synth  # noqa: E402


def _requests(seed, count=24):
    rs = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        kind = ["bf16", "bf16", "fp32", "i8"][i % 4]
        n = int(rs.integers(300, 16000))
        k = int(rs.choice([10, 15, 30]))
        on_device = bool(i % 3 == 0)
        if kind == "i8":
            x = synth.blobs_i8(n, int(rs.choice([64, 256, 512])), int(rs.integers(10, 60)), 0.25, seed=seed * 100 + i)
            prec = E.PREC_FP32
        else:
            x = synth.blobs_f32(n, int(rs.choice([64, 384, 768, 1024])), int(rs.integers(10, 60)), 0.3,
                                seed=seed * 100 + i)
            prec = E.PREC_BF16_RESCORE if kind == "bf16" else E.PREC_FP32
        if on_device:
            x = torch.from_numpy(x).cuda()
        reqs.append((x, E.KnnGraphOptions(n_neighbors=k, precision=prec), 500 + i))
    return reqs


def _alone(reqs):
    ctx = E.Context(0)
    out = []
    for x, o, seed in reqs:
        a = E.knn_graph(x.cpu().numpy() if torch.is_tensor(x) else x, o, E.NumpyRandomState(seed), ctx=ctx)
        out.append(a)
    ctx.close()
    return out


def _served(srv, reqs):
    tickets = []
    for x, o, seed in reqs:
        if torch.is_tensor(x):
            n = x.shape[0]
            k = E.effective_k(o.n_neighbors, n)
            out = (torch.empty((n, k), dtype=torch.int32, device="cuda"),
                   torch.empty((n, k), dtype=torch.float32, device="cuda"))
            tickets.append(srv.submit(x, o, E.NumpyRandomState(seed), out=out))
        else:
            tickets.append(srv.submit(x, o, E.NumpyRandomState(seed)))
    res = []
    for t in tickets:
        r = t.wait(timeout=120.0)
        res.append(tuple(v.cpu().numpy() if torch.is_tensor(v) else v for v in r))
    return res


def _check(alone, served):
    for i, ((ia, da), (ib, db)) in enumerate(zip(alone, served)):
        np.testing.assert_array_equal(ia, ib, err_msg=f"request {i} indices")
        np.testing.assert_array_equal(da, db, err_msg=f"request {i} distances")


def test_batched_results_equal_alone_over_graph_replays():
    reqs = _requests(1)
    print(f"[gpu] batching: {len(reqs)} requests alone", flush=True)
    alone = _alone(reqs)
    srv = E.Server(0, slots=4)
    for rnd in range(3):  # each round's batches replay the captured sequences
        print(f"[gpu] batching: round {rnd}", flush=True)
        _check(alone, _served(srv, reqs))
        # the program's own device work between replays
        y = torch.randn((4096, 512), device="cuda")
        torch.cuda.synchronize()
        del y
    st = srv.batch_stats()
    srv.close()
    print(f"[gpu] batching: {st}", flush=True)
    assert st["requests"] >= 3 * len(reqs) and st["largest"] >= 2
    assert st["graph_replays"] == st["batches"]  # every batch replayed its kind's captured graph


def test_batched_eager_launches_equal_alone(monkeypatch):
    """The same batches launched pass by pass instead of replayed from the graph."""
    monkeypatch.setenv("EVOC_MI355X_BATCH_GRAPH", "0")
    reqs = _requests(2, count=12)
    alone = _alone(reqs)
    srv = E.Server(0, slots=2)
    _check(alone, _served(srv, reqs))
    st = srv.batch_stats()
    srv.close()
    assert st["requests"] >= len(reqs) and st["graph_replays"] == 0


def test_overflowing_request_is_handed_back():
    """30 distinct rows repeated 100 times: every candidate list overflows in the batch;
    the request runs alone instead and is exact; its neighbours in the batch are not
    affected."""
    base = synth.blobs_f32(30, 64, 5, 0.3, seed=3)
    ties = np.ascontiguousarray(np.tile(base, (100, 1)))
    reqs = [(ties, E.KnnGraphOptions(n_neighbors=15, precision=E.PREC_FP32), 7)] + _requests(3, count=6)
    alone = _alone(reqs)
    srv = E.Server(0, slots=2)
    _check(alone, _served(srv, reqs))
    srv.close()


def test_batching_off():
    reqs = _requests(4, count=6)
    alone = _alone(reqs)
    srv = E.Server(0, slots=2)
    srv.set_batching(0)
    _check(alone, _served(srv, reqs))
    assert srv.batch_stats()["requests"] == 0
    srv.close()
