"""kNN-graph benchmark on one MI355X.

Default workload (the metric's configuration): 1,000,000 x 768 f32 rows from 50
Gaussian blobs (centres N(0, I), std 0.25, Zipf(1.1) blob sizes, L2-normalised),
cosine k = 30.  A step is one knn_graph call over the HBM-resident matrix with
device-resident outputs, exact search: BF16 MFMA candidates + FP32 re-score of
k + 16, timed beside the plain FP32-MFMA form.  Recall@30 and the distance
error are checked against FP64 on sampled rows before timing.  The same line
carries the 100,000 x 768 run of the same generator, the serving stream and the
CPU baseline.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "evoc-rs_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "kNN-graph points/sec at recall@k≥0.99 (BF16+rescore vs FP32); p50/p99 latency"
PEAK = {"bf16": 2500.0, "fp32": 157.3, "int8": 5000.0}  # dense TFLOP/s (TOP/s for int8)

# HBM bytes per scoring launch measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, each
# scaled by what the counter reads for a 1 GiB device copy; keyed by (n, d, kernel), with the
# profile summary they come from: the counters cannot be read from inside this process
MEASURED_TRAFFIC = {
    (100000, 768, "pass2_sym_kernel<bf16>"): (15.5e9, "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated "
                                                      "(profiles/pass2_sym_bf16_100k_768_traffic.md)"),
    (1000000, 768, "pass2_sym_kernel<bf16>"): (3.45e12, "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated "
                                                        "(profiles/pass2_sym_bf16_1m_768_traffic.md)"),
    # the NN-descent local join: the mean over a call's 23 launches
    (100000, 768, "nnd_join_mfma_kernel<bf16>"): (2.022e9, "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated, "
                                                           "mean per launch (profiles/nndescent_100k_768_k30_kernel_trace.md)"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def fp64_truth(torch, x, k, chunk=2048):
    """Exact top-k by FP64 similarity with self first (the NN-descent layout)."""
    n = x.shape[0]
    x64 = x.double()
    out = torch.empty((n, k), dtype=torch.int64, device=x.device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        sim = x64[s:e] @ x64.T
        sim[torch.arange(e - s, device=x.device), torch.arange(s, e, device=x.device)] = 2.0
        out[s:e] = torch.topk(sim, k, dim=1).indices
        del sim
    return out


def recall_gpu(torch, idx, truth):
    k = truth.shape[1]
    hits = 0
    for s in range(0, idx.shape[0], 8192):
        a = idx[s:s + 8192].long()
        b = truth[s:s + 8192]
        hits += (a.unsqueeze(2) == b.unsqueeze(1)).any(dim=2).sum().item()
    return hits / (idx.shape[0] * k)


def cpu_baseline(x_np, k, target_s, n_iters=8, delta=0.0, keep=None):
    """The reference's CPU path (oracle restatement): non-deterministic NN-descent,
    T host threads, n_iters iterations (delta = delta_improv as given; 0 runs them
    all), on the first m rows.  With `keep` (a dict) the last run's indices are
    stored in keep["indices"] (for recall against the same truth)."""
    import oracle as O

    host = host_cpus()
    cores = host["threads_used"]
    opts = O.KnnGraphOptions(n_neighbors=k, n_iters=n_iters, delta=delta, delta_improv=delta)

    def run(m):
        t0 = time.perf_counter()
        idx, _, _ = O.knn_graph(x_np[:m], opts, O.NumpyRandomState(42), n_threads=cores, workers=cores)
        if keep is not None:
            keep["indices"] = idx
        return time.perf_counter() - t0

    m = min(2000, x_np.shape[0])
    t = run(m)
    m2 = int(min(x_np.shape[0], max(m, m * target_s / max(t, 1e-3))))
    if m2 > m:
        m, t = m2, run(m2)
    return {"value": m / t, "unit": "points/s", "cores": cores, "kind": "port",
            "sample": f"first {m} rows of the workload, NN-descent k={k}, {n_iters} iterations (delta={delta}), "
                      f"{cores} worker threads, {t:.1f} s",
            "host": host}


def host_cpus():
    """The host's CPUs as lscpu reports them, the CPUs this process may run on, and the
    threads the CPU baseline uses: all of those, at most 16 (the host CPU share of one GPU
    on a shared box)."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("CPU(s)", "Model name", "Thread(s) per core", "Core(s) per socket", "Socket(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"lscpu_cpus": int(info["CPU(s)"]) if info.get("CPU(s)", "").isdigit() else None,
            "model": info.get("Model name"), "affinity_cpus": aff, "threads_used": max(1, min(16, aff))}


def serving_bench(torch, E, n_req, rate, slots, budget_gb, seed=42):
    """Poisson stream of knn_graph requests through the serving loop.

    Request i: n log-uniform in [1e3, 2e5], d in {384, 768, 1024}, k in {10, 15, 30},
    rows from Gaussian blobs (10-100 centres, std 0.25, L2-normalised), BF16 + FP32
    re-score, exact search.  Rows are slices of device-resident pools (one per d and
    blob count), outputs are device tensors: the graph inputs and results stay in
    HBM.  Latency = completion - arrival (queueing included).
    """
    import math

    import numpy as np

    rs = np.random.default_rng(seed)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    pool_n, blob_choices = 200000, (10, 32, 100)
    pools = {}
    for d in (384, 768, 1024):
        for nb in blob_choices:
            c = torch.randn((nb, d), generator=g, device="cuda")
            lab = torch.randint(0, nb, (pool_n,), generator=g, device="cuda")
            x = c[lab] + 0.25 * torch.randn((pool_n, d), generator=g, device="cuda")
            pools[(d, nb)] = torch.nn.functional.normalize(x, dim=1).contiguous()
    torch.cuda.synchronize()
    reqs = []
    for _ in range(n_req):
        n = int(round(math.exp(rs.uniform(math.log(1e3), math.log(2e5)))))
        d = int(rs.choice([384, 768, 1024]))
        k = int(rs.choice([10, 15, 30]))
        nb = blob_choices[int(np.argmin([abs(b - rs.integers(10, 101)) for b in blob_choices]))]
        off = int(rs.integers(0, pool_n - n + 1))
        reqs.append((n, d, k, nb, off))
    gaps = rs.exponential(1.0 / rate, size=n_req)
    srv = E.Server(0, slots=slots, hbm_budget_bytes=int(budget_gb * (1 << 30)))
    # warm the slots' pools and code paths
    for (n, d, k, nb, off) in reqs[:slots]:
        x = pools[(d, nb)][off:off + n]
        srv.submit(x, E.KnnGraphOptions(n_neighbors=k), tau_state=[1, 2, 3], tree_states=None,
                   out=(torch.empty((n, k), dtype=torch.int32, device="cuda"),
                        torch.empty((n, k), dtype=torch.float32, device="cuda"))).wait(timeout=120.0)
    tickets = []
    t0 = time.perf_counter()
    t_next = t0
    for i, (n, d, k, nb, off) in enumerate(reqs):
        t_next += gaps[i]
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        x = pools[(d, nb)][off:off + n]
        out = (torch.empty((n, k), dtype=torch.int32, device="cuda"),
               torch.empty((n, k), dtype=torch.float32, device="cuda"))
        t = srv.submit(x, E.KnnGraphOptions(n_neighbors=k), tau_state=[i + 1, 7, 11], out=out)
        t.arrival = time.perf_counter() - t0
        tickets.append(t)
        if i % 250 == 249:
            log(f"[bench] serving: {i + 1}/{n_req} submitted, {srv.stats()}")
    for t in tickets:
        t.wait(timeout=300.0)
    makespan = time.perf_counter() - t0
    batching = srv.batch_stats()
    srv.close()
    lat = np.array([t.stats.queue_seconds + t.stats.seconds for t in tickets]) * 1e3
    queue = np.array([t.stats.queue_seconds for t in tickets]) * 1e3
    flops = sum(2.0 * n * n * d for (n, d, k, nb, off) in reqs)
    pts = sum(n for (n, d, k, nb, off) in reqs)
    return {"requests": n_req, "rate_per_s": rate, "slots": slots, "hbm_budget_gb": budget_gb,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "p99_queue_ms": float(np.percentile(queue, 99)),
            "makespan_s": makespan, "requests_per_s": n_req / makespan, "points_per_s": pts / makespan,
            "scoring_tflops": flops / makespan / 1e12,
            "mfma_frac": flops / makespan / 1e12 / PEAK["bf16"],
            "batch_lane": {"batches": batching["batches"], "requests_batched": batching["requests"],
                           "largest_batch": batching["largest"], "graph_replays": batching["graph_replays"]},
            "workload": "Poisson arrivals; n log-uniform [1e3, 2e5]; d in {384, 768, 1024}; "
                        "k in {10, 15, 30}; Gaussian blobs (10/32/100 centres, std 0.25), L2-normalised, "
                        "device-resident; exact BF16 + FP32 re-score"}


def fuzzy_bench(torch, E, idx, dist, k, ctx, steps):
    """Fuzzy graph (symmetrized) of the n x k device kNN graph: device time per
    call with the kNN arrays and the CSR in HBM (and the wall time of one call
    that copies the CSR to the host), checked against the oracle's CSR of the
    same arrays (same layout, values within 1e-6), which is timed as the CPU
    baseline of this step (1 thread)."""
    import oracle as O

    n = idx.shape[0]
    g = E.neighbor_graph_matrix(float(k), idx, dist, True, ctx=ctx)
    # HBM-resident: kNN arrays in, CSR out on the device
    cap = 2 * n * k
    dout = (torch.empty(n + 1, dtype=torch.int64, device="cuda"), torch.empty(cap, dtype=torch.int32, device="cuda"),
            torch.empty(cap, dtype=torch.float32, device="cuda"))
    sts = []
    for _ in range(steps + 1):
        st = E.CallStats()
        gd = E.neighbor_graph_matrix(float(k), idx, dist, True, ctx=ctx, stats=st, out=dout)
        sts.append(st)
    sts = sts[1:]
    dev_ms = float(np.median([s.device_seconds for s in sts])) * 1e3
    wall_ms = float(np.median([s.seconds for s in sts])) * 1e3
    same_dev = bool(torch.equal(gd[0].cpu(), torch.from_numpy(g[0])) and torch.equal(gd[1].cpu(), torch.from_numpy(g[1]))
                    and torch.equal(gd[2].cpu(), torch.from_numpy(g[2])))
    host_st = E.CallStats()
    E.neighbor_graph_matrix(float(k), idx, dist, True, ctx=ctx, stats=host_st)
    ind_h, dst_h = idx.cpu().numpy(), dist.cpu().numpy()
    t0 = time.perf_counter()
    o = O.neighbor_graph_matrix(float(k), ind_h, dst_h, True)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    same = bool(np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1]))
    err = float(np.abs(g[2] - o[2]).max()) if same and g[2].size else None
    log(f"[bench] fuzzy graph: {dev_ms:.2f} ms device, {wall_ms:.2f} ms wall, host output "
        f"{host_st.seconds * 1e3:.2f} ms, oracle {cpu_ms:.0f} ms, layout equal {same}, max |delta| {err}")
    return {"device_ms": dev_ms, "wall_ms": wall_ms, "host_output_wall_ms": host_st.seconds * 1e3,
            "nnz": int(g[0][-1]), "n": int(n), "k": int(k), "device_output_equals_host_output": same_dev,
            "layout_equals_oracle": same, "max_abs_diff_vs_oracle": err, "oracle_ms_1_thread": cpu_ms,
            "what": "neighbor_graph_matrix (sigma/rho bisection, membership strengths, fuzzy union, CSR) "
                    "of the BF16 + re-score kNN graph above"}


def _unpack_bits(torch, xb):
    """[m, B] uint8 -> [m, 8B] float32 bits, MSB first (numpy.packbits order)."""
    shifts = torch.arange(7, -1, -1, device=xb.device, dtype=torch.int32)
    return ((xb.to(torch.int32).unsqueeze(2) >> shifts) & 1).reshape(xb.shape[0], -1).float()


def sampled_exact(torch, x, kind, sample, k, chunk=65536):
    """Exact top-k of the sampled rows under the library's ranking: score desc,
    self first (self scores 1.0, the reference's NN-descent seed), lower index
    first.  int8: exact integer dots in FP64; binary: exact counts, IEEE f32
    inter / union (FP64 quotient rounded once to f32: exact for counts <= 2^24).
    Returns (indices [m, k] int64, scores [m, k] float64)."""
    n = x.shape[0]
    m = sample.shape[0]
    lo_bits = (1 << 21) - 1
    if kind == "i8":
        xs = x[sample].double()
    else:
        xs = _unpack_bits(torch, x[sample])
        ps = xs.sum(dim=1)
    best = None
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        idx = torch.arange(s, e, device=x.device)
        if kind == "i8":
            sc = xs @ x[s:e].double().T
            score_key = sc.long() + (1 << 24)
        else:
            xc = _unpack_bits(torch, x[s:e])
            inter = xs @ xc.T
            union = ps[:, None] + xc.sum(dim=1)[None, :] - inter
            q = torch.where(union > 0, inter.double() / union.double().clamp(min=1), torch.zeros_like(inter, dtype=torch.float64))
            score_key = q.float().view(torch.int32).long()
        selfm = sample[:, None] == idx[None, :]
        one = (1 << 24) + 1 if kind == "i8" else int(np.float32(1.0).view(np.int32))
        score_key = torch.where(selfm, torch.full_like(score_key, one), score_key)
        slot = torch.where(selfm, torch.zeros_like(idx)[None, :].expand(m, -1), (idx + 1)[None, :].expand(m, -1))
        key = (score_key << 21) | (lo_bits - slot)
        top = torch.topk(key, k, dim=1).values
        best = top if best is None else torch.topk(torch.cat([best, top], dim=1), k, dim=1).values
    slot = lo_bits - (best & lo_bits)
    ids = torch.where(slot == 0, sample[:, None].expand(-1, k), slot - 1)
    sk = best >> 21
    if kind == "i8":
        scores = (sk - (1 << 24)).double()
    else:
        scores = sk.int().view(torch.float32).double()
    return ids, scores


def _timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn(None)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        import evoc_mi355x as E
        st = E.CallStats()
        fn(st)
        stats.append(st)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, stats


def _scored(stats, n, kind_tag):
    """Row pairs one scoring launch scored (the library reports them: n (n + 1) / 2 in the
    symmetric pass, which scores each unordered pair once, n^2 in the per-split pass) and
    the kernel's name."""
    pairs = float(np.mean([s_.score_pairs for s_ in stats])) or float(n) * n
    name = "pass2_sym_kernel" if pairs < float(n) * n else "pass2_append_kernel"
    return pairs, f"{name}<{kind_tag}>"


def _base_line(args, metric_value, sec, workload, dtype, extra):
    out = {"metric": METRIC, "value": metric_value, "unit": "points/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": workload}
    out.update(extra)
    return out


def run_million(args, torch, E, kind):
    """C3 (1,000,000 x 512 int8, k = 15) or C4 (1,000,000 x 1024 bits, k = 15): exact
    search, checked bit-exact against a GPU brute force on 256 sampled rows."""
    from evoc_mi355x import synth

    n = args.n if args.n != 100000 else 1000000
    k = 15
    if kind == "i8":
        d = 512
        log(f"[bench] generating {n}x{d} int8 blobs on the GPU")
        x = synth.blobs_i8_device(n, d, 200, 0.2, seed=42)
        desc = (f"{n}x{d} int8 (200 Gaussian blobs, std 0.2, L2-normalised, x127, rounded), "
                f"inner-product kNN k={k}, exact i8 MFMA into i32")
        dtype, peak, unit, per_pair = "int8 (i32 accumulate)", PEAK["int8"], "TOP/s", 2.0 * d
    else:
        nbytes = 128
        log(f"[bench] generating {n}x{nbytes * 8}-bit rows on the GPU")
        x = synth.bits_u8_device(n, nbytes * 8, 500, 0.3, 0.08, seed=42)
        desc = (f"{n}x{nbytes * 8}-bit rows (500 Bernoulli(0.3) prototypes, bit flips p=0.08), "
                f"bitwise Jaccard k={k}, VALU popcount")
        # AND + popcount-accumulate per 32-bit word; peak = 256 CU x 64 lanes x 2.4 GHz int32 ops
        dtype, peak, unit, per_pair = "u8 bits (i32 popcount)", 256 * 64 * 2.4e9 / 1e12, "TOP/s", 2.0 * (nbytes / 4)
    ctx = E.Context(0)
    idx = torch.empty((n, k), dtype=torch.int32, device="cuda")
    dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
    opts = E.KnnGraphOptions(n_neighbors=k)
    if kind == "i8":
        log(f"[bench] autotune: {E.autotune(x, opts, ctx=ctx)}")

    def step(st):
        E.knn_graph(x, opts, E.NumpyRandomState(42), ctx=ctx, stats=st, out=(idx, dist))

    step(None)
    torch.cuda.synchronize()
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    sample = torch.randperm(n, generator=g, device="cuda")[:256]
    log("[bench] sampled exact check")
    ti, _ = sampled_exact(torch, x, kind, sample, k)
    exact_rows = int((idx[sample].long() == ti).all(dim=1).sum().item())
    log(f"[bench] {exact_rows}/256 sampled rows bit-exact")
    sec, stats = _timed(torch, step, args.steps, args.warmup)
    score_s = float(np.mean([s.score_seconds for s in stats]))
    pairs, kern = _scored(stats, n, "i8")
    work = per_pair * pairs
    achieved = work / score_s / 1e12
    res = _base_line(args, n / sec, sec, {"workload": desc, "n": n, "k": k, "method": "exact"}, dtype, {
        "sampled_rows_bit_exact": f"{exact_rows}/256",
        "rescore_rounds": max(s_.rescore_rounds for s_ in stats),
        "roofline": {"bound": "mfma" if kind == "i8" else "valu", "kernel":
                     kern if kind == "i8" else ("jac_pass2_sym_kernel" if pairs < float(n) * n else "jac_pass2_kernel"),
                     "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak, "traffic": None,
                     "algorithmic_ops_per_launch": work, "avg_launch_ms": score_s * 1e3},
    })
    if not args.no_cpu_baseline:
        # the reference's defaults at 1M rows: 20 iterations, delta = 0.001
        xs = x[: min(n, 200000)].cpu().numpy()
        res["cpu_baseline"] = cpu_baseline(xs, k, args.cpu_seconds, n_iters=20, delta=0.001)
    ctx.close()
    return res


def fp64_sampled_truth(torch, x, sample, k, chunk=131072):
    """Exact top-k of the sampled rows by FP64 cosine over all n rows, self first (the
    layout the output lists).  Returns indices [m, k] int64."""
    n = x.shape[0]
    xs = x[sample].double()
    best = None
    for s0 in range(0, n, chunk):
        e0 = min(n, s0 + chunk)
        sim = xs @ x[s0:e0].double().T
        selfm = sample[:, None] == torch.arange(s0, e0, device=x.device)[None, :]
        sim = torch.where(selfm, torch.full_like(sim, 2.0), sim)
        v, i = torch.topk(sim, k, dim=1)
        i = i + s0
        if best is None:
            best = (v, i)
        else:
            top = torch.topk(torch.cat([best[0], v], 1), k, dim=1)
            best = (top.values, torch.gather(torch.cat([best[1], i], 1), 1, top.indices))
        del sim
    return best[1]


def distance_error(torch, x, sample, idx, dist, chunk=2048):
    """|reported distance - FP64 distance of the same pair| over the sampled rows' listed
    neighbours: the FP64 distance is max(-log2(dot), 0) of the L2-normalised rows in FP64
    (src/knn.rs:156-171), 0 for the query itself."""
    mx, tot, cnt = 0.0, 0.0, 0
    for s0 in range(0, sample.shape[0], chunk):
        sm = sample[s0:s0 + chunk]
        xs = x[sm].double()
        xs = xs / xs.norm(dim=1, keepdim=True)
        ids = idx[sm].long()
        nb = x[ids].double()
        nb = nb / nb.norm(dim=2, keepdim=True)
        dots = (nb * xs[:, None, :]).sum(dim=2)
        ref = torch.where(dots > 0, torch.clamp(-torch.log2(dots.clamp(min=1e-300)), min=0.0),
                          torch.zeros_like(dots))
        ref = torch.where(ids == sm[:, None], torch.zeros_like(ref), ref)
        err = (dist[sm].double() - ref).abs()
        mx = max(mx, float(err.max().item()))
        tot += float(err.sum().item())
        cnt += int(err.numel())
    return {"max_abs": mx, "mean_abs": tot / max(cnt, 1), "pairs": cnt}


def exact_f32_leg(args, torch, E, x, sample, steps, fp32_steps):
    """Exact search of device rows x, k = args.k: BF16 + FP32 re-score (timed over `steps`
    calls) and plain FP32 (`fp32_steps` calls, none: skipped); recall@k with and without
    the self column and the distance error against FP64 on the sampled rows."""
    n, d = x.shape
    k = args.k
    truth = fp64_sampled_truth(torch, x, sample, k)
    ctx = E.Context(0)
    idx = torch.empty((n, k), dtype=torch.int32, device="cuda")
    dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
    out = {}
    runs = [("bf16", E.PREC_BF16_RESCORE, steps)] + ([("fp32", E.PREC_FP32, fp32_steps)] if fp32_steps else [])
    for name, prec, nsteps in runs:
        opts = E.KnnGraphOptions(n_neighbors=k, precision=prec, margin=args.margin)

        def step(st, opts=opts):
            E.knn_graph(x, opts, E.NumpyRandomState(42), ctx=ctx, stats=st, out=(idx, dist))

        step(None)
        torch.cuda.synchronize()
        got = idx[sample].long()
        hit = (got.unsqueeze(2) == truth.unsqueeze(1)).any(dim=2)
        rec = float(hit.float().mean().item())
        rec_ns = float(hit[:, 1:].float().mean().item()) if k > 1 else rec
        derr = distance_error(torch, x, sample, idx, dist)
        # the checked call above is the first warm-up step
        sec, stats = _timed(torch, step, nsteps, max(0, args.warmup - 1))
        score_s = float(np.mean([s_.score_seconds for s_ in stats]))
        pairs, kern = _scored(stats, n, name)
        flops = 2.0 * pairs * d
        out[name] = {"points_per_s": n / sec, "ms_per_step": sec * 1e3,
                     "device_ms_per_call": float(np.median([s_.device_seconds for s_ in stats])) * 1e3,
                     "recall_at_k": rec, "recall_at_k_without_self": rec_ns, "recall_rows": int(sample.shape[0]),
                     "distance_error_vs_fp64": derr,
                     "kernel": kern, "flop_per_launch": flops,
                     "rescore_rounds": max(s_.rescore_rounds for s_ in stats),
                     "score_tflops": flops / score_s / 1e12,
                     "score_frac": flops / score_s / 1e12 / (PEAK["bf16"] if name == "bf16" else PEAK["fp32"]),
                     "score_ms": score_s * 1e3, "steps": nsteps}
        log(f"[bench] {n} rows {name}: {out[name]}")
    return out, ctx, idx, dist


def roofline_of(leg, n, d):
    traffic, src = MEASURED_TRAFFIC.get((n, d, leg["kernel"]), (None, None))
    return {"bound": "mfma", "kernel": leg["kernel"], "achieved": leg["score_tflops"], "peak": PEAK["bf16"],
            "unit": "TFLOP/s", "frac": leg["score_frac"], "traffic": traffic, "traffic_source": src,
            "algorithmic_flop_per_launch": leg["flop_per_launch"], "avg_launch_ms": leg["score_ms"],
            "avg_launch_ms_source": "HIP events around the scoring launch on the library's stream"}


def run_million_f32(args, torch, E):
    """1,000,000 x 768 f32 (50 Gaussian blobs, std 0.25, Zipf 1.1, L2-normalised),
    cosine k = 30, exact: BF16 MFMA + FP32 re-score of k + 16 against plain FP32
    MFMA; recall@30 and the distance error against FP64 on 1,000 sampled rows."""
    from evoc_mi355x import synth

    n = args.n
    d, k = args.d, args.k
    log(f"[bench] generating {n}x{d} f32 blobs on the GPU")
    x = synth.blobs_f32_device(n, d, 50, 0.25, seed=42, zipf=1.1)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    sample = torch.randperm(n, generator=g, device="cuda")[:1000]
    out, ctx, idx, dist = exact_f32_leg(args, torch, E, x, sample, args.steps, 0 if args.no_fp32 else 1)
    ctx.close()
    del idx, dist, x
    torch.cuda.empty_cache()
    b = out["bf16"]
    f32 = out.get("fp32")
    return _base_line(args, b["points_per_s"], b["ms_per_step"] / 1e3,
                      {"workload": f"{n}x{d} f32 L2-normalised Gaussian blobs (50 centres, std 0.25, Zipf 1.1), "
                                   f"cosine kNN k={k}, exact BF16-MFMA candidates + FP32 re-score of k+{args.margin}",
                       "n": n, "d": d, "k": k, "method": "exact_bf16_rescore"},
                      "bf16 (fp32 accumulate) + fp32 rescore",
                      {"recall_at_k": b["recall_at_k"], "recall_at_k_without_self": b["recall_at_k_without_self"],
                       "recall_sample_rows": int(sample.shape[0]),
                       "distance_error_vs_fp64": b["distance_error_vs_fp64"],
                       "device_ms_per_call": b["device_ms_per_call"],
                       "roofline": roofline_of(b, n, d), "bf16": b,
                       "fp32": f32, "bf16_over_fp32": b["points_per_s"] / f32["points_per_s"] if f32 else None})


def run_c2_exact(args, torch, E, n=100000):
    """100,000 x 768 of the same generator through the same exact search (recall over all
    rows), and the fuzzy graph of its kNN graph against the oracle."""
    from evoc_mi355x import synth

    d, k = args.d, args.k
    log(f"[bench] generating {n}x{d} f32 blobs on the GPU")
    x = synth.blobs_f32_device(n, d, 50, 0.25, seed=42, zipf=1.1)
    sample = torch.arange(n, device="cuda")
    out, ctx, idx, dist = exact_f32_leg(args, torch, E, x, sample, max(args.steps, 10), 0 if args.no_fp32 else 3)
    b = out["bf16"]
    res = {"workload": f"{n}x{d} f32 (same generator), cosine k={k}, exact BF16 + FP32 re-score of k+{args.margin}",
           "points_per_s": b["points_per_s"], "ms_per_step": b["ms_per_step"], "recall_at_k": b["recall_at_k"],
           "roofline": roofline_of(b, n, d), "bf16": b, "fp32": out.get("fp32")}
    E.knn_graph(x, E.KnnGraphOptions(n_neighbors=k, margin=args.margin), E.NumpyRandomState(42), ctx=ctx,
                out=(idx, dist))
    torch.cuda.synchronize()
    res["fuzzy_graph"] = fuzzy_bench(torch, E, idx, dist, k, ctx, args.steps)
    ctx.close()
    return res


def run_nndescent(args, torch, E, n=None, cpu=True):
    """C2 with the reference's algorithm: device NN-descent, 8 iterations,
    delta = delta_improv = 0, FP32 and BF16 + FP32 re-score, recall vs FP64
    (n rows, default --n; cpu: also time the CPU baseline on them)."""
    from evoc_mi355x import synth

    n = args.n if n is None else n
    d, k = args.d, args.k
    x = synth.blobs_f32_device(n, d, 50, 0.25, seed=42, zipf=1.1)
    truth = fp64_truth(torch, x, k)
    ctx = E.Context(0)
    idx = torch.empty((n, k), dtype=torch.int32, device="cuda")
    dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
    out = {}
    # the reference's options for the host threads its CPU run below uses (n_trees from them)
    try:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        threads = max(1, min(16, os.cpu_count() or 1))
    for name, prec in (("bf16", E.PREC_BF16_RESCORE), ("fp32", E.PREC_FP32)):
        opts = E.KnnGraphOptions(n_neighbors=k, n_iters=8, delta=0.0, delta_improv=0.0, mode=E.MODE_NNDESCENT,
                                 precision=prec, margin=args.margin, rayon_threads=threads)

        def step(st, opts=opts):
            E.knn_graph(x, opts, E.NumpyRandomState(42), ctx=ctx, stats=st, out=(idx, dist))

        step(None)
        torch.cuda.synchronize()
        rec = recall_gpu(torch, idx, truth)
        sec, stats = _timed(torch, step, args.steps, args.warmup)
        jb = float(np.mean([s_.join_bytes for s_ in stats]))
        js = float(np.mean([s_.join_seconds for s_ in stats]))
        jl = float(np.mean([s_.join_launches for s_ in stats]))
        out[name] = {"points_per_s": n / sec, "ms_per_step": sec * 1e3, "recall_at_k": rec,
                     "iterations": stats[-1].iterations, "offers_dropped": stats[-1].offers_dropped,
                     "join_bytes_per_call": jb, "join_ms_per_call": js * 1e3, "join_launches_per_call": jl,
                     "join_gb_per_s": jb / js / 1e9 if js > 0 else None}
        log(f"[bench] nndescent {name}: {out[name]}")
    ctx.close()
    best = out["bf16"]
    kern = "nnd_join_mfma_kernel<bf16>"
    traffic, tsrc = MEASURED_TRAFFIC.get((n, d, kern), (None, None))
    res = _base_line(args, best["points_per_s"], best["ms_per_step"] / 1e3,
                     {"workload": f"{n}x{d} f32 blobs (50 centres, std 0.25, Zipf 1.1), cosine k={k}, "
                                  f"device NN-descent 8 iterations (delta=0), BF16 + FP32 re-score of k+{args.margin}",
                      "n": n, "d": d, "k": k, "method": "nndescent"},
                     "bf16 (fp32 accumulate) + fp32 rescore", {"recall_at_k": best["recall_at_k"], "fp32": out["fp32"],
                                                               "bf16": out["bf16"]})
    # the local join is bound by its gathers of candidate rows: algorithmic bytes = each work
    # item's candidate rows once (counted by the kernel), over the join launches' event time
    per_launch = best["join_bytes_per_call"] / max(best["join_launches_per_call"], 1.0)
    res["roofline"] = {"bound": "hbm", "kernel": kern, "achieved": best["join_gb_per_s"], "peak": 8000.0,
                       "unit": "GB/s", "frac": (best["join_gb_per_s"] or 0.0) / 8000.0, "traffic": traffic,
                       "traffic_source": tsrc, "algorithmic_bytes_per_launch": per_launch,
                       "avg_launch_ms": best["join_ms_per_call"] / max(best["join_launches_per_call"], 1.0)}
    if cpu and not args.no_cpu_baseline:
        keep = {}
        res["cpu_baseline"] = cpu_baseline(x.cpu().numpy(), k, args.cpu_seconds, keep=keep)
        ref_idx = keep.get("indices")
        if ref_idx is not None and ref_idx.shape[0] == n:
            # the reference's NN-descent at the same iteration count, against the same FP64 truth
            res["reference_recall_at_k"] = recall_gpu(torch, torch.from_numpy(ref_idx).cuda(), truth)
            log(f"[bench] reference NN-descent recall@{k} at 8 iterations: {res['reference_recall_at_k']:.5f}")
    return res


def copy_calibration(torch):
    """One 1 GiB device-to-device copy (1 GiB read, 1 GiB written): the known traffic that
    scales FETCH_SIZE / WRITE_SIZE counter readings."""
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log("[bench] copied 1 GiB device to device")
    del src, dst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=None, help="rows (default: the workload's)")
    ap.add_argument("--d", type=int, default=768)
    ap.add_argument("--k", type=int, default=30)
    ap.add_argument("--margin", type=int, default=16)
    ap.add_argument("--no-fp32", action="store_true", help="skip timing the plain FP32 form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="1m: skip the 100k line")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", default="1m", choices=["1m", "c2", "c2-nndescent", "c3", "c4"],
                    help="1m: 1M x 768 f32 exact, BF16 + re-score vs FP32, with the 100k line, serving "
                         "stream and CPU baseline (default); c2: 100k x 768 f32 exact only; c2-nndescent: "
                         "100k x 768 through device NN-descent; c3: 1M x 512 int8; c4: 1M x 1024-bit binary")
    ap.add_argument("--serving-requests", type=int, default=2000,
                    help="requests in the serving stream (0: skip the serving leg)")
    ap.add_argument("--serving-rate", type=float, default=60.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy (known traffic, for scaling "
                         "FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for --gpus "
              f"{args.gpus}). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.n is None:
        args.n = 1000000 if args.workload in ("1m", "c3", "c4") else 100000

    import torch

    import evoc_mi355x as E

    if args.workload in ("c3", "c4"):
        print(json.dumps(run_million(args, torch, E, "i8" if args.workload == "c3" else "u8")), flush=True)
        return
    if args.workload == "c2-nndescent":
        res = run_nndescent(args, torch, E)
        if args.copy_calibration:
            copy_calibration(torch)
        print(json.dumps(res), flush=True)
        return
    if args.workload == "c2":
        r = run_c2_exact(args, torch, E, args.n)
        b = r["bf16"]
        res = _base_line(args, b["points_per_s"], b["ms_per_step"] / 1e3,
                         {"workload": r["workload"], "n": args.n, "d": args.d, "k": args.k,
                          "method": "exact_bf16_rescore"}, "bf16 (fp32 accumulate) + fp32 rescore", r)
        print(json.dumps(res), flush=True)
        return

    result = run_million_f32(args, torch, E)
    if not args.no_extra:
        log("[bench] 100k x 768 line")
        result["c2_100k"] = run_c2_exact(args, torch, E, 100000)
        log("[bench] 100k x 768 device NN-descent line")
        nd = run_nndescent(args, torch, E, n=100000, cpu=False)
        result["c2_nndescent"] = {"workload": nd["config"]["workload"], "points_per_s": nd["value"],
                                  "ms_per_step": nd["ms_per_step"], "recall_at_k": nd["recall_at_k"],
                                  "roofline": nd["roofline"], "bf16": nd["bf16"], "fp32": nd["fp32"]}
    if args.serving_requests > 0:
        log(f"[bench] serving: {args.serving_requests} requests at {args.serving_rate}/s")
        result["serving"] = serving_bench(torch, E, args.serving_requests, args.serving_rate,
                                          args.serving_slots, 200.0)
        result["latency_ms"] = {"p50": result["serving"]["p50_ms"], "p99": result["serving"]["p99_ms"],
                                "source": "serving stream"}
    if args.copy_calibration:
        copy_calibration(torch)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        from evoc_mi355x import synth

        # the first 200k rows of the same workload (blob labels are shuffled, so they
        # sample every blob)
        xs = synth.blobs_f32_device(args.n, args.d, 50, 0.25, seed=42, zipf=1.1)[:200000].cpu().numpy()
        torch.cuda.empty_cache()
        result["cpu_baseline"] = cpu_baseline(xs, args.k, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
