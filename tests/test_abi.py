"""C-ABI checks that need no GPU: the library loads, exports every function that
include/evoc_mi355x.h declares, and its host-side contract helpers behave like the
reference (options defaults, effective k, RNG draws, NumPy RandomState)."""
import os
import re
import subprocess

import numpy as np
import pytest

import evoc_mi355x as E
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "evoc_mi355x.h")


def declared_functions():
    with open(HEADER) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(evoc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    E.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", E.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (evoc_[a-z0-9_]+)", out))
    declared = declared_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    # the Python mirror binds every declared function
    assert set(declared) == set(E.SIGNATURES)


def test_library_is_gfx950_only():
    """The offload bundle ids embedded in the library name only gfx950."""
    with open(E.LIB_PATH, "rb") as f:
        blob = f.read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", blob))
    assert targets == {b"gfx950"}


def test_numpy_rng_matches_reference_kat():
    """src/numpy_rng.rs:232-252 through the ABI."""
    r = E.NumpyRandomState(42)
    lo, hi = E.INT32_MIN_P1, E.INT32_MAX_M1
    assert [r.randint(lo, hi) for _ in range(3)] == [-538846105, 1273642420, 1935803229]
    assert [r.randint(lo, hi) for _ in range(5)] == [-1359637233, 996406379, 1201263688, 423734973,
                                                      415968277]
    nrm = [r.normal() for _ in range(3)]
    np.testing.assert_allclose(nrm, [-0.23415337472333597, -0.23413695694918055, 1.5792128155073915],
                               atol=1e-12)


def test_numpy_rng_state_roundtrip_and_oracle_agreement():
    a = E.NumpyRandomState(7)
    b = O.NumpyRandomState(7)
    for _ in range(1000):  # crosses an MT19937 refill
        assert a.randint(0, 1 << 31) == b.randint(0, 1 << 31)
    key, pos, hg, g = a.get_state()
    c = E.NumpyRandomState(0)
    c.set_state(key, pos, hg, g)
    assert [c.randint(-5, 5) for _ in range(20)] == [a.randint(-5, 5) for _ in range(20)]


def test_options_default_mirror_reference():
    """KnnGraphOptions::default(), src/knn.rs:60-76."""
    o = E.KnnGraphOptions()._c()
    assert o.n_neighbors == 30 and o.max_rptree_depth == 200
    assert abs(o.delta - 0.001) < 1e-9 and abs(o.delta_improv - 0.001) < 1e-9
    assert o.use_sorted_updates == 1 and o.brute_force_max_n == 0 and o.deterministic == 0


@pytest.mark.parametrize("k,n,exp", [(15, 200, 15), (15, 1, 1), (15, 2, 1), (50, 5, 4), (1, 10, 1)])
def test_effective_k(k, n, exp):
    """src/knn.rs:402: min(k, n - 1).max(1)."""
    assert E.effective_k(k, n) == exp


def test_rng_draws_contract():
    """src/knn.rs:416 (3 draws) + src/knn.rs:124-129 (3 per tree); 3 on the brute-force branch."""
    assert E.rng_draws(E.DTYPE_F32, 1000, E.KnnGraphOptions(deterministic=True)) == 15
    assert E.rng_draws(E.DTYPE_F32, 1000, E.KnnGraphOptions(rayon_threads=6)) == 21
    assert E.rng_draws(E.DTYPE_F32, 1000, E.KnnGraphOptions(rayon_threads=32)) == 27
    assert E.rng_draws(E.DTYPE_F32, 1000, E.KnnGraphOptions(n_trees=2)) == 9
    bf = E.KnnGraphOptions(brute_force_max_n=5000, rayon_threads=1)
    assert E.rng_draws(E.DTYPE_I8, 1000, bf) == 3
    assert E.rng_draws(E.DTYPE_U8, 1000, bf) == 3
    assert E.rng_draws(E.DTYPE_F32, 1000, bf) == 15  # f32 never takes the brute-force branch


def test_errors_before_device_use():
    """Invalid arguments are reported before any GPU work (src/knn.rs:382-393)."""
    if E.device_count() > 0:
        pytest.skip("a device is visible; covered by the GPU tests")
    with pytest.raises(E.BackendError) as ei:
        E.Context(0)
    assert ei.value.code == E.ERR_NO_DEVICE


def test_rng_draws_default_threads_follow_the_host():
    """rayon_threads = 0 takes rayon's global-pool size: RAYON_NUM_THREADS, else the CPUs this
    process may run on (src/knn.rs:404-407 calls rayon::current_num_threads())."""
    if "RAYON_NUM_THREADS" in os.environ and int(os.environ["RAYON_NUM_THREADS"] or 0) > 0:
        t = int(os.environ["RAYON_NUM_THREADS"])
    else:
        t = len(os.sched_getaffinity(0))
    assert E.rng_draws(E.DTYPE_F32, 1000, E.KnnGraphOptions()) == 3 + 3 * min(max(t, 4), 8)
    # an explicit thread count wins
    assert E.rng_draws(E.DTYPE_F32, 1000, E.KnnGraphOptions(rayon_threads=5)) == 18


@pytest.mark.parametrize("dtype,d,prec", [(E.DTYPE_F32, 768, E.PREC_BF16_RESCORE), (E.DTYPE_F32, 768, E.PREC_FP32),
                                          (E.DTYPE_I8, 512, E.PREC_FP32)])
def test_symmetric_plan_keeps_selection_in_registers(dtype, d, prec):
    """The selection merges at most 64 key lists per query in registers; the symmetric plan
    never asks for more (its splits plus the candidate-side list), at mid sizes where it
    used to cut the matrix into one split per tile, and at the million-row workloads."""
    for n in (3000, 9000, 12000, 16383, 16384, 30000, 65536, 100000, 1000000):
        p = E.exact_plan(dtype, n, d, E.KnnGraphOptions(n_neighbors=30, precision=prec))
        assert p["segments"] <= 64, (n, p)
        if p["symmetric"]:
            assert p["segments"] == p["n_splits"] + 1
            assert p["work_items"] > 0 and p["ccap"] >= 256
        if n > 2048:
            assert p["stride"] >= 1 and p["n_sample"] == -(-n // p["stride"])
        assert p["candidates"] == (46 if prec == E.PREC_BF16_RESCORE and dtype == E.DTYPE_F32 else 30)
    # serving slots never take the symmetric pass
    s = E.exact_plan(E.DTYPE_F32, 100000, 768, E.KnnGraphOptions(n_neighbors=30), serving=True)
    assert s["symmetric"] == 0 and s["segments"] == s["n_splits"] <= 16


def _binom_tail(m, p, j):
    """P(X >= j), X ~ Binomial(m, p), with exact rationals (independent of the library's
    log-domain sum)."""
    from fractions import Fraction
    from math import comb

    q = Fraction(p)
    return sum(comb(m, i) * q ** i * (1 - q) ** (m - i) for i in range(j, m + 1))


@pytest.mark.parametrize("n,d,k,prec", [(1000000, 768, 30, E.PREC_BF16_RESCORE), (100000, 768, 30, E.PREC_BF16_RESCORE),
                                         (100000, 768, 30, E.PREC_FP32), (30000, 384, 15, E.PREC_FP32),
                                         (20000, 128, 10, E.PREC_FP32)])
def test_bound_rank_is_the_smallest_rank_within_the_miss_budget(n, d, k, prec, monkeypatch):
    """The scoring pass tests the J-th largest group maximum of the strided sample.  It
    exceeds a query's K-th best score only if J sampled rows rank among its K - 1 best,
    about Binomial(K - 1, 1 / stride): J is the smallest rank with that tail <= 1e-6
    (EVOC_MI355X_BOUND_EPS), checked here with exact rational arithmetic; eps = 1 gives
    the guaranteed rank K."""
    from fractions import Fraction

    o = E.KnnGraphOptions(n_neighbors=k, precision=prec)
    p = E.exact_plan(E.DTYPE_F32, n, d, o)
    K, st, J = p["candidates"], p["stride"], p["bound_rank"]
    assert st > 1 and 1 <= J <= K
    eps = Fraction(1, 1000000)
    assert _binom_tail(K - 1, Fraction(1, st), J) <= eps
    if J > 1:
        assert _binom_tail(K - 1, Fraction(1, st), J - 1) > eps
    if n == 1000000 and K == 46:
        assert (st, J) == (31, 11)  # rank 46 -> 11: about 4x fewer appended keys
    monkeypatch.setenv("EVOC_MI355X_BOUND_EPS", "1")
    assert E.exact_plan(E.DTYPE_F32, n, d, o)["bound_rank"] == K
    monkeypatch.setenv("EVOC_MI355X_BOUND_EPS", "0.5")
    j_loose = E.exact_plan(E.DTYPE_F32, n, d, o)["bound_rank"]
    assert j_loose < J and _binom_tail(K - 1, Fraction(1, st), j_loose) <= Fraction(1, 2)


def test_bound_rank_without_a_strided_sample():
    # no threshold pass at all
    p = E.exact_plan(E.DTYPE_F32, 2000, 128, E.KnnGraphOptions(n_neighbors=15))
    assert p["stride"] == 0 and p["bound_rank"] == 0
    # every row sampled: only the guaranteed rank K is safe
    p = E.exact_plan(E.DTYPE_F32, 5000, 128, E.KnnGraphOptions(n_neighbors=15, precision=E.PREC_FP32))
    assert p["stride"] == 1 and p["bound_rank"] == p["candidates"] == 15
