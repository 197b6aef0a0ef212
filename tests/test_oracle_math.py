"""Oracle elementary functions, weights and CDF against independent references.

- exp/log: ulp error against numpy float64 / longdouble on dense sweeps
  (bounds: expf, logf <= 1 ulp; log <= 1.5 ulp; log1p <= 3 ulp).
- weights: closed forms (alpha=0, k=0, inv_tau=0) within those bounds.
- CDF: integer-valued and power-of-two weights, where every FP64 partial sum
  is exact, so the result must equal numpy's exact cumulative sum.
"""
import numpy as np
import pytest

from oracle import binding as ob


def ulp_err(y, ref, dtype):
    ref = np.asarray(ref, np.float64)
    sp = np.spacing(np.abs(ref).astype(dtype)).astype(np.float64)
    return np.abs(np.asarray(y, np.float64) - ref) / sp


def test_expf_sweep():
    x = np.linspace(-87.3, 88.7, 1_000_001, dtype=np.float32)
    assert ulp_err(ob.vec("expf", x), np.exp(x.astype(np.float64)), np.float32).max() <= 1.0
    assert ob.vec("expf", np.array([0.0], np.float32))[0] == 1.0
    assert ob.vec("expf", np.array([-200.0], np.float32))[0] == 0.0
    assert np.isinf(ob.vec("expf", np.array([100.0], np.float32))[0])


def test_logf_sweep():
    rng = np.random.default_rng(0)
    x = rng.integers(1, 0x7f7fffff, 1_000_000).astype(np.uint32).view(np.float32)
    assert ulp_err(ob.vec("logf", x), np.log(x.astype(np.float64)), np.float32).max() <= 1.0
    k = np.arange(2, 3_000_000, dtype=np.float32)  # log(k+1) arguments of the weights
    assert ulp_err(ob.vec("logf", k), np.log(k.astype(np.float64)), np.float32).max() <= 1.0
    assert ob.vec("logf", np.array([1.0], np.float32))[0] == 0.0


def test_log_double_sweep():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(200_000), rng.random(50_000) * 1e6, 2.0 ** -np.arange(1, 1000)])
    x = x[x > 0]
    ref = np.log(x.astype(np.longdouble))
    assert ulp_err(ob.vec("log", x), ref, np.float64).max() <= 1.5


def test_log1p_double_sweep():
    x = np.concatenate([-np.logspace(-15, np.log10(0.29), 100_000), np.logspace(-15, np.log10(0.41), 100_000)])
    ref = np.log1p(x.astype(np.longdouble))
    assert ulp_err(ob.vec("log1p", x), ref, np.float64).max() <= 3.0


def graph_arrays(n, seed=3):
    rng = np.random.default_rng(seed)
    in_deg = rng.integers(0, 5000, n).astype(np.int32)
    birth = rng.integers(-30, 0, n).astype(np.int32)
    logfit = (rng.standard_normal(n) * 0.5).astype(np.float32)
    return in_deg, birth, logfit


def weights(in_deg, birth, logfit, year, alpha, inv_tau):
    n = in_deg.size
    g = ob.Graph(np.zeros(n + 1), np.zeros(0), np.zeros(n + 1), np.zeros(0), in_deg, birth, logfit)
    return ob.weights(g, year, alpha, inv_tau)


@pytest.mark.parametrize("alpha,inv_tau", [(0.0, 0.2), (1.3, 0.0), (0.0, 0.0)])
def test_weight_closed_forms(alpha, inv_tau):
    in_deg, birth, logfit = graph_arrays(100_000)
    w = weights(in_deg, birth, logfit, 1, alpha, inv_tau).astype(np.float64)
    age = (1 - birth).astype(np.float64)
    ref = np.exp(alpha * np.log(in_deg + 1.0) - age * np.float32(inv_tau) + logfit.astype(np.float64))
    # FP32 arithmetic of the exponent: a few ulp of the argument, times |x|
    x = np.abs(alpha * np.log(in_deg + 1.0)) + age * inv_tau + np.abs(logfit)
    rel = np.abs(w - ref) / ref
    assert np.all(rel <= 4 * 2.0 ** -24 * (1 + x))


def test_weight_k0_is_fitness_times_decay():
    n = 1000
    in_deg = np.zeros(n, np.int32)
    birth = np.full(n, -4, np.int32)
    logfit = np.zeros(n, np.float32)
    w = weights(in_deg, birth, logfit, 1, 2.5, 0.0)
    assert np.all(w == 1.0)


def test_cdf_exact_for_integer_weights():
    rng = np.random.default_rng(5)
    w = rng.integers(0, 1000, 500_000).astype(np.float32)
    c = ob.cdf(w)
    assert np.array_equal(c, np.cumsum(w.astype(np.int64)).astype(np.float64))


def test_cdf_exact_for_power_of_two_weights():
    rng = np.random.default_rng(6)
    w = (2.0 ** rng.integers(-10, 10, 200_000)).astype(np.float32)
    c = ob.cdf(w)
    exact = np.cumsum((w.astype(np.float64) * 1024).astype(np.int64)).astype(np.float64) / 1024
    assert np.array_equal(c, exact)


def test_upper_bound_semantics():
    c = np.array([1.0, 1.0, 3.0, 6.0])  # zero weight at index 1
    L = ob.lib()
    p = c.ctypes.data_as(ob.P(ob.ctypes.c_double))
    assert L.or_upper_bound(p, 4, 0.0) == 0
    assert L.or_upper_bound(p, 4, 0.999) == 0
    assert L.or_upper_bound(p, 4, 1.0) == 2   # never the zero-weight paper
    assert L.or_upper_bound(p, 4, 5.99) == 3
    assert L.or_upper_bound(p, 4, 6.0) == 3   # clamped


def run_math_xcheck(*args):
    import os
    import subprocess
    tool = os.path.join(os.path.dirname(ob.LIB_PATH), "math_xcheck")
    if not os.path.exists(tool):
        pytest.skip("oracle/math_xcheck not built")
    r = subprocess.run([tool, *map(str, args)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


@pytest.mark.parametrize("fn", ["expf", "logf"])
def test_device_math_header_equals_oracle(fn):
    """The library's det_expf / det_logf (paragraph_amd/csrc/pg_math.h, the
    same source the kernels compile, built here for the host) give the
    oracle's bits on every third float bit pattern (the full 2^32 sweep,
    stride 1, also passes: 40 s on 8 cores)."""
    run_math_xcheck(fn, 3)


def test_device_log_equals_oracle():
    """det_log on random double bit patterns and det_log_pos (the slot pass's
    log of a partition sum) on random positive normals equal or_log."""
    run_math_xcheck("log", 2_000_000)
