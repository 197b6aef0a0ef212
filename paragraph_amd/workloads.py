"""Synthetic workloads of the benchmark and the full-size parity tests.

The growth step's four configurations are fixed by their sizes and seeds
(seed graphs from pg.seed_graph, Philox-keyed); this module holds the one
whose parameters are drawn: the ensemble's per-simulation parameters.
"""
from __future__ import annotations

import math

import numpy as np


def ensemble_params(n_sims: int, m: int = 20) -> list[dict]:
    """The ensemble's simulations: seed sizes log-uniform in [10k, 500k],
    alpha ~ U[0.5, 1.5], tau ~ U[1, 10] years, neighbourhood fraction
    ~ U[0, 0.5] (n_nb = round(f*m), at most m-1), Philox keys 0..n_sims-1.
    Drawn from numpy's PCG64 with seed 3."""
    rng = np.random.default_rng(3)
    out = []
    for s in range(n_sims):
        n0 = int(round(math.exp(rng.uniform(math.log(1e4), math.log(5e5)))))
        alpha = float(rng.uniform(0.5, 1.5))
        tau = float(rng.uniform(1.0, 10.0))
        f_nb = float(rng.uniform(0.0, 0.5))
        out.append(dict(n0=n0, alpha=alpha, tau=tau,
                        n_nb=min(int(math.floor(f_nb * m + 0.5)), m - 1), seed=s))
    return out


def reference_model_inputs(n_bag: int = 10_000) -> dict:
    """Inputs of the reference's scoring model for the benchmark, shaped like
    the reference's own runs (drivers/ablation_parallel_gpu_10y.sh: alpha
    0.5, equal weights, 5% fully random, 12% same-year; an out-degree bag of
    papers with at least five references): bag values 5 + floor(lognormal(
    2.3, 0.8)) capped at 250 (mean about 19), drawn from numpy's PCG64 with
    seed 4; recency probabilities p(d) proportional to d * exp(-d / 4) for
    year differences d = 1..30."""
    rng = np.random.default_rng(4)
    bag = np.minimum(5 + np.floor(rng.lognormal(2.3, 0.8, n_bag)), 250).astype(np.int32)
    d = np.arange(31, dtype=np.float64)
    rec = d * np.exp(-d / 4.0)
    rec /= rec.sum()
    return dict(bag=bag, rec=rec, alpha=0.5, w_pa=0.33, w_rec=0.33, w_fit=0.33,
                fully_random=0.05, same_year=0.12)
