"""Year-by-year GPU-vs-oracle checks shared by the GPU parity tests.

Each check starts the oracle from the exact state the GPU started the year
from (downloaded from the device), so every year is checked on its own:
  - FP32 weights: bit-exact;
  - FP64 CDF: within 1e-12 of Z (blocked vs sequential summation);
  - draws: bit-exact, except papers where a PA uniform fell within 1e-12*Z of
    a CDF boundary (counted by both sides; every mismatching paper must be
    one of them);
  - CSR after the append: bit-exact against the oracle's append of the
    GPU's draws; in-degrees equal the backward row lengths.
"""
import sys

import numpy as np

import paragraph_amd as pg
from oracle import binding as ob


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def new_ctx(g: pg.SeedGraph, sched, m, cap=4096, seed=1):
    n_final = g.n + int(sum(sched))
    e_cap = int(g.src.size + sum(sched) * max(m, 1) + 16)
    ctx = pg.Context(n_final, e_cap, max(max(sched), 1), max_cites=max(m, 1), bfs_cap=cap,
                     seed=seed)
    ctx.load(g)
    return ctx


def oracle_graph(ctx, logfit_full):
    return ob.Graph.from_download(ctx.download(), logfit_full)


def compare_year(G, got: dict, w_gpu, c_gpu, year, yi, D, m, n_nb, alpha, tau, seed, cap):
    """Weights, CDF and draws of one year against the oracle's year from G.
    Returns (oracle result, mismatching papers)."""
    inv_tau = 0.0 if np.isinf(tau) else 1.0 / tau
    w_ref = ob.weights(G, year, alpha, inv_tau)
    assert np.array_equal(w_gpu.view(np.uint32), w_ref.view(np.uint32)), "weights differ"
    if w_ref.size:
        c_ref = ob.cdf(w_ref)
        assert np.all(np.abs(c_gpu - c_ref) <= 1e-12 * c_ref[-1]), "CDF beyond 1e-12 relative"
        assert np.all(np.diff(c_gpu) >= -1e-15 * c_ref[-1])
    ref = ob.pa_year(G, seed, yi, year, D, m, n_nb, alpha, tau, cap)
    same = np.all(got["cit"] == ref["cit"], axis=1) & (got["cnt"] == ref["cnt"]) & \
        np.all(got["kind"] == ref["kind"], axis=1)
    bad = np.nonzero(~same)[0]
    flagged = (got["near"] > 0) | (ref["near"] > 0)
    assert np.all(flagged[bad]), f"unflagged mismatches at papers {bad[~flagged[bad]][:10]}"
    # rows: distinct, existing targets only
    N = G.n
    for i in range(0, D, max(1, D // 200)):
        row = got["cit"][i, :got["cnt"][i]]
        assert np.unique(row).size == row.size
        assert row.size == 0 or (row.min() >= 0 and row.max() < N)
    return ref, bad


def compare_append(G, got: dict, after: dict, logfit_full, year):
    """The CSR after the year equals the oracle's append of the GPU's draws."""
    off, tgt = ob.pa_compact(got)
    G2 = ob.csr_append(G, off, tgt, logfit_full, year)
    for k in ("fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "in_deg", "birth"):
        assert np.array_equal(after[k], getattr(G2, k)), f"{k} differs after append"
    assert np.array_equal(np.diff(after["bwd_off"]), after["in_deg"])
    return G2


def check_pa_year(ctx, logfit_full, year, yi, D, m, n_nb, alpha, tau, seed, cap):
    G = oracle_graph(ctx, logfit_full)
    N = G.n
    log(f"PA year {year} (idx {yi}): N={N} D={D} m={m} n_nb={n_nb} alpha={alpha} tau={tau} cap={cap}")
    st = ctx.year_pa(year, yi, D, m, n_nb, alpha, tau)
    assert st.n_before == N and st.n_after == N + D
    w_gpu, c_gpu = ctx.weights(N)
    got = ctx.citations(D, m)
    ref, bad = compare_year(G, got, w_gpu, c_gpu, year, yi, D, m, n_nb, alpha, tau, seed, cap)
    flagged = (got["near"] > 0) | (ref["near"] > 0)
    log(f"  edges={st.edges_added} near={st.near_boundary} mismatching papers={bad.size} "
        f"(flagged {int(flagged[bad].sum()) if bad.size else 0}) device_ms={st.device_ms:.2f}")
    assert int(got["near"].sum()) == st.near_boundary
    assert st.edges_added == int(got["cnt"].sum())
    assert st.shortfall == D * m - st.edges_added
    compare_append(G, got, ctx.download(), logfit_full, year)
    return st, bad.size
