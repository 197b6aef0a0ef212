"""Batched ensemble against plain runs (needs an MI355X).

Simulation s of an ensemble must reproduce, bit for bit, a plain context
loaded with the same seed graph and stepped with the same parameters:
forward and backward CSR, in-degrees and births after every run, for the
eager path and for the captured hipGraph and its replays.
"""
import sys

import numpy as np
import pytest

import paragraph_amd as pg

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SIMS = [
    dict(n0=3000, alpha=1.0, tau=5.0, n_nb=3, seed=11),
    dict(n0=8000, alpha=0.6, tau=2.0, n_nb=0, seed=12),
    dict(n0=500, alpha=1.4, tau=8.0, n_nb=5, seed=13),
    dict(n0=20000, alpha=1.1, tau=3.0, n_nb=2, seed=14),
    dict(n0=4096, alpha=0.9, tau=6.0, n_nb=4, seed=15),
    dict(n0=12000, alpha=1.2, tau=4.0, n_nb=3, seed=16),
]
M, YEARS, RATE = 10, 3, 0.05


def seed_graphs(sims):
    out = []
    for k, sp in enumerate(sims):
        sched = pg.growth_schedule(sp["n0"], RATE, YEARS)
        g = pg.seed_graph(100 + k, sp["n0"], 8.0, n_final=sp["n0"] + sum(sched))
        out.append((g, sched))
    return out


def plain_run(sp, g, sched):
    n_final = g.n + sum(sched)
    ctx = pg.Context(n_final, g.src.size + sum(sched) * M + 16, max(sched), max_cites=M,
                     seed=sp["seed"])
    ctx.load(g)
    edges = 0
    for yi, d in enumerate(sched):
        st = ctx.year_pa(1 + yi, yi, d, M, sp["n_nb"], sp["alpha"], sp["tau"])
        edges += st.edges_added
    out = ctx.download()
    ctx.close()
    return out, edges


def make_ensemble(sims, graphs):
    spec = [dict(sp, seed_edges=int(g.src.size)) for sp, (g, _) in zip(sims, graphs)]
    ens = pg.Ensemble(spec, YEARS, M, RATE)
    for k, (g, _) in enumerate(graphs):
        ens.load(k, g)
    return ens


def check_same(ens, refs):
    for k, ref in enumerate(refs):
        got = ens.download(k)
        for key in ("fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "in_deg", "birth"):
            assert np.array_equal(got[key], ref[key]), f"sim {k}: {key} differs"


def test_ensemble_equals_plain_runs_eager_and_graph():
    graphs = seed_graphs(SIMS)
    refs, total = [], 0
    for sp, (g, sched) in zip(SIMS, graphs):
        r, e = plain_run(sp, g, sched)
        refs.append(r)
        total += e
    log(f"plain runs: {total} edges")
    ens = make_ensemble(SIMS, graphs)
    st = ens.run(use_graph=False)
    log(f"eager: {st.edges_added} edges, {st.device_ms:.2f} ms")
    assert st.edges_added == total
    check_same(ens, refs)
    for rep in range(3):
        ens.restore()
        st = ens.run(use_graph=True)
        log(f"graph run {rep}: {st.edges_added} edges, {st.device_ms:.2f} ms")
        assert st.edges_added == total
        check_same(ens, refs)


def test_ensemble_tuning_does_not_change_results():
    sims = SIMS[:3]
    graphs = seed_graphs(sims)
    refs = [plain_run(sp, g, sched)[0] for sp, (g, sched) in zip(sims, graphs)]
    ens = make_ensemble(sims, graphs)
    for lanes in (8, 16, 32, 64):
        ens.set_tuning(pa_lanes=lanes)
        ens.restore()
        ens.run(use_graph=True)
        check_same(ens, refs)


def test_ensemble_autotune_keeps_results():
    """The autotuner times 4 PA lane counts, 2 more 2-hop modes, 3 wavefront
    tier limits and 2 merge tiles on whole captured runs, keeps the fastest,
    and leaves the seed state: a run with the chosen shapes reproduces the
    plain runs bit for bit."""
    sims = SIMS[:4]
    graphs = seed_graphs(sims)
    refs = [plain_run(sp, g, sched)[0] for sp, (g, sched) in zip(sims, graphs)]
    ens = make_ensemble(sims, graphs)
    rep = ens.autotune(reps=2)
    log("autotune", rep["chosen"], f"{rep['chosen_ms']:.3f} ms (default {rep['default_ms']:.3f} ms)")
    assert len(rep["trials"]) == 11
    assert rep["chosen_ms"] == min(t["ms"] for t in rep["trials"])
    assert rep["chosen"]["pa_lanes"] in (8, 16, 32, 64) and rep["chosen"]["bfs_mode"] in (0, 1, 2)
    assert rep["default_ms"] > 0
    ens.run(use_graph=True)   # the autotuner left the seed state
    check_same(ens, refs)


def test_ensemble_errors():
    with pytest.raises(pg.PGError, match="EINVAL"):
        pg.Ensemble([dict(n0=100, alpha=1.0, tau=5.0, n_nb=12, seed=1, seed_edges=10)], 2, 10)
    sims = SIMS[:2]
    graphs = seed_graphs(sims)
    spec = [dict(sp, seed_edges=int(g.src.size)) for sp, (g, _) in zip(sims, graphs)]
    ens = pg.Ensemble(spec, YEARS, M, RATE)
    with pytest.raises(pg.PGError, match="ESTATE"):   # out of order
        ens.load(1, graphs[1][0])
    with pytest.raises(pg.PGError, match="ESTATE"):   # not loaded
        ens.run()
    ens.load(0, graphs[0][0])
    ens.load(1, graphs[1][0])
    ens.run(use_graph=False)
    with pytest.raises(pg.PGError, match="ESTATE"):   # must restore first
        ens.run()


def test_device_seed_generation_equals_host_graphs():
    """pg_ensemble_generate builds every simulation's seed graph and
    log-fitness on the device: the same graph as pg_seed_graph with the
    simulation's key (rows, transposed rows, in-degrees, births), and runs
    from it equal runs from the host graphs (which also needs the same
    log-fitness bits: the weights of the last year are compared)."""
    graphs = []
    for sp in SIMS:
        sched = pg.growth_schedule(sp["n0"], RATE, YEARS)
        graphs.append((pg.seed_graph(sp["seed"], sp["n0"], 8.0, n_final=sp["n0"] + sum(sched)),
                       sched))
    host = make_ensemble(SIMS, graphs)
    bound = [dict(sp, seed_edges=int(sp["n0"] * 8 + 10 * np.sqrt(sp["n0"] * 8) + 100))
             for sp in SIMS]
    gen = pg.Ensemble(bound, YEARS, M, RATE)
    log("generating", len(SIMS), "seed graphs on the device")
    gen.generate(8.0)
    check_same(gen, [host.download(k) for k in range(len(SIMS))])
    a, b = host.run(use_graph=False), gen.run(use_graph=False)
    assert a.edges_added == b.edges_added and a.pa_attempts == b.pa_attempts
    check_same(gen, [host.download(k) for k in range(len(SIMS))])
    for k, sp in enumerate(SIMS):
        g, sched = graphs[k]
        n_last = g.n + sum(sched[:-1])
        wa = host.year_draws(k, sched[-1], n_last)["w"]
        wb = gen.year_draws(k, sched[-1], n_last)["w"]
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), f"sim {k}: weights"
    host.close()
    gen.close()
    # refusals: a bound below the draws, and generation after loading
    small = [dict(sp, seed_edges=10) for sp in SIMS]
    e = pg.Ensemble(small, YEARS, M, RATE)
    with pytest.raises(pg.PGError, match="ECAP"):
        e.generate(8.0)
    e.close()
    e = make_ensemble(SIMS[:1], graphs[:1])
    with pytest.raises(pg.PGError, match="ESTATE"):
        e.generate(8.0)
    e.close()
