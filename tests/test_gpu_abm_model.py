"""The reference's scoring model on the GPU against the oracle (needs an MI355X).

Every year starts the oracle from the state the GPU started it from and
checks, bit for bit: the score components P, R, F of every existing paper;
per new paper the generator, out-degree, same-year target and the numbers
of 1-hop, 2-hop and uniform picks; the new rows (ids and order); and the
whole CSR after the append.  Cases: several bag / share / weight settings,
a hub seed whose 2-hop sets hold tens of thousands of papers (selection
buffer refills, multi-chunk row expansion), shares 0 and 1, every paper
same-year, all-uniform citing, a three-paper seed where the uniform fill
runs out of papers, and a 200,000-paper seed over two years.
"""
import numpy as np
import pytest

import paragraph_amd as pg
from gpu_checks import log, oracle_graph
from oracle import binding as ob

pytestmark = pytest.mark.gpu

KEYS = ("fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "in_deg", "birth")
REC = 0.5 ** np.arange(30) * (1.0 - 0.5)          # p(d) for d = 0..29
BAG = [0, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 250]


def make_ctx(g: pg.SeedGraph, sched, seed, max_out=250):
    n_final = g.n + int(sum(sched))
    e_cap = int(g.src.size + sum(sched) * (max_out + 2) + 16)
    ctx = pg.Context(n_final, e_cap, max(max(sched), 1), max_cites=1, bfs_cap=16, seed=seed)
    ctx.load(g)
    return ctx


def check_year(ctx, logfit, m: ob.AbmModel, year, yi, D):
    G = oracle_graph(ctx, logfit)
    N = G.n
    log(f"scoring-model year {year} (idx {yi}): N={N} D={D} E={G.e}")
    st = ctx.year_abm(year, yi, D)
    assert st.n_before == N and st.n_after == N + D
    info = ctx.abm_year_info(D)
    comp = ctx.abm_scores(N)
    ref = ob.abm_year(G, m, year, yi, D)
    assert np.array_equal(comp.view(np.uint32), ref["comp"].view(np.uint32)), "scores differ"
    bad = np.nonzero(np.any(info != ref["info"], axis=1))[0]
    assert bad.size == 0, f"paper info differs at {bad[:8]}: {info[bad[:3]]} vs {ref['info'][bad[:3]]}"
    after = ctx.download()
    fo = after["fwd_off"]
    assert st.edges_added == ref["tgt"].size
    assert np.array_equal(fo[N:N + D + 1] - fo[N], ref["off"])
    assert np.array_equal(after["fwd_adj"][fo[N]:fo[N + D]], ref["tgt"]), "new rows differ"
    G2 = ob.csr_append(G, ref["off"], ref["tgt"], logfit, year)
    for k in KEYS:
        assert np.array_equal(after[k], getattr(G2, k)), f"{k} differs after append"
    return st, info


def run_case(g, sched, seed, m: ob.AbmModel, **kw):
    ctx = make_ctx(g, sched, seed)
    ctx.abm_configure(m.bag, m.rec, alpha=m.alpha, w_pa=m.w[0], w_rec=m.w[1], w_fit=m.w[2],
                      fully_random=m.frc, same_year=m.syp)
    infos = []
    for yi, D in enumerate(sched):
        st, info = check_year(ctx, g.logfit, m, 1 + yi, yi, D)
        infos.append(info)
    ctx.close()
    return infos


def seed_graph(seed, n0, mean, sched):
    return pg.seed_graph(seed, n0, mean, n_final=n0 + int(sum(sched)))


def test_reference_settings_three_years():
    sched = pg.growth_schedule(3000, 0.05, 3)
    g = seed_graph(21, 3000, 8.0, sched)
    m = ob.AbmModel(21, BAG, REC, alpha=0.5, frc=0.05, syp=0.12)
    infos = run_case(g, sched, 21, m)
    i0 = infos[0]
    assert (i0[:, 2] >= 0).sum() == int(np.floor(sched[0] * 0.12))
    assert i0[:, 3].sum() > 0 and i0[:, 4].sum() > 0 and i0[:, 5].sum() > 0


def test_per_paper_share_and_weights():
    sched = pg.growth_schedule(5000, 0.04, 2)
    g = seed_graph(22, 5000, 6.0, sched)
    m = ob.AbmModel(22, BAG, REC[:12], alpha=-1.0, w_pa=-1.0, w_rec=-1.0, w_fit=-1.0,
                    frc=0.3, syp=0.5)
    run_case(g, sched, 22, m)


def hub_graph(n0: int, sched, seed=5):
    """Every paper cites paper 0 and up to three random older papers: 2-hop
    sets through the hub hold most of the graph."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for i in range(1, n0):
        ts = {0} | set(rng.integers(0, i, size=min(3, i)).tolist())
        for t in sorted(ts):
            src.append(i)
            dst.append(t)
    n_final = n0 + int(sum(sched))
    birth = np.sort(rng.integers(-20, 1, size=n0)).astype(np.int32)
    lf = np.zeros(n_final, np.float32)
    return pg.SeedGraph(birth, np.array(src, np.int32), np.array(dst, np.int32), lf)


def test_hub_two_hop_sets():
    sched = [600, 700]
    g = hub_graph(30000, sched)
    m = ob.AbmModel(23, [30, 60, 120, 250], REC, alpha=0.3, frc=0.02, syp=0.1)
    infos = run_case(g, sched, 23, m)
    assert infos[0][:, 4].max() > 100   # large 2-hop pick lists


def test_share_extremes_and_all_same_year():
    sched = pg.growth_schedule(4000, 0.05, 1)
    g = seed_graph(24, 4000, 7.0, sched)
    for alpha, syp, frc in [(0.0, 1.0, 0.0), (1.0, 0.0, 0.0), (0.5, 0.3, 1.0)]:
        m = ob.AbmModel(24, BAG, REC, alpha=alpha, frc=frc, syp=syp)
        infos = run_case(g, sched, 24, m)
        if alpha == 0.0:
            assert infos[0][:, 3].sum() == 0
        if syp == 1.0:
            assert np.all(infos[0][:, 2] >= 0)


def test_tiny_seed_uniform_fill_runs_out():
    birth = np.array([-2, -1, -1], np.int32)
    g = pg.SeedGraph(birth, np.array([1, 2], np.int32), np.array([0, 0], np.int32),
                     np.zeros(5, np.float32))
    m = ob.AbmModel(25, [250], REC, alpha=0.5, frc=0.05, syp=0.0)
    infos = run_case(g, [1, 1], 25, m)
    # the first new paper can cite at most the three seed papers
    assert 1 + infos[0][0, 3] + infos[0][0, 4] + infos[0][0, 5] == 3


def test_two_hundred_thousand_papers_two_years():
    sched = pg.growth_schedule(200_000, 0.03, 2)
    g = seed_graph(26, 200_000, 10.0, sched)
    m = ob.AbmModel(26, BAG, REC, alpha=0.5, frc=0.05, syp=0.12)
    run_case(g, sched, 26, m)


def test_host_paper_draws_match_oracle():
    sched = [10]
    g = seed_graph(27, 1000, 5.0, sched)
    ctx = make_ctx(g, sched, 27)
    for alpha, w in [(0.4, (0.2, 0.3, 0.5)), (-1.0, (-1.0, 0.5, 0.5))]:
        m = ob.AbmModel(27, BAG, REC, alpha=alpha, w_pa=w[0], w_rec=w[1], w_fit=w[2])
        ctx.abm_configure(m.bag, m.rec, alpha=alpha, w_pa=w[0], w_rec=w[1], w_fit=w[2])
        d = ctx.abm_paper_draws(0, 500)
        fit = ob.abm_fitness(27, range(500))
        assert np.array_equal(d["fitness"], fit)
        for i in range(500):
            k, a, ww = ob.abm_paper(m, i)
            assert d["out_degree"][i] == k and d["alpha"][i] == a
            assert np.array_equal(d["weights"][i], ww)
    ctx.close()
