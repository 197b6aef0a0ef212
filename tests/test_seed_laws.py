"""The synthetic seed-graph generator follows its stated laws (CPU).

pg_seed_graph: birth years uniform on [year_lo, year_hi] with ids sorted by
birth; each paper cites min(Poisson(mean), #older) distinct papers chosen
uniformly among the papers born in earlier years.  pg_seed_logfit:
log-fitness Normal(mu, sigma^2).  Checked by goodness-of-fit tests at a
1e-4 significance level on fixed seeds (so the tests are deterministic),
plus the hard invariants (sorted births, targets strictly older, no
duplicate citation, reproducibility).
"""
import numpy as np
from scipy import stats

import paragraph_amd as pg

N, LAM = 200_000, 8.0


def graph(seed=11):
    g = pg.seed_graph(seed, N, LAM, year_lo=-20, year_hi=0, sigma=0.5)
    older = np.searchsorted(g.birth, g.birth, side="left")  # papers born earlier
    return g, older


def test_births_uniform_and_sorted():
    g, _ = graph()
    assert np.all(np.diff(g.birth) >= 0)
    assert g.birth.min() >= -20 and g.birth.max() <= 0
    counts = np.bincount(g.birth + 20, minlength=21)
    assert stats.chisquare(counts).pvalue > 1e-4


def test_citation_counts_poisson_truncated():
    g, older = graph()
    cnt = np.bincount(g.src, minlength=N)
    assert np.all(cnt <= older)
    assert np.all(cnt[older == 0] == 0)
    # papers with many older papers: the truncation never binds in practice
    c = cnt[older >= 1000]
    assert c.size > 150_000
    kmax = 20
    obs = np.bincount(np.minimum(c, kmax), minlength=kmax + 1)
    pmf = stats.poisson.pmf(np.arange(kmax), LAM)
    exp = np.append(pmf, 1.0 - pmf.sum()) * c.size
    assert stats.chisquare(obs, exp).pvalue > 1e-4
    assert abs(c.mean() - LAM) < 4 * np.sqrt(LAM / c.size)


def test_targets_older_distinct_uniform():
    g, older = graph()
    src, dst = g.src, g.dst
    assert np.all(np.diff(src) >= 0)
    assert np.all(dst < older[src])             # cited papers are born in an earlier year
    key = src.astype(np.int64) * N + dst
    assert np.unique(key).size == key.size       # no duplicate citation
    sel = older[src] >= 1000
    u = (dst[sel] + 0.5) / older[src[sel]]
    assert stats.kstest(u, "uniform").pvalue > 1e-4


def test_logfit_normal():
    lf = np.empty(N, np.float32)
    pg.lib().pg_seed_logfit(5, 0, N, 0.0, 0.5, lf.ctypes.data_as(pg.ctypes.POINTER(pg.ctypes.c_float)))
    assert stats.kstest(lf.astype(np.float64), "norm", args=(0.0, 0.5)).pvalue > 1e-4
    # the draws are keyed by paper index: a later range continues the same stream
    part = np.empty(1000, np.float32)
    pg.lib().pg_seed_logfit(5, 5000, 1000, 0.0, 0.5,
                            part.ctypes.data_as(pg.ctypes.POINTER(pg.ctypes.c_float)))
    assert np.array_equal(part, lf[5000:6000])


def test_reproducible_and_keyed():
    a, _ = graph(11)
    b, _ = graph(11)
    c, _ = graph(12)
    assert np.array_equal(a.birth, b.birth) and np.array_equal(a.src, b.src)
    assert np.array_equal(a.dst, b.dst) and np.array_equal(a.logfit, b.logfit)
    assert not (a.dst.size == c.dst.size and np.array_equal(a.dst, c.dst))
