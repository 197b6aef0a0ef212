"""Oracle PA year step against exact probabilities and hand-built cases.

Every new paper of a year draws from the same frozen weights with its own
Philox counters, so the papers of one year are independent samples of the
draw process; their empirical frequencies are compared with exact
without-replacement probabilities by chi-square.
"""
import itertools

import numpy as np
import pytest
from scipy import stats

from oracle import binding as ob


def graph(n, edges, birth=None, logfit=None):
    src = [a for a, b in edges]
    dst = [b for a, b in edges]
    birth = np.full(n, -5, np.int32) if birth is None else np.asarray(birth, np.int32)
    logfit = np.zeros(n + 100000, np.float32) if logfit is None else np.asarray(logfit, np.float32)
    return ob.Graph.from_edges(n, np.array(src, np.int64), np.array(dst, np.int64), birth, logfit)


def seq_prob(w, seq):
    p, rem = 1.0, w.sum()
    for s in seq:
        p *= w[s] / rem
        rem -= w[s]
    return p


def test_pa_draws_match_exact_without_replacement_law():
    n = 5
    logfit = np.concatenate([np.log([1.0, 2.0, 3.0, 0.5, 4.0]), np.zeros(60000)]).astype(np.float32)
    G = graph(n, [(1, 0), (2, 0), (3, 1), (4, 2), (4, 0)], logfit=logfit)
    D, m = 40000, 3
    r = ob.pa_year(G, 17, 0, 1, D, m, 0, 0.7, 5.0)
    assert np.all(r["cnt"] == m) and np.all(r["kind"] == 0)
    w = ob.weights(G, 1, 0.7, 0.2).astype(np.float64)
    perms = list(itertools.permutations(range(n), m))
    idx = {p: i for i, p in enumerate(perms)}
    obs = np.zeros(len(perms))
    for row in r["cit"]:
        obs[idx[tuple(row)]] += 1
    exp = np.array([seq_prob(w, p) for p in perms]) * D
    assert abs(exp.sum() - D) < 1e-6 * D
    assert stats.chisquare(obs, exp).pvalue > 1e-5


def test_dominant_paper_first_and_redraw_budget():
    n = 50
    logfit = np.zeros(n + 1000, np.float32)
    logfit[7] = 6.0   # about 89% of the total weight
    G = graph(n, [(i, i - 1) for i in range(1, n)], logfit=logfit)
    r = ob.pa_year(G, 3, 0, 1, 2000, 4, 0, 0.0, float("inf"))
    assert np.mean(r["cit"][:, 0] == 7) > 0.85
    assert np.all(r["cnt"] == 4)          # redraws still complete every paper
    for row in r["cit"]:
        assert len(set(row)) == 4
    # a paper holding all but 1e-7 of the weight: redraws exhaust the
    # 64*m attempt budget and the shortfall is left unfilled
    logfit[7] = 20.0
    G = graph(n, [(i, i - 1) for i in range(1, n)], logfit=logfit)
    r = ob.pa_year(G, 3, 0, 1, 500, 4, 0, 0.0, float("inf"))
    assert np.all(r["cit"][:, 0] == 7)
    assert np.all(r["cnt"] < 4) and np.all(r["cit"][:, 3] == -1)


def nb_graph(dominance=12.0):
    # 0 dominates the weights, so it is the first pick (and root) of most
    # papers; its 2-hop neighbourhood is {1,2,3,4}; 5..9 are outside.
    n = 10
    logfit = np.zeros(n + 100000, np.float32)
    logfit[0] = dominance
    logfit[1:5] = np.log([1.0, 2.0, 4.0, 8.0])
    edges = [(1, 0), (2, 0), (3, 1), (4, 2), (6, 5), (7, 5), (8, 6), (9, 8)]
    return graph(n, edges, logfit=logfit)


def test_neighbourhood_draws_follow_weights_inside_S():
    G = nb_graph()
    assert list(ob.bfs(G, 0, 4096)) == [1, 2, 3, 4]
    D = 40000
    r = ob.pa_year(G, 5, 0, 1, D, 3, 2, 1.0, float("inf"))
    ok = r["cit"][:, 0] == 0
    assert ok.mean() > 0.99
    assert np.all(r["kind"][ok][:, 1:] == 1)
    pairs = r["cit"][ok][:, 1:]
    w = ob.weights(G, 1, 1.0, 0.0).astype(np.float64)
    ws = np.zeros(5)
    ws[1:5] = w[1:5]
    perms = list(itertools.permutations([1, 2, 3, 4], 2))
    idx = {p: i for i, p in enumerate(perms)}
    obs = np.zeros(len(perms))
    for a, b in pairs:
        obs[idx[(a, b)]] += 1
    exp = np.array([seq_prob(ws, p) for p in perms]) * ok.sum()
    assert stats.chisquare(obs, exp).pvalue > 1e-5


def test_neighbourhood_take_all_and_refill():
    G = nb_graph()
    # |S \ P| = 4 <= n_nb = 4: every neighbour taken in ascending order
    r = ob.pa_year(G, 5, 0, 1, 200, 5, 4, 1.0, float("inf"))
    ok = r["cit"][:, 0] == 0
    assert np.all(r["cit"][ok][:, 1:] == [1, 2, 3, 4])
    assert np.all(r["kind"][ok][:, 1:] == 1)
    # n_nb = 6 > |S| = 4: two refill PA draws, no duplicates (paper 0 holds
    # ~90% of the weight here, so refill redraws stay within the budget)
    G = nb_graph(dominance=5.0)
    r = ob.pa_year(G, 5, 0, 1, 200, 8, 6, 1.0, float("inf"))
    ok = r["cit"][:, 0] == 0
    assert ok.mean() > 0.8
    for row, kinds in zip(r["cit"][ok], r["kind"][ok]):
        P = list(row[:2])
        nb = [s for s in [1, 2, 3, 4] if s not in P]
        assert list(row[2:2 + len(nb)]) == nb and np.all(kinds[2:2 + len(nb)] == 1)
        rest = kinds[2 + len(nb):]
        assert np.all((rest == 2) | (rest == -1))   # refill, or budget spent
        valid = row[row >= 0]
        assert len(set(valid)) == valid.size >= 2 + len(nb)
    assert np.mean(r["cnt"][ok] == 8) > 0.5


def test_isolated_root_refills_everything():
    n = 6
    logfit = np.zeros(n + 1000, np.float32)
    logfit[5] = 3.0
    G = graph(n, [(1, 0), (2, 1)], logfit=logfit)
    r = ob.pa_year(G, 2, 0, 1, 300, 4, 3, 0.0, float("inf"))
    iso = r["cit"][:, 0] == 5          # root 5 has no neighbours
    assert iso.mean() > 0.6
    assert np.all(r["kind"][iso][:, 1:] == 2) and np.all(r["cnt"][iso] == 4)


def test_invariants_and_thread_independence():
    rng = np.random.default_rng(9)
    n = 3000
    edges = set()
    for i in range(1, n):
        for t in rng.choice(i, size=min(i, 6), replace=False):
            edges.add((i, int(t)))
    birth = np.sort(rng.integers(-20, 1, n))
    logfit = (rng.standard_normal(n + 500) * 0.5).astype(np.float32)
    G = graph(n, sorted(edges), birth=birth, logfit=logfit)
    a = ob.pa_year(G, 8, 2, 3, 400, 16, 5, 1.2, 4.0, cap=64, threads=1)
    b = ob.pa_year(G, 8, 2, 3, 400, 16, 5, 1.2, 4.0, cap=64, threads=8)
    for k in a:
        assert np.array_equal(a[k], b[k])
    for row, kinds, c in zip(a["cit"], a["kind"], a["cnt"]):
        assert c == 16
        assert len(set(row)) == 16 and row.min() >= 0 and row.max() < n
        assert np.all(np.diff(kinds.astype(int)) >= 0)  # PA, then neighbourhood, then refill
        assert kinds[0] == 0 and np.sum(kinds == 0) == 11


def test_exhaustion_is_bounded():
    G = graph(3, [(1, 0), (2, 1)], birth=[-3, -2, -1])
    r = ob.pa_year(G, 1, 0, 1, 10, 5, 2, 1.0, 5.0)
    assert np.all(r["cnt"] == 3)
    assert np.all(r["cit"][:, 3:] == -1)


def test_year_without_positive_weight_adds_nothing():
    """Every weight underflows to zero (log-fitness -200): no PA attempt is
    made and no paper gets a citation (no neighbourhood draw or refill
    without a first pick).  One paper with a positive weight is then every
    paper's first pick."""
    lf = np.full(100004, -200.0, np.float32)
    G = graph(4, [(1, 0), (2, 1), (3, 2)], logfit=lf)
    r = ob.pa_year(G, 1, 0, 1, 50, 5, 2, 1.0, 5.0)
    assert np.all(r["cnt"] == 0) and np.all(r["cit"] == -1) and np.all(r["near"] == 0)
    lf[0] = 0.0
    G = graph(4, [(1, 0), (2, 1), (3, 2)], logfit=lf)
    r = ob.pa_year(G, 1, 0, 1, 50, 5, 2, 1.0, 5.0)
    assert np.all(r["cnt"] >= 1) and np.all(r["cit"][:, 0] == 0)
