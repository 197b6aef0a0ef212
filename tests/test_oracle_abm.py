"""Pins the oracle's restatement of the reference's scoring model (CPU).

Each check compares the oracle with something other than its own formulas:
  - score components: exact rational values on small graphs (Python
    integers), the 128-bit sum path against big-integer arithmetic, sums of
    P, R and F equal to 1, R constant within an age class, the 2^21 guard;
  - fitness peak values: chi-square against p(f) ~ f^-1.634 on [1,1000];
  - per-paper draws: out-degree uniform over the bag (chi-square), fixed and
    drawn shares and weights (KS against numpy's law of normalised
    uniforms);
  - selection keys: the top-2 sets of a five-candidate pool against the exact
    successive-sampling probabilities of weighted sampling without
    replacement (chi-square); zero scores come last;
  - neighbourhoods: F1 and F2 of every paper by brute force with Python sets,
    the picks drawn from them, and the counts against a step-by-step Python
    run of the reference's stage arithmetic (src/kernel.cu:244-255, :383-386,
    the early returns of src/abm.cu:461-489 and the fill bound of
    src/abm.cu:386-432);
  - same-year subset: size floor(D * share), never the paper itself,
    membership and target uniform (chi-square);
  - uniform fill: distinct, never the generator, uniform over the others.
"""
import math
from fractions import Fraction

import numpy as np
from scipy import stats

import paragraph_amd as pg
from oracle import binding as ob

REC = 0.5 ** np.arange(30) * 0.5
BAG = [0, 1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 250]


def model(seed=1, **kw):
    args = dict(alpha=0.5, frc=0.05, syp=0.12)
    args.update(kw)
    bag = args.pop("bag", BAG)
    rec = args.pop("rec", REC)
    return ob.AbmModel(seed, bag, rec, **args)


def test_scores_exact_small():
    in_deg = np.array([0, 1, 2, 5, 0, 7], np.int32)
    birth = np.array([-3, -3, -2, -1, -1, 0], np.int32)
    m = model(3, rec=[0.0, 0.5, 0.3, 0.2])
    comp = ob.abm_scores(6, in_deg, birth, 1, m)
    S = sum(int(k) ** 3 + 1 for k in in_deg)
    for c, k in enumerate(in_deg):
        assert comp[c, 0] == np.float32(float(Fraction(int(k) ** 3 + 1, S)))
    fit = ob.abm_fitness(3, range(6))
    SF = sum(int(f) ** 3 + 1 for f in fit)
    for c, f in enumerate(fit):
        assert comp[c, 2] == np.float32(float(Fraction(int(f) ** 3 + 1, SF)))
    # ages 4, 4, 3, 2, 2, 1: ages 1, 2, 3 are in the table, age 4 is not
    U = 0.5 + 0.3 + 0.2
    want = [0.0, 0.0, 0.2 / (1 * U), 0.3 / (2 * U), 0.3 / (2 * U), 0.5 / (1 * U)]
    assert np.allclose(comp[:, 1], want, rtol=1e-7, atol=0)
    assert np.all(comp[:, 3] == 0)


def test_scores_sums_and_classes():
    g = pg.seed_graph(5, 20000, 7.0)
    G = ob.Graph.from_edges(20000, g.src, g.dst, g.birth, g.logfit)
    comp = ob.abm_scores(G.n, G.in_deg, G.birth, 1, model(5))
    for j in range(3):
        assert abs(comp[:, j].astype(np.float64).sum() - 1.0) < 1e-5
    age = 1 - G.birth
    for a in np.unique(age):
        r = comp[age == a, 1]
        assert np.all(r == r[0])


def test_scores_wide_sum_and_guard():
    k = (1 << 21) - 1
    n = 3000                      # the k^3 + 1 sum exceeds 2^64
    in_deg = np.full(n, k, np.int32)
    in_deg[::7] = 3
    birth = np.zeros(n, np.int32)
    comp = ob.abm_scores(n, in_deg, birth, 1, model(6))
    S = sum(int(x) ** 3 + 1 for x in in_deg)
    assert S > 2 ** 64
    for c in (0, 1, 7, 8):
        exact = np.float32(float(Fraction(int(in_deg[c]) ** 3 + 1, S)))
        assert abs(int(comp[c, 0].view(np.uint32)) - int(exact.view(np.uint32))) <= 1
    in_deg[5] = 1 << 21
    try:
        ob.abm_scores(n, in_deg, birth, 1, model(6))
        raise AssertionError("no error for an in-degree of 2^21")
    except RuntimeError:
        pass


def test_fitness_power_law():
    f = ob.abm_fitness(11, range(200_000))
    assert f.min() >= 1 and f.max() <= 1000
    p = np.arange(1, 1001, dtype=np.float64) ** -1.634
    p /= p.sum()
    edges = [1, 2, 3, 4, 5, 6, 8, 10, 15, 20, 30, 50, 100, 200, 1001]
    obs = np.array([np.sum((f >= a) & (f < b)) for a, b in zip(edges[:-1], edges[1:])])
    exp = np.array([p[a - 1:b - 1].sum() for a, b in zip(edges[:-1], edges[1:])]) * f.size
    assert stats.chisquare(obs, exp).pvalue > 1e-4


def test_paper_draws():
    m = model(12, bag=list(range(10, 60, 5)))   # ten distinct values
    ks = np.array([ob.abm_paper(m, i)[0] for i in range(50_000)])
    obs = np.array([np.sum(ks == v) for v in range(10, 60, 5)])
    assert stats.chisquare(obs).pvalue > 1e-4
    k, a, w = ob.abm_paper(model(12, alpha=0.25, w_pa=1.0, w_rec=2.0, w_fit=1.0), 7)
    assert a == 0.25 and np.allclose(w, [0.25, 0.5, 0.25], rtol=0, atol=1e-16)
    mr = model(13, alpha=-1.0, w_pa=-1.0)
    draws = [ob.abm_paper(mr, i) for i in range(20_000)]
    al = np.array([d[1] for d in draws])
    assert stats.kstest(al, "uniform").pvalue > 1e-4
    W = np.array([d[2] for d in draws])
    assert np.allclose(W.sum(axis=1), 1.0, atol=1e-15)
    rng = np.random.default_rng(0)
    u = rng.random((200_000, 3))
    ref = u / u.sum(axis=1, keepdims=True)
    for j in range(3):
        assert stats.ks_2samp(W[:, j], ref[:, j]).pvalue > 1e-4


def test_keys_weighted_sampling_without_replacement():
    s = np.array([1.0, 2.0, 3.0, 4.0, 10.0])
    comp = np.zeros((5, 4), np.float32)
    comp[:, 0] = s / s.sum()                 # P only
    w = np.array([1.0, 0.0, 0.0])
    m = model(14)
    R = 40_000
    pairs = {}
    for v in range(100_000, 100_000 + R):
        keys = [ob.abm_key(m, v, c, 3, comp, w) for c in range(5)]
        top = tuple(sorted(np.argsort(keys)[::-1][:2].tolist()))
        pairs[top] = pairs.get(top, 0) + 1
    S = s.sum()
    obs, exp = [], []
    for a in range(5):
        for b in range(a + 1, 5):
            pa = s[a] / S * s[b] / (S - s[a]) + s[b] / S * s[a] / (S - s[b])
            obs.append(pairs.get((a, b), 0))
            exp.append(pa * R)
    assert stats.chisquare(obs, exp).pvalue > 1e-4
    # a zero score keys at INT_MIN: below every positive-score key
    comp0 = comp.copy()
    comp0[2, :] = 0.0
    k0 = ob.abm_key(m, 5, 2, 3, comp0, w)
    assert k0 == -2147483648.0
    assert all(ob.abm_key(m, 5, c, 3, comp0, w) > k0 for c in (0, 1, 3, 4))


def neighbourhoods(G: ob.Graph, g: int):
    adj = lambda u: set(G.fwd_adj[G.fwd_off[u]:G.fwd_off[u + 1]].tolist()) | \
        set(G.bwd_adj[G.bwd_off[u]:G.bwd_off[u + 1]].tolist())
    F1 = adj(g) - {g}
    F2 = set()
    for u in F1:
        F2 |= adj(u)
    return F1, F2 - F1 - {g}


def reference_counts(k, sy, alpha, frc, n1, n2, N):
    """The reference's stage arithmetic step by step (src/kernel.cu:244-255,
    :383-386; MakePopulateCitations returns early for a count <= 0 and caps it
    at the candidate count; the fill's bound counts existing papers)."""
    reserved = math.floor(np.float32(frc * k))
    remaining = k - 1 - sy - reserved
    inside = math.ceil(np.float32(remaining * alpha))
    inside = min(inside, n1)
    cited = 0
    if inside > 0 and n1 > 0:
        cited += min(inside, n1)
    remaining3 = k - 1 - sy - reserved - inside
    outside = min(remaining3, n2)
    if outside > 0 and n2 > 0:
        cited += min(outside, n2)
    fully = k - 1 - sy - inside - outside
    uni = 0
    if fully > 0:
        uni = min(fully, N - 1 - cited)
    return max(inside, 0), max(outside, 0), uni


def test_neighbourhood_picks_and_counts():
    sched = [400]
    g = pg.seed_graph(15, 3000, 6.0, n_final=3400)
    G = ob.Graph.from_edges(3000, g.src, g.dst, g.birth, g.logfit)
    m = model(15, alpha=0.6, frc=0.1, syp=0.2)
    r = ob.abm_year(G, m, 1, 0, sched[0])
    comp = r["comp"]
    for l in range(sched[0]):
        v = 3000 + l
        gen, k, syt, p1, p2, nu = r["info"][l]
        row = r["tgt"][r["off"][l]:r["off"][l + 1]].tolist()
        F1, F2 = neighbourhoods(G, gen)
        k_, alpha, w = ob.abm_paper(m, v)
        assert k_ == k
        sy = 1 if syt >= 0 else 0
        assert (p1, p2, nu) == reference_counts(k, sy, alpha, 0.1, len(F1), len(F2), 3000)
        # the picks: the best keys of each level, every other entry uniform
        def best(S, t):
            keyed = sorted(S, key=lambda c: (-ob.abm_key(m, v, c, 0, comp, w), c))
            return set(keyed[:t])
        b1, b2 = best(F1, p1), best(F2, p2)
        rs = set(row)
        assert len(rs) == len(row) == 1 + sy + p1 + p2 + nu
        assert gen in rs and b1 <= rs and b2 <= rs
        rest = rs - b1 - b2 - {gen} - ({syt} if sy else set())
        assert len(rest) == nu and all(0 <= c < 3000 for c in rest)
        assert row == sorted(row)


def test_same_year_subset():
    m = model(16, syp=0.3)
    D, N = 50, 1000
    hits = np.zeros(D)
    tgt_rel = []
    for yi in range(4000):
        t = ob.abm_same_year(m, N, D, yi)
        sel = np.nonzero(t >= 0)[0]
        assert sel.size == math.floor(D * 0.3)
        assert np.all(t[sel] != N + sel) and np.all((t[sel] >= N) & (t[sel] < N + D))
        hits[sel] += 1
        tgt_rel.extend(((t[sel] - N - sel - 1) % D).tolist())   # 0..D-2 for other papers
    assert stats.chisquare(hits).pvalue > 1e-4
    obs = np.bincount(tgt_rel, minlength=D - 1)
    assert obs.size == D - 1
    assert stats.chisquare(obs).pvalue > 1e-4
    assert np.all(ob.abm_same_year(m, N, 1, 0) == -1)


def test_uniform_fill_law():
    N, D = 30, 20_000
    G = ob.Graph.from_edges(N, np.zeros(0, np.int32), np.zeros(0, np.int32),
                            np.zeros(N, np.int32), np.zeros(N + D, np.float32))
    m = model(17, bag=[10], frc=0.0, syp=0.0)
    r = ob.abm_year(G, m, 1, 0, D)
    inc = np.zeros(N)
    eligible = np.zeros(N)
    for l in range(D):
        gen = r["info"][l, 0]
        assert tuple(r["info"][l, 3:]) == (0, 0, 9)
        row = r["tgt"][r["off"][l]:r["off"][l + 1]]
        assert row.size == 10 and np.unique(row).size == 10 and gen in row
        for c in row:
            if c != gen:
                inc[c] += 1
        eligible += 1
        eligible[gen] -= 1
    p = 9 / (N - 1)
    z = (inc - eligible * p) / np.sqrt(eligible * p * (1 - p))
    assert np.all(np.abs(z) < 4.5)
    assert stats.chisquare(np.bincount(r["info"][:, 0], minlength=N)).pvalue > 1e-4
