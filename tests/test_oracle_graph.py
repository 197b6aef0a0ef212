"""Oracle BFS, ER and CSR append against brute force and known laws."""
import numpy as np
import pytest
from scipy import stats

from oracle import binding as ob


def rand_graph(n, k, seed, hubs=0):
    rng = np.random.default_rng(seed)
    edges = set()
    for i in range(1, n):
        for t in rng.choice(i, size=min(i, k), replace=False):
            edges.add((i, int(t)))
        for h in range(min(hubs, i)):
            if rng.random() < 0.5:
                edges.add((i, h))
    e = np.array(sorted(edges), np.int64).reshape(-1, 2)
    return ob.Graph.from_edges(n, e[:, 0], e[:, 1], np.zeros(n, np.int32),
                               np.zeros(n, np.float32)), e


def brute_bfs(n, e, root, cap):
    nb = [set() for _ in range(n)]
    for a, b in e:
        nb[a].add(int(b))
        nb[b].add(int(a))
    f1 = sorted(nb[root] - {root})
    f1c = f1[:cap]
    f2 = set()
    for v in f1c:
        f2 |= nb[v]
    f2 -= set(f1)
    f2.discard(root)
    return sorted(set(f1c) | f2)[:cap]


@pytest.mark.parametrize("cap", [4096, 40, 3, 1])
def test_bfs_matches_brute_force(cap):
    n = 400
    G, e = rand_graph(n, 4, 11, hubs=3)
    for root in list(range(0, 6)) + list(range(100, 400, 13)):
        assert list(ob.bfs(G, root, cap)) == brute_bfs(n, e, root, cap), root


def test_bfs_shapes():
    # star: centre 0 cited by 1..20
    e = np.array([(i, 0) for i in range(1, 21)], np.int64)
    G = ob.Graph.from_edges(21, e[:, 0], e[:, 1], np.zeros(21, np.int32), np.zeros(21, np.float32))
    assert list(ob.bfs(G, 0, 4096)) == list(range(1, 21))
    assert list(ob.bfs(G, 5, 4096)) == [0] + [i for i in range(1, 21) if i != 5]
    # path 0-1-2-3-4
    e = np.array([(1, 0), (2, 1), (3, 2), (4, 3)], np.int64)
    G = ob.Graph.from_edges(6, e[:, 0], e[:, 1], np.zeros(6, np.int32), np.zeros(6, np.float32))
    assert list(ob.bfs(G, 2, 4096)) == [0, 1, 3, 4]
    assert list(ob.bfs(G, 5, 4096)) == []            # isolated
    # clique of 6 (all older citations)
    e = np.array([(i, j) for i in range(6) for j in range(i)], np.int64)
    G = ob.Graph.from_edges(6, e[:, 0], e[:, 1], np.zeros(6, np.int32), np.zeros(6, np.float32))
    assert list(ob.bfs(G, 3, 4096)) == [0, 1, 2, 4, 5]


def test_bfs_hub_over_cap_excludes_far_one_hop():
    # root 50 cites 5..9; cap 3 keeps F1c = {5,6,7}.  5 cites 1 and is cited
    # by 8 (8 is in F1 \ F1c, so it is not a 2-hop id); 6 cites 2.
    edges = [(50, t) for t in range(5, 10)] + [(5, 1), (8, 5), (6, 2)]
    e = np.array(edges, np.int64)
    G = ob.Graph.from_edges(51, e[:, 0], e[:, 1], np.zeros(51, np.int32), np.zeros(51, np.float32))
    assert list(ob.bfs(G, 50, 3)) == [1, 2, 5]
    assert list(ob.bfs(G, 50, 4)) == [1, 2, 5, 6]
    assert list(ob.bfs(G, 50, 10)) == [1, 2, 5, 6, 7, 8, 9]
    for cap in (1, 3, 4, 10):
        assert list(ob.bfs(G, 50, cap)) == brute_bfs(51, e, 50, cap)


def test_er_extremes():
    off, tgt = ob.er_year(1000, 50, 0.0, 1, 0)
    assert tgt.size == 0 and np.all(off == 0)
    off, tgt = ob.er_year(40, 5, 1.0, 1, 0)
    assert np.array_equal(tgt, np.tile(np.arange(40), 5))


def test_er_binomial_and_geometric_laws():
    N, D, p = 20000, 2000, 2e-3
    off, tgt = ob.er_year(N, D, p, 3, 1)
    cnt = np.diff(off)
    mu, var = N * p, N * p * (1 - p)
    assert abs(cnt.mean() - mu) < 5 * np.sqrt(var / D)
    assert 0.85 < cnt.var() / var < 1.15
    # targets ascending and distinct within a paper; gaps are Geometric(p).
    # Only gaps that start more than 20/p before N are used, so that
    # censoring at N (probability (1-p)^(20/p) ~ e^-20) cannot bias them.
    limit = N - int(20 / p)
    gaps = []
    for i in range(D):
        row = tgt[off[i]:off[i + 1]]
        assert np.all(np.diff(row) > 0) and (row.size == 0 or (row[0] >= 0 and row[-1] < N))
        starts = np.concatenate([[-1], row])
        g = np.diff(np.concatenate([starts, [N + 10 ** 9]]))[:-1]
        gaps.extend(g[starts[:-1] < limit][: max(0, row.size)])
    gaps = np.array(gaps)
    ks = stats.kstest(gaps, lambda x: 1 - (1 - p) ** np.floor(x))
    assert ks.pvalue > 1e-4
    # every existing paper equally likely: uniform target frequencies
    h = np.bincount(tgt // 1000, minlength=N // 1000)
    assert stats.chisquare(h).pvalue > 1e-4


def test_er_matches_per_pair_bernoulli_in_law():
    N, D, p = 300, 3000, 0.02
    off, tgt = ob.er_year(N, D, p, 5, 0)
    cnt = np.diff(off)
    rng = np.random.default_rng(0)
    bern = (rng.random((D, N)) < p).sum(axis=1)
    assert stats.ks_2samp(cnt, bern).pvalue > 1e-4


def test_csr_append_invariants():
    n = 500
    G, e = rand_graph(n, 3, 4)
    rng = np.random.default_rng(2)
    n_new = 60
    rows = [np.sort(rng.choice(n, size=rng.integers(0, 9), replace=False)) for _ in range(n_new)]
    off = np.zeros(n_new + 1, np.int32)
    off[1:] = np.cumsum([r.size for r in rows])
    tgt = np.concatenate(rows).astype(np.int32)
    G2 = ob.csr_append(G, off, tgt, None, 1)
    src_new = np.repeat(np.arange(n, n + n_new), [r.size for r in rows])
    all_e = np.concatenate([e, np.stack([src_new, tgt], 1)])
    ref = ob.Graph.from_edges(n + n_new, all_e[:, 0], all_e[:, 1], np.zeros(n + n_new, np.int32),
                              np.zeros(n + n_new, np.float32))
    for k in ("fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "in_deg"):
        assert np.array_equal(getattr(G2, k), getattr(ref, k)), k
    assert G2.in_deg.sum() == G2.e
    assert np.all(np.diff(G2.fwd_off) >= 0) and np.all(np.diff(G2.bwd_off) >= 0)
