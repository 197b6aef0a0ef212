"""GPU-vs-oracle parity through the C ABI (needs an MI355X).

Every comparison starts the oracle from the exact state the GPU started the
year from (downloaded from the device), so each year is checked on its own:
  - FP32 weights: bit-exact;
  - FP64 CDF: within 1e-12 relative (blocked vs sequential summation);
  - draws: bit-exact, except papers where a PA uniform fell within 1e-12*Z of
    a CDF boundary (counted by both sides; every mismatching paper must be
    one of them);
  - CSR after the append: bit-exact against the oracle's append of the GPU's
    draws; in-degrees equal the backward row lengths.
"""
import numpy as np
import pytest

import paragraph_amd as pg
from gpu_checks import check_pa_year, log, new_ctx, oracle_graph
from oracle import binding as ob

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("years", [3])
def test_pa_config0_years(years):
    """Config 0: 50,000 papers, Poisson(8), 3 years at 3%, m=10, 30% neighbourhood."""
    sched = pg.growth_schedule(50000, 0.03, years)
    g = pg.seed_graph(1, 50000, 8.0, n_final=50000 + sum(sched))
    ctx = new_ctx(g, sched, 10)
    _, e0 = ctx.size()  # distinct seed citations
    added = 0
    for yi, D in enumerate(sched):
        st, _ = check_pa_year(ctx, g.logfit, 1 + yi, yi, D, 10, 3, 1.0, 5.0, 1, 4096)
        assert st.shortfall == 0
        added += st.edges_added
    n, e = ctx.size()
    assert n == 54637
    assert added == 46370 and e == e0 + 46370


def hub_graph(n=20000, hubs=8, seed=7):
    """Seed graph with a few very highly cited papers (stresses caps/windows)."""
    rng = np.random.default_rng(seed)
    birth = np.sort(rng.integers(-20, 1, n)).astype(np.int32)
    src, dst = [], []
    for i in range(hubs + 1, n):
        k = min(6, i)
        t = rng.choice(i, size=k, replace=False)
        src += [i] * k
        dst += list(t)
        hs = np.nonzero(rng.random(hubs) < 0.45)[0]
        hs = hs[~np.isin(hs, t)]
        src += [i] * hs.size
        dst += list(hs)
    logfit = (rng.standard_normal(n + 4000) * 0.5).astype(np.float32)
    return pg.SeedGraph(birth, np.array(src, np.int32), np.array(dst, np.int32), logfit)


@pytest.mark.parametrize("cap", [4096, 300, 17])
def test_pa_hub_graph_caps(cap):
    g = hub_graph()
    sched = [600, 600]
    ctx = new_ctx(g, sched, 12, cap=cap, seed=5)
    for yi, D in enumerate(sched):
        check_pa_year(ctx, g.logfit, 1 + yi, yi, D, 12, 5, 1.3, 3.0, 5, cap)


@pytest.mark.parametrize("params", [(0.0, float("inf"), 4, 3), (1.5, 1.0, 20, 19),
                                    (0.5, 10.0, 6, 0), (1.0, 5.0, 1, 0)])
def test_pa_param_edges(params):
    alpha, tau, m, n_nb = params
    g = pg.seed_graph(3, 4000, 5.0, n_final=4500)
    ctx = new_ctx(g, [300], m, cap=4096, seed=11)
    check_pa_year(ctx, g.logfit, 1, 0, 300, m, n_nb, alpha, tau, 11, 4096)


def test_first_year_on_dirty_device_memory():
    """The buffers a context allocates in its first PA year with
    neighbourhood draws (root-dedup bitmap, byte counters, staging-slot
    flags) are cleared on the context's stream before the kernels that use
    them: after 1 GiB of device memory full of set bits has been freed back
    to the runtime, the first year still equals the oracle."""
    import torch
    junk = torch.full((1 << 28,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    del junk
    torch.cuda.empty_cache()
    sched = pg.growth_schedule(50000, 0.03, 1)
    g = pg.seed_graph(3, 50000, 8.0, n_final=50000 + sum(sched))
    ctx = new_ctx(g, sched, 10)
    st, bad = check_pa_year(ctx, g.logfit, 1, 0, sched[0], 10, 3, 1.0, 5.0, 1, 4096)
    assert st.shortfall == 0


def test_pa_tiny_graph_exhaustion():
    """Fewer eligible papers than citations: attempt budget ends the draws."""
    g = pg.SeedGraph(np.array([-3, -2, -1], np.int32), np.array([1, 2], np.int32),
                     np.array([0, 1], np.int32), np.zeros(10, np.float32))
    ctx = new_ctx(g, [4], 5, cap=8, seed=2)
    st, _ = check_pa_year(ctx, g.logfit, 1, 0, 4, 5, 2, 1.0, 5.0, 2, 8)
    assert st.shortfall > 0


def test_pa_year_without_positive_weight():
    """Every weight underflows to zero (log-fitness -200): the year adds no
    citations, as in the oracle."""
    g = pg.seed_graph(4, 3000, 5.0, n_final=3200)
    g.logfit[:] = -200.0
    ctx = new_ctx(g, [200], 20, cap=4096, seed=5)
    st, _ = check_pa_year(ctx, g.logfit, 1, 0, 200, 20, 6, 1.0, 5.0, 5, 4096)
    assert st.edges_added == 0 and st.near_boundary == 0


def tier_graph(n=12000, big=6, mid=40, seed=13):
    """Seed graph whose papers spread over the three 2-hop paths: 6 papers
    cited by ~7% of the others (in-degree ~800), 40 by ~2% (~240), each paper
    also citing 3 random earlier ones."""
    rng = np.random.default_rng(seed)
    birth = np.sort(rng.integers(-20, 1, n)).astype(np.int32)
    hubs = np.arange(big + mid)
    pb = np.r_[np.full(big, 0.07), np.full(mid, 0.02)]
    src, dst = [], []
    for i in range(big + mid + 1, n):
        t = rng.choice(i, size=3, replace=False)
        hs = hubs[rng.random(big + mid) < pb]
        hs = hs[~np.isin(hs, t)]
        src += [i] * (3 + hs.size)
        dst += list(t) + list(hs)
    logfit = (rng.standard_normal(n + 2000) * 0.5).astype(np.float32)
    return pg.SeedGraph(birth, np.array(src, np.int32), np.array(dst, np.int32), logfit)


def bfs_tier_counts(d, roots, cap):
    """Distinct roots per 2-hop path: [one wavefront with a 512/2048-entry
    window, one wavefront with 1024/4096, one workgroup], by the 1-hop row
    entries and the 2-hop candidate entries (F1c plus the rows of F1c)."""
    fo, fa, bo, ba = d["fwd_off"], d["fwd_adj"], d["bwd_off"], d["bwd_adj"]
    deg = np.diff(fo) + np.diff(bo)
    out = np.zeros(3, np.int64)
    for r in np.unique(roots):
        f1 = np.unique(np.r_[fa[fo[r]:fo[r + 1]], ba[bo[r]:bo[r + 1]]])
        f1c = f1[f1 != r][:cap]
        d1, tot2 = deg[r], deg[f1c].sum() + f1c.size
        out[0 if (d1 <= 512 and tot2 <= 2048) else 1 if (d1 <= 1024 and tot2 <= 4096) else 2] += 1
    return out


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("cap", [4096, 64])
def test_pa_bfs_tiers(cap, mode):
    """Year steps whose roots take all three 2-hop paths (with cap 64 the
    1024-entry wavefront also sees 1-hop rows beyond the cap)."""
    g = tier_graph()
    sched = [900, 900]
    ctx = new_ctx(g, sched, 10, cap=cap, seed=9)
    ctx.set_tuning(bfs_mode=mode)
    for yi, D in enumerate(sched):
        before = ctx.download()
        check_pa_year(ctx, g.logfit, 1 + yi, yi, D, 10, 4, 1.2, 4.0, 9, cap)
        gres = ctx.citations(D, 10)
        roots = gres["cit"][(gres["cnt"] > 0) & (gres["kind"][:, 0] == 0), 0]
        tiers = bfs_tier_counts(before, roots, cap)
        log(f"  distinct roots per path: {tiers.tolist()}")
        assert tiers[0] > 0 and tiers[1] > 0
        if cap == 4096:
            assert tiers[2] > 0


def cut_graph(n=40000, hubs=4, p=0.1, k=4, seed=17):
    """Seed graph with 4 papers each cited by ~10% of the others (in-degree
    ~4000, citers mostly distinct), every paper citing 4 random earlier
    ones: a root citing two or more of them has 8000+ distinct 2-hop
    candidates below the answer bound."""
    rng = np.random.default_rng(seed)
    birth = np.sort(rng.integers(-20, 1, n)).astype(np.int32)
    src, dst = [], []
    for i in range(hubs + 1, n):
        t = rng.choice(i, size=min(k, i), replace=False)
        hs = np.nonzero(rng.random(hubs) < p)[0]
        hs = hs[~np.isin(hs, t)]
        src += [i] * (t.size + hs.size)
        dst += list(t) + list(hs)
    return pg.SeedGraph(birth, np.array(src, np.int32), np.array(dst, np.int32),
                        np.zeros(n, np.float32))


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("cap", [4096, 1500])
def test_bfs_radix_cut(cap, mode):
    """Workgroup path on lists whose entries below the answer bound exceed
    the in-LDS sort (7168): cut at a histogram bound when the kept ids below
    it reach cap, otherwise (many duplicates, the hub roots) handed to the
    staged or swept path."""
    g = cut_graph()
    ctx = new_ctx(g, [1], 1, cap=cap)
    ctx.set_tuning(bfs_mode=mode)
    G = oracle_graph(ctx, g.logfit)
    src = np.asarray(g.src)
    two_hubs = np.nonzero(np.bincount(src[np.asarray(g.dst) < 4], minlength=g.n) >= 2)[0]
    roots = np.concatenate([np.arange(0, 6), two_hubs[:40]]).astype(np.int32)
    log(f"BFS cut graph cap={cap} mode={mode} roots={roots.size}")
    got = ctx.bfs(roots)
    for r, s in zip(roots, got):
        ref = ob.bfs(G, int(r), cap)
        assert np.array_equal(s, ref), f"root {r}: {s[:8]} vs {ref[:8]}"


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("cap", [4096, 64, 1])
def test_bfs_sets(cap, mode):
    g = hub_graph(n=8000, hubs=5, seed=3)
    ctx = new_ctx(g, [1], 1, cap=cap)
    ctx.set_tuning(bfs_mode=mode)
    G = oracle_graph(ctx, g.logfit)
    rng = np.random.default_rng(0)
    roots = np.concatenate([np.arange(0, 8), rng.integers(0, G.n, 120)]).astype(np.int32)
    log(f"BFS cap={cap} roots={roots.size}")
    got = ctx.bfs(roots)
    for r, s in zip(roots, got):
        ref = ob.bfs(G, int(r), cap)
        assert np.array_equal(s, ref), f"root {r}: {s[:8]} vs {ref[:8]}"


def test_bfs_isolated_and_path():
    n = 12
    birth = np.zeros(n, np.int32)
    src = np.array([1, 2, 3, 4, 6], np.int32)   # path 0-1-2-3-4, 5 isolated, 6-0
    dst = np.array([0, 1, 2, 3, 0], np.int32)
    g = pg.SeedGraph(birth, src, dst, np.zeros(n, np.float32))
    ctx = new_ctx(g, [1], 1, cap=4096)
    G = oracle_graph(ctx, g.logfit)
    roots = np.arange(n, dtype=np.int32)
    for r, s in zip(roots, ctx.bfs(roots)):
        assert np.array_equal(s, ob.bfs(G, int(r), 4096))
    assert ctx.bfs(np.array([5], np.int32))[0].size == 0
    assert list(ctx.bfs(np.array([2], np.int32))[0]) == [0, 1, 3, 4]


@pytest.mark.parametrize("p_link", [5e-4, 0.0, 1.0])
def test_er_edge_sets(p_link):
    n0 = 20000 if 0 < p_link < 1 else 60
    D = 600 if 0 < p_link < 1 else 7
    g = pg.SeedGraph(np.zeros(n0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32),
                     np.zeros(n0 + 2 * D, np.float32))
    cap_e = int(n0 * D * 2 * max(p_link, 1e-9) + 5000) if p_link < 1 else n0 * 2 * D + 2 * D * D
    ctx = pg.Context(n0 + 2 * D, cap_e, D, max_cites=1, bfs_cap=16, seed=9)
    ctx.load(g)
    N = n0
    for yi in range(2):
        log(f"ER year {yi}: N={N} D={D} p={p_link}")
        G = oracle_graph(ctx, g.logfit)
        st = ctx.year_er(1 + yi, yi, D, p_link)
        off, tgt = ob.er_year(N, D, p_link, 9, yi)
        assert st.edges_added == tgt.size
        d = ctx.download()
        fo = d["fwd_off"]
        got = d["fwd_adj"][fo[N]:fo[N + D]]
        assert np.array_equal(got, tgt)
        assert np.array_equal(fo[N:N + D + 1] - fo[N], off)
        G2 = ob.csr_append(G, off, tgt, g.logfit, 1 + yi)
        for k in ("fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "in_deg"):
            assert np.array_equal(d[k], getattr(G2, k)), k
        N += D


def test_snapshot_restore_reproduces():
    sched = [500, 515]
    g = pg.seed_graph(4, 5000, 6.0, n_final=5000 + sum(sched))
    ctx = new_ctx(g, sched, 8, seed=4)
    ctx.snapshot_save()
    runs = []
    for _ in range(2):
        for yi, D in enumerate(sched):
            ctx.year_pa(1 + yi, yi, D, 8, 2, 1.0, 5.0)
        runs.append(ctx.download())
        ctx.snapshot_restore()
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k


def multi_hub_graph(n=100_000, hubs=30, p=0.03, seed=12):
    """Paper n-1 cites `hubs` hubs (ids 0..hubs-1), each cited by ~p*n papers
    spread over the whole id range: its level-2 list has ~90k entries and no
    row longer than the cap, which takes the bitmap-window path."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for h in range(hubs):
        cit = np.nonzero(rng.random(n - 1 - hubs) < p)[0] + hubs
        src.append(cit); dst.append(np.full(cit.size, h))
    src.append(np.full(hubs, n - 1)); dst.append(np.arange(hubs))
    for i in range(hubs + 1, n - 1, 7):   # some ordinary citations too
        src.append(np.array([i])); dst.append(np.array([rng.integers(hubs, i)]))
    src = np.concatenate(src).astype(np.int32)
    dst = np.concatenate(dst).astype(np.int32)
    birth = np.zeros(n, np.int32)
    return pg.SeedGraph(birth, src, dst, np.zeros(n + 10, np.float32))


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("cap", [4096, 1000])
def test_bfs_bitmap_window_path(cap, mode):
    """Large candidate lists (~90k entries): mode 2 stages them in id buckets
    (long hub rows read in pieces), modes 0 and 1 sweep the rows window by
    window; the smaller roots' lists are staged in mode 0."""
    g = multi_hub_graph()
    ctx = new_ctx(g, [1], 1, cap=cap)
    ctx.set_tuning(bfs_mode=mode)
    G = oracle_graph(ctx, g.logfit)
    roots = np.array([g.n - 1, 0, 1, 29, 31, 5000], np.int32)
    log(f"bitmap-path BFS cap={cap} mode={mode}")
    for r, s in zip(roots, ctx.bfs(roots)):
        ref = ob.bfs(G, int(r), cap)
        assert np.array_equal(s, ref), f"root {r}: {s[:8]} vs {ref[:8]} ({s.size} vs {ref.size})"


def test_bfs_stage_overflow():
    """60 hubs of ~5000 citers each: the root's level-2 list, cut at cap+1
    entries per row, exceeds the staging slot (131072 ids), so the staged
    mode falls back to the sweep for it and stages the smaller roots."""
    g = multi_hub_graph(hubs=60, p=0.05, seed=13)
    ctx = new_ctx(g, [1], 1, cap=4096)
    ctx.set_tuning(bfs_mode=2)
    G = oracle_graph(ctx, g.logfit)
    roots = np.array([g.n - 1, 0, 59, 61, 777], np.int32)
    log("staging-overflow BFS")
    for r, s in zip(roots, ctx.bfs(roots)):
        ref = ob.bfs(G, int(r), 4096)
        assert np.array_equal(s, ref), f"root {r}: {s[:8]} vs {ref[:8]} ({s.size} vs {ref.size})"

def long_row_graph(hubs, mids, shared, n=60_000, hub_deg=6000, mid_deg=400, seed=21):
    """Paper n-1 cites `hubs` hubs (ids 0..hubs-1, ~hub_deg citers each) and
    `mids` papers after them (~mid_deg citers each).  With `shared` the hubs'
    citers come from one pool and the mids' from a small one, so the rows
    repeat each other's ids."""
    rng = np.random.default_rng(seed + 7 * hubs + mids)
    lo = hubs + mids
    pool_h = rng.choice(np.arange(lo, n - 1), size=int(hub_deg * 1.25), replace=False)
    pool_m = rng.choice(np.arange(lo, n - 1), size=3 * mid_deg, replace=False)
    src, dst = [], []
    for h in range(lo):
        deg = hub_deg if h < hubs else mid_deg
        if shared:
            cit = rng.choice(pool_h if h < hubs else pool_m, size=deg, replace=False)
        else:
            cit = rng.choice(np.arange(lo, n - 1), size=deg, replace=False)
        src.append(np.sort(cit)); dst.append(np.full(deg, h))
    src.append(np.full(lo, n - 1)); dst.append(np.arange(lo))
    for i in range(lo + 1, n - 1, 5):   # some ordinary citations too
        src.append(np.array([i])); dst.append(np.array([rng.integers(lo, i)]))
    src = np.concatenate(src).astype(np.int32)
    dst = np.concatenate(dst).astype(np.int32)
    return pg.SeedGraph(np.zeros(n, np.int32), src, dst, np.zeros(n + 10, np.float32))


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("cap", [4096, 1000, 64])
@pytest.mark.parametrize("hubs,mids,shared", [(1, 0, False), (3, 0, True), (4, 0, False),
                                              (5, 0, False), (4, 40, False), (4, 40, True)])
def test_bfs_long_rows_merged(hubs, mids, shared, cap, mode):
    """Level-2 lists holding rows longer than the cap: on the in-LDS sort
    path up to 4 such rows are merged into the sorted short entries (one,
    several with repeated ids, exactly 4; 5 take the plain sort), also
    when the short entries alone need the histogram cut, and when that cut
    keeps fewer than cap distinct ids (repeated ids) and hands the list on."""
    g = long_row_graph(hubs, mids, shared)
    ctx = new_ctx(g, [1], 1, cap=cap)
    ctx.set_tuning(bfs_mode=mode)
    G = oracle_graph(ctx, g.logfit)
    roots = np.array([g.n - 1, 0, hubs, hubs + mids + 3, 4321], np.int32)
    log(f"long-row BFS hubs={hubs} mids={mids} shared={shared} cap={cap} mode={mode}")
    for r, s in zip(roots, ctx.bfs(roots)):
        ref = ob.bfs(G, int(r), cap)
        assert np.array_equal(s, ref), f"root {r}: {s[:8]} vs {ref[:8]} ({s.size} vs {ref.size})"


def wide_root_graph(n=5000, seed=5):
    """Paper 2500 cites 700 earlier papers and is cited by 900 later ones;
    paper n-1 cites 600 (paper 2500 among them); every other paper cites
    three earlier ones.  Both 1-hop rows of paper 2500, and the forward row
    of paper n-1, are longer than one wavefront sorts."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for i in range(1, n - 1):
        if i == 2500:
            t = rng.choice(2500, size=700, replace=False)
        else:
            t = rng.choice(i, size=min(3, i), replace=False)
        src.append(np.full(t.size, i)); dst.append(t)
    citers = rng.choice(np.arange(2501, n - 1), size=900, replace=False)
    src.append(citers); dst.append(np.full(citers.size, 2500))
    t = np.unique(np.r_[rng.choice(n - 1, size=599, replace=False), 2500])[:600]
    src.append(np.full(t.size, n - 1)); dst.append(t)
    src = np.concatenate(src).astype(np.int32)
    dst = np.concatenate(dst).astype(np.int32)
    keep = np.unique(src.astype(np.int64) * n + dst, return_index=True)[1]
    return pg.SeedGraph(np.zeros(n, np.int32), src[keep], dst[keep], np.zeros(n + 10, np.float32))


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("cap", [4096, 1000, 300, 1])
def test_bfs_wide_roots_merged(cap, mode):
    """Roots whose 1-hop rows exceed one wavefront's sort: the rows are
    merged (both rows long, only the backward or only the forward one)."""
    g = wide_root_graph()
    ctx = new_ctx(g, [1], 1, cap=cap)
    G = oracle_graph(ctx, g.logfit)
    roots = np.array([2500, g.n - 1, 0, 17, 2501, 4000], np.int32)
    refs = [ob.bfs(G, int(r), cap) for r in roots]
    # default tier limits, then the smallest ones (paper n-1, 600 1-hop
    # entries, takes the workgroup path)
    for small_max, mid_max in [(0, 0), (256, 512)]:
        ctx.set_tuning(bfs_mode=mode, bfs_small_max=small_max, bfs_mid_max=mid_max)
        log(f"wide-root BFS cap={cap} mode={mode} tier limits {small_max}/{mid_max}")
        for r, s, ref in zip(roots, ctx.bfs(roots), refs):
            assert np.array_equal(s, ref), f"root {r}: {s[:8]} vs {ref[:8]} ({s.size} vs {ref.size})"
