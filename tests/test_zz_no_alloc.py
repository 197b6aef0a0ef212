"""pg_bfs_noalloc (static scratch, caller device buffers) against pg_bfs.

The no-allocation entry point runs the year step's three 2-hop tiers
(wavefront small, wavefront medium, workgroup) over caller device buffers
with its scratch in one static device array.  Its neighbourhoods must equal
pg_bfs's bit for bit, on two calls in a row (the second smaller, so stale
scratch from the first would show), on the context's stream and on a torch
stream, with the staging slots in use (mode 2) and not (mode 1).
"""
import sys

import numpy as np
import pytest

import paragraph_amd as pg
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def multi_hub_graph(n=60_000, hubs=30, p=0.03, seed=12):
    """Paper n-1 cites `hubs` hubs, each cited by ~p*n papers across the id
    range (a 2-hop list of ~50k entries: the workgroup tier's staged or swept
    path); every seventh paper also cites one earlier paper, so most roots
    are small (wavefront tiers)."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for h in range(hubs):
        cit = np.nonzero(rng.random(n - 1 - hubs) < p)[0] + hubs
        src.append(cit)
        dst.append(np.full(cit.size, h))
    src.append(np.full(hubs, n - 1))
    dst.append(np.arange(hubs))
    for i in range(hubs + 1, n - 1, 7):
        src.append(np.array([i]))
        dst.append(np.array([rng.integers(hubs, i)]))
    src = np.concatenate(src).astype(np.int32)
    dst = np.concatenate(dst).astype(np.int32)
    return pg.SeedGraph(np.zeros(n, np.int32), src, dst, np.zeros(n + 10, np.float32))


def run_noalloc(torch, ctx, roots, cap, stream=None):
    dev = torch.device("cuda:0")
    r = torch.from_numpy(roots).to(dev)
    out = torch.full((roots.size * cap,), -7, dtype=torch.int32, device=dev)
    ln = torch.full((roots.size,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    ctx.bfs_noalloc(roots.size, r.data_ptr(), out.data_ptr(), ln.data_ptr(),
                    stream=None if stream is None else stream.cuda_stream)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1, cap), ln.cpu().numpy()


def check_same(ctx, roots, out, ln, cap):
    valid = (roots >= 0) & (roots < ctx.size()[0])
    assert np.all(ln[~valid] == 0)
    ref = ctx.bfs(roots[valid])
    for k, i in enumerate(np.nonzero(valid)[0]):
        assert ln[i] == ref[k].size, f"root {roots[i]}: length {ln[i]} vs {ref[k].size}"
        assert np.array_equal(out[i, :ln[i]], ref[k]), f"root {roots[i]} differs"
        assert np.all(out[i, ln[i]:] == -7), f"root {roots[i]}: written past its length"


@pytest.mark.parametrize("mode", [2, 1])
def test_noalloc_bfs_equals_pg_bfs(mode):
    torch = pytest.importorskip("torch")
    cap = 4096
    g = multi_hub_graph()
    ctx = pg.Context(g.n + 1, g.src.size + 16, 1, max_cites=1, bfs_cap=cap, seed=3)
    ctx.load(g)
    ctx.set_tuning(bfs_mode=mode)
    rng = np.random.default_rng(5)
    big = np.concatenate([[g.n - 1], np.arange(0, 32), rng.integers(0, g.n, 400),
                          [-1, g.n, 2**31 - 1]]).astype(np.int32)
    small = np.concatenate([rng.integers(0, g.n, 50), [g.n - 1, 3]]).astype(np.int32)
    log(f"no-alloc BFS mode={mode}: {big.size} roots, then {small.size}")
    out, ln = run_noalloc(torch, ctx, big, cap)
    check_same(ctx, big, out, ln, cap)
    assert ln[0] > 0 and ln.max() == cap
    out, ln = run_noalloc(torch, ctx, small, cap, stream=torch.cuda.Stream())
    check_same(ctx, small, out, ln, cap)
    # the oracle agrees on the largest root and a small one
    G = ob.Graph.from_download(ctx.download(), g.logfit)
    for i in (small.size - 2, small.size - 1):
        assert np.array_equal(out[i, :ln[i]], ob.bfs(G, int(small[i]), cap))
    ctx.close()


def test_noalloc_bfs_limits():
    torch = pytest.importorskip("torch")
    g = multi_hub_graph(n=2000, hubs=3, p=0.05)
    ctx = pg.Context(g.n + 1, g.src.size + 16, 1, max_cites=1, bfs_cap=64, seed=3)
    ctx.load(g)
    d = torch.zeros(16, dtype=torch.int32, device="cuda:0")
    ctx.bfs_noalloc(0, 0, 0, 0)  # nothing to do
    with pytest.raises(pg.PGError) as e:
        ctx.bfs_noalloc((1 << 20) + 1, d.data_ptr(), d.data_ptr(), d.data_ptr())
    assert e.value.code == 4 and "1048576" in str(e.value)
    with pytest.raises(pg.PGError) as e:
        ctx.bfs_noalloc(4, 0, d.data_ptr(), d.data_ptr())
    assert e.value.code == 1
    # the context's own entry points still work after the refusals
    roots = np.array([0, 1, g.n - 1], np.int32)
    out, ln = run_noalloc(torch, ctx, roots, 64)
    check_same(ctx, roots, out, ln, 64)
    ctx.close()
