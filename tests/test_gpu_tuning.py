"""Launch shapes change speed, never results (needs an MI355X).

The one-wavefront 2-hop tiers' candidate limits (which roots take one
wavefront and which a 512-thread workgroup), the 2-hop modes, the PA lanes
and the backward-merge tile are varied over three PA years on a graph with
hubs: the CSR and every year's draws must equal the default shapes' bit for
bit.  The scoring model's year is checked the same way over the merge tile.
pg_year_autotune times the candidates on the caller's schedule, keeps the
state it was called in, and its choice reproduces the default run.
"""
import hashlib

import numpy as np
import pytest

import paragraph_amd as pg
from gpu_checks import log, new_ctx
from paragraph_amd.workloads import reference_model_inputs

pytestmark = pytest.mark.gpu

KEYS = ("fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "in_deg", "birth")
M, NNB = 20, 6


def digest(d: dict) -> str:
    h = hashlib.sha1()
    for k in KEYS:
        h.update(np.ascontiguousarray(d[k]).tobytes())
    return h.hexdigest()


def graph():
    n0 = 60_000
    sched = pg.growth_schedule(n0, 0.05, 3)
    g = pg.seed_graph(31, n0, 12.0, n_final=n0 + sum(sched))
    return g, sched


def run_pa(ctx, sched, **tuning):
    ctx.set_tuning(**tuning)
    ctx.snapshot_restore()
    cits = []
    for yi, d in enumerate(sched):
        ctx.year_pa(1 + yi, yi, d, M, NNB, 1.2, 4.0)
        cits.append(ctx.citations(d, M)["cit"].copy())
    return digest(ctx.download()), cits


SHAPES = [dict(bfs_small_max=256, bfs_mid_max=512, merge_out=256),
          dict(bfs_small_max=512, bfs_mid_max=1024, merge_out=512),
          dict(bfs_small_max=1024, bfs_mid_max=2048, merge_out=2048, pa_lanes=32),
          dict(bfs_small_max=2048, bfs_mid_max=4096, merge_out=1024, bfs_mode=1),
          dict(bfs_small_max=256, bfs_mid_max=4096, merge_out=512, bfs_mode=2)]


def test_shapes_keep_pa_results():
    g, sched = graph()
    ctx = new_ctx(g, sched, M, cap=4096, seed=31)
    ctx.snapshot_save()
    ref, ref_cit = run_pa(ctx, sched)
    for sh in SHAPES:
        got, cit = run_pa(ctx, sched, **sh)
        log("shape", sh, "equal" if got == ref else "DIFFERENT")
        assert got == ref
        for a, b in zip(cit, ref_cit):
            assert np.array_equal(a, b)
    ctx.close()


def test_year_autotune_keeps_state_and_results():
    g, sched = graph()
    ctx = new_ctx(g, sched, M, cap=4096, seed=31)
    ctx.snapshot_save()
    ref, _ = run_pa(ctx, sched)
    ctx.snapshot_restore()
    before = digest(ctx.download())
    years = [dict(year=1 + yi, year_index=yi, n_new=d, m=M, n_nb=NNB, alpha=1.2, tau=4.0)
             for yi, d in enumerate(sched)]
    rep = ctx.year_autotune(years, reps=1)
    log("year autotune", rep["chosen"], f"{rep['chosen_ms']:.3f} ms (default {rep['default_ms']:.3f} ms)")
    assert len(rep["trials"]) == 11
    assert rep["chosen_ms"] == min(t["ms"] for t in rep["trials"])
    assert digest(ctx.download()) == before          # the state of the call
    for yi, d in enumerate(sched):                    # with the chosen shapes
        ctx.year_pa(1 + yi, yi, d, M, NNB, 1.2, 4.0)
    assert digest(ctx.download()) == ref
    ctx.close()


def test_merge_tile_keeps_scoring_model_results():
    inp = reference_model_inputs(2000)
    n0 = 40_000
    sched = pg.growth_schedule(n0, 0.05, 2)
    g = pg.seed_graph(32, n0, 8.0, n_final=n0 + sum(sched))
    ctx = pg.Context(n0 + sum(sched), int(g.src.size + sum(sched) * 260), max(sched),
                     max_cites=1, bfs_cap=16, seed=32)
    ctx.load(g)
    ctx.abm_configure(inp["bag"], inp["rec"], alpha=0.5, fully_random=0.05, same_year=0.12)
    ctx.snapshot_save()
    out = []
    for tile in (0, 256, 512, 2048):
        ctx.set_tuning(merge_out=tile)
        ctx.snapshot_restore()
        for yi, d in enumerate(sched):
            ctx.year_abm(1 + yi, yi, d)
        out.append(digest(ctx.download()))
    assert len(set(out)) == 1
    ctx.close()


def test_bad_shapes_are_refused():
    g, sched = graph()
    ctx = new_ctx(g, sched, M, cap=4096, seed=31)
    for bad in (dict(bfs_small_max=300), dict(bfs_mid_max=8192), dict(merge_out=4096),
                dict(merge_out=128)):
        with pytest.raises(pg.PGError, match="EINVAL"):
            ctx.set_tuning(**bad)
    ctx.close()
