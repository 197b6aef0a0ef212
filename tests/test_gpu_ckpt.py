"""Exact year-granular checkpoint and resume (needs an MI355X).

A context that loads a checkpoint written after year k must continue bit
for bit like the context that wrote it: both CSRs, degrees, births and every
draw of the later years.  Also the file checks (magic, checksum, seed,
capacities) and `abm --checkpoint-file / --resume-from` end to end.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import paragraph_amd as pg

pytestmark = pytest.mark.gpu

KEYS = ("fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "in_deg", "birth", "logfit")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_ctx(g, sched, m, seed=3):
    n_final = g.n + sum(sched)
    ctx = pg.Context(n_final, g.src.size + sum(sched) * m + 16, max(sched), max_cites=m, seed=seed)
    return ctx


def pa_years(ctx, sched, y0, y1, m=10, n_nb=3):
    out = []
    for y in range(y0, y1):
        st = ctx.year_pa(1 + y, y, sched[y], m, n_nb, 1.1, 4.0)
        out.append((st.edges_added, ctx.citations(sched[y], m)["cit"].copy()))
    return out


def same_state(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


def test_resume_continues_bit_for_bit(tmp_path):
    n0, m = 20000, 10
    sched = pg.growth_schedule(n0, 0.05, 5)
    g = pg.seed_graph(21, n0, 8.0, n_final=n0 + sum(sched))
    path = str(tmp_path / "run.pgck")
    a = make_ctx(g, sched, m)
    a.load(g)
    pa_years(a, sched, 0, 2)
    a.checkpoint_save(path, b"years 0-1")
    assert a.progress() == (2, 2, 1)
    log("checkpoint written:", os.path.getsize(path), "bytes")
    rest_a = pa_years(a, sched, 2, 5)
    b = make_ctx(g, sched, m)          # no graph loaded: the checkpoint supplies it
    assert b.checkpoint_load(path) == b"years 0-1"
    assert b.progress() == (2, 2, 1)
    rest_b = pa_years(b, sched, 2, 5)
    for (ea, ca), (eb, cb) in zip(rest_a, rest_b):
        assert ea == eb and np.array_equal(ca, cb)
    same_state(a.download(), b.download())
    assert b.progress() == (5, 5, 4)
    # a snapshot keeps the progress too
    b.snapshot_save()
    pa_years(b, sched, 5, 5)
    b.snapshot_restore()
    assert b.progress() == (5, 5, 4)


def test_resume_er_years(tmp_path):
    n0, p = 30000, 2e-4
    sched = pg.growth_schedule(n0, 0.03, 4)
    none = np.zeros(0, np.int32)
    g = pg.SeedGraph(np.zeros(n0, np.int32), none, none, np.zeros(n0 + sum(sched), np.float32))
    cap = int(n0 * sum(sched) * p * 2 + 10000)
    mk = lambda: pg.Context(n0 + sum(sched), cap, max(sched), max_cites=1, bfs_cap=16, seed=5)
    a = mk()
    a.load(g)
    for y in range(2):
        a.year_er(1 + y, y, sched[y], p)
    path = str(tmp_path / "er.pgck")
    a.checkpoint_save(path)
    b = mk()
    b.checkpoint_load(path)
    for ctx in (a, b):
        for y in range(2, 4):
            ctx.year_er(1 + y, y, sched[y], p)
    same_state(a.download(), b.download())


def test_checkpoint_errors(tmp_path):
    n0, m = 3000, 6
    sched = pg.growth_schedule(n0, 0.05, 2)
    g = pg.seed_graph(22, n0, 6.0, n_final=n0 + sum(sched))
    a = make_ctx(g, sched, m)
    with pytest.raises(pg.PGError, match="ESTATE"):
        a.checkpoint_save(str(tmp_path / "none.pgck"))        # nothing loaded
    a.load(g)
    pa_years(a, sched, 0, 1, m=m, n_nb=2)
    path = str(tmp_path / "a.pgck")
    a.checkpoint_save(path, b"x" * 100)
    with pytest.raises(pg.PGError, match="EIO"):
        make_ctx(g, sched, m).checkpoint_load(str(tmp_path / "missing.pgck"))
    with pytest.raises(pg.PGError, match="EINVAL"):           # another Philox seed
        make_ctx(g, sched, m, seed=4).checkpoint_load(path)
    with pytest.raises(pg.PGError, match="ECAP"):             # blob larger than the buffer
        make_ctx(g, sched, m).checkpoint_load(path, user_cap=10)
    small = pg.Context(n0 + 10, g.src.size + 16, 16, max_cites=m, seed=3)
    with pytest.raises(pg.PGError, match="ECAP"):             # does not fit the context
        small.checkpoint_load(path)
    raw = bytearray(open(path, "rb").read())
    bad = str(tmp_path / "bad.pgck")
    raw2 = bytearray(raw)
    raw2[200] ^= 0xFF                                         # payload byte: checksum
    open(bad, "wb").write(raw2)
    with pytest.raises(pg.PGError, match="EIO"):
        make_ctx(g, sched, m).checkpoint_load(bad)
    open(bad, "wb").write(raw[:len(raw) // 2])                # truncated
    with pytest.raises(pg.PGError, match="EIO"):
        make_ctx(g, sched, m).checkpoint_load(bad)
    raw2 = bytearray(raw)
    raw2[0:8] = b"NOTACKPT"                                   # magic
    open(bad, "wb").write(raw2)
    with pytest.raises(pg.PGError, match="EIO"):
        make_ctx(g, sched, m).checkpoint_load(bad)


def test_abm_checkpoint_and_resume(tmp_path):
    d = str(tmp_path)
    n0, rate, m = 2500, 0.05, 8
    g = pg.seed_graph(9, n0, 6.0)
    with open(os.path.join(d, "nodes.csv"), "w") as f:
        for i in range(n0):
            f.write(f"{i + 10},{g.birth[i] + 2000}\n")
    with open(os.path.join(d, "edges.csv"), "w") as f:
        for s, t in zip(g.src, g.dst):
            f.write(f"{s + 10},{t + 10}\n")
    open(os.path.join(d, "bag.csv"), "w").write(f"0,{m}\n")
    open(os.path.join(d, "rec.csv"), "w").write("0,1\n")
    j = lambda f: os.path.join(d, f)

    def abm(years, out, extra):
        args = [pg.ABM_PATH, "--edgelist", j("edges.csv"), "--nodelist", j("nodes.csv"),
                "--out-degree-bag", j("bag.csv"), "--recency-probabilities", j("rec.csv"),
                "--growth-rate", str(rate), "--num-cycles", str(years),
                "--same-year-proportion", "0", "--output-file", j(out + ".csv"),
                "--auxiliary-information-file", j(out + "_aux.csv"), "--log-file", j(out + ".log"),
                "--alpha", "0.3", "--seed", "11"] + extra
        log("abm", years, extra)
        r = subprocess.run(args, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        return r

    abm(4, "full", [])
    abm(2, "half", ["--checkpoint-file", j("ck.pgck")])
    r = abm(4, "resumed", ["--resume-from", j("ck.pgck")])
    assert "Growth: 2 years" in r.stdout
    assert open(j("full.csv")).read() == open(j("resumed.csv")).read()
    assert open(j("full_aux.csv")).read() == open(j("resumed_aux.csv")).read()
    # a checkpoint of another run is refused
    r = subprocess.run([pg.ABM_PATH, "--edgelist", j("edges.csv"), "--nodelist", j("nodes.csv"),
                        "--out-degree-bag", j("bag.csv"), "--recency-probabilities", j("rec.csv"),
                        "--growth-rate", str(rate), "--num-cycles", "1", "--same-year-proportion",
                        "0", "--output-file", j("x.csv"), "--auxiliary-information-file",
                        j("x_aux.csv"), "--log-file", j("x.log"), "--seed", "11",
                        "--resume-from", j("ck.pgck")], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
