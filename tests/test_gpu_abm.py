"""`abm` end to end on the GPU: reference file formats, and the graph it
writes equals the oracle's replay of the same years."""
import os
import subprocess

import numpy as np
import pytest

import paragraph_amd as pg
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def test_abm_writes_reference_formats_and_matches_oracle(tmp_path):
    d = str(tmp_path)
    n0, years, rate, m, f_nb = 3000, 3, 0.05, 8, 0.25
    g = pg.seed_graph(5, n0, 6.0)
    ids = np.arange(n0) * 3 + 100  # sparse original ids
    rng = np.random.default_rng(0)
    order = rng.permutation(n0)
    with open(os.path.join(d, "nodes.csv"), "w") as f:
        f.write("#node,year\n")
        for i in order:
            f.write(f"{ids[i]},{g.birth[i] + 2000}\n")
    with open(os.path.join(d, "edges.tsv"), "w") as f:
        for s, t in zip(g.src, g.dst):
            f.write(f"{ids[s]}\t{ids[t]}\n")
        f.write("999999\t100\n")  # unknown id: dropped
    open(os.path.join(d, "bag.csv"), "w").write(f"0,{m}\n")
    open(os.path.join(d, "rec.csv"), "w").write("0,1\n")
    j = lambda f: os.path.join(d, f)
    args = [pg.ABM_PATH, "--edgelist", j("edges.tsv"), "--nodelist", j("nodes.csv"),
            "--out-degree-bag", j("bag.csv"), "--recency-probabilities", j("rec.csv"),
            "--growth-rate", str(rate), "--num-cycles", str(years), "--same-year-proportion", "0",
            "--output-file", j("out.csv"), "--auxiliary-information-file", j("aux.csv"),
            "--log-file", j("log.txt"), "--alpha", str(f_nb), "--pa-exponent", "1.1",
            "--age-decay", "4", "--seed", "7"]
    print("running abm", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = open(j("out.csv")).read().splitlines()
    assert lines[0] == "#source,target"
    e = np.array([list(map(int, l.split(","))) for l in lines[1:]], np.int64)
    # oracle replay on dense ids (ascending original ids = dense order here)
    birth = (g.birth + 2000).astype(np.int32)
    sched = pg.growth_schedule(n0, rate, years)
    lf = np.empty(n0 + sum(sched), np.float32)
    pg.lib().pg_seed_logfit(7, 0, lf.size, 0.0, 0.5, lf.ctypes.data_as(pg.ctypes.POINTER(pg.ctypes.c_float)))
    G = ob.Graph.from_edges(n0, g.src, g.dst, birth, lf)
    near = 0
    n_nb = min(int(np.floor(f_nb * m + 0.5)), m - 1)
    for y, D in enumerate(sched):
        res = ob.pa_year(G, 7, y, int(birth.max()) + 1 + y, D, m, n_nb, 1.1, 4.0)
        near += int(res["near"].sum())
        off, tgt = ob.pa_compact(res)
        G = ob.csr_append(G, off, tgt, lf, int(birth.max()) + 1 + y)
    orig = np.concatenate([ids, ids[-1] + 1 + np.arange(G.n - n0)])
    src = np.repeat(np.arange(G.n), np.diff(G.fwd_off))
    ref = np.stack([orig[src], orig[G.fwd_adj]], 1)
    # this fixed input has no PA attempt near a CDF boundary, so the whole
    # replay is exact (with one, the years after it could diverge)
    assert near == 0
    assert np.array_equal(e, ref)
    aux = open(j("aux.csv")).read().splitlines()
    assert aux[0].startswith("node_id,type,year,pa_weight")
    assert len(aux) == 1 + G.n
    row = aux[1].split(",")
    assert row[1] == "seed" and row[-1] == "no_generators" and len(row) == 15
    row = aux[-1].split(",")
    assert row[1] == "agent" and len(row) == 15 and int(row[11]) == m


def test_abm_input_forms_equal_clean_input(tmp_path):
    """CRLF line ends, space delimiters, a repeated node (its last year
    counts), a self-citation and an unknown id read as the clean files do:
    both runs write the same bytes."""
    d = str(tmp_path)
    n0 = 2000
    g = pg.seed_graph(8, n0, 5.0)
    ids = np.arange(n0) + 50
    j = lambda f: os.path.join(d, f)
    with open(j("nodes.csv"), "w") as f:
        for i in range(n0):
            f.write(f"{ids[i]},{g.birth[i] + 2000}\n")
    with open(j("edges.csv"), "w") as f:
        for s, t in zip(g.src, g.dst):
            f.write(f"{ids[s]},{ids[t]}\n")
    with open(j("nodes_q.txt"), "w", newline="") as f:
        f.write("# node year\r\n")
        f.write(f"{ids[0]} 1990\r\n")                    # overridden below
        for i in range(n0):
            f.write(f"{ids[i]} {g.birth[i] + 2000}\r\n")
    with open(j("edges_q.txt"), "w", newline="") as f:
        f.write(f"{ids[10]} {ids[10]}\r\n")              # self-citation: dropped
        for s, t in zip(g.src, g.dst):
            f.write(f"{ids[s]} {ids[t]}\r\n")
        f.write(f"5 {ids[0]}\r\n")                       # unknown id: dropped
    open(j("bag.csv"), "w").write("0,6\n")
    open(j("rec.csv"), "w").write("0,1\n")

    def run(nodes, edges, tag):
        args = [pg.ABM_PATH, "--edgelist", j(edges), "--nodelist", j(nodes),
                "--out-degree-bag", j("bag.csv"), "--recency-probabilities", j("rec.csv"),
                "--growth-rate", "0.05", "--num-cycles", "2", "--same-year-proportion", "0",
                "--output-file", j(f"out_{tag}.csv"), "--auxiliary-information-file",
                j(f"aux_{tag}.csv"), "--log-file", j(f"log_{tag}.txt"), "--seed", "3"]
        print("running abm", tag, flush=True)
        r = subprocess.run(args, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert "I/O: read" in r.stdout

    run("nodes.csv", "edges.csv", "clean")
    run("nodes_q.txt", "edges_q.txt", "quirky")
    for f in ("out", "aux"):
        assert open(j(f"{f}_clean.csv"), "rb").read() == open(j(f"{f}_quirky.csv"), "rb").read(), f


def test_abm_nodelist_output_chains_runs(tmp_path):
    """--output-nodelist writes `node,year` for every paper: with the edgelist
    it is the input of a chained run.  A zero-year run on those files writes
    the same bytes back, and a run that grows on from them keeps the first
    run's graph as the prefix of its edgelist."""
    d = str(tmp_path)
    j = lambda f: os.path.join(d, f)
    n0 = 1500
    g = pg.seed_graph(9, n0, 5.0)
    ids = np.arange(n0) * 2 + 7
    with open(j("nodes.csv"), "w") as f:
        for i in range(n0):
            f.write(f"{ids[i]},{g.birth[i] + 2000}\n")
    with open(j("edges.csv"), "w") as f:
        for s, t in zip(g.src, g.dst):
            f.write(f"{ids[s]},{ids[t]}\n")
    open(j("bag.csv"), "w").write("0,6\n")
    open(j("rec.csv"), "w").write("0,1\n")

    def run(nodes, edges, years, tag):
        args = [pg.ABM_PATH, "--edgelist", j(edges), "--nodelist", j(nodes),
                "--out-degree-bag", j("bag.csv"), "--recency-probabilities", j("rec.csv"),
                "--growth-rate", "0.05", "--num-cycles", str(years), "--same-year-proportion", "0",
                "--output-file", j(f"out_{tag}.csv"), "--auxiliary-information-file",
                j(f"aux_{tag}.csv"), "--log-file", j(f"log_{tag}.txt"), "--seed", "5",
                "--output-nodelist", j(f"nodes_{tag}.csv")]
        print("running abm", tag, flush=True)
        r = subprocess.run(args, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr

    run("nodes.csv", "edges.csv", 2, "a")
    lines = open(j("nodes_a.csv")).read().splitlines()
    assert lines[0] == "#node,year"
    rows = np.array([list(map(int, l.split(","))) for l in lines[1:]], np.int64)
    sched = pg.growth_schedule(n0, 0.05, 2)
    assert rows.shape[0] == n0 + sum(sched)
    assert np.array_equal(rows[:n0, 0], ids) and np.array_equal(rows[:n0, 1], g.birth + 2000)
    new_ids = rows[n0:, 0]
    assert np.array_equal(new_ids, ids[-1] + 1 + np.arange(sum(sched)))
    y0 = int(g.birth.max()) + 2000 + 1
    assert np.array_equal(rows[n0:, 1], np.repeat([y0, y0 + 1], sched))
    # zero years on the first run's outputs: the same bytes back
    run("nodes_a.csv", "out_a.csv", 0, "b")
    for f in ("out", "nodes"):
        assert open(j(f"{f}_a.csv"), "rb").read() == open(j(f"{f}_b.csv"), "rb").read(), f
    # growing on: the first run's graph is the prefix of the edgelist
    run("nodes_a.csv", "out_a.csv", 1, "c")
    a_txt = open(j("out_a.csv")).read()
    c_txt = open(j("out_c.csv")).read()
    assert c_txt.startswith(a_txt) and len(c_txt) > len(a_txt)


def test_abm_reference_model_matches_oracle(tmp_path):
    """`--model reference` with the flags of drivers/ablation_parallel_gpu_10y.sh:
    a multi-value out-degree bag (values above 250 clamped), a recency table,
    equal weights, alpha 0.5, 5% fully random, 12% same-year: the edgelist
    equals the oracle's replay of the scoring model, and the auxiliary CSV
    carries each paper's out-degree, generator, weights, share and fitness."""
    d = str(tmp_path)
    n0, years, rate, seed = 4000, 3, 0.04, 9
    g = pg.seed_graph(8, n0, 6.0)
    ids = np.arange(n0) * 2 + 50
    with open(os.path.join(d, "nodes.csv"), "w") as f:
        f.write("#node,year\n")
        for i in range(n0):
            f.write(f"{ids[i]},{g.birth[i] + 1990}\n")
    with open(os.path.join(d, "edges.csv"), "w") as f:
        for s, t in zip(g.src, g.dst):
            f.write(f"{ids[s]},{ids[t]}\n")
    bag = [5, 6, 8, 12, 20, 33, 47, 90, 251, 400]
    with open(os.path.join(d, "bag.csv"), "w") as f:
        f.write("#idx,outdeg\n")
        for i, k in enumerate(bag):
            f.write(f"{i},{k}\n")
    rec = {1: 0.3, 2: 0.25, 3: 0.15, 5: 0.1, 10: 0.05, 20: 0.02}
    with open(os.path.join(d, "rec.csv"), "w") as f:
        f.write("#year_diff,probability\n")
        for k, v in rec.items():
            f.write(f"{k},{v}\n")
    j = lambda f: os.path.join(d, f)
    args = [pg.ABM_PATH, "--edgelist", j("edges.csv"), "--nodelist", j("nodes.csv"),
            "--out-degree-bag", j("bag.csv"), "--recency-probabilities", j("rec.csv"),
            "--alpha", "0.5", "--preferential-weight", "0.33", "--recency-weight", "0.33",
            "--fitness-weight", "0.33", "--growth-rate", str(rate), "--fully-random-citations",
            "0.05", "--num-cycles", str(years), "--same-year-proportion", "0.12",
            "--output-file", j("out.csv"), "--auxiliary-information-file", j("aux.csv"),
            "--log-file", j("log.txt"), "--num-processors", "16", "--log-level", "1",
            "--use-warp-bfs", "true", "--use-batching", "true", "--max-batch-size", "20000",
            "--use-multistage-kernel", "true", "--model", "reference", "--seed", str(seed)]
    print("running abm --model reference", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "2 out-degree bag values above 250" in r.stderr
    lines = open(j("out.csv")).read().splitlines()
    e = np.array([list(map(int, l.split(","))) for l in lines[1:]], np.int64)
    birth = (g.birth + 1990).astype(np.int32)
    sched = pg.growth_schedule(n0, rate, years)
    lf = np.zeros(n0 + sum(sched), np.float32)
    G = ob.Graph.from_edges(n0, g.src, g.dst, birth, lf)
    table = np.zeros(21)
    for k, v in rec.items():
        table[k] = v
    m = ob.AbmModel(seed, bag, table, alpha=0.5, w_pa=0.33, w_rec=0.33, w_fit=0.33, frc=0.05,
                    syp=0.12)
    infos = []
    for y, D in enumerate(sched):
        yr = int(birth.max()) + 1 + y
        res = ob.abm_year(G, m, yr, y, D)
        infos.append(res["info"])
        G = ob.csr_append(G, res["off"], res["tgt"], lf, yr)
    orig = np.concatenate([ids, ids[-1] + 1 + np.arange(G.n - n0)])
    src = np.repeat(np.arange(G.n), np.diff(G.fwd_off))
    ref = np.stack([orig[src], orig[G.fwd_adj]], 1)
    assert np.array_equal(e, ref)
    info = np.concatenate(infos)
    aux = [l.split(",") for l in open(j("aux.csv")).read().splitlines()[1:]]
    assert len(aux) == G.n
    fit = ob.abm_fitness(seed, range(G.n))
    for v in range(G.n):
        row = aux[v]
        assert int(row[7]) == fit[v] and row[6] == "0" and row[8] == "1000"
        assert int(row[10]) == G.in_deg[v] and int(row[11]) == G.fwd_off[v + 1] - G.fwd_off[v]
        if v < n0:
            assert row[1] == "seed" and row[14] == "no_generators"
        else:
            gen, k = info[v - n0, 0], info[v - n0, 1]
            assert row[1] == "agent" and int(row[12]) == k and int(row[14]) == orig[gen]
            assert [float(x) for x in row[3:6]] == [float(f"{1 / 3:.10g}")] * 3
            assert float(row[9]) == 0.5
