"""Benchmark: citation edges/s of the 10-year PA growth on one MI355X, and
likelihood-gradient evaluations/s on a recorded 30-year sequence.

Workload (one step): the 2,000,000-paper synthetic seed graph (birth years
uniform in [-20,0], Poisson(10) citations to uniformly chosen older papers,
lognormal(0,0.5) fitness), grown for 10 years at 3% per year with 20
citations per new paper (14 PA + 6 two-hop neighbourhood draws, frontier cap
4096), alpha=1, tau=5 years, Philox seed 1: 687,837 new papers and
13,756,740 citation edges.  Each step starts from the same seed graph
(device snapshot restored between steps, outside the timed region).

Prints one JSON line on stdout.  Details go to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import paragraph_amd as pg  # noqa: E402
from paragraph_amd.workloads import ensemble_params  # noqa: E402

PEAK_HBM_GBS = 8000.0

# Kernels of each timed group, as they appear in the rocprofv3 summaries.
GROUP_KERNELS = {
    "bfs": ["k_bfs_small", "k_bfs_mid", "k_bfs_root"],
    "lik_grad": ["k_lik_papers<false>", "k_lik_year_z", "k_lik_cite<true, 4, 5>", "k_lik_year_ev",
                 "k_lik_segsum", "k_lik_papers<true>", "k_lik_final"],
    "lik_batch16": ["k_likb_papers<false, 16, false>", "k_likb_year_z<16>",
                    "k_likb_cite<4, 5, 16, false>", "k_likb_year_ev<16>", "k_likb_segjob<16>",
                    "k_likb_bigsum<16>", "k_likb_papers<true, 16, false>", "k_likb_final<16>"],
    "lik_batch16_fast": ["k_likb_papers<false, 16, true>", "k_likb_year_z<16>",
                         "k_likb_cite<4, 5, 16, true>", "k_likb_year_ev<16>", "k_likb_segjob<16>",
                         "k_likb_bigsum<16>", "k_likb_papers<true, 16, true>", "k_likb_final<16>"],
    "er": ["k_er_count", "k_er_write"],
    "ens_bfs": ["k_bfs_small", "k_bfs_mid", "k_bfs_root"],
    "abm_select": ["k_abm_select"],
}


def measured_traffic(summary: str, group: str):
    """Memory-side bytes per launch of a kernel group: the sum over its
    kernels of the read and write bytes per dispatch in the committed
    rocprofv3 summary profiles/<summary>: L2-to-fabric requests counted by
    size (TCC_EA0_RDREQ_32B/64B/128B, TCC_EA0_WRREQ/_64B; two --pmc runs of
    the same workload).  None when the summary or a kernel row is missing."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    rows = {}
    for line in open(path):
        c = [x.strip() for x in line.strip().strip("|").split("|")]
        if len(c) == 4 and c[1].isdigit():
            try:
                rows[c[0]] = (float(c[2]) + float(c[3])) * 1e6
            except ValueError:
                pass
    ks = GROUP_KERNELS.get(group, [])
    if not ks or any(k not in rows for k in ks):
        return None
    return sum(rows[k] for k in ks)


def roofline(kernel: str, achieved: float, summary: str, group: str) -> dict:
    t = measured_traffic(summary, group)
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": t,
            "traffic_note": None if t is None else
            f"bytes per launch of {', '.join(GROUP_KERNELS[group])}: L2-to-fabric read and "
            f"write requests by size per dispatch from profiles/{summary} (rocprofv3 --pmc)"}


def state_digest(ctx) -> str:
    """SHA-1 of the context's graph (both CSRs, in-degrees, births): every
    timed step starts from the same snapshot, so all must agree."""
    import hashlib
    d = ctx.download()
    h = hashlib.sha1()
    for k in ("fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "in_deg", "birth"):
        h.update(np.ascontiguousarray(d[k]).tobytes())
    return h.hexdigest()


def eprint(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(args):
    n0, years, rate, m, f_nb = args.n0, args.years, 0.03, 20, 0.3
    sched = pg.growth_schedule(n0, rate, years)
    n_nb = min(int(math.floor(f_nb * m + 0.5)), m - 1)
    return n0, sched, m, n_nb


def cpu_baseline(G0, sched, m, n_nb, alpha, tau, seed, cap):
    """The whole workload (all years, draws + CSR append) on the C oracle,
    OpenMP over the papers of a year, on this host's cores."""
    from oracle import binding as ob
    cores = len(os.sched_getaffinity(0))
    G = G0
    edges = 0
    t0 = time.perf_counter()
    for yi, D in enumerate(sched):
        res = ob.pa_year(G, seed, yi, 1 + yi, D, m, n_nb, alpha, tau, cap, threads=cores)
        off, tgt = ob.pa_compact(res)
        G = ob.csr_append(G, off, tgt, G0.logfit, 1 + yi)
        edges += int(res["cnt"].sum())
    dt = time.perf_counter() - t0
    # one thread on a bounded sample: the first two years
    G1, e1 = G0, 0
    t1 = time.perf_counter()
    for yi, D in enumerate(sched[:2]):
        res = ob.pa_year(G1, seed, yi, 1 + yi, D, m, n_nb, alpha, tau, cap, threads=1)
        off, tgt = ob.pa_compact(res)
        G1 = ob.csr_append(G1, off, tgt, G0.logfit, 1 + yi)
        e1 += int(res["cnt"].sum())
    dt1 = time.perf_counter() - t1
    return {"value": edges / dt, "unit": "citation edges/s", "cores": cores, "kind": "port",
            "sample": f"the full workload ({len(sched)} years, {edges} edges) on the C oracle "
                      f"with {cores} OpenMP threads: {dt:.2f} s",
            "host_cpu": host_cpu(),
            "one_thread": {"value": e1 / dt1, "unit": "citation edges/s", "cores": 1,
                           "sample": f"years 0-1 ({e1} edges, draws + CSR append) on one "
                                     f"thread: {dt1:.2f} s"}}, G


def host_cpu() -> str:
    """The host CPU model (lscpu's model name) and its logical CPU count."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs, {len(os.sched_getaffinity(0))} usable"


def degree_ks(gpu_in_deg, cpu_in_deg) -> dict:
    """Two-sample Kolmogorov-Smirnov statistic between the final in-degree
    distributions of the GPU run and the oracle's run of the same workload,
    plus the number of papers whose in-degree differs (PA draws differ only
    where a uniform fell within 1e-12 relative of a CDF boundary)."""
    from scipy import stats
    r = stats.ks_2samp(gpu_in_deg, cpu_in_deg)
    return {"statistic": float(r.statistic), "pvalue": float(r.pvalue),
            "papers": int(gpu_in_deg.size),
            "papers_differing_in_degree": int(np.count_nonzero(gpu_in_deg != cpu_in_deg))}


def abm_bench(args) -> dict:
    """`abm` end to end on CSV files: a 2,000,000-paper Philox seed graph
    (Poisson(10) citations) written as `node,year` and `citing,cited` CSVs,
    then `abm` grows it 10 years at 3% (m=20 from the out-degree bag, 30%
    neighbourhood) and writes the `#source,target` edgelist and the
    15-column auxiliary CSV.  Reported: the process wall time and abm's own
    read, growth and write times."""
    import subprocess
    import tempfile
    import pandas as pd
    with tempfile.TemporaryDirectory() as d:
        g = pg.seed_graph(1, 2_000_000, 10.0)
        pd.DataFrame({"a": np.arange(g.n) + 100, "b": g.birth.astype(np.int64) + 2000}).to_csv(
            f"{d}/nodes.csv", header=False, index=False)
        with open(f"{d}/edges.csv", "w") as f:
            f.write("#source,target\n")
        pd.DataFrame({"a": g.src.astype(np.int64) + 100, "b": g.dst.astype(np.int64) + 100}).to_csv(
            f"{d}/edges.csv", header=False, index=False, mode="a")
        open(f"{d}/bag.csv", "w").write("0,20\n")
        open(f"{d}/rec.csv", "w").write("0,1\n")
        cmd = [pg.ABM_PATH, "--edgelist", f"{d}/edges.csv", "--nodelist", f"{d}/nodes.csv",
               "--out-degree-bag", f"{d}/bag.csv", "--recency-probabilities", f"{d}/rec.csv",
               "--growth-rate", "0.03", "--num-cycles", "10", "--same-year-proportion", "0",
               "--output-file", f"{d}/out.csv", "--auxiliary-information-file", f"{d}/aux.csv",
               "--log-file", f"{d}/log.txt", "--seed", "1"]
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode:
            raise RuntimeError(f"abm failed: {r.stderr[-500:]}")
        out = {"metric": "abm wall time (2M-paper CSV seed, 10y PA, CSV outputs)", "value": wall,
               "unit": "s", "higher_is_better": False,
               "seed_edges": int(g.src.size), "edgelist_bytes": os.path.getsize(f"{d}/out.csv"),
               "aux_bytes": os.path.getsize(f"{d}/aux.csv")}
        for line in r.stdout.splitlines():
            if line.startswith("Growth:") or line.startswith("I/O:"):
                eprint(line)
                out["growth_line" if line.startswith("Growth") else "io_line"] = line
        return out


def er_bench(args) -> dict:
    """ER G(n,p) growth: 1,000,000 isolated seed papers, 10 years at 3% per
    year, each new paper linked to each existing paper with p = 1e-5
    (geometric skips), Philox seed 1.  Timed: the 10 device year steps (edge
    formation + CSR append), the seed restored from a device snapshot between
    steps outside the timed region.  Year 0's edge set is compared with the
    oracle's, and the oracle times the whole workload on the host cores."""
    n0, years, rate, p, seed = 1_000_000, 10, 0.03, 1e-5, 1
    sched = pg.growth_schedule(n0, rate, years)
    n_final = n0 + sum(sched)
    none = np.zeros(0, np.int32)
    g = pg.SeedGraph(np.zeros(n0, np.int32), none, none, np.zeros(n_final, np.float32))
    mean_e, n = 0.0, n0
    for d in sched:
        mean_e += n * d * p
        n += d
    e_cap = int(mean_e * 1.2 + 200_000)
    ctx = pg.Context(n_final, e_cap, max(sched), max_cites=1, bfs_cap=16, seed=seed)
    ctx.load(g)
    ctx.snapshot_save()

    def run():
        e = 0
        for yi, d in enumerate(sched):
            e += ctx.year_er(1 + yi, yi, d, p).edges_added
        return e

    run()                      # warm-up
    ctx.snapshot_restore()
    # year 0 against the oracle: the new papers' forward rows
    from oracle import binding as ob
    ctx.year_er(1, 0, sched[0], p)
    G = ctx.download()
    off_ref, tgt_ref = ob.er_year(n0, sched[0], p, seed, 0)
    fo = G["fwd_off"][n0:n0 + sched[0] + 1].astype(np.int64)
    tgt_gpu = G["fwd_adj"][fo[0]:fo[-1]]
    ok = (np.array_equal(fo - fo[0], off_ref.astype(np.int64)) and
          np.array_equal(tgt_gpu, tgt_ref))
    if not ok:
        raise RuntimeError("ER year 0 differs from the oracle")
    ctx.snapshot_restore()
    ctx.profile_reset()
    ctx.profile(True)
    walls, edges = [], 0
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        edges = run()
        walls.append(time.perf_counter() - t0)
        ctx.snapshot_restore()
    prof = ctx.profile_read()
    ctx.profile(False)
    ctx.close()
    wall = float(np.median(walls))
    er = prof["er"]
    achieved = (er["bytes"] / max(er["launches"], 1)) / (er["ms"] / max(er["launches"], 1) * 1e-3) / 1e9
    eprint(f"ER: {edges} edges in {wall * 1e3:.2f} ms ({edges / wall:.3e} edges/s); "
           f"er kernels {er['ms'] / len(walls):.2f} ms, csr_append "
           f"{prof['csr_append']['ms'] / len(walls):.2f} ms per run; year 0 equals the oracle")
    res = {"metric": "ER citation edges/s (10y, 1M isolated seed, p=1e-5)", "value": edges / wall,
           "unit": "citation edges/s", "ms_per_step": wall * 1e3, "edges_per_step": edges,
           "year0_equals_oracle": ok,
           "kernel_ms_per_step": {k: v["ms"] / len(walls) for k, v in prof.items() if v["launches"]},
           "roofline": dict(roofline("er (k_er_count + k_er_write)", achieved,
                                     "er_10y_1M_kernels.md", "er"),
                            bound_note="4 B per target + 4 B per new paper: about 16 MB per "
                                       "run; the kernels are latency- and FP64-log-bound")}
    if not args.no_cpu_baseline:
        cores = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        ce, n = 0, n0
        for yi, d in enumerate(sched):
            _, tgt = ob.er_year(n, d, p, seed, yi, threads=cores)
            ce += tgt.size
            n += d
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": ce / dt, "unit": "citation edges/s", "cores": cores,
                               "kind": "port",
                               "sample": f"the edge formation of all {years} years ({ce} edges) on "
                                         f"the C oracle with {cores} OpenMP threads: {dt:.2f} s "
                                         f"(no CSR append)"}
        if ce != edges:
            raise RuntimeError(f"ER oracle {ce} edges, GPU {edges}")
    return res


def ref_model_bench(args) -> dict:
    """The reference's scoring model (PG_MODEL_ABM, `abm --model reference`):
    the headline workload's 2,000,000-paper seed graph grown 10 years at 3%
    per year with the inputs of workloads.reference_model_inputs (out-degree
    bag with mean about 19, recency table, alpha 0.5, equal weights, 5% fully
    random, 12% same-year), Philox seed 1.  Timed: the 10 device year steps
    (scores, same-year subset, 1-hop/2-hop selection, rows, CSR append), the
    seed restored from a device snapshot between steps outside the timed
    region.  Year 0 is compared bit for bit with the oracle (scores, per-paper
    draws and counts, rows), and the oracle times the workload on the host."""
    from oracle import binding as ob
    from paragraph_amd.workloads import reference_model_inputs
    n0, years, rate, seed = args.n0, args.years, 0.03, 1
    sched = pg.growth_schedule(n0, rate, years)
    n_final = n0 + sum(sched)
    g = pg.seed_graph(seed, n0, 10.0, n_final=n_final)
    inp = reference_model_inputs()
    e_cap = int(g.src.size + sum(sched) * (int(inp["bag"].max()) + 2) + 16)
    ctx = pg.Context(n_final, e_cap, max(sched), max_cites=1, bfs_cap=16, seed=seed)
    ctx.load(g)
    cfg = dict(alpha=inp["alpha"], w_pa=inp["w_pa"], w_rec=inp["w_rec"], w_fit=inp["w_fit"],
               fully_random=inp["fully_random"], same_year=inp["same_year"])
    ctx.abm_configure(inp["bag"], inp["rec"], **cfg)
    ctx.snapshot_save()
    year0 = 1   # calendar year of the first new papers (seed births are -20..0)

    def run():
        e, cand = 0, [0, 0]
        for yi, d in enumerate(sched):
            e += ctx.year_abm(year0 + yi, yi, d).edges_added
            c1, c2 = ctx.abm_year_counts()
            cand[0] += c1
            cand[1] += c2
        return e, cand

    run()                      # warm-up
    ctx.snapshot_restore()
    # year 0 against the oracle, bit for bit
    G0 = ob.Graph.from_download(ctx.download(), g.logfit)
    m = ob.AbmModel(seed, inp["bag"], inp["rec"], alpha=inp["alpha"], w_pa=inp["w_pa"],
                    w_rec=inp["w_rec"], w_fit=inp["w_fit"], frc=inp["fully_random"],
                    syp=inp["same_year"])
    cores = len(os.sched_getaffinity(0))
    ctx.year_abm(year0, 0, sched[0])
    ref = ob.abm_year(G0, m, year0, 0, sched[0], threads=cores)
    after = ctx.download()
    fo = after["fwd_off"]
    ok = (np.array_equal(ctx.abm_year_info(sched[0]), ref["info"]) and
          np.array_equal(ctx.abm_scores(n0).view(np.uint32), ref["comp"].view(np.uint32)) and
          np.array_equal(fo[n0:n0 + sched[0] + 1] - fo[n0], ref["off"]) and
          np.array_equal(after["fwd_adj"][fo[n0]:fo[n0 + sched[0]]], ref["tgt"]))
    if not ok:
        raise RuntimeError("scoring-model year 0 differs from the oracle")
    ctx.snapshot_restore()
    ctx.profile_reset()
    ctx.profile(True)
    walls, edges, cand = [], 0, [0, 0]
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        edges, cand = run()
        walls.append(time.perf_counter() - t0)
        ctx.snapshot_restore()
    prof = ctx.profile_read()
    ctx.profile(False)
    ctx.close()
    wall = float(np.median(walls))
    steps = len(walls)
    groups = ["abm_scores", "abm_select", "abm_rows", "csr_append"]
    dom = max(groups, key=lambda k: prof[k]["ms"])
    dv = prof[dom]
    achieved = (dv["bytes"] / dv["launches"]) / (dv["ms"] / dv["launches"] * 1e-3) / 1e9
    eprint(f"reference model: {edges} edges in {wall * 1e3:.2f} ms ({edges / wall:.3e} edges/s); "
           + ", ".join(f"{k} {prof[k]['ms'] / steps:.2f} ms" for k in groups) +
           f" per run; candidates keyed {cand[0]} (1-hop) + {cand[1]} (2-hop); year 0 equals "
           f"the oracle")
    res = {"metric": "citation edges/s (reference scoring model, 10y, 2M seed)",
           "value": edges / wall, "unit": "citation edges/s", "ms_per_step": wall * 1e3,
           "edges_per_step": edges, "year0_equals_oracle": ok,
           "candidates_per_step": {"one_hop": cand[0], "two_hop": cand[1]},
           "config": {"workload": f"reference scoring model {years}y at 3%/y from a {n0}-paper "
                                  f"seed, bag mean {float(inp['bag'].mean()):.1f}, alpha 0.5, "
                                  f"weights 0.33, 5% fully random, 12% same-year"},
           "kernel_ms_per_step": {k: v["ms"] / steps for k, v in prof.items() if v["launches"]},
           "roofline": dict(roofline(dom, achieved, "reference_10y_2M_kernels.md", dom),
                            bound_note="selection: 4 B per candidate entry plus its 16 B score "
                                       "record (and 16 B row offsets per 1-hop entry); each "
                                       "entry also costs a Philox draw, a float log and an "
                                       "FP64 divide")}
    if not args.no_cpu_baseline:
        G = G0
        ce = 0
        t0 = time.perf_counter()
        for yi, d in enumerate(sched):
            r = ob.abm_year(G, m, year0 + yi, yi, d, threads=cores)
            G = ob.csr_append(G, r["off"], r["tgt"], g.logfit, year0 + yi)
            ce += r["tgt"].size
        dt = time.perf_counter() - t0
        t1 = time.perf_counter()
        r1 = ob.abm_year(G0, m, year0, 0, sched[0], threads=1)
        dt1 = time.perf_counter() - t1
        res["cpu_baseline"] = {
            "value": ce / dt, "unit": "citation edges/s", "cores": cores, "kind": "port",
            "sample": f"the full workload ({years} years, {ce} edges, selection + CSR append) on "
                      f"the C oracle with {cores} OpenMP threads: {dt:.2f} s",
            "host_cpu": host_cpu(),
            "one_thread": {"value": r1["tgt"].size / dt1, "unit": "citation edges/s", "cores": 1,
                           "sample": f"year 0 ({r1['tgt'].size} edges, no append) on one thread: "
                                     f"{dt1:.2f} s"}}
        if ce != edges:
            raise RuntimeError(f"scoring model: oracle {ce} edges, GPU {edges}")
    return res


def likelihood_bench(args) -> dict:
    """Likelihood-gradient evaluations per second on the 30-year sequence.

    Sequence: a 1,000,000-paper seed (Poisson(8) citations, births uniform
    in [-20,0], lognormal(0,0.5) fitness) grown 30 years at 3% per year with
    20 citations per paper (14 PA + 6 neighbourhood), alpha=1.2, tau=4,
    Philox seed 4, recorded on the device: 1,427,286 citing papers,
    28,545,720 slots, 2,427,286 log-fitness parameters.  Timed: the fused
    pg_pa_loglik_grad (log L, dL/dalpha, dL/dtau, dL/dphi) at alpha=1,
    tau=5, back to back on one stream; beside it the same gradient with stock
    PyTorch autograd (FP64) on the same GPU, and the FP64 C oracle on one
    host core.  Also checks the full-size gradient against the oracle."""
    import torch
    from paragraph_amd.likelihood import (PALogLikelihood, PlainSequence, fit_adam,
                                          loglik_plain_torch)
    n0, years, rate, m, n_nb = 1_000_000, 30, 0.03, 20, 6
    alpha_g, tau_g, seed = 1.2, 4.0, 4
    sched = pg.growth_schedule(n0, rate, years)
    n_final = n0 + sum(sched)
    t0 = time.perf_counter()
    g = pg.seed_graph(seed, n0, 8.0, n_final=n_final)
    ctx = pg.Context(n_final, g.src.size + sum(sched) * m + 16, max(sched), max_cites=m,
                     bfs_cap=4096, seed=seed)
    ctx.load(g)
    paper_years, n = 0, n0
    for d in sched:
        paper_years += n
        n += d
    ctx.seq_reserve(m, paper_years, sum(sched))
    ctx.seq_record(True)
    for yi, d in enumerate(sched):
        ctx.year_pa(1 + yi, yi, d, m, n_nb, alpha_g, tau_g)
    ctx.seq_record(False)
    info = ctx.seq_info()
    eprint(f"likelihood sequence: {info.n_years} years, {info.paper_years} paper-years, "
           f"{info.new_papers} citing papers, {info.n_papers} papers "
           f"({time.perf_counter() - t0:.1f} s)")
    dev = torch.device("cuda", 0)
    phi = torch.from_numpy(np.ascontiguousarray(g.logfit[:n_final])).to(dev)
    grad = torch.empty(n_final, dtype=torch.float64, device=dev)
    sc = torch.empty(3, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    a0, t0_ = 1.0, 5.0

    tune = None
    if not args.no_tune:
        tune = ctx.lik_autotune(a0, t0_, phi.data_ptr(), reps=5)
        eprint(f"likelihood autotune: chosen {tune['chosen']['lik_lanes']} lanes "
               f"{tune['chosen_ms']:.3f} ms (default {tune['default_ms']:.3f} ms, results within "
               f"{tune['max_rel_dev']:.1e}); " +
               ", ".join(f"{t['lik_lanes']}: {t['ms']:.3f}" for t in tune["trials"]))

    def fused():
        ctx.pa_loglik_grad(a0, t0_, phi.data_ptr(), grad.data_ptr(), scalars_dev=sc.data_ptr(),
                           stream=stream, host=False)

    for _ in range(3):
        fused()
    torch.cuda.synchronize()
    K = args.lik_evals
    t1 = time.perf_counter()
    for _ in range(K):
        fused()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    evals_s = K / dt
    # device time and algorithmic bytes per evaluation (events around the passes)
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(20):
        fused()
    prof = ctx.profile_read()["lik_grad"]
    ctx.profile(False)
    dev_ms = prof["ms"] / prof["launches"]
    bytes_eval = prof["bytes"] / prof["launches"]
    achieved = bytes_eval / (dev_ms * 1e-3) / 1e9
    ll, ga, gt = (float(x) for x in sc.cpu())
    eprint(f"fused gradient: {evals_s:.1f} evals/s ({1e3 / evals_s:.3f} ms wall, {dev_ms:.3f} ms "
           f"device, {bytes_eval / 1e9:.3f} GB algorithmic, {achieved:.0f} GB/s); "
           f"log L {ll:.10g} dL/dalpha {ga:.8g} dL/dtau {gt:.8g}")
    batched, g_b0 = batched_likelihood(ctx, phi, n_final, a0, t0_, stream, args,
                                       (ll, ga, gt), grad)
    seq = ctx.seq_download()
    autograd_s = plain_s = None
    if not args.lik_no_torch:
        # torch.autograd.Function (forward + backward, device-resident parameters)
        at = torch.tensor([a0, t0_], dtype=torch.float64, device=dev, requires_grad=True)
        phr = phi.clone().requires_grad_(True)
        K2 = max(K // 4, 5)
        for i in range(K2 + 2):
            if i == 2:
                torch.cuda.synchronize()
                t2 = time.perf_counter()
            at.grad = None
            phr.grad = None
            PALogLikelihood.apply(ctx, at, phr).backward()
        torch.cuda.synchronize()
        autograd_s = K2 / (time.perf_counter() - t2)
        # plain PyTorch autograd, FP64
        ps = PlainSequence(seq, dev)
        at2 = torch.tensor([a0, t0_], dtype=torch.float64, device=dev, requires_grad=True)
        ph2 = phi.to(torch.float64).requires_grad_(True)
        K3 = 3
        for i in range(K3 + 1):
            if i == 1:
                torch.cuda.synchronize()
                t3 = time.perf_counter()
            at2.grad = None
            ph2.grad = None
            ll2 = loglik_plain_torch(ps, at2, ph2)
            ll2.backward()
        torch.cuda.synchronize()
        plain_s = K3 / (time.perf_counter() - t3)
        g_plain = ph2.grad.cpu().numpy()
        ga2, gt2 = at2.grad.cpu().numpy()
        eprint(f"autograd.Function: {autograd_s:.1f} evals/s; plain PyTorch autograd (FP64): "
               f"{plain_s:.2f} evals/s; plain log L {float(ll2.detach()):.10g}")
    g_gpu = grad.cpu().numpy()
    res = {
        "metric": "likelihood-grad evals/s (30y sequence, 1M seed)",
        "value": evals_s,
        "unit": "gradient evaluations/s",
        "ms_per_eval": 1e3 / evals_s,
        "device_ms_per_eval": dev_ms,
        "vs_plain": evals_s / plain_s if plain_s else None,
        "plain_torch_evals_per_s": plain_s,
        "autograd_function_evals_per_s": autograd_s,
        "dtype": "f32 weights, f64 partition sums and gradients",
        "config": {"workload": "30y PA sequence at 3%/y from a 1,000,000-paper seed, m=20 "
                               "(14 PA + 6 neighbourhood), generated at alpha=1.2, tau=4; "
                               "gradient at alpha=1, tau=5",
                   "paper_years": info.paper_years, "citing_papers": info.new_papers,
                   "slots": info.new_papers * m, "params": info.n_papers + 2},
        "roofline": roofline("pg_pa_loglik_grad (all passes)", achieved,
                             "likelihood_30y_1M_kernels.md", "lik_grad"),
        "batched": batched,
        "vs_plain_torch_rel_err": None if plain_s is None else {
            "dalpha": abs(ga - ga2) / abs(ga2), "dtau": abs(gt - gt2) / abs(gt2),
            "dphi_l2": float(np.linalg.norm(g_gpu - g_plain) / np.linalg.norm(g_plain))},
    }
    if not args.no_cpu_baseline:
        from oracle import binding as ob
        kw = dict(n_exist=seq["n_exist"], n_new=seq["n_new"], year_of=seq["year_of"],
                  deg_list=seq["deg"], cit_list=seq["cit"], kind_list=seq["kind"],
                  cnt_list=seq["cnt"], m=m, birth=seq["birth"], logfit=g.logfit[:n_final],
                  alpha=a0, tau=t0_)
        t4 = time.perf_counter()
        d = ob.loglik_grad(**kw, w32=False)
        cpu_s = time.perf_counter() - t4
        w = ob.loglik_grad(**kw, w32=True)
        res["cpu_baseline"] = {"value": 1.0 / cpu_s, "unit": "gradient evaluations/s",
                               "cores": 1, "kind": "port",
                               "sample": f"one FP64 gradient evaluation of the full sequence on "
                                         f"the C oracle, one thread: {cpu_s:.2f} s"}
        res["vs_oracle_fp64_rel_err"] = {
            "loglik": abs(ll - d[0]) / abs(d[0]), "dalpha": abs(ga - d[1]) / abs(d[1]),
            "dtau": abs(gt - d[2]) / abs(d[2]),
            "dphi_l2": float(np.linalg.norm(g_gpu - d[3]) / np.linalg.norm(d[3]))}
        res["vs_oracle_same_arith_rel_err"] = {
            "loglik": abs(ll - w[0]) / abs(w[0]), "dalpha": abs(ga - w[1]) / abs(w[1]),
            "dtau": abs(gt - w[2]) / abs(w[2]),
            "dphi_max": float(np.max(np.abs(g_gpu - w[3]) / np.maximum(1.0, np.abs(w[3]))))}
        # the batched calls' point 0 is this evaluation's (alpha, tau, phi):
        # both arithmetics against the FP64 oracle, the simulator's also
        # against the same-arithmetic oracle
        for key, h0, gb0 in (("point0", batched["point0"], g_b0[True]),
                             ("sim_arith_point0", batched["sim_arith"]["point0"], g_b0[False])):
            batched[key + "_vs_oracle_fp64_rel_err"] = {
                "loglik": abs(h0[0] - d[0]) / abs(d[0]), "dalpha": abs(h0[1] - d[1]) / abs(d[1]),
                "dtau": abs(h0[2] - d[2]) / abs(d[2]),
                "dphi_l2": float(np.linalg.norm(gb0 - d[3]) / np.linalg.norm(d[3]))}
        h0 = batched["sim_arith"]["point0"]
        batched["sim_arith_point0_vs_oracle_same_arith_rel_err"] = {
            "loglik": abs(h0[0] - w[0]) / abs(w[0]), "dalpha": abs(h0[1] - w[1]) / abs(w[1]),
            "dtau": abs(h0[2] - w[2]) / abs(w[2]),
            "dphi_max": float(np.max(np.abs(g_b0[False] - w[3]) / np.maximum(1.0, np.abs(w[3]))))}
        fe = batched["point0_vs_oracle_fp64_rel_err"]
        if max(fe["dalpha"], fe["dtau"], fe["dphi_l2"]) > 1e-5:
            raise RuntimeError(f"fast batched gradient beyond 1e-5 of the FP64 oracle: {fe}")
        eprint(f"oracle: {cpu_s:.2f} s per FP64 evaluation; errors {res['vs_oracle_fp64_rel_err']}"
               f" (FP64), {res['vs_oracle_same_arith_rel_err']} (same arithmetic)")
    # Adam from alpha=1, tau=5 (log-fitness held at its true values)
    n_ev = int(sum(((np.arange(m)[None, :] < c[:, None]) & (k != 1)).sum()
                   for c, k in zip(seq["cnt"], seq["kind"])))
    torch.cuda.synchronize()
    t5 = time.perf_counter()
    at_fit, _, hist = fit_adam(ctx, phi, 1.0, 5.0, steps=args.fit_steps, lr=0.02,
                               fit_logfit=False, per_event=1.0 / n_ev)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t5
    af, tf = (float(x) for x in at_fit.cpu())
    h = hist.cpu().numpy()
    res["fit"] = {"steps": args.fit_steps, "seconds": fit_s, "alpha": af, "tau": tf,
                  "start": [1.0, 5.0], "generating": [alpha_g, tau_g],
                  "loglik_first": float(h[0]), "loglik_last": float(h[-1])}
    eprint(f"Adam fit ({args.fit_steps} steps, {fit_s:.2f} s): alpha {af:.4f} tau {tf:.4f}, "
           f"log L {h[0]:.10g} -> {h[-1]:.10g}")
    res["autotune"] = tune
    ctx.close()
    return res


def batched_likelihood(ctx, phi, n_final, a0, t0_, stream, args, plain3, plain_grad):
    """pg_pa_loglik_grad_batch on the same sequence at B = 4, 8, 16 points
    per pass (point b at alpha = a0 + 0.02 b, tau = t0 - 0.1 b, the same
    log-fitness), in both arithmetics: evaluations/s, device time and
    algorithmic bytes per call.  Point 0 is the single-point evaluation's:
    in the simulator's arithmetic it must agree to FP64 summation order
    (1e-12), in the fast one within 1e-5 (the gradient-error bound against
    the plain path)."""
    import torch
    dev = phi.device
    rows = {False: [], True: []}
    best, g0 = {}, {}
    gref = plain_grad.cpu().numpy()
    for fast in (False, True):
        for B in (4, 8, 16):
            pb = phi.repeat_interleave(B).contiguous()
            gb = torch.empty(n_final * B, dtype=torch.float64, device=dev)
            sb = torch.empty(3 * B, dtype=torch.float64, device=dev)
            at = [(a0 + 0.02 * b, t0_ - 0.1 * b) for b in range(B)]

            def call():
                ctx.pa_loglik_grad_batch(at, pb.data_ptr(), gb.data_ptr(),
                                         scalars_dev=sb.data_ptr(), stream=stream, host=False,
                                         fast=fast)

            for _ in range(2):
                call()
            torch.cuda.synchronize()
            K = max(args.lik_evals // B, 5)
            t1 = time.perf_counter()
            for _ in range(K):
                call()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t1
            ctx.profile(True)
            ctx.profile_reset()
            for _ in range(10):
                call()
            prof = ctx.profile_read()["lik_batch"]
            ctx.profile(False)
            ms = prof["ms"] / prof["launches"]
            by = prof["bytes"] / prof["launches"]
            ach = by / (ms * 1e-3) / 1e9
            h = sb.cpu().numpy().reshape(B, 3)
            gcol = gb.view(n_final, B)[:, 0].cpu().numpy()
            rel = max(abs(h[0, c] - plain3[c]) / max(abs(plain3[c]), 1e-300) for c in range(3))
            gerr = float(np.max(np.abs(gcol - gref) / np.maximum(1.0, np.abs(gref))))
            gl2 = float(np.linalg.norm(gcol - gref) / np.linalg.norm(gref))
            row = {"points": B, "arith": "fast" if fast else "sim",
                   "evals_per_s": B * K / dt, "ms_per_call": 1e3 * dt / K,
                   "device_ms_per_call": ms, "gb_per_call": by / 1e9, "achieved_GBs": ach,
                   "frac": ach / PEAK_HBM_GBS, "point0_vs_single_max_rel": rel,
                   "point0_dphi_vs_single_max": gerr, "point0_dphi_vs_single_l2": gl2}
            eprint(f"batched B={B} ({row['arith']}): {row['evals_per_s']:.1f} evals/s ({ms:.3f} ms "
                   f"device per call, {by / 1e9:.3f} GB algorithmic, {ach:.0f} GB/s = "
                   f"{ach / PEAK_HBM_GBS:.3f} of peak); point 0 vs single-point: scalars {rel:.1e}, "
                   f"dL/dphi {gerr:.1e} (max), {gl2:.1e} (2-norm)")
            # the fast arithmetic's bound is on the gradient vector (2-norm), as
            # against the FP64 oracle; single components that nearly cancel
            # carry its relative rounding times their terms' size
            bad = (rel > 1e-5 or gl2 > 1e-5) if fast else (rel > 1e-12 or gerr > 1e-10)
            if bad:
                raise RuntimeError(f"batched point 0 differs from the single-point evaluation "
                                   f"(B={B}, {row['arith']})")
            rows[fast].append(row)
            if B == 16:
                best[fast] = (row, h[0].tolist(), ach)
                g0[fast] = gcol
            del pb, gb
    row16, h0, ach16 = best[True]
    rsim, h0sim, achsim = best[False]
    return {"metric": "likelihood-grad evals/s, 16 parameter points per pass, fast arithmetic "
                      "(30y sequence)",
            "value": row16["evals_per_s"], "unit": "gradient evaluations/s",
            "arith": "fast: hardware exp2/log2/reciprocal, FP64 partition sums",
            "per_points": rows[True], "per_points_sim_arith": rows[False], "point0": h0,
            "sim_arith": {"evals_per_s": rsim["evals_per_s"], "point0": h0sim,
                          "roofline": roofline("pg_pa_loglik_grad_batch, 16 points, simulator "
                                               "arithmetic (all passes)", achsim,
                                               "likelihood_30y_1M_kernels.md", "lik_batch16")},
            "roofline": roofline("pg_pa_loglik_grad_batch, 16 points, fast arithmetic "
                                 "(all passes)", ach16, "likelihood_30y_1M_kernels.md",
                                 "lik_batch16_fast")}, g0


def ensemble_bench(args) -> dict:
    """Batched ensemble: n_sims PA simulations (seed graphs as the 2M-paper
    workload's: Poisson(10) citations, births uniform in [-20,0],
    lognormal(0,0.5) fitness), each grown 10 years at 3% with 20 citations
    per paper, run together (one hipGraph replay for the whole run) and, for
    comparison, one after another as plain runs (their 10 year steps timed;
    context setup and loading excluded).  Both must add the same edges."""
    m, years, rate = 20, 10, 0.03
    sims = ensemble_params(args.ens_sims, m)
    t0 = time.perf_counter()
    graphs = []
    for s, sp in enumerate(sims):
        sched = pg.growth_schedule(sp["n0"], rate, years)
        g = pg.seed_graph(sp["seed"], sp["n0"], 10.0, n_final=sp["n0"] + sum(sched))
        graphs.append((g, sched))
        sp["seed_edges"] = int(g.src.size)
    n_seed = sum(sp["n0"] for sp in sims)
    n_new = sum(sum(sc) for _, sc in graphs)
    eprint(f"ensemble: {len(sims)} simulations, {n_seed} seed papers, {n_new} new papers "
           f"(seed graphs {time.perf_counter() - t0:.1f} s)")
    host_gen_s = time.perf_counter() - t0
    # the ensemble's seed graphs and log-fitness generated on the device (the
    # same bits as the host graphs above, which the plain runs load)
    t1 = time.perf_counter()
    ens = pg.Ensemble(sims, years, m, rate)
    t1b = time.perf_counter()
    ens.generate(10.0)
    gen_s = time.perf_counter() - t1b
    eprint(f"ensemble created ({t1b - t1:.1f} s) and its seed graphs generated on the device "
           f"({gen_s:.2f} s; the host generator took {host_gen_s:.1f} s)")
    tune = None
    if args.ens_shape:
        lanes, mode = (int(x) for x in args.ens_shape.split(","))
        ens.set_tuning(pa_lanes=lanes, bfs_mode=mode)
    elif not args.no_tune:
        t4 = time.perf_counter()
        tune = ens.autotune(reps=args.tune_reps)   # leaves the seed state
        eprint(f"autotune ({time.perf_counter() - t4:.1f} s): chosen {tune['chosen']} "
               f"{tune['chosen_ms']:.1f} ms, defaults {tune['default_ms']:.1f} ms; " +
               ", ".join(f"{t['pa_lanes']}/{t['bfs_mode']}: {t['ms']:.1f}" for t in tune["trials"]))
    st = ens.run(use_graph=False)
    eager_ms = st.device_ms
    ens.restore()
    st = ens.run(use_graph=True)       # captures the graph
    capture_ms = st.device_ms
    walls, devs = [], []
    for _ in range(args.ens_reps):
        ens.restore()
        t2 = time.perf_counter()
        st = ens.run(use_graph=True)
        walls.append(time.perf_counter() - t2)
        devs.append(st.device_ms)
    edges = st.edges_added
    wall = float(np.median(walls))
    dev = float(np.median(devs))
    eprint(f"ensemble run: {edges} edges, {wall * 1e3:.1f} ms wall ({dev:.1f} ms device; eager "
           f"{eager_ms:.1f} ms, first graph run {capture_ms:.1f} ms), {edges / wall:.3e} edges/s, "
           f"near-boundary {st.near_boundary}, shortfall {st.shortfall}")
    # the dominant group (2-hop kernels) timed in eager year steps
    ens.restore()
    bfs_ms = bfs_bytes = step_ms = 0.0
    for _ in range(years):
        sst = ens.step()
        bfs_ms += sst.bfs_ms
        bfs_bytes += sst.bfs_bytes
        step_ms += sst.device_ms
    ens_ach = bfs_bytes / years / (bfs_ms / years * 1e-3) / 1e9
    eprint(f"ensemble 2-hop kernels: {bfs_ms:.1f} of {step_ms:.1f} ms over the {years} eager "
           f"steps, {bfs_bytes / 1e9:.2f} GB algorithmic, {ens_ach:.0f} GB/s")
    ens.close()
    # the same simulations as plain runs, one after another
    plain_s, plain_edges = 0.0, 0
    for sp, (g, sched) in ([] if args.ens_no_plain else zip(sims, graphs)):
        ctx = pg.Context(g.n + sum(sched), g.src.size + sum(sched) * m + 16, max(sched),
                         max_cites=m, bfs_cap=4096, seed=sp["seed"])
        ctx.load(g)
        t3 = time.perf_counter()
        for yi, d in enumerate(sched):
            plain_edges += ctx.year_pa(1 + yi, yi, d, m, sp["n_nb"], sp["alpha"], sp["tau"]).edges_added
        plain_s += time.perf_counter() - t3
        ctx.close()
    if not args.ens_no_plain:
        eprint(f"plain runs: {plain_edges} edges in {plain_s:.3f} s "
               f"({plain_edges / plain_s:.3e} edges/s)")
        if plain_edges != edges:
            raise RuntimeError(f"ensemble added {edges} edges, plain runs {plain_edges}")
    plain_rate = plain_edges / plain_s if plain_s else None
    return {
        "metric": f"ensemble citation edges/s ({len(sims)} simulations, 10y)",
        "value": edges / wall,
        "unit": "citation edges/s",
        "ms_per_run": wall * 1e3,
        "device_ms_per_run": dev,
        "plain_sequential_edges_per_s": plain_rate,
        "plain_sequential_s": plain_s or None,
        "vs_plain": (edges / wall) / plain_rate if plain_rate else None,
        "config": {"workload": f"{len(sims)} PA simulations, seeds log-uniform in [1e4, 5e5] "
                               f"papers, alpha~U[0.5,1.5], tau~U[1,10], f_nb~U[0,0.5], m=20, "
                               f"10y at 3%/y, Philox keys 0..{len(sims) - 1}",
                   "seed_papers": n_seed, "new_papers": n_new, "edges_added": edges,
                   "near_boundary_attempts": st.near_boundary, "hipgraph": True},
        "bfs_ms_per_run": bfs_ms, "eager_step_ms_per_run": step_ms,
        "seed_generation_s": {"device": gen_s, "host": host_gen_s},
        "roofline": roofline("2-hop kernels per year launch (k_bfs_small, k_bfs_mid, "
                             "k_bfs_root)", ens_ach, "ensemble_256x10y_kernels.md", "ens_bfs"),
        "autotune": tune,
    }


def copy_calibration(nbytes: int = 1 << 30) -> None:
    """One plain device-to-device copy of nbytes (HIP runtime copy kernel):
    the reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE
    counters to bytes (the copy reads and writes nbytes each)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n0", type=int, default=2_000_000)
    ap.add_argument("--years", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--lik-evals", type=int, default=200,
                    help="timed likelihood-gradient evaluations")
    ap.add_argument("--fit-steps", type=int, default=100)
    ap.add_argument("--lik-no-torch", action="store_true",
                    help="skip the autograd.Function and plain-PyTorch comparisons")
    ap.add_argument("--no-lik", action="store_true", help="skip the likelihood part")
    ap.add_argument("--no-ens", action="store_true", help="skip the ensemble part")
    ap.add_argument("--no-er", action="store_true", help="skip the ER part")
    ap.add_argument("--er-only", action="store_true", help="only the ER part")
    ap.add_argument("--no-ref", action="store_true", help="skip the reference scoring model part")
    ap.add_argument("--ref-only", action="store_true",
                    help="only the reference scoring model part")
    ap.add_argument("--ens-only", action="store_true",
                    help="only the ensemble part (its JSON object on stdout)")
    ap.add_argument("--ens-sims", type=int, default=256)
    ap.add_argument("--ens-shape", default=None,
                    help="ensemble launch shape 'PA_LANES,BFS_MODE' instead of the autotuner")
    ap.add_argument("--ens-no-plain", action="store_true",
                    help="skip the ensemble's sequential plain-run comparison")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuners (ensemble, likelihood)")
    ap.add_argument("--tune-reps", type=int, default=2)
    ap.add_argument("--bfs-mode", type=int, default=0,
                    help="2-hop workgroup path: 0 staged where it fits, 1 sweep only")
    ap.add_argument("--ens-reps", type=int, default=3)
    ap.add_argument("--abm-only", action="store_true",
                    help="only abm end to end on CSV files (its JSON object on stdout)")
    ap.add_argument("--lik-only", action="store_true",
                    help="only the likelihood part (its JSON object on stdout)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.lik_only:
        print(json.dumps(likelihood_bench(args)), flush=True)
        return
    if args.ens_only:
        print(json.dumps(ensemble_bench(args)), flush=True)
        return
    if args.er_only:
        print(json.dumps(er_bench(args)), flush=True)
        return
    if args.ref_only:
        print(json.dumps(ref_model_bench(args)), flush=True)
        return
    if args.abm_only:
        print(json.dumps(abm_bench(args)), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    alpha, tau, seed, cap = 1.0, 5.0, 1, 4096
    n0, sched, m, n_nb = workload(args)
    n_final = n0 + sum(sched)
    t0 = time.perf_counter()
    g = pg.seed_graph(seed, n0, 10.0, n_final=n_final)
    eprint(f"seed graph: {n0} papers, {g.src.size} edges ({time.perf_counter() - t0:.1f} s)")
    e_cap = g.src.size + sum(sched) * m + 16
    ctx = pg.Context(n_final, e_cap, max(sched), max_cites=m, bfs_cap=cap, seed=seed)
    ctx.set_tuning(bfs_mode=args.bfs_mode)
    ctx.load(g)
    ctx.snapshot_save()
    tune = None
    if not args.no_tune:
        # launch shapes timed on this workload (results identical by check)
        years = [dict(year=1 + yi, year_index=yi, n_new=D, m=m, n_nb=n_nb, alpha=alpha, tau=tau)
                 for yi, D in enumerate(sched)]
        tune = ctx.year_autotune(years, reps=args.tune_reps)
        eprint(f"year autotune: chosen {tune['chosen']} {tune['chosen_ms']:.2f} ms per 10 years "
               f"(default shapes {tune['default_ms']:.2f} ms)")
    G0 = None
    if not args.no_cpu_baseline:
        from oracle import binding as ob
        G0 = ob.Graph.from_download(ctx.download(), g.logfit)

    def one_step():
        edges = 0
        near = 0
        for yi, D in enumerate(sched):
            st = ctx.year_pa(1 + yi, yi, D, m, n_nb, alpha, tau)
            edges += st.edges_added
            near += st.near_boundary
        return edges, near

    for w in range(args.warmup):
        one_step()
        ctx.snapshot_restore()
        eprint(f"warmup {w} done")
    # the timed steps record events around the 2-hop group only (each event
    # pair between kernels costs the stream a few microseconds); an untimed
    # step afterwards times every group for the breakdown
    ctx.profile(True)
    ctx.profile_groups(["bfs"])
    ctx.profile_reset()
    total_edges, total_near, total_s = 0, 0, 0.0
    digests = []
    for k in range(args.steps):
        t1 = time.perf_counter()
        e, nr = one_step()
        t2 = time.perf_counter()
        total_edges += e
        total_near += nr
        total_s += t2 - t1
        digests.append(state_digest(ctx))  # after the timed region
        eprint(f"step {k}: {e} edges in {t2 - t1:.3f} s ({e / (t2 - t1):.3e} edges/s), "
               f"near-boundary attempts {nr}, state {digests[-1][:12]}")
        ctx.snapshot_restore()
    if len(set(digests)) != 1:
        eprint(f"WARNING: the timed steps ended in {len(set(digests))} different graphs")
    prof_timed = ctx.profile_read()
    ctx.profile_groups(None)
    ctx.profile_reset()
    one_step()  # untimed, every group timed; its final graph feeds the degree comparison
    prof = ctx.profile_read()
    ctx.profile(False)
    gpu_in_deg = None if args.no_cpu_baseline else ctx.download()["in_deg"].copy()
    for name, v in prof.items():
        if v["launches"]:
            eprint(f"  {name:12s} {v['ms']:9.2f} ms over {v['launches']} launches, "
                   f"{v['bytes'] / max(v['ms'], 1e-9) / 1e6:8.1f} GB/s algorithmic")
    dom = max(prof, key=lambda k: prof[k]["ms"])
    dv = prof_timed[dom] if prof_timed[dom]["launches"] else prof[dom]
    avg_ms = dv["ms"] / dv["launches"]
    avg_bytes = dv["bytes"] / dv["launches"]
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    value = total_edges / total_s
    out = {
        "metric": "citation edges/s (PA 10y, 2M seed)",
        "value": value,
        "unit": "citation edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total_s / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 weights, f64 CDF, int32 CSR",
        "data": "synthetic (Philox seed graph)",
        "config": {"workload": f"PA {args.years}y at 3%/y from a {n0}-paper seed, m=20 "
                               f"(14 PA + 6 neighbourhood), cap 4096, alpha=1, tau=5",
                   "edges_per_step": total_edges // max(args.steps, 1),
                   "near_boundary_attempts": total_near,
                   "steps_identical_graphs": len(set(digests)) == 1,
                   "autotune": None if tune is None else {
                       "chosen": tune["chosen"], "chosen_ms": tune["chosen_ms"],
                       "default_ms": tune["default_ms"], "trials": len(tune["trials"])},
                   "kernel_ms_per_step": {k: v["ms"] for k, v in prof.items()
                                          if v["launches"]}},
        "roofline": roofline(dom, achieved, "pa_10y_2M_kernels.md", dom),
    }
    ctx.close()
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"], G = cpu_baseline(G0, sched, m, n_nb, alpha, tau, seed, cap)
            out["degree_ks"] = degree_ks(gpu_in_deg, G.in_deg)
        except Exception as exc:  # reported, never fatal for the GPU number
            eprint(f"cpu baseline failed: {exc}")
            out["cpu_baseline"] = None
    if not args.no_lik:
        out["likelihood"] = likelihood_bench(args)
    if not args.no_ens:
        out["ensemble"] = ensemble_bench(args)
    if not args.no_er:
        out["er"] = er_bench(args)
    if not args.no_ref:
        out["reference_model"] = ref_model_bench(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
