"""Benchmark: citation edges/s of the 10-year PA growth on one MI355X.

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

PEAK_HBM_GBS = 8000.0


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
    return {"value": edges / dt, "unit": "citation edges/s", "cores": cores, "kind": "port",
            "sample": f"the full workload ({len(sched)} years, {edges} edges) on the C oracle "
                      f"with {cores} OpenMP threads: {dt:.2f} s"}


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
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(every GPU box here has one GPU).", file=sys.stderr)
        sys.exit(2)
    alpha, tau, seed, cap = 1.0, 5.0, 1, 4096
    n0, sched, m, n_nb = workload(args)
    n_final = n0 + sum(sched)
    t0 = time.perf_counter()
    g = pg.seed_graph(seed, n0, 10.0, n_final=n_final)
    eprint(f"seed graph: {n0} papers, {g.src.size} edges ({time.perf_counter() - t0:.1f} s)")
    e_cap = g.src.size + sum(sched) * m + 16
    ctx = pg.Context(n_final, e_cap, max(sched), max_cites=m, bfs_cap=cap, seed=seed)
    ctx.load(g)
    ctx.snapshot_save()
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
    ctx.profile(True)
    ctx.profile_reset()
    total_edges, total_near, total_s = 0, 0, 0.0
    for k in range(args.steps):
        t1 = time.perf_counter()
        e, nr = one_step()
        t2 = time.perf_counter()
        total_edges += e
        total_near += nr
        total_s += t2 - t1
        eprint(f"step {k}: {e} edges in {t2 - t1:.3f} s ({e / (t2 - t1):.3e} edges/s), "
               f"near-boundary attempts {nr}")
        ctx.snapshot_restore()
    prof = ctx.profile_read()
    ctx.profile(False)
    for name, v in prof.items():
        if v["launches"]:
            eprint(f"  {name:12s} {v['ms']:9.2f} ms over {v['launches']} launches, "
                   f"{v['bytes'] / max(v['ms'], 1e-9) / 1e6:8.1f} GB/s algorithmic")
    dom = max(prof, key=lambda k: prof[k]["ms"])
    dv = prof[dom]
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
                   "kernel_ms_per_step": {k: v["ms"] / args.steps for k, v in prof.items()
                                          if v["launches"]}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "traffic": None},
    }
    if not args.no_cpu_baseline:
        ctx.close()
        try:
            out["cpu_baseline"] = cpu_baseline(G0, sched, m, n_nb, alpha, tau, seed, cap)
        except Exception as exc:  # reported, never fatal for the GPU number
            eprint(f"cpu baseline failed: {exc}")
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
