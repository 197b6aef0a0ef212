"""Oracle PA log-likelihood and gradient: closed form, brute force, and
central finite differences."""
import math

import numpy as np
import pytest

from oracle import binding as ob


def simulate(n0=300, years=3, rate=0.1, m=5, n_nb=2, alpha=1.2, tau=4.0, seed=7):
    """Citation sequence from the oracle's own year step, with per-year
    in-degree snapshots (the likelihood's inputs)."""
    rng = np.random.default_rng(seed)
    edges = set()
    for i in range(1, n0):
        for t in rng.choice(i, size=min(i, 4), replace=False):
            edges.add((i, int(t)))
    e = np.array(sorted(edges), np.int64)
    birth = np.sort(rng.integers(-10, 1, n0)).astype(np.int32)
    n_final = n0
    sched = []
    for _ in range(years):
        d = math.ceil(n_final * rate)
        sched.append(d)
        n_final += d
    logfit = (np.round(rng.standard_normal(n_final) * 0.5 * 4096) / 4096).astype(np.float32)
    G = ob.Graph.from_edges(n0, e[:, 0], e[:, 1], birth, logfit)
    seq = dict(n_exist=[], n_new=[], year_of=[], deg=[], cit=[], kind=[], cnt=[])
    for yi, D in enumerate(sched):
        r = ob.pa_year(G, seed, yi, 1 + yi, D, m, n_nb, alpha, tau, cap=64)
        seq["n_exist"].append(G.n); seq["n_new"].append(D); seq["year_of"].append(1 + yi)
        seq["deg"].append(G.in_deg.copy()); seq["cit"].append(r["cit"]); seq["kind"].append(r["kind"])
        seq["cnt"].append(r["cnt"])
        off, tgt = ob.pa_compact(r)
        G = ob.csr_append(G, off, tgt, logfit, 1 + yi)
    return seq, G.birth, logfit, m


def loglik(seq, birth, logfit, m, alpha, tau):
    return ob.loglik_grad(seq["n_exist"], seq["n_new"], seq["year_of"], seq["deg"], seq["cit"],
                          seq["kind"], seq["cnt"], m, birth, logfit, alpha, tau)


def test_uniform_weights_closed_form():
    seq, birth, logfit, m = simulate()
    ll, _, _, _ = loglik(seq, birth, np.zeros_like(logfit), m, 0.0, float("inf"))
    ref = 0.0
    for N, cit, kind, cnt in zip(seq["n_exist"], seq["cit"], seq["kind"], seq["cnt"]):
        for row, kr, c in zip(cit, kind, cnt):
            for j in range(c):
                if kr[j] != 1:
                    ref -= math.log(N - j)
    assert abs(ll - ref) <= 1e-12 * abs(ref)


def test_brute_force_sequence_probability():
    seq, birth, logfit, m = simulate(n0=40, years=2, rate=0.2, m=4, n_nb=1)
    alpha, tau = 0.8, 3.0
    ll, _, _, _ = loglik(seq, birth, logfit, m, alpha, tau)
    ref = 0.0
    for N, y, deg, cit, kind, cnt in zip(seq["n_exist"], seq["year_of"], seq["deg"], seq["cit"],
                                         seq["kind"], seq["cnt"]):
        w = [math.exp(alpha * math.log(deg[j] + 1) - (y - birth[j]) / tau + float(logfit[j]))
             for j in range(N)]
        Z = math.fsum(w)
        for row, kr, c in zip(cit, kind, cnt):
            taken = []
            for j in range(c):
                t = int(row[j])
                if kr[j] != 1:
                    ref += math.log(w[t] / (Z - math.fsum(w[s] for s in taken)))
                taken.append(t)
    assert abs(ll - ref) <= 1e-9 * abs(ref)


def test_gradient_matches_central_differences():
    seq, birth, logfit, m = simulate()
    alpha, tau = 1.1, 4.5
    ll, ga, gt, gphi = loglik(seq, birth, logfit, m, alpha, tau)
    h = 1e-5
    fa = (loglik(seq, birth, logfit, m, alpha + h, tau)[0] -
          loglik(seq, birth, logfit, m, alpha - h, tau)[0]) / (2 * h)
    ft = (loglik(seq, birth, logfit, m, alpha, tau + h)[0] -
          loglik(seq, birth, logfit, m, alpha, tau - h)[0]) / (2 * h)
    assert abs(ga - fa) <= 1e-5 * abs(fa)
    assert abs(gt - ft) <= 1e-5 * abs(ft)
    hp = 2.0 ** -12
    n = birth.size
    for j in [0, 1, 5, 50, 200, n - 1]:
        lp = logfit.copy(); lp[j] += np.float32(hp)
        lm = logfit.copy(); lm[j] -= np.float32(hp)
        fd = (loglik(seq, birth, lp, m, alpha, tau)[0] - loglik(seq, birth, lm, m, alpha, tau)[0]) / (2 * hp)
        assert abs(gphi[j] - fd) <= 1e-5 * max(1.0, abs(fd)), j


def test_w32_mode_closed_form_and_fp64_agreement():
    """The FP32-weight mode (the GPU's arithmetic) gives exact unit weights at
    alpha=0, 1/tau -> 0, phi=0, and stays within 1e-5 of the FP64 gradient."""
    seq, birth, logfit, m = simulate()
    ref = 0.0
    for N, cit, kind, cnt in zip(seq["n_exist"], seq["cit"], seq["kind"], seq["cnt"]):
        for row, kr, c in zip(cit, kind, cnt):
            for j in range(c):
                if kr[j] != 1:
                    ref -= math.log(N - j)
    ll, ga, gt, _ = ob.loglik_grad(seq["n_exist"], seq["n_new"], seq["year_of"], seq["deg"],
                                   seq["cit"], seq["kind"], seq["cnt"], m, birth,
                                   np.zeros_like(logfit), 0.0, 1e300, w32=True)
    assert abs(ll - ref) <= 1e-13 * abs(ref)
    for alpha, tau in [(1.25, 4.0), (1.0, 5.0), (0.5, 2.5)]:
        a = ob.loglik_grad(seq["n_exist"], seq["n_new"], seq["year_of"], seq["deg"], seq["cit"],
                           seq["kind"], seq["cnt"], m, birth, logfit, alpha, tau, w32=True)
        d = loglik(seq, birth, logfit, m, alpha, tau)
        assert abs(a[0] - d[0]) <= 1e-6 * abs(d[0])
        assert abs(a[1] - d[1]) <= 1e-5 * abs(d[1])
        assert abs(a[2] - d[2]) <= 1e-5 * abs(d[2])
        assert np.linalg.norm(a[3] - d[3]) <= 1e-5 * np.linalg.norm(d[3])
