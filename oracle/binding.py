"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker of the HIP path.  The product
(paragraph_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

P = ctypes.POINTER
_i32p, _i64p, _f32p, _f64p, _i8p = (P(ctypes.c_int32), P(ctypes.c_int64), P(ctypes.c_float),
                                    P(ctypes.c_double), P(ctypes.c_int8))


class Csr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("fwd_off", _i32p), ("fwd_adj", _i32p),
                ("bwd_off", _i32p), ("bwd_adj", _i32p)]


class PaParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("year_index", ctypes.c_int32),
                ("year", ctypes.c_int32), ("n_exist", ctypes.c_int64),
                ("n_new", ctypes.c_int64), ("m", ctypes.c_int32), ("n_nb", ctypes.c_int32),
                ("cap", ctypes.c_int32), ("alpha", ctypes.c_float), ("inv_tau", ctypes.c_float)]


class ErParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("year_index", ctypes.c_int32),
                ("n_exist", ctypes.c_int64), ("n_new", ctypes.c_int64), ("p", ctypes.c_double)]


class AbmModelC(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("bag", _i32p), ("n_bag", ctypes.c_int32),
                ("rec_prob", _f64p), ("n_rec", ctypes.c_int32), ("alpha", ctypes.c_double),
                ("w_pa", ctypes.c_double), ("w_rec", ctypes.c_double), ("w_fit", ctypes.c_double),
                ("frc", ctypes.c_double), ("syp", ctypes.c_double)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "or_philox4x32_10": (None, [P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]),
            "or_u53": (ctypes.c_double, [ctypes.c_uint32, ctypes.c_uint32]),
            "or_u53_open0": (ctypes.c_double, [ctypes.c_uint32, ctypes.c_uint32]),
            "or_expf": (ctypes.c_float, [ctypes.c_float]),
            "or_logf": (ctypes.c_float, [ctypes.c_float]),
            "or_log": (ctypes.c_double, [ctypes.c_double]),
            "or_log1p": (ctypes.c_double, [ctypes.c_double]),
            "or_vec_expf": (None, [ctypes.c_int64, _f32p, _f32p]),
            "or_vec_logf": (None, [ctypes.c_int64, _f32p, _f32p]),
            "or_vec_log": (None, [ctypes.c_int64, _f64p, _f64p]),
            "or_vec_log1p": (None, [ctypes.c_int64, _f64p, _f64p]),
            "or_weights": (None, [ctypes.c_int64, _i32p, _i32p, _f32p, ctypes.c_int32,
                                  ctypes.c_float, ctypes.c_float, _f32p]),
            "or_cdf": (None, [ctypes.c_int64, _f32p, _f64p]),
            "or_upper_bound": (ctypes.c_int64, [_f64p, ctypes.c_int64, ctypes.c_double]),
            "or_bfs": (ctypes.c_int32, [P(Csr), ctypes.c_int32, ctypes.c_int32, _i32p]),
            "or_pa_year": (ctypes.c_int, [P(Csr), _i32p, _i32p, _f32p, P(PaParams), _i32p,
                                          _i8p, _i32p, _i32p, ctypes.c_int]),
            "or_er_year": (ctypes.c_int64, [P(ErParams), _i32p, _i32p, ctypes.c_int64,
                                            ctypes.c_int]),
            "or_csr_append": (ctypes.c_int, [P(Csr), ctypes.c_int64, _i32p, _i32p, _i32p, _i32p,
                                             _i32p, _i32p, _i32p]),
            "or_pa_loglik_grad": (ctypes.c_int, [ctypes.c_int32, _i64p, _i64p, _i32p, _i32p,
                                                 _i64p, _i32p, _i8p, _i64p, _i32p, _i64p,
                                                 ctypes.c_int32, _i32p, _f32p, ctypes.c_double,
                                                 ctypes.c_double, _f64p, _f64p, _f64p, _f64p]),
            "or_pa_loglik_grad_w32": (ctypes.c_int, [ctypes.c_int32, _i64p, _i64p, _i32p, _i32p,
                                                     _i64p, _i32p, _i8p, _i64p, _i32p, _i64p,
                                                     ctypes.c_int32, _i32p, _f32p, ctypes.c_double,
                                                     ctypes.c_double, _f64p, _f64p, _f64p, _f64p]),
            "or_abm_fitness_cdf": (None, [_f64p]),
            "or_abm_fitness": (ctypes.c_int32, [ctypes.c_uint64, ctypes.c_int64, _f64p]),
            "or_abm_paper": (None, [P(AbmModelC), ctypes.c_int64, _i32p, _f64p, _f64p]),
            "or_abm_scores": (ctypes.c_int, [ctypes.c_int64, _i32p, _i32p, ctypes.c_int32,
                                             P(AbmModelC), _f32p]),
            "or_abm_key": (ctypes.c_double, [P(AbmModelC), ctypes.c_uint32, ctypes.c_int32,
                                             ctypes.c_uint32, _f32p, _f64p]),
            "or_abm_same_year": (ctypes.c_int, [P(AbmModelC), ctypes.c_int64, ctypes.c_int64,
                                                ctypes.c_uint32, _i32p]),
            "or_abm_year": (ctypes.c_int64, [P(Csr), _i32p, _i32p, P(AbmModelC), ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int64, _i32p, _i32p,
                                             ctypes.c_int64, _f32p, _i32p, ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P(t))


def philox(ctr, key) -> list[int]:
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


class Graph:
    """Host CSR (sorted rows) the oracle works on."""

    def __init__(self, fwd_off, fwd_adj, bwd_off, bwd_adj, in_deg, birth, logfit):
        i32 = lambda a: np.ascontiguousarray(a, np.int32)
        self.fwd_off, self.fwd_adj = i32(fwd_off), i32(fwd_adj)
        self.bwd_off, self.bwd_adj = i32(bwd_off), i32(bwd_adj)
        self.in_deg, self.birth = i32(in_deg), i32(birth)
        self.logfit = np.ascontiguousarray(logfit, np.float32)

    @property
    def n(self) -> int:
        return int(self.fwd_off.shape[0] - 1)

    @property
    def e(self) -> int:
        return int(self.fwd_off[-1])

    @classmethod
    def from_edges(cls, n, src, dst, birth, logfit):
        src = np.asarray(src, np.int64)
        dst = np.asarray(dst, np.int64)
        if src.size:
            key = np.unique(src * (n + 1) + dst)
            src, dst = key // (n + 1), key % (n + 1)
        fo = np.zeros(n + 1, np.int64)
        np.add.at(fo, src + 1, 1)
        fo = np.cumsum(fo)
        order = np.lexsort((src, dst))
        bo = np.zeros(n + 1, np.int64)
        np.add.at(bo, dst + 1, 1)
        bo = np.cumsum(bo)
        indeg = np.diff(bo)
        return cls(fo, dst, bo, src[order], indeg, birth, logfit)

    @classmethod
    def from_download(cls, d: dict, logfit_full: np.ndarray):
        return cls(d["fwd_off"], d["fwd_adj"], d["bwd_off"], d["bwd_adj"], d["in_deg"],
                   d["birth"], logfit_full)

    def csr(self) -> Csr:
        return Csr(self.n, _p(self.fwd_off, ctypes.c_int32), _p(self.fwd_adj, ctypes.c_int32),
                   _p(self.bwd_off, ctypes.c_int32), _p(self.bwd_adj, ctypes.c_int32))


def weights(g: Graph, year: int, alpha: float, inv_tau: float) -> np.ndarray:
    n = g.n
    w = np.empty(n, np.float32)
    lib().or_weights(n, _p(g.in_deg, ctypes.c_int32), _p(g.birth, ctypes.c_int32),
                     _p(g.logfit, ctypes.c_float), year, alpha, inv_tau, _p(w, ctypes.c_float))
    return w


def cdf(w: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(w, np.float32)
    c = np.empty(w.shape[0], np.float64)
    lib().or_cdf(w.shape[0], _p(w, ctypes.c_float), _p(c, ctypes.c_double))
    return c


def bfs(g: Graph, root: int, cap: int) -> np.ndarray:
    out = np.empty(cap, np.int32)
    c = g.csr()
    k = lib().or_bfs(ctypes.byref(c), root, cap, _p(out, ctypes.c_int32))
    return out[:k].copy()


def pa_year(g: Graph, seed: int, year_index: int, year: int, n_new: int, m: int, n_nb: int,
            alpha: float, tau: float, cap: int = 4096, threads: int = 0) -> dict:
    inv_tau = 0.0 if math.isinf(tau) else 1.0 / tau
    p = PaParams(seed, year_index, year, g.n, n_new, m, n_nb, cap, alpha, inv_tau)
    cit = np.empty(n_new * m, np.int32)
    kind = np.empty(n_new * m, np.int8)
    cnt = np.empty(n_new, np.int32)
    near = np.empty(n_new, np.int32)
    c = g.csr()
    rc = lib().or_pa_year(ctypes.byref(c), _p(g.in_deg, ctypes.c_int32),
                          _p(g.birth, ctypes.c_int32), _p(g.logfit, ctypes.c_float),
                          ctypes.byref(p), _p(cit, ctypes.c_int32), _p(kind, ctypes.c_int8),
                          _p(cnt, ctypes.c_int32), _p(near, ctypes.c_int32), threads)
    if rc:
        raise RuntimeError(f"or_pa_year failed: {rc}")
    return {"cit": cit.reshape(n_new, m), "kind": kind.reshape(n_new, m), "cnt": cnt,
            "near": near}


def er_year(n_exist: int, n_new: int, p_link: float, seed: int, year_index: int,
            threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    prm = ErParams(seed, year_index, n_exist, n_new, p_link)
    off = np.empty(n_new + 1, np.int32)
    cap = max(16, int(n_exist * n_new * p_link * 1.5 + 64 * math.sqrt(n_exist * n_new * p_link + 1)))
    if p_link >= 1.0:
        cap = n_exist * n_new
    while True:
        tgt = np.empty(cap, np.int32)
        e = lib().or_er_year(ctypes.byref(prm), _p(off, ctypes.c_int32), _p(tgt, ctypes.c_int32),
                             cap, threads)
        if e >= 0:
            return off, tgt[:e].copy()
        cap *= 2


def csr_append(g: Graph, new_off: np.ndarray, new_tgt: np.ndarray, logfit_full=None,
               year: int | None = None) -> Graph:
    n_new = new_off.shape[0] - 1
    n, e = g.n, g.e
    ea = int(new_off[-1])
    new_off = np.ascontiguousarray(new_off, np.int32)
    new_tgt = np.ascontiguousarray(new_tgt, np.int32)
    fo = np.empty(n + n_new + 1, np.int32)
    fa = np.empty(e + ea, np.int32)
    bo = np.empty(n + n_new + 1, np.int32)
    ba = np.empty(e + ea, np.int32)
    deg = np.empty(n + n_new, np.int32)
    c = g.csr()
    rc = lib().or_csr_append(ctypes.byref(c), n_new, _p(new_off, ctypes.c_int32),
                             _p(new_tgt, ctypes.c_int32), _p(fo, ctypes.c_int32),
                             _p(fa, ctypes.c_int32), _p(bo, ctypes.c_int32),
                             _p(ba, ctypes.c_int32), _p(deg, ctypes.c_int32))
    if rc:
        raise RuntimeError(f"or_csr_append failed: {rc}")
    birth = np.concatenate([g.birth, np.full(n_new, year if year is not None else 0, np.int32)])
    lf = g.logfit if logfit_full is None else logfit_full
    return Graph(fo, fa, bo, ba, deg, birth, lf)


def pa_compact(res: dict) -> tuple[np.ndarray, np.ndarray]:
    """Per-paper citation slots -> (offsets, targets) in draw order."""
    cnt = res["cnt"].astype(np.int64)
    off = np.zeros(cnt.shape[0] + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    cit = res["cit"]
    mask = np.arange(cit.shape[1])[None, :] < cnt[:, None]
    return off.astype(np.int32), cit[mask].astype(np.int32)


def loglik_grad(n_exist, n_new, year_of, deg_list, cit_list, kind_list, cnt_list, m, birth,
                logfit, alpha, tau, w32=False):
    """Log-likelihood and gradient over per-year arrays (lists of numpy).

    w32=False: FP64 throughout.  w32=True: the GPU likelihood's arithmetic
    (FP32 weights, FP64 sums)."""
    Y = len(n_exist)
    ne = np.ascontiguousarray(n_exist, np.int64)
    nn = np.ascontiguousarray(n_new, np.int64)
    yo = np.ascontiguousarray(year_of, np.int32)
    deg = np.ascontiguousarray(np.concatenate(deg_list), np.int32)
    deg_off = np.zeros(Y, np.int64)
    deg_off[1:] = np.cumsum([len(d) for d in deg_list])[:-1]
    cit = np.ascontiguousarray(np.concatenate([c.reshape(-1) for c in cit_list]), np.int32)
    kind = np.ascontiguousarray(np.concatenate([k.reshape(-1) for k in kind_list]), np.int8)
    ev_off = np.zeros(Y, np.int64)
    ev_off[1:] = np.cumsum([c.size for c in cit_list])[:-1]
    cnt = np.ascontiguousarray(np.concatenate(cnt_list), np.int32)
    cnt_off = np.zeros(Y, np.int64)
    cnt_off[1:] = np.cumsum([len(c) for c in cnt_list])[:-1]
    nmax = int(max(ne + nn))
    gphi = np.zeros(nmax, np.float64)
    ll, ga, gt = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    birth = np.ascontiguousarray(birth, np.int32)
    logfit = np.ascontiguousarray(logfit, np.float32)
    fn = lib().or_pa_loglik_grad_w32 if w32 else lib().or_pa_loglik_grad
    rc = fn(Y, _p(ne, ctypes.c_int64), _p(nn, ctypes.c_int64),
            _p(yo, ctypes.c_int32), _p(deg, ctypes.c_int32),
            _p(deg_off, ctypes.c_int64), _p(cit, ctypes.c_int32),
            _p(kind, ctypes.c_int8), _p(ev_off, ctypes.c_int64),
            _p(cnt, ctypes.c_int32), _p(cnt_off, ctypes.c_int64), m,
            _p(birth, ctypes.c_int32), _p(logfit, ctypes.c_float),
            alpha, tau, ctypes.byref(ll), ctypes.byref(ga),
            ctypes.byref(gt), _p(gphi, ctypes.c_double))
    if rc:
        raise RuntimeError(f"or_pa_loglik_grad failed: {rc}")
    return ll.value, ga.value, gt.value, gphi


def vec(fn: str, x: np.ndarray) -> np.ndarray:
    """Elementwise oracle function: expf/logf (float32) or log/log1p (float64)."""
    if fn in ("expf", "logf"):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty_like(x)
        getattr(lib(), "or_vec_" + fn)(x.size, _p(x, ctypes.c_float), _p(y, ctypes.c_float))
    else:
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        getattr(lib(), "or_vec_" + fn)(x.size, _p(x, ctypes.c_double), _p(y, ctypes.c_double))
    return y


# ---- the reference's scoring model (abm_model.c) ----

class AbmModel:
    """Parameters of the scoring model; keeps the bag and recency arrays alive."""

    def __init__(self, seed: int, bag, rec_prob, alpha: float = 0.5, w_pa: float = 1 / 3,
                 w_rec: float = 1 / 3, w_fit: float = 1 / 3, frc: float = 0.05,
                 syp: float = 0.12):
        self.bag = np.ascontiguousarray(np.minimum(np.asarray(bag, np.int64), 250), np.int32)
        self.rec = np.ascontiguousarray(rec_prob, np.float64)
        self.seed, self.alpha, self.frc, self.syp = seed, alpha, frc, syp
        self.w = (w_pa, w_rec, w_fit)
        self.c = AbmModelC(seed, _p(self.bag, ctypes.c_int32), self.bag.size,
                           _p(self.rec, ctypes.c_double), self.rec.size, alpha, w_pa, w_rec,
                           w_fit, frc, syp)


def abm_fitness_cdf() -> np.ndarray:
    c = np.empty(1000, np.float64)
    lib().or_abm_fitness_cdf(_p(c, ctypes.c_double))
    return c


def abm_fitness(seed: int, ids) -> np.ndarray:
    c = abm_fitness_cdf()
    return np.array([lib().or_abm_fitness(seed, int(i), _p(c, ctypes.c_double)) for i in ids],
                    np.int32)


def abm_paper(m: AbmModel, pid: int) -> tuple[int, float, np.ndarray]:
    k = ctypes.c_int32()
    a = ctypes.c_double()
    w = np.empty(3, np.float64)
    lib().or_abm_paper(ctypes.byref(m.c), pid, ctypes.byref(k), ctypes.byref(a),
                       _p(w, ctypes.c_double))
    return k.value, a.value, w


def abm_scores(n: int, in_deg: np.ndarray, birth: np.ndarray, year: int,
               m: AbmModel) -> np.ndarray:
    comp = np.empty((n, 4), np.float32)
    rc = lib().or_abm_scores(n, _p(np.ascontiguousarray(in_deg[:n], np.int32), ctypes.c_int32),
                             _p(np.ascontiguousarray(birth[:n], np.int32), ctypes.c_int32),
                             year, ctypes.byref(m.c), _p(comp, ctypes.c_float))
    if rc:
        raise RuntimeError(f"or_abm_scores failed: {rc}")
    return comp


def abm_key(m: AbmModel, v: int, c: int, yi: int, comp: np.ndarray, w: np.ndarray) -> float:
    comp = np.ascontiguousarray(comp, np.float32)
    w = np.ascontiguousarray(w, np.float64)
    return lib().or_abm_key(ctypes.byref(m.c), v, c, yi, _p(comp, ctypes.c_float),
                            _p(w, ctypes.c_double))


def abm_same_year(m: AbmModel, n_exist: int, n_new: int, yi: int) -> np.ndarray:
    t = np.empty(max(n_new, 1), np.int32)
    if lib().or_abm_same_year(ctypes.byref(m.c), n_exist, n_new, yi, _p(t, ctypes.c_int32)):
        raise RuntimeError("or_abm_same_year failed")
    return t[:n_new].copy()


def abm_year(g: Graph, m: AbmModel, year: int, year_index: int, n_new: int,
             threads: int = 0) -> dict:
    """One year of the scoring model: rows (off, tgt), per-paper info, scores."""
    off = np.empty(n_new + 1, np.int32)
    info = np.empty((max(n_new, 1), 6), np.int32)
    comp = np.empty((g.n, 4), np.float32)
    cap = int(n_new) * 252 + 1
    tgt = np.empty(cap, np.int32)
    c = g.csr()
    e = lib().or_abm_year(ctypes.byref(c), _p(g.in_deg, ctypes.c_int32),
                          _p(g.birth, ctypes.c_int32), ctypes.byref(m.c), year, year_index,
                          n_new, _p(off, ctypes.c_int32), _p(tgt, ctypes.c_int32), cap,
                          _p(comp, ctypes.c_float), _p(info, ctypes.c_int32), threads)
    if e < 0:
        raise RuntimeError(f"or_abm_year failed: {e}")
    return {"off": off, "tgt": tgt[:e].copy(), "info": info[:n_new], "comp": comp}
