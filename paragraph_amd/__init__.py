"""ParAGraph graph-growth step on MI355X: Python binding of libparagraph.so.

The product is the C ABI in ``include/paragraph.h`` (HIP kernels for gfx950
behind it).  This module is a thin ctypes mirror of that ABI used by the
tests, the benchmark and the likelihood-fitting harness; it holds no compute
of its own and raises if the native library is missing.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparagraph.so")
ABM_PATH = os.path.join(_HERE, "abm")

PG_MODEL_PA = 0
PG_MODEL_ER = 1
PG_MODEL_ABM = 2
PG_LIK_ARITH_SIM = 0
PG_LIK_ARITH_FAST = 1

_ERRORS = {1: "EINVAL", 2: "ENOMEM", 3: "EHIP", 4: "ECAP", 5: "ESTATE", 6: "ENODEV", 7: "EIO"}


class PGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("max_papers", ctypes.c_int64), ("max_edges", ctypes.c_int64),
                ("max_new_per_year", ctypes.c_int64), ("max_cites", ctypes.c_int32),
                ("bfs_cap", ctypes.c_int32), ("seed", ctypes.c_uint64)]


class YearParams(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("model", ctypes.c_int32),
                ("year", ctypes.c_int32), ("year_index", ctypes.c_int32),
                ("n_new", ctypes.c_int64), ("m", ctypes.c_int32), ("n_nb", ctypes.c_int32),
                ("alpha", ctypes.c_float), ("inv_tau", ctypes.c_float),
                ("er_p", ctypes.c_double)]


class AbmModel(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("bag", ctypes.POINTER(ctypes.c_int32)),
                ("n_bag", ctypes.c_int32), ("rec_prob", ctypes.POINTER(ctypes.c_double)),
                ("n_rec", ctypes.c_int32), ("alpha", ctypes.c_double), ("w_pa", ctypes.c_double),
                ("w_rec", ctypes.c_double), ("w_fit", ctypes.c_double),
                ("fully_random", ctypes.c_double), ("same_year", ctypes.c_double)]


class YearStats(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("n_before", ctypes.c_int64),
                ("n_after", ctypes.c_int64), ("edges_added", ctypes.c_int64),
                ("near_boundary", ctypes.c_int64), ("shortfall", ctypes.c_int64),
                ("pa_attempts", ctypes.c_int64), ("device_ms", ctypes.c_double)]


class SeqInfo(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("n_years", ctypes.c_int32), ("m", ctypes.c_int32),
                ("n_papers", ctypes.c_int64), ("paper_years", ctypes.c_int64),
                ("new_papers", ctypes.c_int64)]


class LikParams(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("alpha", ctypes.c_double), ("tau", ctypes.c_double),
                ("alpha_tau_dev", ctypes.c_void_p), ("logfit_dev", ctypes.c_void_p),
                ("stream", ctypes.c_void_p), ("stream_given", ctypes.c_int32)]


class LikOut(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("scalars_dev", ctypes.c_void_p),
                ("grad_logfit_dev", ctypes.c_void_p), ("host3", ctypes.c_void_p)]


class LikBatchParams(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("n_points", ctypes.c_int32),
                ("alpha_tau", ctypes.c_void_p), ("alpha_tau_dev", ctypes.c_void_p),
                ("logfit_dev", ctypes.c_void_p), ("stream", ctypes.c_void_p),
                ("stream_given", ctypes.c_int32), ("arith", ctypes.c_int32)]


class LikBatchOut(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("scalars_dev", ctypes.c_void_p),
                ("grad_logfit_dev", ctypes.c_void_p), ("host", ctypes.c_void_p)]


class EnsembleConfig(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("device", ctypes.c_int32), ("n_sims", ctypes.c_int32),
                ("n_years", ctypes.c_int32), ("m", ctypes.c_int32), ("bfs_cap", ctypes.c_int32),
                ("first_year", ctypes.c_int32), ("growth_rate", ctypes.c_double)]


class SimParams(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("n_nb", ctypes.c_int32), ("n0", ctypes.c_int64),
                ("seed_edges", ctypes.c_int64), ("alpha", ctypes.c_float),
                ("inv_tau", ctypes.c_float), ("seed", ctypes.c_uint64)]


class EnsembleStats(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("edges_added", ctypes.c_int64),
                ("near_boundary", ctypes.c_int64), ("shortfall", ctypes.c_int64),
                ("pa_attempts", ctypes.c_int64), ("device_ms", ctypes.c_double),
                ("bfs_ms", ctypes.c_double), ("bfs_bytes", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("pa_lanes", ctypes.c_int32),
                ("lik_lanes", ctypes.c_int32), ("bfs_mode", ctypes.c_int32),
                ("bfs_small_max", ctypes.c_int32), ("bfs_mid_max", ctypes.c_int32),
                ("merge_out", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {"pa_lanes": self.pa_lanes, "lik_lanes": self.lik_lanes, "bfs_mode": self.bfs_mode,
                "bfs_small_max": self.bfs_small_max, "bfs_mid_max": self.bfs_mid_max,
                "merge_out": self.merge_out}


TUNE_MAX_TRIALS = 16


class TuneReport(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("n_trials", ctypes.c_int32),
                ("trial", Tuning * TUNE_MAX_TRIALS), ("trial_ms", ctypes.c_double * TUNE_MAX_TRIALS),
                ("chosen", Tuning), ("chosen_ms", ctypes.c_double), ("default_ms", ctypes.c_double),
                ("max_rel_dev", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {"trials": [dict(self.trial[k].as_dict(), ms=self.trial_ms[k])
                           for k in range(self.n_trials)],
                "chosen": self.chosen.as_dict(), "chosen_ms": self.chosen_ms,
                "default_ms": self.default_ms, "max_rel_dev": self.max_rel_dev}


KERNEL_GROUPS = ["weights_cdf", "pa_draw", "bfs", "nb_draw", "er", "csr_append", "lik",
                 "lik_grad", "lik_batch", "abm_scores", "abm_select", "abm_rows"]


_lib = None


def lib() -> ctypes.CDLL:
    """Loads libparagraph.so (built in-tree); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # One HIP runtime per process.  PyTorch-ROCm ships its own
        # libamdhip64/libhsa-runtime64; loaded first, it satisfies this
        # library's libamdhip64.so.7 dependency, so torch tensors and this
        # library share one runtime.  Loaded the other way round, the process
        # would map a second runtime from /opt/rocm and torch would find no
        # device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        i32p, i64p = P(ctypes.c_int32), P(ctypes.c_int64)
        f32p, f64p, i8p = P(ctypes.c_float), P(ctypes.c_double), P(ctypes.c_int8)
        vp = ctypes.c_void_p
        sig = {
            "pg_version": (ctypes.c_int, []),
            "pg_ctx_create": (ctypes.c_int, [P(Config), P(vp)]),
            "pg_ctx_destroy": (None, [vp]),
            "pg_last_error": (ctypes.c_char_p, [vp]),
            "pg_graph_load": (ctypes.c_int, [vp, ctypes.c_int64, i32p, f32p, ctypes.c_int64,
                                             ctypes.c_int64, i32p, i32p]),
            "pg_year_step": (ctypes.c_int, [vp, P(YearParams), P(YearStats)]),
            "pg_abm_configure": (ctypes.c_int, [vp, P(AbmModel)]),
            "pg_abm_year_info": (ctypes.c_int, [vp, i32p]),
            "pg_abm_scores": (ctypes.c_int, [vp, f32p]),
            "pg_abm_year_counts": (ctypes.c_int, [vp, i64p]),
            "pg_abm_paper_draws": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int64, i32p, f64p,
                                                  f64p, i32p]),
            "pg_abm_fitness_table": (None, [f64p]),
            "pg_graph_size": (ctypes.c_int, [vp, i64p, i64p]),
            "pg_graph_download": (ctypes.c_int, [vp, i32p, i32p, i32p, i32p, i32p, i32p, f32p]),
            "pg_year_citations": (ctypes.c_int, [vp, i32p, i8p, i32p, i32p]),
            "pg_year_weights": (ctypes.c_int, [vp, f32p, f64p]),
            "pg_bfs": (ctypes.c_int, [vp, ctypes.c_int64, i32p, i32p, i32p]),
            "pg_bfs_timed": (ctypes.c_int, [vp, ctypes.c_int64, i32p, i32p, i32p, i64p]),
            "pg_bfs_noalloc": (ctypes.c_int, [vp, ctypes.c_int64, vp, vp, vp, vp,
                                              ctypes.c_int32]),
            "pg_growth_count": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_double]),
            "pg_snapshot_save": (ctypes.c_int, [vp]),
            "pg_snapshot_restore": (ctypes.c_int, [vp]),
            "pg_get_progress": (ctypes.c_int, [vp, i32p, i32p, i32p]),
            "pg_checkpoint_save": (ctypes.c_int, [vp, ctypes.c_char_p, vp, ctypes.c_int64]),
            "pg_checkpoint_load": (ctypes.c_int, [vp, ctypes.c_char_p, vp, ctypes.c_int64, i64p]),
            "pg_profile_enable": (ctypes.c_int, [vp, ctypes.c_int]),
            "pg_profile_groups": (ctypes.c_int, [vp, ctypes.c_uint32]),
            "pg_profile_read": (ctypes.c_int, [vp, ctypes.c_int, f64p, f64p, i64p]),
            "pg_profile_reset": (ctypes.c_int, [vp]),
            "pg_seed_graph": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_double, i32p, i64p, i32p,
                                             i32p]),
            "pg_seed_logfit": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                                              ctypes.c_double, ctypes.c_double, f32p]),
            "pg_seq_reserve": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64,
                                              ctypes.c_int64]),
            "pg_seq_record": (ctypes.c_int, [vp, ctypes.c_int]),
            "pg_seq_append_year": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int64,
                                                  ctypes.c_int64, i32p, i32p, i8p, i32p]),
            "pg_seq_set_birth": (ctypes.c_int, [vp, ctypes.c_int64, i32p]),
            "pg_seq_get_info": (ctypes.c_int, [vp, P(SeqInfo)]),
            "pg_seq_download": (ctypes.c_int, [vp, i64p, i64p, i32p, i32p, i32p, i8p, i32p,
                                               i32p]),
            "pg_ensemble_create": (ctypes.c_int, [P(EnsembleConfig), P(SimParams), P(vp)]),
            "pg_ensemble_destroy": (None, [vp]),
            "pg_ensemble_error": (ctypes.c_char_p, [vp]),
            "pg_ensemble_load": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int64, i32p, f32p,
                                                ctypes.c_int64, i32p, i32p]),
            "pg_ensemble_run": (ctypes.c_int, [vp, ctypes.c_int32, P(EnsembleStats)]),
            "pg_ensemble_generate": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32,
                                                    ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double]),
            "pg_ensemble_restore": (ctypes.c_int, [vp]),
            "pg_ensemble_step": (ctypes.c_int, [vp, P(EnsembleStats)]),
            "pg_ensemble_year_draws": (ctypes.c_int, [vp, ctypes.c_int32, i32p, i8p, i32p, i32p,
                                                      f32p, f64p]),
            "pg_ensemble_sim_size": (ctypes.c_int, [vp, ctypes.c_int32, i64p, i64p]),
            "pg_ensemble_download": (ctypes.c_int, [vp, ctypes.c_int32, i32p, i32p, i32p, i32p,
                                                    i32p, i32p]),
            "pg_ensemble_set_tuning": (ctypes.c_int, [vp, P(Tuning)]),
            "pg_set_tuning": (ctypes.c_int, [vp, P(Tuning)]),
            "pg_ensemble_autotune": (ctypes.c_int, [vp, ctypes.c_int32, P(TuneReport)]),
            "pg_lik_autotune": (ctypes.c_int, [vp, P(LikParams), ctypes.c_int32, P(TuneReport)]),
            "pg_year_autotune": (ctypes.c_int, [vp, P(YearParams), ctypes.c_int32, ctypes.c_int32,
                                                P(TuneReport)]),
            "pg_pa_loglik": (ctypes.c_int, [vp, P(LikParams), P(LikOut)]),
            "pg_pa_loglik_grad": (ctypes.c_int, [vp, P(LikParams), P(LikOut)]),
            "pg_pa_loglik_grad_batch": (ctypes.c_int, [vp, P(LikBatchParams), P(LikBatchOut)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray | None, ctype):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def growth_count(n: int, rate: float) -> int:
    """ceil(n*rate) in double precision, as the reference grows the graph."""
    return int(lib().pg_growth_count(int(n), float(rate)))


def growth_schedule(n0: int, rate: float, years: int) -> list[int]:
    out, n = [], n0
    for _ in range(years):
        d = growth_count(n, rate)
        out.append(d)
        n += d
    return out


def abm_fitness_table() -> np.ndarray:
    """The 1000 cumulative fitness-peak weights the scoring model draws from."""
    c = np.empty(1000, np.float64)
    lib().pg_abm_fitness_table(_p(c, ctypes.c_double))
    return c


@dataclass
class SeedGraph:
    birth: np.ndarray      # int32[n]
    src: np.ndarray        # int32[e] citing
    dst: np.ndarray        # int32[e] cited
    logfit: np.ndarray     # float32[n_final]

    @property
    def n(self) -> int:
        return int(self.birth.shape[0])


def seed_graph(seed: int, n: int, mean_cites: float, n_final: int | None = None,
               year_lo: int = -20, year_hi: int = 0, sigma: float = 0.5) -> SeedGraph:
    """Synthetic seed graph plus log-fitness for n_final papers (host, Philox)."""
    L = lib()
    birth = np.empty(n, np.int32)
    ne = ctypes.c_int64(0)
    rc = L.pg_seed_graph(seed, n, year_lo, year_hi, mean_cites, _p(birth, ctypes.c_int32),
                         ctypes.byref(ne), None, None)
    if rc:
        raise PGError(rc, "pg_seed_graph")
    src = np.empty(ne.value, np.int32)
    dst = np.empty(ne.value, np.int32)
    rc = L.pg_seed_graph(seed, n, year_lo, year_hi, mean_cites, _p(birth, ctypes.c_int32),
                         ctypes.byref(ne), _p(src, ctypes.c_int32), _p(dst, ctypes.c_int32))
    if rc:
        raise PGError(rc, "pg_seed_graph")
    nf = n if n_final is None else n_final
    logfit = np.empty(nf, np.float32)
    rc = L.pg_seed_logfit(seed, 0, nf, 0.0, sigma, _p(logfit, ctypes.c_float))
    if rc:
        raise PGError(rc, "pg_seed_logfit")
    return SeedGraph(birth, src, dst, logfit)


class Context:
    """One simulation on one GPU (wraps pg_ctx)."""

    def __init__(self, max_papers: int, max_edges: int, max_new_per_year: int,
                 max_cites: int = 32, bfs_cap: int = 4096, seed: int = 1, device: int = 0):
        L = lib()
        cfg = Config(ctypes.sizeof(Config), device, max_papers, max_edges, max_new_per_year,
                     max_cites, bfs_cap, seed)
        h = ctypes.c_void_p()
        rc = L.pg_ctx_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise PGError(rc, L.pg_last_error(None).decode())
        self._h = h
        self.cfg = cfg

    def close(self):
        if getattr(self, "_h", None):
            lib().pg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc:
            raise PGError(rc, f"{what}: {lib().pg_last_error(self._h).decode()}")

    def load(self, g: SeedGraph):
        birth = np.ascontiguousarray(g.birth, np.int32)
        logfit = np.ascontiguousarray(g.logfit, np.float32)
        src = np.ascontiguousarray(g.src, np.int32)
        dst = np.ascontiguousarray(g.dst, np.int32)
        self._check(lib().pg_graph_load(self._h, birth.shape[0], _p(birth, ctypes.c_int32),
                                        _p(logfit, ctypes.c_float), logfit.shape[0],
                                        src.shape[0], _p(src, ctypes.c_int32),
                                        _p(dst, ctypes.c_int32)), "pg_graph_load")

    def year_pa(self, year: int, year_index: int, n_new: int, m: int, n_nb: int,
                alpha: float, tau: float) -> YearStats:
        inv_tau = 0.0 if math.isinf(tau) else 1.0 / tau
        p = YearParams(ctypes.sizeof(YearParams), PG_MODEL_PA, year, year_index, n_new, m,
                       n_nb, alpha, inv_tau, 0.0)
        st = YearStats(ctypes.sizeof(YearStats))
        self._check(lib().pg_year_step(self._h, ctypes.byref(p), ctypes.byref(st)),
                    "pg_year_step")
        return st

    def year_er(self, year: int, year_index: int, n_new: int, p_link: float) -> YearStats:
        p = YearParams(ctypes.sizeof(YearParams), PG_MODEL_ER, year, year_index, n_new, 0, 0,
                       0.0, 0.0, p_link)
        st = YearStats(ctypes.sizeof(YearStats))
        self._check(lib().pg_year_step(self._h, ctypes.byref(p), ctypes.byref(st)),
                    "pg_year_step")
        return st

    # ---- the reference's scoring model ----
    def abm_configure(self, bag, rec_prob, alpha: float = 0.5, w_pa: float = 1 / 3,
                      w_rec: float = 1 / 3, w_fit: float = 1 / 3, fully_random: float = 0.05,
                      same_year: float = 0.12):
        """pg_abm_configure: out-degree bag, recency table p(d) for d = 0..,
        neighbourhood share (< 0: per paper), weights (any -1: per paper)."""
        bag = np.ascontiguousarray(bag, np.int32)
        rec = np.ascontiguousarray(rec_prob, np.float64)
        m = AbmModel(ctypes.sizeof(AbmModel), _p(bag, ctypes.c_int32), bag.shape[0],
                     _p(rec, ctypes.c_double) if rec.size else None, rec.shape[0], alpha, w_pa,
                     w_rec, w_fit, fully_random, same_year)
        self._check(lib().pg_abm_configure(self._h, ctypes.byref(m)), "pg_abm_configure")

    def year_abm(self, year: int, year_index: int, n_new: int) -> YearStats:
        p = YearParams(ctypes.sizeof(YearParams), PG_MODEL_ABM, year, year_index, n_new, 0, 0,
                       0.0, 0.0, 0.0)
        st = YearStats(ctypes.sizeof(YearStats))
        self._check(lib().pg_year_step(self._h, ctypes.byref(p), ctypes.byref(st)),
                    "pg_year_step")
        return st

    def abm_year_info(self, n_new: int) -> np.ndarray:
        """Per new paper of the last scoring-model year: generator, out-degree,
        same-year target, 1-hop picks, 2-hop picks, uniform picks."""
        info = np.empty((max(n_new, 1), 6), np.int32)
        self._check(lib().pg_abm_year_info(self._h, _p(info, ctypes.c_int32)),
                    "pg_abm_year_info")
        return info[:n_new]

    def abm_year_counts(self) -> tuple[int, int]:
        """1-hop and 2-hop candidate entries keyed in the last scoring-model year."""
        c = np.zeros(2, np.int64)
        self._check(lib().pg_abm_year_counts(self._h, _p(c, ctypes.c_int64)), "pg_abm_year_counts")
        return int(c[0]), int(c[1])

    def abm_scores(self, n_exist: int) -> np.ndarray:
        comp = np.empty((max(n_exist, 1), 4), np.float32)
        self._check(lib().pg_abm_scores(self._h, _p(comp, ctypes.c_float)), "pg_abm_scores")
        return comp[:n_exist]

    def abm_paper_draws(self, first: int, count: int) -> dict:
        k = np.empty(count, np.int32)
        a = np.empty(count, np.float64)
        w = np.empty((count, 3), np.float64)
        f = np.empty(count, np.int32)
        self._check(lib().pg_abm_paper_draws(self._h, first, count, _p(k, ctypes.c_int32),
                                             _p(a, ctypes.c_double), _p(w, ctypes.c_double),
                                             _p(f, ctypes.c_int32)), "pg_abm_paper_draws")
        return {"out_degree": k, "alpha": a, "weights": w, "fitness": f}

    def size(self) -> tuple[int, int]:
        n, e = ctypes.c_int64(), ctypes.c_int64()
        self._check(lib().pg_graph_size(self._h, ctypes.byref(n), ctypes.byref(e)),
                    "pg_graph_size")
        return n.value, e.value

    def download(self) -> dict:
        n, e = self.size()
        out = {k: np.empty(s, np.int32) for k, s in
               [("fwd_off", n + 1), ("fwd_adj", e), ("bwd_off", n + 1), ("bwd_adj", e),
                ("in_deg", n), ("birth", n)]}
        out["logfit"] = np.empty(n, np.float32)
        i32 = ctypes.c_int32
        self._check(lib().pg_graph_download(
            self._h, _p(out["fwd_off"], i32), _p(out["fwd_adj"], i32), _p(out["bwd_off"], i32),
            _p(out["bwd_adj"], i32), _p(out["in_deg"], i32), _p(out["birth"], i32),
            _p(out["logfit"], ctypes.c_float)), "pg_graph_download")
        return out

    def citations(self, n_new: int, m: int) -> dict:
        cit = np.empty(n_new * m, np.int32)
        kind = np.empty(n_new * m, np.int8)
        cnt = np.empty(n_new, np.int32)
        near = np.empty(n_new, np.int32)
        self._check(lib().pg_year_citations(self._h, _p(cit, ctypes.c_int32),
                                            _p(kind, ctypes.c_int8), _p(cnt, ctypes.c_int32),
                                            _p(near, ctypes.c_int32)), "pg_year_citations")
        return {"cit": cit.reshape(n_new, m), "kind": kind.reshape(n_new, m), "cnt": cnt,
                "near": near}

    def weights(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        w = np.empty(n, np.float32)
        c = np.empty(n, np.float64)
        self._check(lib().pg_year_weights(self._h, _p(w, ctypes.c_float),
                                          _p(c, ctypes.c_double)), "pg_year_weights")
        return w, c

    def bfs_timed(self, roots: np.ndarray):
        """Neighbourhoods plus per-root (ticks to 1-hop end, ticks total,
        2-hop candidate entries) from the workgroup path."""
        roots = np.ascontiguousarray(roots, np.int32)
        cap = self.cfg.bfs_cap
        out = np.empty(roots.shape[0] * cap, np.int32)
        ln = np.empty(roots.shape[0], np.int32)
        tk = np.empty(roots.shape[0] * 3, np.int64)
        self._check(lib().pg_bfs_timed(self._h, roots.shape[0], _p(roots, ctypes.c_int32),
                                       _p(out, ctypes.c_int32), _p(ln, ctypes.c_int32),
                                       _p(tk, ctypes.c_int64)), "pg_bfs_timed")
        return ln, tk.reshape(-1, 3)

    def bfs(self, roots: np.ndarray) -> list[np.ndarray]:
        roots = np.ascontiguousarray(roots, np.int32)
        cap = self.cfg.bfs_cap
        out = np.empty(roots.shape[0] * cap, np.int32)
        ln = np.empty(roots.shape[0], np.int32)
        self._check(lib().pg_bfs(self._h, roots.shape[0], _p(roots, ctypes.c_int32),
                                 _p(out, ctypes.c_int32), _p(ln, ctypes.c_int32)), "pg_bfs")
        out = out.reshape(-1, cap)
        return [out[i, :ln[i]].copy() for i in range(roots.shape[0])]

    def bfs_noalloc(self, count: int, roots_dev: int, out_dev: int, out_len_dev: int,
                    stream: int | None = None):
        """pg_bfs_noalloc on device buffers (pointers as ints): enqueues the
        neighbourhoods of roots_dev[count] into out_dev[count*cap] and
        out_len_dev[count] without allocating; stream None = the context's."""
        self._check(lib().pg_bfs_noalloc(self._h, int(count), roots_dev or None, out_dev or None,
                                         out_len_dev or None, stream or None,
                                         0 if stream is None else 1), "pg_bfs_noalloc")

    # ---- citation sequence and PA likelihood ----
    def seq_reserve(self, m: int, max_paper_years: int, max_new_papers: int,
                    max_papers: int | None = None):
        mp = self.cfg.max_papers if max_papers is None else max_papers
        self._check(lib().pg_seq_reserve(self._h, m, max_paper_years, max_new_papers, mp),
                    "pg_seq_reserve")

    def seq_record(self, on: bool = True):
        self._check(lib().pg_seq_record(self._h, 1 if on else 0), "pg_seq_record")

    def seq_append_year(self, year: int, deg: np.ndarray, cit: np.ndarray, kind: np.ndarray,
                        cnt: np.ndarray):
        deg = np.ascontiguousarray(deg, np.int32)
        cit = np.ascontiguousarray(cit, np.int32)
        kind = np.ascontiguousarray(kind, np.int8)
        cnt = np.ascontiguousarray(cnt, np.int32)
        self._check(lib().pg_seq_append_year(self._h, year, deg.shape[0], cnt.shape[0],
                                             _p(deg, ctypes.c_int32), _p(cit, ctypes.c_int32),
                                             _p(kind, ctypes.c_int8), _p(cnt, ctypes.c_int32)),
                    "pg_seq_append_year")

    def seq_set_birth(self, birth: np.ndarray):
        birth = np.ascontiguousarray(birth, np.int32)
        self._check(lib().pg_seq_set_birth(self._h, birth.shape[0], _p(birth, ctypes.c_int32)),
                    "pg_seq_set_birth")

    def seq_info(self) -> SeqInfo:
        info = SeqInfo(ctypes.sizeof(SeqInfo))
        self._check(lib().pg_seq_get_info(self._h, ctypes.byref(info)), "pg_seq_get_info")
        return info

    def seq_download(self) -> dict:
        """The sequence as per-year lists (the oracle likelihood's inputs)."""
        info = self.seq_info()
        Y, m = info.n_years, info.m
        n_exist = np.zeros(Y, np.int64)
        n_new = np.zeros(Y, np.int64)
        year_of = np.zeros(Y, np.int32)
        deg = np.zeros(info.paper_years, np.int32)
        cit = np.zeros(info.new_papers * m, np.int32)
        kind = np.zeros(info.new_papers * m, np.int8)
        cnt = np.zeros(info.new_papers, np.int32)
        birth = np.zeros(info.n_papers, np.int32)
        self._check(lib().pg_seq_download(self._h, _p(n_exist, ctypes.c_int64),
                                          _p(n_new, ctypes.c_int64), _p(year_of, ctypes.c_int32),
                                          _p(deg, ctypes.c_int32), _p(cit, ctypes.c_int32),
                                          _p(kind, ctypes.c_int8), _p(cnt, ctypes.c_int32),
                                          _p(birth, ctypes.c_int32)), "pg_seq_download")
        out = dict(n_exist=list(n_exist), n_new=list(n_new), year_of=list(year_of), deg=[],
                   cit=[], kind=[], cnt=[], birth=birth, m=m)
        d0 = e0 = 0
        for y in range(Y):
            N, D = int(n_exist[y]), int(n_new[y])
            out["deg"].append(deg[d0:d0 + N])
            out["cit"].append(cit[e0 * m:(e0 + D) * m].reshape(D, m))
            out["kind"].append(kind[e0 * m:(e0 + D) * m].reshape(D, m))
            out["cnt"].append(cnt[e0:e0 + D])
            d0 += N
            e0 += D
        return out

    def _lik(self, fn: str, alpha: float, tau: float, logfit_dev: int, *, grad_dev: int = 0,
             scalars_dev: int = 0, alpha_tau_dev: int = 0, stream: int | None = None,
             host: bool = True):
        """stream: a hipStream_t handle (0 = the null stream), or None for the
        context's own stream."""
        p = LikParams(ctypes.sizeof(LikParams), float(alpha), float(tau), alpha_tau_dev or None,
                      logfit_dev or None, stream or None, 0 if stream is None else 1)
        h3 = (ctypes.c_double * 3)()
        o = LikOut(ctypes.sizeof(LikOut), scalars_dev or None, grad_dev or None,
                   ctypes.cast(h3, ctypes.c_void_p) if host else None)
        self._check(getattr(lib(), fn)(self._h, ctypes.byref(p), ctypes.byref(o)), fn)
        return (h3[0], h3[1], h3[2]) if host else None

    def pa_loglik(self, alpha: float, tau: float, logfit_dev: int, **kw):
        """log L (device pointers as ints); returns (log L, 0, 0) when host=True."""
        return self._lik("pg_pa_loglik", alpha, tau, logfit_dev, **kw)

    def pa_loglik_grad(self, alpha: float, tau: float, logfit_dev: int, grad_dev: int, **kw):
        """log L, dL/dalpha, dL/dtau (host=True) and dL/dphi into grad_dev (FP64)."""
        return self._lik("pg_pa_loglik_grad", alpha, tau, logfit_dev, grad_dev=grad_dev, **kw)

    def pa_loglik_grad_batch(self, alpha_tau, logfit_dev: int, grad_dev: int, *,
                             alpha_tau_dev: int = 0, scalars_dev: int = 0,
                             stream: int | None = None, host: bool = True, fast: bool = False):
        """log L and its gradient at B = len(alpha_tau) points in one pass.
        logfit_dev: device float[n_papers*B] (point b of paper j at j*B+b);
        grad_dev: device double[n_papers*B].  Returns a (B, 3) array of
        {log L, dL/dalpha, dL/dtau} when host=True.  fast: the hardware
        exp/log/reciprocal (PG_LIK_ARITH_FAST) instead of the simulator's
        arithmetic."""
        at = np.ascontiguousarray(np.asarray(alpha_tau, np.float64).reshape(-1, 2))
        B = at.shape[0]
        p = LikBatchParams(ctypes.sizeof(LikBatchParams), B, at.ctypes.data,
                           alpha_tau_dev or None, logfit_dev or None, stream or None,
                           0 if stream is None else 1,
                           PG_LIK_ARITH_FAST if fast else PG_LIK_ARITH_SIM)
        h = np.zeros((B, 3), np.float64)
        o = LikBatchOut(ctypes.sizeof(LikBatchOut), scalars_dev or None, grad_dev or None,
                        h.ctypes.data if host else None)
        self._check(lib().pg_pa_loglik_grad_batch(self._h, ctypes.byref(p), ctypes.byref(o)),
                    "pg_pa_loglik_grad_batch")
        return h if host else None

    def lik_autotune(self, alpha: float, tau: float, logfit_dev: int, reps: int = 3) -> dict:
        """Times the likelihood's lanes-per-citing-paper candidates on this
        sequence and keeps the fastest; returns the report."""
        p = LikParams(ctypes.sizeof(LikParams), float(alpha), float(tau), None,
                      logfit_dev or None, None, 0)
        r = TuneReport()
        r.size = ctypes.sizeof(TuneReport)
        self._check(lib().pg_lik_autotune(self._h, ctypes.byref(p), int(reps), ctypes.byref(r)),
                    "pg_lik_autotune")
        return r.as_dict()

    def set_tuning(self, pa_lanes: int = 0, lik_lanes: int = 0, bfs_mode: int = 0,
                   bfs_small_max: int = 0, bfs_mid_max: int = 0, merge_out: int = 0):
        """Launch shapes (speed only): PA lanes per paper, likelihood lanes
        per citing paper, 2-hop workgroup path mode for lists above 4096
        candidate entries (0 radix-sorted in LDS when they fit, else staged
        up to 65536 entries, else swept; 1 always swept; 2 staged whenever
        it fits), the one-wavefront 2-hop tiers' candidate limits, the
        merge tile."""
        t = Tuning(ctypes.sizeof(Tuning), pa_lanes, lik_lanes, bfs_mode, bfs_small_max,
                   bfs_mid_max, merge_out)
        self._check(lib().pg_set_tuning(self._h, ctypes.byref(t)), "pg_set_tuning")

    def year_autotune(self, years: list[dict], reps: int = 2) -> dict:
        """pg_year_autotune over PA year steps given as dicts (year,
        year_index, n_new, m, n_nb, alpha, tau); the context's state is kept,
        the chosen shapes are set; returns the report."""
        arr = (YearParams * len(years))()
        for k, y in enumerate(years):
            tau = y["tau"]
            arr[k] = YearParams(ctypes.sizeof(YearParams), PG_MODEL_PA, y["year"], y["year_index"],
                                y["n_new"], y["m"], y["n_nb"], y["alpha"],
                                0.0 if math.isinf(tau) else 1.0 / tau, 0.0)
        r = TuneReport()
        r.size = ctypes.sizeof(TuneReport)
        self._check(lib().pg_year_autotune(self._h, arr, len(years), int(reps), ctypes.byref(r)),
                    "pg_year_autotune")
        return r.as_dict()

    def snapshot_save(self):
        self._check(lib().pg_snapshot_save(self._h), "pg_snapshot_save")

    def snapshot_restore(self):
        self._check(lib().pg_snapshot_restore(self._h), "pg_snapshot_restore")

    def progress(self) -> tuple[int, int, int]:
        """(year steps done, last year, last year index)."""
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._check(lib().pg_get_progress(self._h, ctypes.byref(a), ctypes.byref(b),
                                          ctypes.byref(c)), "pg_get_progress")
        return a.value, b.value, c.value

    def checkpoint_save(self, path: str, user: bytes = b""):
        """Writes the graph state, progress and `user` to `path`."""
        buf = ctypes.create_string_buffer(user, len(user)) if user else None
        self._check(lib().pg_checkpoint_save(self._h, os.fsencode(path), buf, len(user)),
                    "pg_checkpoint_save")

    def checkpoint_load(self, path: str, user_cap: int = 1 << 20) -> bytes:
        """Replaces the graph state by the checkpoint's; returns its user blob."""
        buf = ctypes.create_string_buffer(max(user_cap, 1))
        nb = ctypes.c_int64()
        self._check(lib().pg_checkpoint_load(self._h, os.fsencode(path), buf, user_cap,
                                             ctypes.byref(nb)), "pg_checkpoint_load")
        return buf.raw[:nb.value]

    def profile(self, on: bool = True):
        self._check(lib().pg_profile_enable(self._h, 1 if on else 0), "pg_profile_enable")

    def profile_groups(self, names=None):
        """Time only the named kernel groups while profiling (None: all)."""
        mask = 0xFFFFFFFF if names is None else sum(1 << KERNEL_GROUPS.index(n) for n in names)
        self._check(lib().pg_profile_groups(self._h, mask), "pg_profile_groups")

    def profile_reset(self):
        self._check(lib().pg_profile_reset(self._h), "pg_profile_reset")

    def profile_read(self) -> dict:
        k = len(KERNEL_GROUPS)
        ms = np.zeros(k, np.float64)
        by = np.zeros(k, np.float64)
        n = np.zeros(k, np.int64)
        self._check(lib().pg_profile_read(self._h, k, _p(ms, ctypes.c_double),
                                          _p(by, ctypes.c_double), _p(n, ctypes.c_int64)),
                    "pg_profile_read")
        return {name: {"ms": float(ms[i]), "bytes": float(by[i]), "launches": int(n[i])}
                for i, name in enumerate(KERNEL_GROUPS)}


class Ensemble:
    """Many independent PA simulations advanced together (wraps pg_ensemble).

    sims: list of dicts with n0, alpha, tau, n_nb, seed and seed_edges (an
    upper bound of the distinct seed citations that load() will receive).
    """

    def __init__(self, sims: list, n_years: int, m: int, rate: float = 0.03,
                 bfs_cap: int = 4096, first_year: int = 1, device: int = 0):
        L = lib()
        cfg = EnsembleConfig(ctypes.sizeof(EnsembleConfig), device, len(sims), n_years, m,
                             bfs_cap, first_year, rate)
        arr = (SimParams * len(sims))()
        for k, sp in enumerate(sims):
            tau = sp["tau"]
            arr[k] = SimParams(ctypes.sizeof(SimParams), int(sp["n_nb"]), int(sp["n0"]),
                               int(sp["seed_edges"]), float(sp["alpha"]),
                               0.0 if math.isinf(tau) else 1.0 / tau, int(sp["seed"]))
        h = ctypes.c_void_p()
        rc = L.pg_ensemble_create(ctypes.byref(cfg), arr, ctypes.byref(h))
        if rc:
            raise PGError(rc, L.pg_ensemble_error(None).decode())
        self._h = h
        self.cfg = cfg
        self.sims = sims

    def close(self):
        if getattr(self, "_h", None):
            lib().pg_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, what: str):
        if rc:
            raise PGError(rc, f"{what}: {lib().pg_ensemble_error(self._h).decode()}")

    def load(self, sim: int, g: SeedGraph):
        birth = np.ascontiguousarray(g.birth, np.int32)
        logfit = np.ascontiguousarray(g.logfit, np.float32)
        src = np.ascontiguousarray(g.src, np.int32)
        dst = np.ascontiguousarray(g.dst, np.int32)
        self._check(lib().pg_ensemble_load(self._h, sim, birth.shape[0], _p(birth, ctypes.c_int32),
                                           _p(logfit, ctypes.c_float), src.shape[0],
                                           _p(src, ctypes.c_int32), _p(dst, ctypes.c_int32)),
                    "pg_ensemble_load")

    def generate(self, mean_cites: float, year_lo: int = -20, year_hi: int = 0,
                 mu: float = 0.0, sigma: float = 0.5):
        """Every simulation's seed graph and log-fitness generated on the
        device (instead of load): what seed_graph(sim seed, n0, mean_cites,
        n_final) gives, bit for bit."""
        self._check(lib().pg_ensemble_generate(self._h, year_lo, year_hi, float(mean_cites),
                                               float(mu), float(sigma)), "pg_ensemble_generate")

    def run(self, use_graph: bool = True) -> EnsembleStats:
        st = EnsembleStats(ctypes.sizeof(EnsembleStats))
        self._check(lib().pg_ensemble_run(self._h, 1 if use_graph else 0, ctypes.byref(st)),
                    "pg_ensemble_run")
        return st

    def restore(self):
        self._check(lib().pg_ensemble_restore(self._h), "pg_ensemble_restore")

    def step(self) -> EnsembleStats:
        """The next growth year for every simulation (eager)."""
        st = EnsembleStats(ctypes.sizeof(EnsembleStats))
        self._check(lib().pg_ensemble_step(self._h, ctypes.byref(st)), "pg_ensemble_step")
        return st

    def year_draws(self, sim: int, n_new: int, n_exist: int) -> dict:
        """The last year's draws of `sim` (local ids) and its weights and CDF."""
        m = self.cfg.m
        cit = np.empty(n_new * m, np.int32)
        kind = np.empty(n_new * m, np.int8)
        cnt = np.empty(n_new, np.int32)
        near = np.empty(n_new, np.int32)
        w = np.empty(n_exist, np.float32)
        cdf = np.empty(n_exist, np.float64)
        self._check(lib().pg_ensemble_year_draws(
            self._h, sim, _p(cit, ctypes.c_int32), _p(kind, ctypes.c_int8),
            _p(cnt, ctypes.c_int32), _p(near, ctypes.c_int32), _p(w, ctypes.c_float),
            _p(cdf, ctypes.c_double)), "pg_ensemble_year_draws")
        return {"cit": cit.reshape(n_new, m), "kind": kind.reshape(n_new, m), "cnt": cnt,
                "near": near, "w": w, "cdf": cdf}

    def set_tuning(self, pa_lanes: int = 0, lik_lanes: int = 0, bfs_mode: int = 0,
                   bfs_small_max: int = 0, bfs_mid_max: int = 0, merge_out: int = 0):
        """Launch shapes (speed only): PA lanes per paper, 2-hop path mode,
        the one-wavefront 2-hop tiers' candidate limits, the merge tile."""
        t = Tuning(ctypes.sizeof(Tuning), pa_lanes, lik_lanes, bfs_mode, bfs_small_max,
                   bfs_mid_max, merge_out)
        self._check(lib().pg_ensemble_set_tuning(self._h, ctypes.byref(t)),
                    "pg_ensemble_set_tuning")

    def autotune(self, reps: int = 2) -> dict:
        """Times PA lanes and 2-hop modes on whole graph-captured runs of this
        ensemble, keeps the fastest and returns the report (seed state kept)."""
        r = TuneReport()
        r.size = ctypes.sizeof(TuneReport)
        self._check(lib().pg_ensemble_autotune(self._h, int(reps), ctypes.byref(r)),
                    "pg_ensemble_autotune")
        return r.as_dict()

    def download(self, sim: int) -> dict:
        n, e = ctypes.c_int64(), ctypes.c_int64()
        self._check(lib().pg_ensemble_sim_size(self._h, sim, ctypes.byref(n), ctypes.byref(e)),
                    "pg_ensemble_sim_size")
        n, e = n.value, e.value
        out = {k: np.empty(sz, np.int32) for k, sz in
               [("fwd_off", n + 1), ("fwd_adj", e), ("bwd_off", n + 1), ("bwd_adj", e),
                ("in_deg", n), ("birth", n)]}
        i32 = ctypes.c_int32
        self._check(lib().pg_ensemble_download(
            self._h, sim, _p(out["fwd_off"], i32), _p(out["fwd_adj"], i32),
            _p(out["bwd_off"], i32), _p(out["bwd_adj"], i32), _p(out["in_deg"], i32),
            _p(out["birth"], i32)), "pg_ensemble_download")
        return out
