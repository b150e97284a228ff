"""MI355X period scan of PeriodSearch (convex lightcurve inversion).

Python mirror of the reference's GPU layer interface (start_CUDA.h:1-18):
``Scan.prepare`` / ``Scan.precalc`` / ``Scan.start`` / ``Scan.close`` map to
CUDAPrepare / CUDAPrecalc / CUDAStart / CUDAUnprepare, on top of the C ABI
(include/ps_mi355x.h) and the hand-written gfx950 kernels.  Errors are raised
as ``PsError`` carrying the library's status code and message (the reference
exits the process instead).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _abi

__all__ = ["PsError", "Scan", "Line", "format_line", "write_output"]


class PsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ps_mi355x error {code}: {msg}")
        self.code = code


@dataclass
class Line:
    period_h: float
    rms: float
    chi2: float
    dark_or_conw: float
    lambda_deg: float
    beta_deg: float


def format_line(L: Line) -> str:
    """period_search_out line (start_CUDA.cu:1222-1226)."""
    return "%.8f  %.6f  %.6f %4.1f %4.0f %4.0f\n" % (
        L.period_h, L.rms, L.chi2, L.dark_or_conw, _c_round(L.lambda_deg), _c_round(L.beta_deg))


def _c_round(x: float) -> float:
    # C round(): halves away from zero
    return float(np.sign(x) * np.floor(abs(x) + 0.5))


def write_output(path: str, lines: Sequence[Line]) -> None:
    with open(path, "w") as f:
        for L in lines:
            if not np.isnan(L.period_h):
                f.write(format_line(L))


_RESULT_FIELDS = ("reported", "niter", "dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "n_vis")


RESULT_DTYPE = np.dtype([(f, "f8") for f in _RESULT_FIELDS])
# ps_pair_result as laid out in memory (include/ps_mi355x.h)
_C_RESULT = np.dtype([("reported", "i4"), ("niter", "i4"), ("dev", "f8"), ("dark", "f8"), ("period", "f8"),
                      ("lambda_deg", "f8"), ("beta_deg", "f8"), ("chisq", "f8"), ("n_vis", "i8")])


def _results(out, n: int) -> np.ndarray:
    """The first n ps_pair_result entries of a ctypes array, as a float64
    record array."""
    assert C.sizeof(_abi.PsPairResult) == _C_RESULT.itemsize
    raw = np.frombuffer(out, dtype=_C_RESULT, count=n)
    res = np.zeros(n, dtype=RESULT_DTYPE)
    for f in _RESULT_FIELDS:
        res[f] = raw[f]
    return res


class Scan:
    """One device context holding one prepared problem (or a batch of
    problems sharing a shape basis)."""

    def __init__(self, device: int = 0):
        self._lib = _abi.load()
        h = C.c_void_p()
        rc = self._lib.ps_open(device, C.byref(h))
        if rc != _abi.PS_OK:
            raise PsError(rc, f"cannot open HIP device {device}")
        self._h = h
        self.info = None

    # -- helpers -----------------------------------------------------------
    def _check(self, rc: int):
        if rc != _abi.PS_OK:
            raise PsError(rc, self._lib.ps_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.ps_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- reference interface ------------------------------------------------
    def prepare_file(self, path: str):
        """Parse period_search_in and upload the problem (CUDAPrepare)."""
        self._check(self._lib.ps_prepare_file(self._h, path.encode()))
        self._load_info()
        return self

    def _load_info(self):
        ints = np.zeros(8, np.int32)
        dbls = np.zeros(4, np.float64)
        self._check(self._lib.ps_info(self._h, _abi.iptr(ints), _abi.dptr(dbls)))
        keys = ("ndata", "n_curves", "nfac", "ncoef", "ma", "mfit", "n_max", "n_iter_max")
        self.info = {k: int(v) for k, v in zip(keys, ints)}
        self.info.update(f_start=float(dbls[0]), f_step=float(dbls[1]), escl=float(dbls[2]), conw=float(dbls[3]))

    def precalc(self, conw_r: float):
        """One convexity-weight round (CUDAPrecalc): returns (conw_r, new_conw)."""
        cw = C.c_double(conw_r)
        nc = C.c_int(0)
        self._check(self._lib.ps_precalc(self._h, C.byref(cw), C.byref(nc)))
        return cw.value, nc.value

    def conw_search(self) -> float:
        cw = C.c_double(0.0)
        self._check(self._lib.ps_conw_search(self._h, C.byref(cw)))
        return cw.value

    def start(self, conw_r: float, n_start: int = 1, n_count: Optional[int] = None) -> List[Line]:
        """Scan n_start..n_start+n_count-1 (CUDAStart); one Line per frequency."""
        if n_count is None:
            n_count = self.info["n_max"] - n_start + 1
        arr = (_abi.PsLine * max(n_count, 1))()
        self._check(self._lib.ps_start(self._h, n_start, n_count, conw_r, arr))
        return [Line(a.period_h, a.rms, a.chi2, a.dark_or_conw, a.lambda_deg, a.beta_deg)
                for a in arr[:n_count]]

    def fit_pairs(self, n_list, m_list, conw_r: float, conw_r_dev: Optional[float] = None):
        """Raw per-pair results as a structured numpy array."""
        n_arr = np.ascontiguousarray(n_list, np.int32)
        m_arr = np.ascontiguousarray(m_list, np.int32)
        out = (_abi.PsPairResult * max(len(n_arr), 1))()
        dev = conw_r if conw_r_dev is None else conw_r_dev
        self._check(self._lib.ps_fit_pairs(self._h, len(n_arr), _abi.iptr(n_arr), _abi.iptr(m_arr),
                                           conw_r, dev, out))
        return _results(out, len(n_arr))

    # -- batches of asteroids (one persistent launch for all) ---------------
    def prepare_files(self, paths: Sequence[str]):
        """Upload several period_search_in files sharing one shape basis."""
        arr = (C.c_char_p * len(paths))(*[p.encode() for p in paths])
        self._check(self._lib.ps_prepare_files(self._h, len(paths), arr))
        self._load_info()
        self.batch_info = [self._problem_info(i) for i in range(self._lib.ps_n_problems(self._h))]
        return self

    def _problem_info(self, i: int) -> dict:
        ints = np.zeros(8, np.int32)
        dbls = np.zeros(4, np.float64)
        self._check(self._lib.ps_problem_info(self._h, i, _abi.iptr(ints), _abi.dptr(dbls)))
        keys = ("ndata", "n_curves", "nfac", "ncoef", "ma", "mfit", "n_max", "n_iter_max")
        d = {k: int(v) for k, v in zip(keys, ints)}
        d.update(f_start=float(dbls[0]), f_step=float(dbls[1]), escl=float(dbls[2]), conw=float(dbls[3]))
        return d

    def conw_search_batch(self) -> np.ndarray:
        out = np.zeros(len(self.batch_info))
        self._check(self._lib.ps_conw_search_batch(self._h, _abi.dptr(out)))
        return out

    def start_batch(self, conw_r, n_start=None, n_count=None) -> List[List[Line]]:
        """Scan every problem of the batch in one launch; lines per problem."""
        n = len(self.batch_info)
        ns = np.ascontiguousarray(np.ones(n) if n_start is None else n_start, np.int32)
        nc = np.ascontiguousarray([b["n_max"] - int(s) + 1 for b, s in zip(self.batch_info, ns)]
                                  if n_count is None else n_count, np.int32)
        cw = np.ascontiguousarray(conw_r, np.float64)
        total = int(nc.sum())
        arr = (_abi.PsLine * max(total, 1))()
        self._check(self._lib.ps_start_batch(self._h, _abi.iptr(ns), _abi.iptr(nc), _abi.dptr(cw), arr))
        lines, k = [], 0
        for i in range(n):
            lines.append([Line(a.period_h, a.rms, a.chi2, a.dark_or_conw, a.lambda_deg, a.beta_deg)
                          for a in arr[k:k + int(nc[i])]])
            k += int(nc[i])
        return lines

    def fit_items(self, prob, n_list, m_list, conw_r, conw_r_dev=None):
        """Pairs of several problems in one launch (per-item conwR)."""
        p_arr = np.ascontiguousarray(prob, np.int32)
        n_arr = np.ascontiguousarray(n_list, np.int32)
        m_arr = np.ascontiguousarray(m_list, np.int32)
        cw = np.ascontiguousarray(np.broadcast_to(conw_r, p_arr.shape), np.float64)
        cd = np.ascontiguousarray(np.broadcast_to(conw_r if conw_r_dev is None else conw_r_dev, p_arr.shape),
                                  np.float64)
        out = (_abi.PsPairResult * max(len(p_arr), 1))()
        self._check(self._lib.ps_fit_items(self._h, len(p_arr), _abi.iptr(p_arr), _abi.iptr(n_arr),
                                           _abi.iptr(m_arr), _abi.dptr(cw), _abi.dptr(cd), out))
        return _results(out, len(p_arr))

    # -- warm starts -----------------------------------------------------------
    def warm_features(self, prob, n_list, m_list) -> np.ndarray:
        p_arr = np.ascontiguousarray(prob, np.int32)
        n_arr = np.ascontiguousarray(n_list, np.int32)
        m_arr = np.ascontiguousarray(m_list, np.int32)
        nf = self._lib.ps_warm_n_features()
        out = np.zeros((len(p_arr), nf), np.float32)
        self._check(self._lib.ps_warm_features(self._h, len(p_arr), _abi.iptr(p_arr), _abi.iptr(n_arr),
                                               _abi.iptr(m_arr), _abi.fptr(out)))
        return out

    def screen_features(self, n_start: int, n_count: int, prob: int = 0) -> np.ndarray:
        """Fold features of the learned frequency screen for frequencies
        n_start .. n_start + n_count - 1 of problem `prob`
        ([n_count][PS_SCREEN_FEATURES], FP32; ps_screen_features)."""
        nf = self._lib.ps_screen_n_features()
        out = np.zeros((max(n_count, 0), nf), np.float32)
        self._check(self._lib.ps_screen_features(self._h, prob, n_start, n_count, _abi.fptr(out)))
        return out

    def initial_params(self, prob, n_list, m_list) -> np.ndarray:
        """The reference's initial parameter vectors [items][ma]."""
        p_arr = np.ascontiguousarray(prob, np.int32)
        n_arr = np.ascontiguousarray(n_list, np.int32)
        m_arr = np.ascontiguousarray(m_list, np.int32)
        out = np.zeros((len(p_arr), self.info["ma"]))
        self._check(self._lib.ps_initial_params(self._h, len(p_arr), _abi.iptr(p_arr), _abi.iptr(n_arr),
                                                _abi.iptr(m_arr), _abi.dptr(out)))
        return out

    def free_params(self) -> np.ndarray:
        ia = np.zeros(self.info["ma"], np.int32)
        self._check(self._lib.ps_free_params(self._h, _abi.iptr(ia)))
        return ia

    def fit_items_from(self, prob, n_list, m_list, conw_r, conw_r_dev, a0=None, want_final: bool = False,
                       n_iter: int = 0):
        """ps_fit_items_from: fits starting from a0 [items][ma] (None: the
        reference's initialisation), n_iter > 0 = exactly that many LM
        iterations; returns (results, final parameters or None)."""
        p_arr = np.ascontiguousarray(prob, np.int32)
        n_arr = np.ascontiguousarray(n_list, np.int32)
        m_arr = np.ascontiguousarray(m_list, np.int32)
        k = len(p_arr)
        cw = np.ascontiguousarray(np.broadcast_to(conw_r, (k,)), np.float64)
        cd = np.ascontiguousarray(np.broadcast_to(conw_r_dev, (k,)), np.float64)
        a0p = None if a0 is None else np.ascontiguousarray(a0, np.float64)
        af = np.zeros((k, self.info["ma"])) if want_final else None
        out = (_abi.PsPairResult * max(k, 1))()
        self._check(self._lib.ps_fit_items_from(self._h, k, _abi.iptr(p_arr), _abi.iptr(n_arr), _abi.iptr(m_arr),
                                                _abi.dptr(cw), _abi.dptr(cd),
                                                _abi.dptr(a0p) if a0p is not None else None, n_iter, out,
                                                _abi.dptr(af) if want_final else None))
        return _results(out, k), af

    # -- launch configuration ------------------------------------------------
    def set_launch(self, waves_per_pair: int = 4, wg_per_cu: int = 0):
        self._check(self._lib.ps_set_launch(self._h, waves_per_pair, wg_per_cu))

    def get_launch(self) -> dict:
        w = C.c_int(0)
        g = C.c_int(0)
        t = C.c_int(0)
        self._check(self._lib.ps_get_launch(self._h, C.byref(w), C.byref(g)))
        self._check(self._lib.ps_get_tile_points(self._h, C.byref(t)))
        return dict(waves_per_pair=w.value, wg_per_cu=g.value, tile_points=t.value)

    def set_tile_points(self, points: int = 16):
        """Points per facet-pass claim: 16, 64, 128 or 256."""
        self._check(self._lib.ps_set_tile_points(self._h, points))

    def launch_options(self) -> dict:
        """{waves per pair: maximum resident workgroups per CU} of the
        prepared problem."""
        w = np.zeros(8, np.int32)
        g = np.zeros(8, np.int32)
        n = self._lib.ps_launch_options(self._h, _abi.iptr(w), _abi.iptr(g), 8)
        if n < 0:
            self._check(n)
        return {int(w[i]): int(g[i]) for i in range(min(n, 8))}

    def set_graphs(self, on: bool = True):
        """Capture each fit launch into a HIP graph and replay it while the
        launch arguments stay the same (results are bit-identical)."""
        self._check(self._lib.ps_set_graphs(self._h, 1 if on else 0))

    def graph_stats(self) -> dict:
        c = C.c_int(0)
        r = C.c_int(0)
        self._check(self._lib.ps_graph_stats(self._h, C.byref(c), C.byref(r)))
        return dict(captures=c.value, replays=r.value)

    def autotune(self, prob, n_list, m_list, conw_r) -> dict:
        """Times every launch configuration on these items and keeps the
        fastest (results do not depend on the configuration)."""
        p_arr = np.ascontiguousarray(prob, np.int32)
        n_arr = np.ascontiguousarray(n_list, np.int32)
        m_arr = np.ascontiguousarray(m_list, np.int32)
        cw = np.ascontiguousarray(np.broadcast_to(conw_r, p_arr.shape), np.float64)
        w = C.c_int(0)
        g = C.c_int(0)
        ms = C.c_double(0.0)
        self._check(self._lib.ps_autotune(self._h, len(p_arr), _abi.iptr(p_arr), _abi.iptr(n_arr), _abi.iptr(m_arr),
                                          _abi.dptr(cw), C.byref(w), C.byref(g), C.byref(ms)))
        t = C.c_int(0)
        self._check(self._lib.ps_get_tile_points(self._h, C.byref(t)))
        return dict(waves_per_pair=w.value, wg_per_cu=g.value, tile_points=t.value, kernel_ms=ms.value)

    def last_stats(self) -> dict:
        """Device time and work counters of the last fit call."""
        ms = C.c_double(0.0)
        launches = C.c_int(0)
        v = [C.c_longlong(0) for _ in range(4)]
        self._check(self._lib.ps_last_stats(self._h, C.byref(ms), C.byref(launches), *[C.byref(x) for x in v]))
        return dict(kernel_ms=ms.value, launches=launches.value, pairs=v[0].value, vis_pairs=v[1].value,
                    iterations=v[2].value, evaluations=v[3].value)

    def unprepare(self):
        self._check(self._lib.ps_unprepare(self._h))

    # -- kernel-level entry points (parity tests) ---------------------------
    def eval(self, a: np.ndarray, conw_r2_rows: float, jrows: bool = False):
        info = self.info
        mf, F = info["mfit"], info["nfac"]
        a = np.ascontiguousarray(a, np.float64)
        alpha = np.zeros((mf, mf))
        beta = np.zeros(mf)
        chisq = np.zeros(1)
        area = np.zeros(F)
        jr = np.zeros((info["ndata"] + 3, info["ma"] + 1)) if jrows else None
        self._check(self._lib.ps_eval(self._h, _abi.dptr(a), conw_r2_rows, _abi.dptr(alpha), _abi.dptr(beta),
                                      _abi.dptr(chisq), _abi.dptr(area),
                                      _abi.dptr(jr) if jrows else None))
        return alpha, beta, float(chisq[0]), area, jr

    def gaussj(self, alpha: np.ndarray, beta: np.ndarray, lam: float):
        mf = len(beta)
        alpha = np.ascontiguousarray(alpha, np.float64)
        beta = np.ascontiguousarray(beta, np.float64)
        x = np.zeros(mf)
        err = np.zeros(1, np.int32)
        self._check(self._lib.ps_gaussj(self._h, mf, _abi.dptr(alpha), _abi.dptr(beta), lam, _abi.dptr(x),
                                        _abi.iptr(err)))
        return x, int(err[0])

    def math_probe(self, x: np.ndarray):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(3 * len(x))
        out2 = np.zeros(len(x))
        self._check(self._lib.ps_math_probe(self._h, len(x), _abi.dptr(x), _abi.dptr(out), _abi.dptr(out2)))
        o = out.reshape(-1, 3)
        return o[:, 0], o[:, 1], o[:, 2], out2


def run_file(path_in: str, path_out: str, device: int = 0) -> float:
    """The whole application on one input file; returns the final conwR."""
    with Scan(device) as s:
        s.prepare_file(path_in)
        cw = s.conw_search()
        write_output(path_out, s.start(cw))
        return cw
