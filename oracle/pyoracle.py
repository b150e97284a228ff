"""ctypes binding of the CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() (as the checker) and
bench.py's cpu_baseline leg.  The product (ps_claude_opt_amd) never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libps_oracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Order(C.Structure):
    _fields_ = [("n_waves", C.c_int), ("tile", C.c_int)]


class PairResult(C.Structure):
    _fields_ = [("reported", C.c_int), ("niter", C.c_int),
                ("dev", C.c_double), ("dark", C.c_double), ("period", C.c_double),
                ("lambda_deg", C.c_double), ("beta_deg", C.c_double), ("chisq", C.c_double),
                ("n_evals", C.c_longlong), ("n_vis", C.c_longlong)]


DEFAULT_ORDER = Order(4, 16)

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} missing: run make -C oracle")
    L = C.CDLL(LIB_PATH)
    L.pso_load.restype = C.c_void_p
    L.pso_load.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.pso_free.argtypes = [C.c_void_p]
    L.pso_info.argtypes = [C.c_void_p, _ip, _dp]
    L.pso_table.restype = C.c_int
    L.pso_table.argtypes = [C.c_void_p, C.c_char_p, _dp, C.c_int]
    L.pso_mrqcof.argtypes = [C.c_void_p, _dp, C.c_double, Order, _dp, _dp, _dp, _dp, _dp]
    L.pso_gaussj.restype = C.c_int
    L.pso_gaussj.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
    L.pso_exp.restype = C.c_double
    L.pso_exp.argtypes = [C.c_double]
    L.pso_exp2.restype = C.c_double
    L.pso_exp2.argtypes = [C.c_double]
    L.pso_sincos.argtypes = [C.c_double, _dp, _dp]
    L.pso_fit_pair.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, Order,
                               C.POINTER(PairResult), _dp, C.c_int]
    L.pso_run.argtypes = [C.c_void_p, C.c_char_p, C.c_int, Order, C.c_int, _dp]
    L.pso_conw_search.argtypes = [C.c_void_p, C.c_int, Order, _dp]
    L.pso_fit_pairs.restype = C.c_double
    L.pso_fit_pairs.argtypes = [C.c_void_p, C.c_int, _ip, _ip, C.c_double, C.c_double, Order, C.c_int,
                                C.POINTER(PairResult)]
    L.pso_set_stop.argtypes = [C.c_void_p, C.c_double, C.c_int]
    L.pso_fit_pairs_from.restype = C.c_double
    L.pso_fit_pairs_from.argtypes = [C.c_void_p, C.c_int, _ip, _ip, C.c_double, C.c_double, C.c_int, _dp,
                                     C.c_int, C.POINTER(PairResult), _dp]
    L.pso_fit_params.restype = C.c_double
    L.pso_fit_params.argtypes = [C.c_void_p, C.c_int, _dp, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.POINTER(PairResult), _dp]
    L.pso_finish.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, _dp, _dp, _dp]
    L.pso_conw_decide.argtypes = [C.POINTER(PairResult), C.c_int, _dp, _ip]
    _lib = L
    return L


INFO_INTS = ("ndata", "l_curves", "nfac", "ncoef", "ma", "mfit", "lastma", "lastone", "n_max",
             "lmax", "mmax", "nrows", "n_iter_max", "n_iter_min", "onlyrel")
INFO_DBLS = ("f_start", "f_end", "f_step", "escl", "conw", "jd0", "phi0", "alamda_start",
             "alamda_incr", "alamda_incrr", "iter_diff_max", "cl", "jd_min", "jd_max")


class Problem:
    def __init__(self, path: str):
        self.L = load()
        err = C.create_string_buffer(256)
        h = self.L.pso_load(path.encode(), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self.h = C.c_void_p(h)
        ints = np.zeros(15, np.int32)
        dbls = np.zeros(14)
        self.L.pso_info(self.h, ints.ctypes.data_as(_ip), dbls.ctypes.data_as(_dp))
        self.info = dict(zip(INFO_INTS, map(int, ints)))
        self.info.update(zip(INFO_DBLS, map(float, dbls)))

    def __del__(self):
        try:
            self.L.pso_free(self.h)
        except Exception:
            pass

    def table(self, name: str) -> np.ndarray:
        n = self.L.pso_table(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1))
        self.L.pso_table(self.h, name.encode(), out.ctypes.data_as(_dp), n)
        return out[:n]

    def mrqcof(self, a, conw_r2_rows, order=DEFAULT_ORDER, jrows=False):
        mf, ma, F, N = self.info["mfit"], self.info["ma"], self.info["nfac"], self.info["ndata"]
        a = np.ascontiguousarray(a, np.float64)
        alpha = np.zeros((mf, mf))
        beta = np.zeros(mf)
        chisq = np.zeros(1)
        area = np.zeros(F)
        jr = np.zeros((N + 3, ma + 1)) if jrows else None
        self.L.pso_mrqcof(self.h, a.ctypes.data_as(_dp), conw_r2_rows, order, alpha.ctypes.data_as(_dp),
                          beta.ctypes.data_as(_dp), chisq.ctypes.data_as(_dp), area.ctypes.data_as(_dp),
                          jr.ctypes.data_as(_dp) if jrows else None)
        return alpha, beta, float(chisq[0]), area, jr

    def initial_params(self, n: int, m: int) -> np.ndarray:
        """Parameter vector of pair (n, m) as the reference initialises it
        (Start.cu:4346-4418)."""
        C_ = self.info["ncoef"]
        a = np.zeros(self.info["ma"])
        a[:C_] = self.table("cg_first")
        lam = [0, 90, 180, 270, 60, 180, 300, 60, 180, 300][m - 1]
        bet = [0, 0, 0, 0, 60, 60, 60, -60, -60, -60][m - 1]
        d2r = np.pi / 180.0
        period = 1.0 / (self.info["f_start"] - (n - 1) * self.info["f_step"])
        a[C_] = d2r * 90 - d2r * bet
        a[C_ + 1] = d2r * lam
        a[C_ + 2] = (24.0 * 2.0 * np.pi) / period
        a[C_ + 3:C_ + 8] = self.table("par")  # A, D, k, log(cl), Lommel-Seeliger 1
        return a

    def fit_pair(self, n, m, conw_r, conw_r_dev, order=DEFAULT_ORDER, traj_cap=0):
        r = PairResult()
        traj = np.zeros(max(traj_cap, 1) * 4)
        self.L.pso_fit_pair(self.h, n, m, conw_r * conw_r, conw_r_dev, order, C.byref(r),
                            traj.ctypes.data_as(_dp) if traj_cap else None, traj_cap)
        return r, traj.reshape(-1, 4)[:traj_cap]

    def fit_pairs(self, n_list, m_list, conw_r, conw_r_dev, threads=0, order=DEFAULT_ORDER):
        n_arr = np.ascontiguousarray(n_list, np.int32)
        m_arr = np.ascontiguousarray(m_list, np.int32)
        out = (PairResult * max(len(n_arr), 1))()
        wall = self.L.pso_fit_pairs(self.h, len(n_arr), n_arr.ctypes.data_as(_ip), m_arr.ctypes.data_as(_ip),
                                    conw_r, conw_r_dev, order, threads, out)
        return list(out[:len(n_arr)]), wall

    def fit_pairs_from(self, n_list, m_list, conw_r, conw_r_dev, a0=None, threads=0, want_final=False, n_iter=0):
        """fit_pairs from given initial parameters a0 [pairs][ma] (None: the
        reference's initialisation); returns (results, final params or None)."""
        n_arr = np.ascontiguousarray(n_list, np.int32)
        m_arr = np.ascontiguousarray(m_list, np.int32)
        k = len(n_arr)
        out = (PairResult * max(k, 1))()
        a0p = None if a0 is None else np.ascontiguousarray(a0, np.float64)
        af = np.zeros((k, self.info["ma"])) if want_final else None
        self.L.pso_fit_pairs_from(self.h, k, n_arr.ctypes.data_as(_ip), m_arr.ctypes.data_as(_ip), conw_r,
                                  conw_r_dev, threads, a0p.ctypes.data_as(_dp) if a0p is not None else None, n_iter, out,
                                  af.ctypes.data_as(_dp) if want_final else None)
        return list(out[:k]), af

    def fit_params(self, a0, conw_r, conw_r_dev, threads=0, n_iter=0):
        """Fits starting from full parameter vectors a0 [pairs][ma] (pole
        and rotation rate included); returns (results, final params)."""
        a0 = np.ascontiguousarray(np.atleast_2d(a0), np.float64)
        k = a0.shape[0]
        out = (PairResult * k)()
        af = np.zeros((k, self.info["ma"]))
        self.L.pso_fit_params(self.h, k, a0.ctypes.data_as(_dp), conw_r, conw_r_dev, threads, n_iter, out,
                              af.ctypes.data_as(_dp))
        return list(out), af

    def finish(self, a, chisq, conw_r_dev):
        """(dev, chck, dark) of Iter2 + FinishPole at parameters a."""
        a = np.ascontiguousarray(a, np.float64)
        d, c, k = C.c_double(), C.c_double(), C.c_double()
        self.L.pso_finish(self.h, a.ctypes.data_as(_dp), chisq, conw_r_dev, C.byref(d), C.byref(c), C.byref(k))
        return d.value, c.value, k.value

    def conw_search(self, threads=0, order=DEFAULT_ORDER) -> float:
        cw = C.c_double(0)
        self.L.pso_conw_search(self.h, threads, order, C.byref(cw))
        return cw.value

    def run(self, out_path: str, threads=0, freq_limit=0, order=DEFAULT_ORDER) -> float:
        cw = C.c_double(0)
        self.L.pso_run(self.h, out_path.encode(), threads, order, freq_limit, C.byref(cw))
        return cw.value

    def set_stop(self, stop_condition: float, n_iter_min: int = 0):
        self.L.pso_set_stop(self.h, stop_condition, n_iter_min)
        self.info["n_iter_max"] = 60 if stop_condition < 1 else int(stop_condition)


def gaussj(alpha, beta, lam):
    L = load()
    mf = len(beta)
    alpha = np.ascontiguousarray(alpha, np.float64)
    beta = np.ascontiguousarray(beta, np.float64)
    x = np.zeros(mf)
    err = L.pso_gaussj(mf, alpha.ctypes.data_as(_dp), beta.ctypes.data_as(_dp), lam, x.ctypes.data_as(_dp))
    return x, err


def conw_decide(results, conw_r):
    """One conw round's decision on len(results) // 10 periods x 10 poles of
    PairResult: returns (new conw_r, new_conw)."""
    n = len(results) // 10
    arr = (PairResult * (10 * n))(*results[:10 * n])
    cw = C.c_double(conw_r)
    nc = C.c_int(0)
    load().pso_conw_decide(arr, n, C.byref(cw), C.byref(nc))
    return cw.value, nc.value


def exp(x):
    return load().pso_exp(x)


def exp2(x):
    return load().pso_exp2(x)


def sincos(x):
    s, c = C.c_double(), C.c_double()
    load().pso_sincos(x, C.byref(s), C.byref(c))
    return s.value, c.value
