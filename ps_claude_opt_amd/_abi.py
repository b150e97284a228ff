"""ctypes bindings of the C ABI in include/ps_mi355x.h.

The shared library is built in-tree (ps_claude_opt_amd/lib/libps_mi355x.so)
by ``__graft_entry__.build()`` or ``make -C ps_claude_opt_amd/csrc``.  There is
no fallback: if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libps_mi355x.so")

PS_OK = 0
PS_E_ARG = -1
PS_E_INPUT = -2
PS_E_HIP = -3
PS_E_UNSUPPORTED = -4
PS_E_STATE = -5
PS_N_POLES = 10

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class PsProblem(C.Structure):
    _fields_ = [
        ("ndata", C.c_int), ("n_curves", C.c_int),
        ("l_points", _ip), ("in_rel", _ip),
        ("tim", _dp), ("ee", _dp), ("ee0", _dp), ("brightness", _dp),
        ("sig", _dp), ("sigr2", _dp), ("weight", _dp),
        ("nfac", C.c_int), ("lmax", C.c_int), ("mmax", C.c_int), ("ncoef", C.c_int),
        ("normal", _dp), ("d_area", _dp), ("pleg", _dp), ("fc", _dp), ("fs", _dp),
        ("d_sphere", _dp), ("cg_first", _dp), ("ia", _ip),
        ("par", C.c_double * 3), ("cl", C.c_double), ("phi0", C.c_double),
        ("alamda_start", C.c_double), ("alamda_incr", C.c_double), ("alamda_incrr", C.c_double),
        ("stop_condition", C.c_double), ("n_iter_min", C.c_int),
        ("f_start", C.c_double), ("f_end", C.c_double), ("f_step", C.c_double),
        ("escl", C.c_double),
        ("beta_pole", C.c_double * PS_N_POLES), ("lambda_pole", C.c_double * PS_N_POLES),
    ]


class PsLine(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("period_h", "rms", "chi2", "dark_or_conw", "lambda_deg", "beta_deg")]


class PsPairResult(C.Structure):
    _fields_ = [("reported", C.c_int), ("niter", C.c_int)] + [
        (n, C.c_double) for n in ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq")] + [
        ("n_vis", C.c_longlong)]


# every entry point of include/ps_mi355x.h with its signature
SIGNATURES = {
    "ps_open": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "ps_last_error": (C.c_char_p, [C.c_void_p]),
    "ps_close": (C.c_int, [C.c_void_p]),
    "ps_prepare": (C.c_int, [C.c_void_p, C.POINTER(PsProblem)]),
    "ps_prepare_file": (C.c_int, [C.c_void_p, C.c_char_p]),
    "ps_info": (C.c_int, [C.c_void_p, _ip, _dp]),
    "ps_precalc": (C.c_int, [C.c_void_p, _dp, _ip]),
    "ps_conw_search": (C.c_int, [C.c_void_p, _dp]),
    "ps_start": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.POINTER(PsLine)]),
    "ps_fit_pairs": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, C.c_double, C.c_double,
                               C.POINTER(PsPairResult)]),
    "ps_last_stats": (C.c_int, [C.c_void_p, _dp, _ip] + [C.POINTER(C.c_longlong)] * 4),
    "ps_unprepare": (C.c_int, [C.c_void_p]),
    "ps_eval": (C.c_int, [C.c_void_p, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp]),
    "ps_gaussj": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, C.c_double, _dp, _ip]),
    "ps_math_probe": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "ps_input_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_char_p, C.c_int]),
    "ps_input_view": (C.c_int, [C.c_void_p, C.POINTER(PsProblem), _dp, _ip]),
    "ps_input_free": (None, [C.c_void_p]),
    "ps_select_best_pole": (C.c_int, [C.POINTER(PsPairResult)]),
    "ps_prepare_batch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(PsProblem)]),
    "ps_prepare_files": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_char_p)]),
    "ps_n_problems": (C.c_int, [C.c_void_p]),
    "ps_problem_info": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp]),
    "ps_fit_items": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp, C.POINTER(PsPairResult)]),
    "ps_conw_search_batch": (C.c_int, [C.c_void_p, _dp]),
    "ps_start_batch": (C.c_int, [C.c_void_p, _ip, _ip, _dp, C.POINTER(PsLine)]),
    "ps_initial_params": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp]),
    "ps_free_params": (C.c_int, [C.c_void_p, _ip]),
    "ps_warm_n_features": (C.c_int, []),
    "ps_warm_features": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip, C.POINTER(C.c_float)]),
    "ps_screen_n_features": (C.c_int, []),
    "ps_screen_features": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "ps_fit_items_from": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, C.c_int,
                                    C.POINTER(PsPairResult), _dp]),
    "ps_mlp_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong,
                                C.POINTER(C.c_void_p)]),
    "ps_mlp_n_params": (C.c_longlong, [C.c_void_p]),
    "ps_mlp_get_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "ps_mlp_set_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "ps_mlp_forward": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int]),
    "ps_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "ps_mlp_train": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.c_int,
                               C.c_float, C.c_ulonglong, C.POINTER(C.c_float)]),
    "ps_mlp_free": (C.c_int, [C.c_void_p]),
    "ps_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "ps_get_launch": (C.c_int, [C.c_void_p, _ip, _ip]),
    "ps_fit_items_static": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "ps_set_tile_points": (C.c_int, [C.c_void_p, C.c_int]),
    "ps_get_tile_points": (C.c_int, [C.c_void_p, _ip]),
    "ps_launch_options": (C.c_int, [C.c_void_p, _ip, _ip, C.c_int]),
    "ps_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "ps_graph_stats": (C.c_int, [C.c_void_p, _ip, _ip]),
    "ps_autotune": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _ip, _ip, _dp]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} not built: run __graft_entry__.build() or make -C ps_claude_opt_amd/csrc")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)
