"""ps_fit_items_static, the fit entry point that allocates nothing, against
ps_fit_items: caller-owned device buffers (from the HIP runtime), two calls
in a row on the same prepared problem, the second with fewer items, each
bit-identical to the allocating entry point.  The problem (2,700 points,
2 MB scratch slabs) is large enough that the 256 MiB static scratch, not
the GPU, caps the number of workgroups of the first call."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "niter", "n_vis")


class _Hip:
    """Device buffers from the HIP runtime the library itself links (a
    second runtime in the process, such as torch's bundled one, cannot see
    the GPU once this one is initialised)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _static_fit(scan, n, m, cw):
    from ps_claude_opt_amd import _abi

    lib = _abi.load()
    k = len(n)
    hip = _Hip()
    try:
        d_n = hip.upload(np.asarray(n, np.int32))
        d_m = hip.upload(np.asarray(m, np.int32))
        cw = np.asarray(cw, np.float64)
        d_cw = hip.upload(np.stack([cw * cw, cw], axis=1))  # (conw_r^2 of the rows, conw_r of dev)
        d_out = hip.upload(np.full((k, 10), np.nan))
        rc = lib.ps_fit_items_static(scan._h, k, None, d_n, d_m, d_cw, d_out)
        assert rc == 0, lib.ps_last_error(scan._h)
        o = hip.download(d_out, (k, 10), np.float64)
    finally:
        hip.free()
    return {"dev": o[:, 0], "dark": o[:, 1], "period": o[:, 2], "lambda_deg": o[:, 3], "beta_deg": o[:, 4],
            "chisq": o[:, 5], "niter": o[:, 6], "reported": o[:, 7], "n_vis": o[:, 8]}


def test_static_entry_point_matches_allocating_one(inputs):
    import ps_claude_opt_amd as ps

    path, _ = inputs("c2")
    rng = np.random.default_rng(71)
    with ps.Scan(0) as s:
        s.prepare_file(path)
        for count in (300, 40):
            n = rng.integers(1, s.info["n_max"] + 1, count)
            m = rng.integers(1, 11, count)
            cw = np.full(count, 0.7)
            ref = s.fit_pairs(n, m, 0.7)
            got = _static_fit(s, n, m, cw)
            assert np.all(got["reported"] == 1.0)
            for f in FIELDS:
                a = np.asarray(ref[f], np.float64)
                b = got[f]
                assert np.array_equal(a, b, equal_nan=True), (count, f, np.nanmax(np.abs(a - b)))
