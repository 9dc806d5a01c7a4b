"""heom_solve_dev_noalloc (the entry point that allocates nothing) against
heom_solve_dev through the C ABI.

The no-allocation entry launches every stage kernel directly, so its results
must equal heom_solve_dev with graphs and resident integrators off bit for bit,
and the default heom_solve_dev (graph replay; resident integrators for small
problems) to the FP64 bar of 1e-10 with the same step counts.  Two calls in a
row share the device scratch; the second is smaller than the first.
"""
import ctypes

import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import _capi, configs

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _run(entry, s, rho0, t, kw):
    L = _capi.load()
    rho = np.ascontiguousarray(rho0, np.complex128).copy()
    rs = np.zeros((t.size, s.qme.n_level, s.qme.n_level), np.complex128)
    d = ctypes.c_void_p()
    _capi.check(L.heom_dev_alloc(ctypes.byref(d), rho.nbytes))
    try:
        _capi.check(L.heom_memcpy_h2d(d, _p(rho), rho.nbytes))
        rc = getattr(L, entry)(s._s, d, _p(t), t.size, kw["dt"], kw.get("atol", 0.0),
                               kw.get("rtol", 0.0), _p(rs))
        _capi.check(L.heom_memcpy_d2h(_p(rho), d, rho.nbytes))
        _capi.check(rc)
    finally:
        L.heom_dev_free(d)
    st = s.stats()
    return rho, rs, (st["n_accepted"], st["n_rejected"], st["n_rhs"])


def _case(m, method, kw):
    t = np.ascontiguousarray(m["t_list"], np.float64)
    out = {}
    for name, entry, graphs in (("noalloc", "heom_solve_dev_noalloc", True),
                                ("direct", "heom_solve_dev", False),
                                ("default", "heom_solve_dev", True)):
        q = H.heom_hilb(m["max_depth"]).load(m, H.hip(0))
        s = H.qme_solver(H.hip(0), q, method())
        s.set_graphs(graphs)
        print("case", m["name"], name, flush=True)
        out[name] = _run(entry, s, m["rho0"], t, kw)
    return out


def test_noalloc_two_calls_in_a_row_second_smaller():
    big = configs.exciton_chain7(max_depth=2, t_final=100.0)
    small = configs.spin_boson(max_depth=4)
    small["t_list"] = small["t_list"][:21]
    for m, method, kw in ((big, H.rkdp, {"dt": big["dt"], "atol": big["atol"],
                                         "rtol": big["rtol"]}),
                          (small, H.rk4, {"dt": small["dt"]})):
        r = _case(m, method, kw)
        rho, rs, st = r["noalloc"]
        rho_d, rs_d, st_d = r["direct"]
        assert np.array_equal(rs, rs_d) and np.array_equal(rho, rho_d), m["name"]
        assert st == st_d
        rho_g, rs_g, st_g = r["default"]
        assert np.abs(rs - rs_g).max() <= TOL
        assert np.abs(rho - rho_g).max() <= TOL * max(1.0, np.abs(rho_g).max())
        assert st[2] == st_g[2]


def test_noalloc_refuses_rho_s_beyond_the_scratch():
    m = configs.dense64(max_depth=0, t_final=0.01)
    q = H.heom_hilb(0).load(m, H.hip(0))
    s = H.qme_solver(H.hip(0), q, H.rk4())
    L = _capi.load()
    n_t = 400  # 400 x 64^2 x 16 bytes = 26 MB of rho_S > 6 MiB
    t = np.linspace(0.0, 0.01, n_t)
    rho = np.zeros(64 * 64, np.complex128)
    d = ctypes.c_void_p()
    _capi.check(L.heom_dev_alloc(ctypes.byref(d), rho.nbytes))
    try:
        rc = L.heom_solve_dev_noalloc(s._s, d, _p(t), n_t, 1e-3, 0.0, 0.0, None)
    finally:
        L.heom_dev_free(d)
    assert rc == _capi.HEOM_ERR_ARG
    assert b"scratch" in L.heom_last_error()
