"""Regression tests of fixed defects, each run on the fixed code path.

- Zero-length output intervals under Dormand-Prince (t_list with repeated
  times): no attempt is taken and the step size is kept, on the resident
  integrator, the stage kernels with graph replay and with direct launches.
  The reference (adaptive_step_size_solver.h:32-56 with rkdp.h:104-106) would
  clamp the step to 0 there and never finish the next interval; the oracle
  skips such intervals the same way, and the GPU must match it to 1e-10 with
  its accept / reject counts.
- A qme whose set_param is called again while a solver holds it: the solver
  follows the qme (qme_solver.h:18-53 keeps a pointer to it), refuses a qme
  whose size changed (HEOM_ERR_STATE), and keeps a destroyed handle's QME alive.
- qme_solver.calc_diff: the reference forwards to an empty stub
  (qme_solver.h:37-42 -> solver_base.h:55-62); here it computes drho = -D(rho).
"""
import ctypes

import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import _capi, configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _solve(m, t_list, kw, graphs):
    q = H.heom_hilb(m["max_depth"]).load(m, H.hip(0))
    s = H.qme_solver(H.hip(0), q, H.rkdp())
    s.set_graphs(graphs)
    rho, rs = s.solve_rho_s(m["rho0"], t_list, kw)
    return rho, rs, s.stats()


@pytest.mark.parametrize("case", ["resident", "graphs", "direct", "chain7_depth2"])
@pytest.mark.parametrize("t_list", [[0.0, 0.0, 1.0], [0.0, 0.5, 0.5, 1.0, 1.0],
                                    [0.0, 0.3, 0.3, 0.3, 0.8]])
def test_rkdp_zero_length_intervals(case, t_list):
    if case == "chain7_depth2":
        m = configs.exciton_chain7(max_depth=2)
        t_list = [10.0 * t for t in t_list]  # fs
    else:
        m = configs.spin_boson(max_depth=6)
    kw = {"dt": 0.05, "atol": 1e-10, "rtol": 1e-8}
    t = np.array(t_list)
    print("zero-length intervals", case, t_list, flush=True)
    rho, rs, st = _solve(m, t, kw, graphs=case != "direct")
    o = O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=8).load(m)
    rho_o, rs_o, st_o, _, rc = o.solve(m["rho0"], t, O.RKDP, kw["dt"], kw["atol"], kw["rtol"])
    assert rc == 0
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    assert np.abs(rs - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL * max(1.0, np.abs(rho_o).max())
    # repeated output times see the same state
    for i in range(len(t) - 1):
        if t[i + 1] == t[i]:
            assert np.array_equal(rs[i], rs[i + 1])


def test_rkdp_only_zero_length_intervals_keep_dt_save():
    m = configs.spin_boson(max_depth=4)
    q = H.heom_hilb(m["max_depth"]).load(m, H.hip(0))
    s = H.qme_solver(H.hip(0), q, H.rkdp())
    kw = {"dt": 0.05, "atol": 1e-10, "rtol": 1e-8}
    s.solve_rho_s(m["rho0"], [0.0, 0.4], kw)
    dt_save = s.stats()["last_dt"]
    rho, rs = s.solve_rho_s(m["rho0"], [0.4, 0.4, 0.4], kw)
    assert np.array_equal(rho, m["rho0"])
    assert s.stats()["last_dt"] == dt_save


def test_solver_follows_set_param_in_place():
    """Same sizes, new H: the existing solver propagates the new model."""
    m = configs.spin_boson(max_depth=5)
    q = H.heom_hilb(m["max_depth"]).load(m, H.hip(0))
    h0 = q._h.value
    s = H.qme_solver(H.hip(0), q, H.rk4())
    t = m["t_list"][:11]
    s.solve_rho_s(m["rho0"], t, {"dt": m["dt"]})
    q.H = m["H"] * 0.5
    q.set_param(H.hip(0))
    assert q._h.value == h0  # set up again in place
    _, rs = s.solve_rho_s(m["rho0"], t, {"dt": m["dt"]})
    o = O.Oracle(O.HILB, 2, m["max_depth"]).load(dict(m, H=m["H"] * 0.5))
    _, rs_o, _, _, _ = o.solve(m["rho0"], t, O.RK4, m["dt"])
    assert np.abs(rs - rs_o).max() <= TOL


def test_solver_rebinds_when_the_handle_is_replaced():
    """A create-time option changed (dtype): set_param makes a new handle and
    the solver rebinds to it on its next solve."""
    m = configs.spin_boson(max_depth=5)
    q = H.heom_hilb(m["max_depth"]).load(m, H.hip(0))
    s = H.qme_solver(H.hip(0), q, H.rk4())
    t = m["t_list"][:11]
    _, rs64 = s.solve_rho_s(m["rho0"], t, {"dt": m["dt"]})
    q.dtype = np.dtype(np.complex64)
    q.set_param(H.hip(0))
    _, rs32 = s.solve_rho_s(m["rho0"], t, {"dt": m["dt"]})
    err = np.abs(rs32 - rs64).max()
    assert 0.0 < err <= 1e-5  # now the complex64 operator


def test_c_abi_solver_refuses_a_resized_qme_and_outlives_heom_destroy():
    L = _capi.load()
    m = configs.spin_boson(max_depth=4)
    q = H.heom_hilb(m["max_depth"]).load(m, H.hip(0))
    h = q._h
    s = ctypes.c_void_p()
    _capi.check(L.heom_solver_create(h, _capi.HEOM_RK4, ctypes.byref(s)))
    t = np.ascontiguousarray(m["t_list"][:3])
    rs = np.zeros((t.size, 2, 2), np.complex128)
    rho = m["rho0"].copy()
    _capi.check(L.heom_solve(s, _p(rho), _p(t), t.size, _capi.CALLBACK(), None, 0.01, 0.0,
                             0.0, 0))
    # a filter that shrinks the hierarchy, set up in place on the same handle
    keep = _capi.FILTER(lambda idx, n, depth, user: 1 if idx[0] <= 1 else 0)
    _capi.check(L.heom_set_hrchy_filter(h, keep, None))
    _capi.check(L.heom_set_param(h))
    rho = m["rho0"].copy()
    rc = L.heom_solve(s, _p(rho), _p(t), t.size, _capi.CALLBACK(), None, 0.01, 0.0, 0.0, 0)
    assert rc == _capi.HEOM_ERR_STATE
    # back to the full hierarchy: the solver works again; then destroy the handle
    _capi.check(L.heom_set_hrchy_filter(h, _capi.FILTER(), None))
    _capi.check(L.heom_set_param(h))
    q._h = None  # the qme object no longer owns it
    _capi.check(L.heom_destroy(h))
    rho = m["rho0"].copy()
    _capi.check(L.heom_solve(s, _p(rho), _p(t), t.size, _capi.CALLBACK(), None, 0.01, 0.0,
                             0.0, 0))
    o = O.Oracle(O.HILB, 2, m["max_depth"]).load(m)
    rho_o, _, _, _, _ = o.solve(m["rho0"], t, O.RK4, 0.01)
    assert np.abs(rho - rho_o).max() <= TOL
    _capi.check(L.heom_solver_destroy(s))


@pytest.mark.parametrize("form", ["heom_hilb", "heom_liou", "heom_ado"])
def test_qme_solver_calc_diff(form):
    m = configs.random_model(4, 2, 2, 3, seed=12, dense_v=True)
    q = getattr(H, form)(m["max_depth"]).load(m, H.hip(0))
    s = H.qme_solver(H.hip(0), q, H.rkdp())
    rng = np.random.default_rng(1)
    rho = (rng.normal(size=q.main_size()) + 1j * rng.normal(size=q.main_size())) * 0.1
    drho = np.zeros_like(rho)
    got = s.calc_diff(drho, rho)
    kind = {"heom_hilb": O.HILB, "heom_liou": O.LIOU, "heom_ado": O.ADO}[form]
    ref = O.Oracle(kind, 4, m["max_depth"]).load(m).calc_diff(rho, 1.0, 0.0)
    assert np.abs(got - ref).max() <= TOL * max(1.0, np.abs(ref).max())
    # drho_dt is filled in place as well (the reference's signature writes it)
    assert np.array_equal(drho, got)
