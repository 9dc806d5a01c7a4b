"""Optional FP64 renormalisation of rho_0 (heom_options.renorm_rho0).

After every RK step / Dormand-Prince attempt rho_0 is replaced by
(rho_0 + rho_0^+)/2 scaled to its initial trace.  The exact HEOM dynamics keeps
rho_0 Hermitian with a constant trace, so on the FP64 form the projection only
removes rounding drift: the result must still meet the FP64 bar (1e-10) against
the oracle, which does not project.  The outputs are exactly Hermitian and keep
the initial trace to rounding; the complex64 variants state their error with
and without the projection.
"""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10
ENG = None


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def check_projected(rs):
    for r in rs:
        assert np.array_equal(r, r.conj().T)
        assert abs(np.trace(r) - 1.0) <= 1e-13


def test_rkdp_chain7_fp64_with_projection():
    m = configs.exciton_chain7(max_depth=2, t_final=200.0)
    kw = {"dt": m["dt"], "atol": m["atol"], "rtol": m["rtol"]}
    g = H.heom_hilb(2, renorm_rho0=True).load(m, eng())
    s = H.qme_solver(eng(), g, H.rkdp())
    _, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    o = O.Oracle(O.HILB, 7, 2, n_threads=8).load(m)
    _, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"], kw["atol"], kw["rtol"])
    st = s.stats()
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    assert np.abs(rs - rs_o).max() <= TOL
    check_projected(rs[1:])


@pytest.mark.parametrize("solver", ["rk4", "lsrk4"])
def test_fixed_step_spin_boson_with_projection(solver):
    m = configs.spin_boson(max_depth=8)
    g = H.heom_hilb(8, renorm_rho0=True).load(m, eng())
    s = H.qme_solver(eng(), g, getattr(H, solver)())
    _, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": m["dt"]})
    o = O.Oracle(O.HILB, 2, 8, n_threads=8).load(m)
    _, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], getattr(O, solver.upper()), m["dt"])
    assert np.abs(rs - rs_o).max() <= TOL
    check_projected(rs[1:])


@pytest.mark.parametrize("solver", ["rk4", "rkdp"])
def test_complex64_with_and_without_projection(solver):
    m = configs.exciton_chain7(max_depth=2, t_final=200.0)
    kw = {"dt": 0.1} if solver == "rk4" else {"dt": m["dt"], "atol": m["atol"], "rtol": m["rtol"]}
    o = O.Oracle(O.HILB, 7, 2, n_threads=8).load(m)
    _, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], getattr(O, solver.upper()), kw["dt"],
                               kw.get("atol", 0.0), kw.get("rtol", 0.0))
    err, tr = {}, {}
    for proj in (False, True):
        g = H.heom_hilb(2, dtype="complex64", renorm_rho0=proj).load(m, eng())
        s = H.qme_solver(eng(), g, getattr(H, solver)())
        _, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
        err[proj] = float(np.abs(rs - rs_o).max())
        tr[proj] = float(np.abs(np.trace(rs, axis1=1, axis2=2) - 1.0).max())
        if proj:
            check_projected(rs[1:])
    print("complex64 %s rho_S error vs FP64 oracle: %.3e plain, %.3e projected; trace error "
          "%.3e plain, %.3e projected" % (solver, err[False], err[True], tr[False], tr[True]))
    assert err[False] <= 1e-6 and err[True] <= 1e-6


def test_projection_graphs_bitwise():
    m = configs.spin_boson(max_depth=6)
    m["t_list"] = np.linspace(0.0, 3.0, 8)
    g = H.heom_hilb(6, dtype="complex64", renorm_rho0=True).load(m, eng())
    out = []
    for graphs in (False, True):
        s = H.qme_solver(eng(), g, H.rk4())
        s.set_graphs(graphs)
        out.append(s.solve_rho_s(m["rho0"], m["t_list"], {"dt": 0.03}))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_redfield_with_projection():
    m = configs.spin_boson(max_depth=0)
    m["rho0"] = m["rho0"][:4].copy()
    g = H.redfield_hilb(renorm_rho0=True).load(m, eng())
    s = H.qme_solver(eng(), g, H.rk4())
    _, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": m["dt"]})
    o = O.Oracle(O.REDFIELD_HILB, 2, 0).load(m)
    _, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], O.RK4, m["dt"])
    assert np.abs(rs - rs_o).max() <= TOL
    check_projected(rs[1:])


def test_projection_refused_for_batches():
    ms = configs.batched_sweep(n_inst=2, seed=3, n_steps=10, out_every=5)
    with pytest.raises(H.HeomError):
        b = H.heom_batch([H.heom_hilb(m["max_depth"], renorm_rho0=True).load(m, eng())
                          for m in ms])
        H.qme_solver(eng(), b, H.rk4())
