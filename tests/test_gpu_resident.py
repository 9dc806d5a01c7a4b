"""Resident integrators: small heom_hilb problems (diagonal V, FP64, at most
1024 elements -- 960 for Dormand-Prince -- and at most 8 links per ADO) and
small redfield_hilb problems (diagonal V) run a whole solve in one launch with
the state and stage buffers in LDS (RK4 / lsrk4 through heom_solve_dev and per
interval otherwise; Dormand-Prince through heom_solve_dev with the step control
on the device).  They must meet the FP64 bar (1e-10) against the oracle with the
oracle's step / accept / reject counts, and agree with the stage-kernel path
(set_graphs(False)) to rounding."""
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


def run(m, solver, kw, graphs=True):
    g = H.heom_hilb(m["max_depth"]).load(m, eng())
    s = H.qme_solver(eng(), g, getattr(H, solver)())
    s.set_graphs(graphs)
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    return rho, rs, s.stats()


def oracle(m, solver, dt):
    o = O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=8).load(m)
    return o.solve(m["rho0"], m["t_list"], getattr(O, solver.upper()), dt)


@pytest.mark.parametrize("solver", ["rk4", "lsrk4"])
def test_spin_boson_full_config(solver):
    """Spin-boson, 66 ADOs (264 elements), dt=0.01 x 2000, rho_S every 10 steps."""
    m = configs.spin_boson()
    rho, rs, st = run(m, solver, {"dt": m["dt"]})
    rho_o, rs_o, st_o, _, _ = oracle(m, solver, m["dt"])
    assert st["n_rhs"] == st_o.n_rhs
    assert np.abs(rs - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL
    rho_d, rs_d, _ = run(m, solver, {"dt": m["dt"]}, graphs=False)
    assert np.abs(rho - rho_d).max() <= 1e-12


def test_shortened_last_steps():
    m = configs.spin_boson(max_depth=6)
    m["t_list"] = np.linspace(0.0, 3.0, 8)
    rho, rs, st = run(m, "rk4", {"dt": 0.03})
    rho_o, rs_o, st_o, _, _ = oracle(m, "rk4", 0.03)
    assert st["n_rhs"] == st_o.n_rhs
    assert np.abs(rs - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL


@pytest.mark.parametrize("complex_diag", [False, True])
def test_eight_links_per_ado(complex_diag):
    """N=3, 2 noises x 2 modes (up to 8 links per ADO), depth 3: 35 ADOs, 315 elements."""
    m = configs.random_model(3, 2, 2, 3, seed=9, dense_v=False)
    rng = np.random.default_rng(2)
    if complex_diag:
        for z in m["noises"]:
            z["V"] = np.diag(rng.normal(size=3) + 1j * rng.normal(size=3))
    m["t_list"] = np.linspace(0.0, 0.5, 6)
    rho, rs, st = run(m, "rk4", {"dt": 0.01})
    rho_o, rs_o, st_o, _, _ = oracle(m, "rk4", 0.01)
    assert st["n_rhs"] == st_o.n_rhs
    scale = max(1.0, np.abs(rho_o).max())
    assert np.abs(rho - rho_o).max() <= TOL * scale
    assert np.abs(rs - rs_o.reshape(rs.shape)).max() <= TOL * scale


def test_largest_resident_problem():
    """N=4, 1 noise x 2 modes, depth 9: 55 ADOs of 16 elements (880 of at most 1024)."""
    m = configs.random_model(4, 1, 2, 9, seed=4, dense_v=False)
    m["t_list"] = np.linspace(0.0, 0.2, 3)
    assert m["rho0"].size <= 1024
    rho, rs, st = run(m, "lsrk4", {"dt": 0.005})
    rho_o, rs_o, st_o, _, _ = oracle(m, "lsrk4", 0.005)
    scale = max(1.0, np.abs(rho_o).max())
    assert np.abs(rho - rho_o).max() <= TOL * scale


@pytest.mark.parametrize("case", ["spin_boson_rejections", "eight_links"])
def test_dopri_resident_matches_oracle(case):
    if case == "spin_boson_rejections":
        m = configs.spin_boson(max_depth=6, T=0.5)
        m["t_list"] = np.linspace(0.0, 5.0, 11)
        kw = {"dt": 0.8, "atol": 1e-10, "rtol": 1e-8}
    else:
        m = configs.random_model(3, 2, 2, 3, seed=9, dense_v=False)
        m["t_list"] = np.linspace(0.0, 0.5, 6)
        kw = {"dt": 0.05, "atol": 1e-10, "rtol": 1e-8}
    rho, rs, st = run(m, "rkdp", kw)
    o = O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=8).load(m)
    rho_o, rs_o, st_o, dts, _ = o.solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"], kw["atol"],
                                        kw["rtol"])
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    assert st["n_rhs"] == 7 * (st_o.n_accepted + st_o.n_rejected)
    scale = max(1.0, np.abs(rho_o).max())
    assert np.abs(rho - rho_o).max() <= TOL * scale
    assert np.abs(rs - rs_o.reshape(rs.shape)).max() <= TOL
    rho_d, _, st_d = run(m, "rkdp", kw, graphs=False)
    assert (st_d["n_accepted"], st_d["n_rejected"]) == (st["n_accepted"], st["n_rejected"])
    assert np.abs(rho - rho_d).max() <= 1e-12 * scale


def test_dopri_resident_dt_save_carries_over():
    """dt_save of one solve seeds the next (adaptive_step_size_solver.h:39-41, 55)."""
    m = configs.spin_boson(max_depth=4)
    kw = {"dt": 0.05, "atol": 1e-10, "rtol": 1e-8}
    g = H.heom_hilb(4).load(m, eng())
    s = H.qme_solver(eng(), g, H.rkdp())
    t1, t2 = np.linspace(0.0, 1.0, 3), np.linspace(1.0, 2.0, 3)
    rho1, _ = s.solve_rho_s(m["rho0"], t1, kw)
    rho2, _ = s.solve_rho_s(rho1, t2, kw)
    o = O.Oracle(O.HILB, 2, 4).load(m)
    r, _, _, dts, _ = o.solve(m["rho0"], t1, O.RKDP, kw["dt"], kw["atol"], kw["rtol"])
    r, _, _, _, _ = o.solve(r, t2, O.RKDP, kw["dt"], kw["atol"], kw["rtol"], dt_save=dts)
    assert np.abs(rho2 - r).max() <= TOL


@pytest.mark.parametrize("solver", ["rk4", "lsrk4", "rkdp"])
def test_redfield_hilb_resident_against_stage_kernels_and_oracle(solver):
    """Redfield 7-site chain (7 site baths, diagonal V) to 100 fs, rho_S every 10 fs:
    the resident integrator (default) and the stage kernels (set_graphs(False))
    both against the oracle, with its step / accept / reject counts."""
    m = configs.exciton_chain7(max_depth=0, t_final=100.0)
    m["rho0"] = m["rho0"][:49].copy()
    kw = ({"dt": m["dt"], "atol": m["atol"], "rtol": m["rtol"]} if solver == "rkdp"
          else {"dt": 0.1})
    out = {}
    for graphs in (True, False):
        g = H.redfield_hilb().load(m, eng())
        s = H.qme_solver(eng(), g, getattr(H, solver)())
        s.set_graphs(graphs)
        rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
        out[graphs] = (rho, rs, s.stats())
    o = O.Oracle(O.REDFIELD_HILB, 7, 0, n_threads=8).load(m)
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], getattr(O, solver.upper()),
                                      kw["dt"], kw.get("atol", 0.0), kw.get("rtol", 0.0))
    for rho, rs, st in out.values():
        assert np.abs(rs - rs_o).max() <= TOL
        assert np.abs(rho - rho_o).max() <= TOL
        if solver == "rkdp":
            assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
        else:
            assert st["n_rhs"] == st_o.n_rhs
    assert np.abs(out[True][0] - out[False][0]).max() <= 1e-12
