"""Resident fixed-step integrator: small heom_hilb problems (diagonal V, FP64,
at most 1024 elements, at most 8 links per ADO) run every RK4 / lsrk4 step of
an output interval in one launch, with the state and stage buffers in LDS.
It must meet the FP64 bar (1e-10) against the oracle with the oracle's step
count, shortened last steps included, and agree with the stage-kernel path
(direct launches, set_graphs(False)) to rounding."""
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
