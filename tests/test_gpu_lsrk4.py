"""lsrk4 (solver/lsrk4.h:33-48): the four-stage low-storage Runge-Kutta scheme.

Each stage is one fused pass: buff = beta*buff - dt*D(rho), rho += c*buff with
(beta, c) = (0, 1/3), (-1, 3/4), (-1, 2/3), (-1, 1/4).  Oracle: the CPU
restatement (pinned by the degree-4 Taylor polynomial of its amplification
factor).  Bar: FP64 <= 1e-10 max-abs.  The HEOM right-hand side is linear and
time independent, so lsrk4 and classical RK4 apply the same polynomial of
dt*L: their trajectories agree to rounding, which the GPU must show too.
"""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10
ENG = None
FORMS = {"hilb": (H.heom_hilb, O.HILB), "liou": (H.heom_liou, O.LIOU),
         "ado": (H.heom_ado, O.ADO)}


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def pair(form, m, dtype="complex128", n_threads=8):
    cls, okind = FORMS[form]
    g = cls(m["max_depth"], dtype=dtype).load(m, eng())
    o = O.Oracle(okind, m["N"], m["max_depth"], n_threads=n_threads).load(m)
    return g, o


@pytest.mark.parametrize("form", ["hilb", "liou", "ado"])
def test_lsrk4_spin_boson_full_config(form):
    """Spin-boson, depth 10, dt=0.01 x 2000 with an output every 10 steps."""
    m = configs.spin_boson()
    g, o = pair(form, m)
    s = H.qme_solver(eng(), g, H.lsrk4())
    rho = m["rho0"].copy()
    seen = []
    s.solve(rho, m["t_list"], lambda t: seen.append(t), {"dt": m["dt"]})
    rho_o, _, st_o, _, _ = o.solve(m["rho0"], m["t_list"], O.LSRK4, m["dt"])
    assert np.allclose(seen, m["t_list"])
    assert s.stats()["n_rhs"] == st_o.n_rhs
    assert np.abs(rho - rho_o).max() <= TOL


def test_lsrk4_chain7_reduced_depth_rho_s():
    """7-site chain at depth 2 (276 ADOs), dt=0.1 fs to 200 fs: rho_S every 10 fs."""
    m = configs.exciton_chain7(max_depth=2, t_final=200.0)
    g, o = pair("hilb", m)
    s = H.qme_solver(eng(), g, H.lsrk4())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": 0.1})
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], O.LSRK4, 0.1)
    assert s.stats()["n_rhs"] == st_o.n_rhs
    assert np.abs(rs - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL


def test_lsrk4_equals_rk4_on_the_linear_heom():
    """Same polynomial of dt*L: GPU lsrk4 and GPU RK4 agree to rounding."""
    m = configs.spin_boson(max_depth=8)
    g = H.heom_hilb(8).load(m, eng())
    out = []
    for solver in (H.lsrk4, H.rk4):
        s = H.qme_solver(eng(), g, solver())
        out.append(s.solve_rho_s(m["rho0"], m["t_list"], {"dt": m["dt"]}))
    assert np.abs(out[0][0] - out[1][0]).max() <= 1e-12
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-12


def test_lsrk4_shortened_last_steps_and_graphs_bitwise():
    """Output intervals that are not whole multiples of dt end in a shortened
    step (as the oracle does); graph replay gives the same bits as direct launches."""
    m = configs.spin_boson(max_depth=6)
    m["t_list"] = np.linspace(0.0, 3.0, 8)
    g, o = pair("hilb", m)
    res = []
    for graphs in (False, True):
        s = H.qme_solver(eng(), g, H.lsrk4())
        s.set_graphs(graphs)
        res.append(s.solve_rho_s(m["rho0"], m["t_list"], {"dt": 0.03}))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    rho_o, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], O.LSRK4, 0.03)
    assert np.abs(res[1][1] - rs_o).max() <= TOL
    assert np.abs(res[1][0] - rho_o).max() <= TOL


def test_lsrk4_complex64_error_vs_fp64():
    """FP32 operator + FP64 stage combination: rho_S within 1e-6 of the FP64 oracle."""
    m = configs.spin_boson(max_depth=8)
    m["t_list"] = np.linspace(0.0, 5.0, 11)
    g, o = pair("hilb", m, dtype="complex64")
    s = H.qme_solver(eng(), g, H.lsrk4())
    _, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": 0.01})
    _, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], O.LSRK4, 0.01)
    err = np.abs(rs - rs_o).max()
    print("complex64 lsrk4 rho_S max-abs error vs FP64 oracle: %.3e" % err)
    assert err <= 1e-6


def test_lsrk4_matrix_core_dense64():
    """N=64 (matrix-core commutator), depth 1, dt=0.01 to t=1."""
    m = configs.dense64(max_depth=1, t_final=1.0, out_every=0.5)
    g, o = pair("hilb", m, n_threads=16)
    s = H.qme_solver(eng(), g, H.lsrk4())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": 0.01})
    rho_o, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], O.LSRK4, 0.01)
    assert np.abs(rs - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL * max(1.0, np.abs(rho_o).max())


def test_lsrk4_batch_members_match_single_solves():
    """A batch of 6 sweep instances: every member as if propagated alone."""
    ms = configs.batched_sweep(n_inst=6, seed=6, n_steps=200, out_every=50)
    b = H.heom_batch([H.heom_hilb(m["max_depth"]).load(m, eng()) for m in ms])
    s = H.qme_solver(eng(), b, H.lsrk4())
    t = ms[0]["t_list"]
    rho, rs = s.solve_rho_s(b.pack([m["rho0"] for m in ms]), t, {"dt": ms[0]["dt"]})
    for m, got_rho, got_rs in zip(ms, b.unpack(rho), b.split_rho_s(rs)):
        o = O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=8).load(m)
        rho_o, rs_o, _, _, _ = o.solve(m["rho0"], t, O.LSRK4, m["dt"])
        if not np.isfinite(rho_o).all() or np.abs(rho_o).max() > 1e30:
            continue  # outside the scheme's stability region at this dt (as for RK4)
        scale = max(1.0, np.abs(rho_o).max())
        assert np.abs(got_rs - rs_o).max() <= TOL * max(1.0, np.abs(rs_o).max()), m["name"]
        assert np.abs(got_rho - rho_o).max() <= TOL * scale, m["name"]
