"""Parity of the HIP engine (through the C ABI) against the CPU oracle.

Bar: FP64 GPU vs FP64 oracle <= 1e-10 max-abs (calc_diff outputs, full
states and rho_S at every output time); FP32 (complex64) variants state their
own error against the FP64 oracle.
"""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10
FORMS = {"hilb": (H.heom_hilb, O.HILB), "liou": (H.heom_liou, O.LIOU),
         "ado": (H.heom_ado, O.ADO)}
ENG = None


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def pair(form, m, parity_q1=True, sqrt_norm=False, dtype="complex128", order=H.row_major):
    cls, okind = FORMS[form]
    g = cls(m["max_depth"], dtype=dtype, parity_q1=parity_q1, sqrt_normalization=sqrt_norm,
            order=order).load(m, eng())
    o = O.Oracle(okind, m["N"], m["max_depth"], n_threads=4, parity_q1=parity_q1,
                 sqrt_norm=sqrt_norm).load(m)
    return g, o


@pytest.mark.parametrize("form", ["hilb", "liou", "ado"])
@pytest.mark.parametrize("dense_v", [True, False])
@pytest.mark.parametrize("variant", [(True, False), (False, False), (False, True)])
@pytest.mark.parametrize("N,n_noise,ln,L", [(2, 1, 2, 5), (3, 2, 2, 3), (5, 2, 1, 3)])
def test_calc_diff_parity(form, dense_v, variant, N, n_noise, ln, L):
    q1, sq = variant
    print("calc_diff", form, dense_v, variant, N, n_noise, ln, L, flush=True)
    m = configs.random_model(N, n_noise, ln, L, seed=N * 100 + L, dense_v=dense_v,
                             offdiag_gamma=True)
    g, o = pair(form, m, q1, sq)
    assert g.n_hrchy == o.n_hrchy
    rng = np.random.default_rng(3)
    rho = m["rho0"]
    d0 = rng.normal(size=rho.size) + 1j * rng.normal(size=rho.size)
    for alpha, beta in [(1.0, 0.0), (0.3 - 0.2j, 0.7 + 0.1j)]:
        ref = o.calc_diff(rho, alpha, beta, d0)
        got = g.calc_diff_impl(None, d0.copy(), rho, alpha, beta)
        err = np.abs(got - ref).max()
        assert err <= TOL * max(1.0, np.abs(ref).max()), (err, alpha, beta)


@pytest.mark.parametrize("N", list(range(1, 17)))
def test_calc_diff_every_n_level(N):
    """Every compiled n_level (1..16) of the Hilbert kernel, both V paths."""
    print("n_level", N, flush=True)
    for dense_v in (True, False):
        m = configs.random_model(N, 2, 1, 2, seed=N, dense_v=dense_v)
        g, o = pair("hilb", m)
        ref = o.calc_diff(m["rho0"], 1.0, 0.0)
        got = g.calc_diff_impl(None, np.zeros_like(ref), m["rho0"], 1.0, 0.0)
        assert np.abs(got - ref).max() <= TOL * max(1.0, np.abs(ref).max())


def test_col_major_matches_transposed_problem():
    m = configs.random_model(3, 2, 2, 3, seed=11, dense_v=True)
    g_r, o = pair("hilb", m)
    g_c = H.heom_hilb(m["max_depth"], order=H.col_major).load(m, eng())
    N = m["N"]
    rho = m["rho0"].reshape(-1, N, N)
    rho_c = np.ascontiguousarray(rho.transpose(0, 2, 1)).reshape(-1)
    ref = o.calc_diff(m["rho0"], 1.0, 0.0).reshape(-1, N, N)
    got = g_c.calc_diff_impl(None, np.zeros_like(rho_c), rho_c, 1.0, 0.0).reshape(-1, N, N)
    assert np.abs(got.transpose(0, 2, 1) - ref).max() <= TOL


@pytest.mark.parametrize("K,L", [(1, 0), (1, 7), (2, 10), (3, 4), (5, 3), (21, 2)])
def test_hierarchy_tables(K, L):
    m = configs.random_model(2, K, 1, L, seed=1, dense_v=False)
    g, o = pair("hilb", m)
    for a, b in zip(g.hrchy_tables(), o.tables()):
        assert np.array_equal(a, b)


def test_hierarchy_chain7_full_size():
    m = configs.exciton_chain7()
    g = H.heom_hilb(m["max_depth"]).load(m, eng())
    o = O.Oracle(O.HILB, 7, 4).load(m)
    assert g.n_hrchy == o.n_hrchy == 12650
    for a, b in zip(g.hrchy_tables(), o.tables()):
        assert np.array_equal(a, b)


def test_calc_diff_chain7_full_size():
    """The 7-site, 21-mode, depth-4 operator (12650 ADOs) at full size."""
    m = configs.exciton_chain7()
    g, o = pair("hilb", m)
    rng = np.random.default_rng(5)
    rho = (rng.normal(size=g.main_size()) + 1j * rng.normal(size=g.main_size())) * 1e-2
    ref = o.calc_diff(rho, 1.0, 0.0)
    got = g.calc_diff_impl(None, np.zeros_like(ref), rho, 1.0, 0.0)
    assert np.abs(got - ref).max() <= TOL * max(1.0, np.abs(ref).max())


def _solve_both(form, m, solver, kw, dtype="complex128", n_threads=8):
    g, o = pair(form, m, dtype=dtype)
    s = H.qme_solver(eng(), g, H.rk4() if solver == "rk4" else H.rkdp())
    rho = m["rho0"].copy()
    seen = []
    s.solve(rho, m["t_list"], lambda t: seen.append(t), kw)
    o.set_threads(n_threads)
    okind = O.RK4 if solver == "rk4" else O.RKDP
    rho_o, rs_o, st_o, dts, rc = o.solve(m["rho0"], m["t_list"], okind, kw["dt"],
                                         kw.get("atol", 0.0), kw.get("rtol", 0.0))
    return rho, seen, rho_o, rs_o, st_o, s.stats()


@pytest.mark.parametrize("form", ["hilb", "liou", "ado"])
def test_rk4_spin_boson_full_config(form):
    """Spin-boson, N=2, 2 modes, depth 10, RK4 dt=0.01 x 2000, rho_S every 10 steps."""
    m = configs.spin_boson()
    rho, seen, rho_o, rs_o, st_o, st = _solve_both(form, m, "rk4", {"dt": m["dt"]})
    assert np.allclose(seen, m["t_list"])
    assert st["n_rhs"] == st_o.n_rhs
    assert np.abs(rho - rho_o).max() <= TOL


def test_rk4_rho_s_every_output():
    m = configs.spin_boson()
    g, o = pair("hilb", m)
    s = H.qme_solver(eng(), g, H.rk4())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": m["dt"]})
    _, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], O.RK4, m["dt"])
    assert np.abs(rs - rs_o).max() <= TOL


@pytest.mark.parametrize("form", ["hilb", "liou", "ado"])
def test_rkdp_small_parity(form):
    m = configs.spin_boson(max_depth=6, T=0.5)
    m["t_list"] = np.linspace(0.0, 5.0, 11)
    kw = {"dt": 0.05, "atol": 1e-10, "rtol": 1e-8}
    rho, seen, rho_o, rs_o, st_o, st = _solve_both(form, m, "rkdp", kw)
    assert st["n_accepted"] == st_o.n_accepted
    assert st["n_rejected"] == st_o.n_rejected
    assert np.abs(rho - rho_o).max() <= TOL


def test_rkdp_chain7_reduced_depth_rho_s():
    """7-site chain at depth 2 (276 ADOs) over 200 fs: rho_S at every 10 fs."""
    m = configs.exciton_chain7(max_depth=2, t_final=200.0)
    g, o = pair("hilb", m)
    s = H.qme_solver(eng(), g, H.rkdp())
    kw = {"dt": m["dt"], "atol": m["atol"], "rtol": m["rtol"]}
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    o.set_threads(8)
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"], kw["atol"],
                                      kw["rtol"])
    st = s.stats()
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    assert np.abs(rs - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL


def test_rkdp_dt_save_carries_over_between_solves():
    m = configs.spin_boson(max_depth=4)
    kw = {"dt": 0.05, "atol": 1e-10, "rtol": 1e-8}
    g, o = pair("hilb", m)
    s = H.qme_solver(eng(), g, H.rkdp())
    t1 = np.linspace(0.0, 1.0, 3)
    t2 = np.linspace(1.0, 2.0, 3)
    rho = m["rho0"].copy()
    s.solve(rho, t1, None, kw)
    s.solve(rho, t2, None, kw)
    rho_o, _, _, dts, _ = o.solve(m["rho0"], t1, O.RKDP, kw["dt"], kw["atol"], kw["rtol"])
    rho_o, _, _, _, _ = o.solve(rho_o, t2, O.RKDP, kw["dt"], kw["atol"], kw["rtol"],
                                dt_save=dts)
    assert np.abs(rho - rho_o).max() <= TOL


def test_complex64_variant_error_vs_fp64():
    """FP32 operator + FP64 stage combination: rho_S within 1e-6 of the FP64 oracle."""
    m = configs.spin_boson(max_depth=8)
    m["t_list"] = np.linspace(0.0, 5.0, 11)
    g, o = pair("hilb", m, dtype="complex64")
    s = H.qme_solver(eng(), g, H.rk4())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": 0.01})
    _, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], O.RK4, 0.01)
    err = np.abs(rs - rs_o).max()
    print("complex64 rk4 rho_S max-abs error vs FP64 oracle: %.3e" % err)
    assert err <= 1e-6


def test_errors_are_status_codes():
    m = configs.spin_boson(max_depth=2)
    g = H.heom_hilb(2).load(m, eng())
    s = H.qme_solver(eng(), g, H.rkdp())
    rho = m["rho0"].copy()
    with pytest.raises(KeyError):
        s.solve(rho, [0.0, 1.0], None, {"dt": 0.1})
    with pytest.raises(H.HeomError) as e:
        s.solve(rho, [0.0, 1.0], None, {"dt": 0.1, "atol": 0.0, "rtol": 0.0})
    assert e.value.code == 1
    s4 = H.qme_solver(eng(), g, H.rk4())
    with pytest.raises(H.HeomError) as e:
        s4.solve(rho, [0.0, 1.0], None, {"dt": -0.1})
    assert e.value.code == 1


def test_rkdp_gives_up_on_nan_state():
    """A NaN state makes every attempt fail; after 1000 attempts the interval is
    skipped (rkdp.h:95-102) and the status reports it instead of staying silent."""
    m = configs.spin_boson(max_depth=2)
    g = H.heom_hilb(2).load(m, eng())
    s = H.qme_solver(eng(), g, H.rkdp())
    rho = m["rho0"].copy()
    rho[1] = np.nan
    with pytest.raises(H.HeomError) as e:
        s.solve(rho, [0.0, 1.0], None, {"dt": 0.1, "atol": 1e-8, "rtol": 1e-6})
    assert e.value.code == 4
    st = s.stats()
    assert st["gave_up"] == 1 and st["n_rejected"] == 1000


def diag_model(N, n_noise, ln, L, seed, complex_diag=False):
    m = configs.random_model(N, n_noise, ln, L, seed=seed, dense_v=False)
    if complex_diag:
        rng = np.random.default_rng(seed + 1)
        for z in m["noises"]:
            z["V"] = np.diag(rng.normal(size=N) + 1j * rng.normal(size=N))
    return m


@pytest.mark.parametrize("N", [17, 20, 33, 45, 46, 64])
@pytest.mark.parametrize("complex_diag", [False, True])
def test_calc_diff_large_n(N, complex_diag):
    """heom_hilb for N > 16 (register-tiled large-N kernel), diagonal V."""
    print("large N", N, complex_diag, flush=True)
    m = diag_model(N, 2, 1, 2, seed=N, complex_diag=complex_diag)
    g, o = pair("hilb", m)
    rng = np.random.default_rng(N)
    d0 = rng.normal(size=g.main_size()) + 1j * rng.normal(size=g.main_size())
    for alpha, beta in [(1.0, 0.0), (0.5 - 0.25j, 0.3 + 0.2j)]:
        ref = o.calc_diff(m["rho0"], alpha, beta, d0)
        got = g.calc_diff_impl(None, d0.copy(), m["rho0"], alpha, beta)
        assert np.abs(got - ref).max() <= TOL * max(1.0, np.abs(ref).max())


def test_large_n_complex64_calc_diff():
    m = diag_model(24, 2, 2, 2, seed=3)
    g = H.heom_hilb(2, dtype="complex64").load(m, eng())
    o = O.Oracle(O.HILB, 24, 2).load(m)
    ref = o.calc_diff(m["rho0"], 1.0, 0.0)
    got = g.calc_diff_impl(None, np.zeros_like(ref), m["rho0"], 1.0, 0.0)
    assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_large_n_rkdp_parity():
    m = diag_model(20, 2, 1, 2, seed=21)
    m["rho0"] = m["rho0"] * 0.0
    m["rho0"][0] = 1.0
    m["t_list"] = np.linspace(0.0, 0.5, 3)
    kw = {"dt": 0.01, "atol": 1e-10, "rtol": 1e-8}
    rho, seen, rho_o, rs_o, st_o, st = _solve_both("hilb", m, "rkdp", kw)
    assert st["n_accepted"] == st_o.n_accepted
    assert np.abs(rho - rho_o).max() <= TOL
