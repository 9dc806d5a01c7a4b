"""Pins the oracle's integrators (rk4.h:40-65, rkdp.h:21-179,
fixed_step_size_solver.h:29-52, adaptive_step_size_solver.h:32-56,
errnrm1 of linalg_engine_mkl.h.j2:187-212).

Checks: one RK4 step on the linear HEOM equals the degree-4 Taylor polynomial
of the dense generator; one Dormand-Prince attempt equals scipy's RK45
tableau (an independent library copy of DP5(4)) and the step-size rule;
hand-computed control cases; the RMS error norm; end-to-end agreement with
scipy.integrate.solve_ivp on the dense generator.
"""
import math

import numpy as np
import pytest
from scipy.integrate import RK45, solve_ivp

from libheom_amd import configs
from oracle import oracle as O
from tests.brute import dense_generator


def small_model(seed=3):
    m = configs.random_model(2, 1, 2, 3, seed=seed, dense_v=True, complex_params=False)
    for z in m["noises"]:
        z["V"] = 0.5 * (z["V"] + z["V"].conj().T)
    return m


def setup(m):
    o = O.Oracle(O.HILB, m["N"], m["max_depth"]).load(m)
    n, _, _ = o.tables()
    M, _ = dense_generator(m, order=n)
    return o, -M  # f(y) = -M y


def test_errnrm1_textbook():
    rng = np.random.default_rng(0)
    e, x1, x2 = (rng.normal(size=50) + 1j * rng.normal(size=50) for _ in range(3))
    atol, rtol = 1e-3, 1e-2
    want = np.sqrt(np.mean((np.abs(e) / (atol + rtol * np.maximum(np.abs(x1), np.abs(x2)))) ** 2))
    assert abs(O.errnrm1(e, x1, x2, atol, rtol) - want) <= 1e-13 * want


def test_rk4_step_is_taylor_polynomial():
    m = small_model()
    o, A = setup(m)
    h = 0.05
    y0 = m["rho0"]
    hA = h * A
    want = y0.copy()
    term = y0.copy()
    for j in range(1, 5):
        term = hA @ term / j
        want = want + term
    got = o.rk4_step(y0, h)
    assert np.abs(got - want).max() <= 1e-13


def test_lsrk4_step_is_taylor_polynomial():
    """lsrk4.h:33-48 (four stages, buff = beta*buff - dt*D, rho += c*buff) is
    fourth order: on a linear generator one step is the degree-4 polynomial."""
    m = small_model()
    o, A = setup(m)
    h = 0.05
    y0 = m["rho0"]
    hA = h * A
    want = y0.copy()
    term = y0.copy()
    for j in range(1, 5):
        term = hA @ term / j
        want = want + term
    got = o.lsrk4_step(y0, h)
    assert np.abs(got - want).max() <= 1e-13


def dopri_attempt(A, y, h, atol, rtol):
    """One DP5(4) attempt with scipy's RK45 coefficients (k = h f)."""
    a, b, e = RK45.A, RK45.B, RK45.E
    K = np.zeros((7, y.size), np.complex128)
    for i in range(6):
        yi = y + (a[i, :i] @ K[:i] if i else 0)
        K[i] = h * (A @ yi)
    ynew = y + b @ K[:6]
    K[6] = h * (A @ ynew)
    err = e @ K
    scale = atol + rtol * np.maximum(np.abs(y), np.abs(ynew))
    return ynew, np.sqrt(np.mean(np.abs(err / scale) ** 2))


@pytest.mark.parametrize("h,atol,rtol", [(0.02, 1e-8, 1e-6), (0.3, 1e-10, 1e-8), (1.5, 1e-6, 1e-4)])
def test_rkdp_attempt_matches_scipy_tableau(h, atol, rtol):
    m = small_model()
    o, A = setup(m)
    y0 = m["rho0"]
    ynew, en = dopri_attempt(A, y0, h, atol, rtol)
    got, t, dt, st, rc = o.rkdp_step(y0, 0.0, 100.0, h, atol, rtol)
    if en < 1.0:
        assert st.n_rejected == 0 and st.n_accepted == 1
        assert np.abs(got - ynew).max() <= 1e-12
        assert t == h
        assert dt == pytest.approx(h * 0.9 * min(10.0, en ** -0.2), rel=1e-12)
    else:
        # first attempt rejected: h shrinks by 0.9*max(0.2, err^-1/5) and retries
        h2 = max(h * 0.9 * max(0.2, en ** -0.2), 1e-16)
        ynew2, en2 = dopri_attempt(A, y0, h2, atol, rtol)
        assert st.n_rejected >= 1
        if st.n_rejected == 1:
            assert en2 < 1.0
            assert np.abs(got - ynew2).max() <= 1e-12


def test_rkdp_zero_error_grows_step_tenfold_times_safety():
    m = small_model()
    for z in m["noises"]:
        for key in ("S", "A", "gamma"):
            z[key] = np.zeros_like(z[key])
        z["sigma"] = np.zeros_like(z["sigma"])
        z["s_delta"] = 0.0
    m["H"] = np.zeros_like(m["H"])
    o = O.Oracle(O.HILB, m["N"], m["max_depth"]).load(m)
    got, t, dt, st, rc = o.rkdp_step(m["rho0"], 0.0, 10.0, 0.1, 1e-8, 1e-6)
    assert st.n_accepted == 1 and t == 0.1
    assert dt == pytest.approx(0.1 * 0.9 * 10.0, rel=1e-15)
    assert np.array_equal(got, m["rho0"])


def test_rkdp_nan_gives_up_after_1000_attempts():
    m = small_model()
    o = O.Oracle(O.HILB, m["N"], m["max_depth"]).load(m)
    y0 = m["rho0"].copy()
    y0[0] = np.nan
    got, t, dt, st, rc = o.rkdp_step(y0, 0.0, 5.0, 0.1, 1e-8, 1e-6)
    assert rc == 1 and st.gave_up == 1
    assert st.n_rejected == 1000 and t == 5.0
    assert dt == 1e-16  # 0.1 * 0.18^k clamps at dt_min
    assert np.isnan(got[0]) and np.array_equal(got[1:], y0[1:])


def test_rkdp_clips_to_interval_end():
    m = small_model()
    o = O.Oracle(O.HILB, m["N"], m["max_depth"]).load(m)
    _, t, dt, st, _ = o.rkdp_step(m["rho0"], 0.0, 0.01, 0.5, 1e-6, 1e-4)
    assert t == 0.01


def test_fixed_step_last_step_shortened():
    """fixed_step_size_solver.h:39-51: steps of dt, the last one ends at t1."""
    m = small_model()
    o = O.Oracle(O.HILB, m["N"], m["max_depth"]).load(m)
    for t1, dt in [(0.1, 0.01), (0.1, 0.03), (1.0, 0.3)]:
        _, _, st, _, _ = o.solve(m["rho0"], [0.0, t1], O.RK4, dt)
        # the first step whose end comes within machine epsilon of t1 is the last
        t, n = 0.0, 0
        while True:
            n += 1
            if t + dt - t1 > -np.finfo(float).eps:
                break
            t += dt
        assert st.n_rhs == 4 * n
        assert n == math.ceil(t1 / dt - 1e-9)


def test_end_to_end_against_solve_ivp():
    m = configs.spin_boson(max_depth=5)
    o = O.Oracle(O.HILB, 2, 5).load(m)
    n, _, _ = o.tables()
    M, _ = dense_generator(m, order=n)
    t_list = np.linspace(0.0, 4.0, 9)
    sol = solve_ivp(lambda t, y: -M @ y, (0.0, 4.0), m["rho0"], method="DOP853",
                    t_eval=t_list, rtol=1e-12, atol=1e-14)
    ref = sol.y.T.reshape(len(t_list), -1, 2, 2)[:, 0]
    _, rs4, _, _, _ = o.solve(m["rho0"], t_list, O.RK4, 0.005)
    assert np.abs(rs4 - ref).max() <= 1e-8
    _, rsd, st, _, _ = o.solve(m["rho0"], t_list, O.RKDP, 0.01, 1e-12, 1e-10)
    assert np.abs(rsd - ref).max() <= 1e-8
    assert st.n_accepted > 0
