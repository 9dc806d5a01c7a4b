"""The oracle's Redfield QMEs (redfield.h, redfield_hilb.h, redfield_liou.h).

Pinned independently of the oracle's method (eigendecomposition of H, a
correlation value per transition frequency from an LU solve, basis changes):
  - Lambda_u against its definition int_0^inf C(t) e^{-iHt} V e^{iHt} dt
    + s_delta V, computed as one linear solve without diagonalising H
    (tests/brute.py), for dense / diagonal V, complex parameters, off-diagonal
    gamma and a degenerate spectrum;
  - the eigenvalues against numpy;
  - both forms' calc_diff against the dense Redfield generator;
  - the pure-dephasing closed form rho_01(t) = rho_01(0) exp(-i eps t - 4 Re Ct(0) t);
  - trace preservation and hermiticity of D(rho) for Hermitian rho;
  - an RK4 solve against exp(-M t) rho_0, converging at fourth order.
A flipped transition-frequency sign (omega = w_r - w_c) is shown to move Lambda
far outside the tolerance, so the definition test has teeth.
"""
import numpy as np
import pytest
import scipy.linalg

from libheom_amd import configs
from oracle import oracle as O
from tests import brute

KINDS = {"hilb": O.REDFIELD_HILB, "liou": O.REDFIELD_LIOU}


def model(N, n_noise, ln, seed, dense_v=True, offdiag_gamma=False, complex_params=True):
    return configs.random_model(N, n_noise, ln, 0, seed=seed, dense_v=dense_v,
                                offdiag_gamma=offdiag_gamma, complex_params=complex_params)


CASES = [
    dict(N=2, n_noise=1, ln=2, seed=1),
    dict(N=3, n_noise=2, ln=3, seed=2, offdiag_gamma=True),
    dict(N=4, n_noise=2, ln=2, seed=3, dense_v=False),
    dict(N=5, n_noise=3, ln=1, seed=4, complex_params=False),
    dict(N=7, n_noise=7, ln=3, seed=5, dense_v=False),
    dict(N=6, n_noise=2, ln=0 + 2, seed=6, offdiag_gamma=True, dense_v=False),
]


def case_id(c):
    return "N%d_u%d_k%d%s%s%s" % (c["N"], c["n_noise"], c["ln"],
                                   "" if c.get("dense_v", True) else "_diagV",
                                   "_offg" if c.get("offdiag_gamma") else "",
                                   "_real" if c.get("complex_params") is False else "")


@pytest.mark.parametrize("c", CASES, ids=case_id)
def test_lambda_matches_definition(c):
    m = model(c["N"], c["n_noise"], c["ln"], c["seed"], c.get("dense_v", True),
              c.get("offdiag_gamma", False), c.get("complex_params", True))
    o = O.Oracle(O.REDFIELD_HILB, m["N"], 0).load(m)
    for u, z in enumerate(m["noises"]):
        lam, w = o.redfield_lambda(u)
        ref = brute.redfield_lambda(m["H"], z)
        assert np.abs(lam - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    assert np.abs(np.sort(w) - np.linalg.eigvalsh(m["H"])).max() <= 1e-12 * max(
        1.0, np.abs(w).max())


def test_lambda_degenerate_spectrum():
    """Lambda does not depend on the basis chosen inside a degenerate eigenspace."""
    m = model(4, 2, 2, 11)
    Q, _ = np.linalg.qr(np.random.default_rng(0).normal(size=(4, 4)) +
                        1j * np.random.default_rng(1).normal(size=(4, 4)))
    m["H"] = Q @ np.diag([0.3, 0.3, -1.0, 2.0]) @ Q.conj().T
    o = O.Oracle(O.REDFIELD_HILB, 4, 0).load(m)
    for u, z in enumerate(m["noises"]):
        lam, _ = o.redfield_lambda(u)
        ref = brute.redfield_lambda(m["H"], z)
        assert np.abs(lam - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_flipped_frequency_sign_is_detected():
    m = model(3, 1, 2, 21)
    z = m["noises"][0]
    o = O.Oracle(O.REDFIELD_HILB, 3, 0).load(m)
    lam, _ = o.redfield_lambda(0)
    # omega = w_r - w_c is the definition with H -> -H inside the propagators
    wrong = brute.redfield_lambda(-np.asarray(m["H"]), z)
    assert np.abs(lam - wrong).max() > 1e-3 * np.abs(lam).max()


@pytest.mark.parametrize("form", ["hilb", "liou"])
@pytest.mark.parametrize("c", CASES, ids=case_id)
def test_calc_diff_matches_generator(form, c):
    m = model(c["N"], c["n_noise"], c["ln"], c["seed"], c.get("dense_v", True),
              c.get("offdiag_gamma", False), c.get("complex_params", True))
    o = O.Oracle(KINDS[form], m["N"], 0).load(m)
    M = brute.redfield_generator(m)
    rng = np.random.default_rng(c["seed"])
    rho = rng.normal(size=m["N"] ** 2) + 1j * rng.normal(size=m["N"] ** 2)
    d0 = rng.normal(size=rho.size) + 1j * rng.normal(size=rho.size)
    alpha, beta = 0.7 - 0.2j, 0.3 + 0.4j
    got = o.calc_diff(rho, alpha, beta, drho=d0)
    ref = beta * d0 - alpha * (M @ rho)
    assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_forms_agree():
    m = model(5, 2, 3, 31)
    rho = m["rho0"]
    a = O.Oracle(O.REDFIELD_HILB, 5, 0).load(m).calc_diff(rho)
    b = O.Oracle(O.REDFIELD_LIOU, 5, 0).load(m).calc_diff(rho)
    assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(a).max())


def test_pure_dephasing_closed_form():
    """H = diag(e0, e1), V = sigma_z: rho_01(t) = rho_01(0) exp(-i(e0-e1)t - 4 Re Ct(0) t)."""
    z = {"V": np.diag([1.0, -1.0]).astype(np.complex128),
         "gamma": np.diag([1.0, 2.5]).astype(np.complex128),
         "phi_0": np.array([1.0, 0.5], np.complex128), "sigma": np.array([0.3, 0.2], np.complex128),
         "S": np.diag([0.4, 0.1]).astype(np.complex128), "A": np.diag([0.2, 0.0]).astype(np.complex128),
         "s_delta": 0.05}
    H = np.diag([0.7, -0.4]).astype(np.complex128)
    m = {"N": 2, "H": H, "noises": [z], "max_depth": 0}
    C0 = z["sigma"] @ (z["S"] + 1j * z["A"]) @ np.linalg.solve(z["gamma"], z["phi_0"]) + z["s_delta"]
    rho0 = np.array([[0.6, 0.3 - 0.2j], [0.3 + 0.2j, 0.4]], np.complex128).reshape(-1)
    t = np.linspace(0.0, 2.0, 5)
    for kind in KINDS.values():
        o = O.Oracle(kind, 2, 0).load(m)
        _, rs, _, _, _ = o.solve(rho0, t, O.RK4, 1e-3)
        ref01 = rho0[1] * np.exp(-1j * 1.1 * t - 4.0 * C0.real * t)
        assert np.abs(rs[:, 0, 1] - ref01).max() <= 1e-10
        assert np.abs(rs[:, 0, 0] - rho0[0]).max() <= 1e-12


@pytest.mark.parametrize("form", ["hilb", "liou"])
def test_trace_and_hermiticity(form):
    m = model(4, 3, 2, 41)
    o = O.Oracle(KINDS[form], 4, 0).load(m)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    rho = (X @ X.conj().T).reshape(-1)
    D = o.calc_diff(rho).reshape(4, 4)
    assert abs(np.trace(D)) <= 1e-12 * np.abs(D).max()
    assert np.abs(D - D.conj().T).max() <= 1e-12 * np.abs(D).max()


@pytest.mark.parametrize("form", ["hilb", "liou"])
def test_rk4_solve_matches_matrix_exponential(form):
    m = model(3, 2, 2, 51)
    m["H"] = m["H"] * 0.5
    M = brute.redfield_generator(m)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    rho0 = (X @ X.conj().T)
    rho0 = (rho0 / np.trace(rho0)).reshape(-1)
    t = np.linspace(0.0, 1.0, 5)
    o = O.Oracle(KINDS[form], 3, 0).load(m)
    ref = np.array([scipy.linalg.expm(-M * tt) @ rho0 for tt in t])
    err = []
    for dt in (1e-3, 5e-4):
        _, rs, _, _, _ = o.solve(rho0, t, O.RK4, dt)
        err.append(np.abs(rs.reshape(t.size, -1) - ref).max())
    assert err[1] <= 1e-9
    assert 12.0 < err[0] / err[1] < 20.0  # fourth order: halving dt divides the error by 16
