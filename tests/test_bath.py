"""Bath decomposition (Drude-Lorentz, Matsubara and [N-1/N] Pade).

Pinned by: the exact coth(x/2) for the Pade expansion, the analytic Matsubara
residues, and a direct numerical quadrature of
C(t) = (1/pi) int J(w) [coth(beta w/2) cos wt - i sin wt] dw.
"""
import numpy as np
import pytest
from scipy.integrate import quad

from libheom_amd import bath


@pytest.mark.parametrize("n,x_max,tol", [(1, 1.0, 1e-5), (2, 2.0, 1e-7), (3, 5.0, 2e-7),
                                          (5, 10.0, 2e-9)])
def test_pade_expansion_matches_coth(n, x_max, tol):
    xi, eta = bath.pade_poles_residues(n)
    x = np.linspace(0.05, x_max, 200)
    exact = 1.0 / np.tanh(x / 2)
    assert np.abs(bath.coth_half_expansion(x, xi, eta) / exact - 1).max() <= tol


def test_pade_converges_faster_than_matsubara():
    x = np.array([3.0, 6.0])
    xi, eta = bath.pade_poles_residues(3)
    pade = np.abs(bath.coth_half_expansion(x, xi, eta) * np.tanh(x / 2) - 1).max()
    mats = np.abs(bath.coth_half_expansion(x, 2 * np.pi * np.arange(1, 4), np.ones(3)) *
                  np.tanh(x / 2) - 1).max()
    assert pade < 1e-3 * mats


def test_matsubara_residues_closed_form():
    lam, gam, beta = 0.5, 1.0, 1.0
    nu, c, delta = bath.drude_modes(lam, gam, beta, 3, "matsubara")
    k = np.arange(1, 4)
    vk = 2 * np.pi * k / beta
    assert np.allclose(nu[1:], vk)
    assert np.allclose(c[1:], 4 * lam * gam / beta * vk / (vk ** 2 - gam ** 2))
    assert np.isclose(c[0], lam * gam * (1 / np.tan(beta * gam / 2) - 1j))


def correlation_quad(t, lam, gam, beta):
    def jcoth(w):  # J(w) coth(beta w/2), finite at w -> 0
        if w < 1e-8:
            return 4 * lam / (beta * gam)
        return 2 * lam * gam * w / (w * w + gam * gam) / np.tanh(beta * w / 2)
    re = quad(jcoth, 0, np.inf, weight="cos", wvar=t, limlst=200)[0] / np.pi
    im = -quad(lambda w: 2 * lam * gam * w / (w * w + gam * gam), 0, np.inf, weight="sin",
               wvar=t, limlst=200)[0] / np.pi
    return re + 1j * im


@pytest.mark.parametrize("t", [0.3, 1.0, 2.5])
def test_drude_correlation_against_quadrature(t):
    lam, gam, beta = 0.5, 1.0, 2.0
    nu, c, _ = bath.drude_modes(lam, gam, beta, 8, "pade")
    got = bath.correlation(np.array([t]), nu, c)[0]
    want = correlation_quad(t, lam, gam, beta)
    assert abs(got - want) <= 1e-6 * max(1.0, abs(want))


def test_residual_delta_accounts_for_dropped_modes():
    lam, gam, beta = 0.5, 1.0, 1.0
    full = bath.drude_modes(lam, gam, beta, 400, "matsubara")
    one = bath.drude_modes(lam, gam, beta, 1, "matsubara")
    # delta + sum_kept Re c/nu == 2 lam/(beta gam) == int_0^inf Re C dt
    for nu, c, d in (full, one):
        assert np.isclose(d + np.sum(c.real / nu), 2 * lam / (beta * gam))
    assert abs(full[2]) < abs(one[2])


def test_noise_params_layout():
    nu, c, d = bath.drude_modes(0.3, 1.0, 1.0, 2)
    z = bath.noise_params(np.diag([1.0, -1.0]), nu, c, d)
    assert z["gamma"].shape == (3, 3) and np.allclose(np.diag(z["gamma"]), nu)
    assert np.allclose(z["S"] @ z["phi_0"], c.real)
    assert np.allclose(z["A"] @ z["phi_0"], c.imag)
