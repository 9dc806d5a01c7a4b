"""Bath-correlation decompositions that produce libheom's noise parameters.

libheom consumes, for every noise u, an exponential basis
``(gamma_u, sigma_u, phi_0_u, S_u, A_u, s_delta_u)`` (qme_base.h:32-48) and
builds the hierarchy from it (heom.h:57-139); the decomposition itself is the
caller's job (the upstream pyheom package, not part of the reference).  This
module supplies it for the Drude-Lorentz spectral density

    J(w) = 2 lam gam w / (w^2 + gam^2),
    C(t) = (1/pi) int_0^inf J(w) [coth(beta w/2) cos wt - i sin wt] dw
         = sum_k c_k exp(-nu_k t)  (+ white-noise residual),

with the Bose factor expanded either in Matsubara poles or in the
[N-1/N] Pade spectrum decomposition (Hu, Xu & Yan, J. Chem. Phys. 133,
101106 (2010); 134, 244106 (2011)).

The mapping to libheom's parameters is diagonal: gamma = diag(nu_k),
sigma = phi_0 = 1, S = diag(Re c_k), A = diag(Im c_k), so that
s = S phi_0 = Re c and a = A phi_0 = Im c enter heom_hilb.h:184-198 as the
standard HEOM couplings  -i n_k Re(c_k)[V, rho] + n_k Im(c_k){V, rho}.
s_delta is the Ishizaki-Tanimura residual
    delta = int_0^inf Re C(t) dt - sum_k Re(c_k)/nu_k = 2 lam/(beta gam) - ...
which enters as +delta [V,[V,rho]] in D (heom_hilb.h:113-118).
"""
import numpy as np

# hbar in cm^-1 fs and k_B in cm^-1/K (CODATA)
HBAR_CM_FS = 5308.837458876145
KB_CM_K = 0.69503476


def pade_poles_residues(n_pade):
    """[N-1/N] Pade decomposition of the Bose function.

    coth(x/2) ~= 2/x + sum_j 4 eta_j x / (x^2 + xi_j^2), j = 1..n_pade.
    Returns (xi, eta).  Poles come from the eigenvalues of the tridiagonal
    matrices of the continued-fraction form of the Bose function.
    """
    if n_pade <= 0:
        return np.zeros(0), np.zeros(0)
    N = n_pade
    # poles: eigenvalues of the (2N x 2N) matrix with off-diagonals
    # 1/sqrt((2m+3)(2m+5)), m = 0..2N-2
    off = np.array([1.0 / np.sqrt((2 * m + 3) * (2 * m + 5)) for m in range(2 * N - 1)])
    lam = np.linalg.eigvalsh(np.diag(off, 1) + np.diag(off, -1))
    lam = np.sort(lam[lam > 0])[::-1]
    xi = 2.0 / lam[:N]
    # zeros of the numerator: (2N-1 x 2N-1) matrix, off-diagonals
    # 1/sqrt((2m+5)(2m+7)), m = 0..2N-3
    if N > 1:
        off2 = np.array([1.0 / np.sqrt((2 * m + 5) * (2 * m + 7)) for m in range(2 * N - 2)])
        lam2 = np.linalg.eigvalsh(np.diag(off2, 1) + np.diag(off2, -1))
        lam2 = np.sort(lam2[lam2 > 0])[::-1]
        zeta = 2.0 / lam2[:N - 1]
    else:
        zeta = np.zeros(0)
    pref = 0.5 * N * (2 * N + 3)
    eta = np.empty(N)
    for j in range(N):
        v = pref
        for k in range(N - 1):
            v *= (zeta[k] ** 2 - xi[j] ** 2)
        for k in range(N):
            if k != j:
                v /= (xi[k] ** 2 - xi[j] ** 2)
        eta[j] = v
    return xi, eta


def coth_half_expansion(x, xi, eta):
    """2/x + sum_j 4 eta_j x/(x^2+xi_j^2), the expansion of coth(x/2)."""
    x = np.asarray(x, dtype=np.float64)
    out = 2.0 / x
    for a, b in zip(xi, eta):
        out = out + 4.0 * b * x / (x * x + a * a)
    return out


def drude_modes(lam, gam, beta, n_extra, scheme="pade"):
    """Exponents nu_k and complex weights c_k of the Drude-Lorentz C(t).

    Mode 0 is the Drude pole (nu = gam, c = lam gam (cot(beta gam/2) - i));
    modes 1..n_extra are Bose-function poles (Matsubara or Pade).
    Returns (nu, c, delta) with delta the white-noise residual.
    """
    nu = [gam]
    c = [lam * gam * (1.0 / np.tan(beta * gam / 2.0) - 1.0j)]
    if scheme == "matsubara":
        xi = 2.0 * np.pi * np.arange(1, n_extra + 1)
        eta = np.ones(n_extra)
    elif scheme == "pade":
        xi, eta = pade_poles_residues(n_extra)
    else:
        raise ValueError(scheme)
    for a, b in zip(xi, eta):
        v = a / beta
        nu.append(v)
        c.append(4.0 * lam * gam / beta * b * v / (v * v - gam * gam))
    nu = np.array(nu, dtype=np.float64)
    c = np.array(c, dtype=np.complex128)
    delta = 2.0 * lam / (beta * gam) - float(np.sum(c.real / nu))
    return nu, c, delta


def noise_params(V, nu, c, delta, use_delta=True):
    """libheom noise-u parameter dict from exponents/weights."""
    n = len(nu)
    return {
        "V": np.asarray(V, dtype=np.complex128),
        "gamma": np.diag(np.asarray(nu, dtype=np.complex128)),
        "phi_0": np.ones(n, dtype=np.complex128),
        "sigma": np.ones(n, dtype=np.complex128),
        "S": np.diag(np.asarray(c).real.astype(np.complex128)),
        "A": np.diag(np.asarray(c).imag.astype(np.complex128)),
        "s_delta": complex(delta if use_delta else 0.0),
    }


def correlation(t, nu, c):
    """Exponential-basis correlation sum_k c_k exp(-nu_k t)."""
    t = np.asarray(t, dtype=np.float64)
    return np.sum(np.asarray(c)[:, None] * np.exp(-np.outer(nu, t)), axis=0)
