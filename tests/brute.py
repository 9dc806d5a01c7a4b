"""Brute-force dense HEOM generator, built independently of the oracle.

The hierarchy is enumerated with itertools (not breadth-first), neighbours
are found by dictionary lookup of multi-indices (not through +-1 tables), and
every term of the equation is assembled as a dense Kronecker superoperator
with row-major vectorisation  vec(A X B) = (A kron B^T) vec(X):

  D_n = i[H, rho_n] + Gamma_n rho_n + sum_u delta_u [V_u, [V_u, rho_n]]
      + sum_u sum_{j != k} f_jk gamma_u[j][k] rho_{n - 1_uj + 1_uk}
      + sum_u ( i[V_u, Phi_nu] + {V_u, Psi_nu} ),
  Phi_nu = sum_k sigma_k g_+ rho_{n+1_uk} + sum_k s_k g_- rho_{n-1_uk},
  Psi_nu = - sum_k a_k g_- rho_{n-1_uk},
  s = S phi_0, a = A phi_0, Gamma_n = sum n_uk gamma_u[k][k],
  g_+ = 1 | sqrt(n+1), g_- = n | sqrt(n)   (plain | sqrt normalisation),
  f_jk = n_j | sqrt(n_j (n_k + 1)).

The dense matrix M satisfies  calc_diff(rho, alpha=1, beta=0) = -M @ rho.
Only usable for n_hrchy * N^2 up to a few thousand.
"""
import itertools

import numpy as np


def multi_indices(K, L):
    out = []
    for tot in range(L + 1):
        for c in itertools.product(range(tot + 1), repeat=K):
            if sum(c) == tot:
                out.append(tuple(c))
    return out


def _comm(A):
    n = A.shape[0]
    I = np.eye(n)
    return np.kron(A, I) - np.kron(I, A.T)


def _acomm(A):
    n = A.shape[0]
    I = np.eye(n)
    return np.kron(A, I) + np.kron(I, A.T)


def dense_generator(model, order=None, offdiag_gamma=False, sqrt_norm=False):
    """Return (M, indices) with M the dense D-operator in the given ADO order.

    order: list of multi-index tuples fixing the block order (e.g. the
    oracle's breadth-first table); default is the itertools order.
    """
    N = model["N"]
    H = np.asarray(model["H"], np.complex128)
    noises = model["noises"]
    lens = [np.atleast_2d(z["gamma"]).shape[0] for z in noises]
    K = sum(lens)
    L = model["max_depth"]
    idx = multi_indices(K, L) if order is None else [tuple(int(v) for v in r) for r in order]
    pos = {v: i for i, v in enumerate(idx)}
    nh = len(idx)
    N2 = N * N
    M = np.zeros((nh * N2, nh * N2), np.complex128)
    lk = []
    c = 0
    for ln in lens:
        lk.append(list(range(c, c + ln)))
        c += ln
    base = 1j * _comm(H)
    cV = [_comm(np.asarray(z["V"], np.complex128)) for z in noises]
    aV = [_acomm(np.asarray(z["V"], np.complex128)) for z in noises]
    for u, z in enumerate(noises):
        base = base + complex(z["s_delta"]) * cV[u] @ cV[u]
    s = [np.atleast_2d(z["S"]) @ np.asarray(z["phi_0"]) for z in noises]
    a = [np.atleast_2d(z["A"]) @ np.asarray(z["phi_0"]) for z in noises]
    I2 = np.eye(N2)
    for n in idx:
        r = pos[n]
        blk = lambda m: (slice(r * N2, (r + 1) * N2), slice(m * N2, (m + 1) * N2))
        gam_n = 0.0
        for u, z in enumerate(noises):
            g = np.atleast_2d(z["gamma"])
            for k in range(lens[u]):
                gam_n = gam_n + n[lk[u][k]] * g[k, k]
        M[blk(r)] += base + gam_n * I2
        for u, z in enumerate(noises):
            g = np.atleast_2d(z["gamma"])
            sig = np.asarray(z["sigma"])
            for k in range(lens[u]):
                nk = n[lk[u][k]]
                p = list(n)
                p[lk[u][k]] += 1
                p = tuple(p)
                if p in pos:
                    gp = np.sqrt(nk + 1) if sqrt_norm else 1.0
                    M[blk(pos[p])] += 1j * sig[k] * gp * cV[u]
                if nk > 0:
                    m = list(n)
                    m[lk[u][k]] -= 1
                    m = tuple(m)
                    gm = np.sqrt(nk) if sqrt_norm else float(nk)
                    M[blk(pos[m])] += gm * (1j * s[u][k] * cV[u] - a[u][k] * aV[u])
            if offdiag_gamma:
                for j in range(lens[u]):
                    for k in range(lens[u]):
                        if j == k or g[j, k] == 0 or n[lk[u][j]] == 0:
                            continue
                        q = list(n)
                        q[lk[u][j]] -= 1
                        q[lk[u][k]] += 1
                        q = tuple(q)
                        if q not in pos:
                            continue
                        nj, nk = n[lk[u][j]], n[lk[u][k]]
                        f = np.sqrt(nj * (nk + 1)) if sqrt_norm else float(nj)
                        M[blk(pos[q])] += f * g[j, k] * I2
    return M, idx


def redfield_lambda(H, z):
    """Redfield relaxation operator from its definition, without diagonalising H:

        Lambda = int_0^inf C(t) e^{-iHt} V e^{iHt} dt + s_delta V,
        C(t) = sigma^T (S + iA) e^{-gamma t} phi_0.

    Y(t) = (e^{-gamma t} phi_0) kron vec(e^{-iHt} V e^{iHt}) obeys
    dY/dt = -(gamma kron 1 + 1 kron i ad_H) Y with ad_H X = HX - XH, so its
    integral is one linear solve in the (bath mode) x (N x N) space.
    """
    H = np.asarray(H, np.complex128)
    N = H.shape[0]
    V = np.asarray(z["V"], np.complex128)
    g = np.atleast_2d(np.asarray(z["gamma"], np.complex128))
    K = g.shape[0]
    lam = complex(z["s_delta"]) * V
    if K == 0:
        return lam
    M = np.kron(g, np.eye(N * N)) + np.kron(np.eye(K), 1j * _comm(H))
    rhs = np.kron(np.asarray(z["phi_0"], np.complex128), V.reshape(-1))
    Y = np.linalg.solve(M, rhs).reshape(K, N * N)
    c = np.asarray(z["sigma"], np.complex128) @ (np.atleast_2d(z["S"]) + 1j * np.atleast_2d(z["A"]))
    return lam + (c @ Y).reshape(N, N)


def redfield_generator(model):
    """Dense M with D(rho) = i[H, rho] + sum_u [V_u, Lambda_u rho - rho Lambda_u^+] = M vec(rho)
    (row-major vectorisation); calc_diff(rho, alpha=1, beta=0) = -M @ rho."""
    H = np.asarray(model["H"], np.complex128)
    N = H.shape[0]
    I = np.eye(N)
    M = 1j * _comm(H)
    for z in model["noises"]:
        lam = redfield_lambda(H, z)
        T = np.kron(lam, I) - np.kron(I, lam.conj().T.T)
        M = M + _comm(np.asarray(z["V"], np.complex128)) @ T
    return M
