"""The five benchmark workloads, as libheom parameter sets.

Each builder returns a plain dict:
  N, H (N x N), noises (list of noise dicts from bath.noise_params),
  max_depth, rho0 (main_size complex, only ADO 0 non-zero), t_list,
  solver ("rk4" | "rkdp"), dt, atol, rtol, and a human-readable name.
Units are hbar = 1 throughout: the 7-site chain is converted from cm^-1 / fs
to rad/fs (times stay in fs).
"""
import numpy as np

from . import bath


def _rho0_projector(N, n_hrchy, site=0):
    rho = np.zeros((n_hrchy, N, N), np.complex128)
    rho[0, site, site] = 1.0
    return rho.reshape(-1)


def hrchy_count(n_dim, depth):
    """C(K+L, L), the ADO count (hrchy_space.cc:68-71)."""
    from math import comb
    return comb(n_dim + depth, depth)


def spin_boson(max_depth=10, n_pade=1, dt=0.01, n_steps=2000, out_every=10,
               lam=0.5, gam=1.0, T=1.0, eps=0.0, Delta=1.0, scheme="pade"):
    """Spin-boson, N=2: H = eps/2 sz + Delta sx, V = sz, Drude bath (1 + n_pade modes)."""
    sx = np.array([[0, 1], [1, 0]], np.complex128)
    sz = np.array([[1, 0], [0, -1]], np.complex128)
    H = 0.5 * eps * sz + Delta * sx
    nu, c, delta = bath.drude_modes(lam, gam, 1.0 / T, n_pade, scheme)
    noises = [bath.noise_params(sz, nu, c, delta)]
    K = len(nu)
    n_t = n_steps // out_every + 1
    t_list = dt * out_every * np.arange(n_t)
    return {
        "name": "spin-boson N=2 K=%d L=%d" % (K, max_depth),
        "N": 2, "H": H, "noises": noises, "max_depth": max_depth,
        "rho0": _rho0_projector(2, hrchy_count(K, max_depth), 0),
        "t_list": t_list, "solver": "rk4", "dt": dt, "atol": 0.0, "rtol": 0.0,
    }


def exciton_chain7(max_depth=4, seed=7, t_final=1000.0, out_every=10.0, rtol=1e-8,
                   atol=None, dt0=0.1, n_pade=2, lam_cm=35.0, tau_fs=50.0, T_K=300.0, N=7):
    """7-site exciton chain, per-site Drude bath (1 Drude + 2 Pade), DoPri.

    Site energies ~ N(0, 100 cm^-1), couplings ~ U(-100, 100) cm^-1 from
    numpy default_rng(seed); V_u = |u><u|.  Energies are converted to rad/fs.
    """
    rng = np.random.default_rng(seed)
    e = rng.normal(0.0, 100.0, N)
    J = rng.uniform(-100.0, 100.0, (N, N))
    J = np.triu(J, 1)
    Hcm = np.diag(e) + J + J.T
    H = (Hcm / bath.HBAR_CM_FS).astype(np.complex128)
    lam = lam_cm / bath.HBAR_CM_FS
    gam = 1.0 / tau_fs
    beta = bath.HBAR_CM_FS / (bath.KB_CM_K * T_K)  # fs
    nu, c, delta = bath.drude_modes(lam, gam, beta, n_pade, "pade")
    noises = []
    for u in range(N):
        V = np.zeros((N, N), np.complex128)
        V[u, u] = 1.0
        noises.append(bath.noise_params(V, nu, c, delta))
    K = N * len(nu)
    n_t = int(round(t_final / out_every)) + 1
    t_list = out_every * np.arange(n_t)
    return {
        "name": "7-site chain N=%d K=%d L=%d" % (N, K, max_depth),
        "N": N, "H": H, "noises": noises, "max_depth": max_depth,
        "rho0": _rho0_projector(N, hrchy_count(K, max_depth), 0),
        "t_list": t_list, "solver": "rkdp", "dt": dt0,
        "atol": rtol * 1e-3 if atol is None else atol, "rtol": rtol,
    }


def batched_sweep(n_inst=512, seed=512, dt=0.005, n_steps=4000, out_every=100):
    """512 independent spin-boson / multi-level instances of uneven N, K and L."""
    rng = np.random.default_rng(seed)
    models = []
    for _ in range(n_inst):
        N = int(rng.choice([2, 4, 8]))
        X = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        H = 0.5 * (X + X.conj().T)
        signs = np.where(rng.random(N) < 0.5, 1.0, -1.0)
        signs[0], signs[-1] = 1.0, -1.0
        V = np.diag(signs).astype(np.complex128)
        lam = rng.uniform(0.05, 2.0)
        gam = rng.uniform(0.2, 5.0)
        T = rng.uniform(0.1, 5.0)
        n_pade = int(rng.integers(1, 4))
        L = int(rng.integers(4, 13))
        nu, c, delta = bath.drude_modes(lam, gam, 1.0 / T, n_pade, "pade")
        K = len(nu)
        n_t = n_steps // out_every + 1
        models.append({
            "name": "inst N=%d K=%d L=%d" % (N, K, L),
            "N": N, "H": H, "noises": [bath.noise_params(V, nu, c, delta)], "max_depth": L,
            "rho0": _rho0_projector(N, hrchy_count(K, L), 0),
            "t_list": dt * out_every * np.arange(n_t), "solver": "rk4", "dt": dt,
            "atol": 0.0, "rtol": 0.0,
        })
    return models


def dense64(N=64, max_depth=3, seed=64, t_final=20.0, out_every=1.0, rtol=1e-6, atol=None,
            dt0=0.01, n_baths=4, n_pade=3, lam=0.2, gam=1.0, T=0.5):
    """N=64 GOE Hamiltonian (spectral width 2), 4 random diagonal projector baths."""
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(N, N))
    H = 0.5 * (X + X.T)
    w = np.linalg.eigvalsh(H)
    H = (2.0 * (H - w.mean() * np.eye(N)) / (w[-1] - w[0])).astype(np.complex128)
    nu, c, delta = bath.drude_modes(lam, gam, 1.0 / T, n_pade, "pade")
    noises = []
    for _ in range(n_baths):
        d = (rng.random(N) < 0.5).astype(np.float64)
        noises.append(bath.noise_params(np.diag(d).astype(np.complex128), nu, c, delta))
    K = n_baths * len(nu)
    n_t = int(round(t_final / out_every)) + 1
    return {
        "name": "dense N=%d K=%d L=%d" % (N, K, max_depth),
        "N": N, "H": H, "noises": noises, "max_depth": max_depth,
        "rho0": _rho0_projector(N, hrchy_count(K, max_depth), 0),
        "t_list": out_every * np.arange(n_t), "solver": "rkdp", "dt": dt0,
        "atol": rtol * 1e-3 if atol is None else atol, "rtol": rtol,
    }


def chain4(seed, max_depth=3, N=4, t_final=10.0, dt=0.01, out_every=10):
    """4-site chain for the learned truncation closure (one Drude+Pade bath per site)."""
    rng = np.random.default_rng(seed)
    e = rng.normal(0.0, 0.5, N)
    Jc = rng.uniform(-1.0, 1.0, N - 1)
    H = np.diag(e) + np.diag(Jc, 1) + np.diag(Jc, -1)
    noises = []
    lam = rng.uniform(0.1, 1.0)
    T = rng.uniform(0.5, 2.0)
    nu, c, delta = bath.drude_modes(lam, 1.0, 1.0 / T, 1, "pade")
    for u in range(N):
        V = np.zeros((N, N), np.complex128)
        V[u, u] = 1.0
        noises.append(bath.noise_params(V, nu, c, delta))
    K = N * len(nu)
    n_steps = int(round(t_final / dt))
    n_t = n_steps // out_every + 1
    return {
        "name": "chain4 seed=%d L=%d" % (seed, max_depth),
        "N": N, "H": H.astype(np.complex128), "noises": noises, "max_depth": max_depth,
        "rho0": _rho0_projector(N, hrchy_count(K, max_depth), 0),
        "t_list": dt * out_every * np.arange(n_t), "solver": "rk4", "dt": dt,
        "atol": 0.0, "rtol": 0.0, "lam": lam, "T": T,
    }


def random_model(N, n_noise, len_gamma, max_depth, seed, dense_v=True, offdiag_gamma=False,
                 complex_params=True):
    """Random parameter set for parity tests (dense or diagonal V, general gamma)."""
    rng = np.random.default_rng(seed)

    def cm(*shape):
        return rng.normal(size=shape) + (1j * rng.normal(size=shape) if complex_params else 0)

    X = cm(N, N)
    H = 0.5 * (X + X.conj().T)
    noises = []
    for _ in range(n_noise):
        if dense_v:
            Y = cm(N, N)
            V = 0.5 * (Y + Y.conj().T)
        else:
            V = np.diag(rng.normal(size=N)).astype(np.complex128)
        g = np.diag(rng.uniform(0.5, 2.0, len_gamma)).astype(np.complex128)
        if offdiag_gamma:
            g = g + 0.1 * cm(len_gamma, len_gamma) * (1 - np.eye(len_gamma))
        noises.append({
            "V": V.astype(np.complex128), "gamma": g,
            "phi_0": cm(len_gamma), "sigma": cm(len_gamma),
            "S": cm(len_gamma, len_gamma), "A": cm(len_gamma, len_gamma),
            "s_delta": complex(cm(1)[0]) * 0.1,
        })
    K = n_noise * len_gamma
    nh = hrchy_count(K, max_depth)
    rho = cm(nh, N, N) * 0.1
    return {"name": "random N=%d noises=%d len=%d L=%d" % (N, n_noise, len_gamma, max_depth),
            "N": N, "H": H.astype(np.complex128), "noises": noises, "max_depth": max_depth,
            "rho0": rho.reshape(-1).astype(np.complex128)}
