"""Redfield QMEs on the GPU (redfield_hilb.h, redfield_liou.h) against the oracle.

Lambda_u from the library's set-up (its own Hermitian eigensolver and LU
solves) must equal the oracle's to 1e-12; calc_diff of both forms (Hilbert
form: one workgroup with every table in LDS when n_level <= 16 and they fit,
for diagonal and dense V, else two passes through a scratch block; Liouville
form: the wavefront-per-row superoperator product) and RK4 / rkdp /
lsrk4 solves must meet the FP64 bar of 1e-10 (relative to the magnitude of
the result where it exceeds 1).  complex64 states its error.  col_major
storage is checked as the transpose of the row-major result.
"""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10
ENG = None
FORMS = {"hilb": (H.redfield_hilb, O.REDFIELD_HILB), "liou": (H.redfield_liou, O.REDFIELD_LIOU)}


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def pair(form, m, dtype="complex128", order=H.row_major):
    cls, okind = FORMS[form]
    g = cls(dtype=dtype, order=order).load(m, eng())
    o = O.Oracle(okind, m["N"], 0, n_threads=8).load(m)
    return g, o


def rel(got, ref):
    return np.abs(got - ref).max() / max(1.0, np.abs(ref).max())


def physical_model(N, dense_v, seed):
    """Random H (spectral radius 1) and two Drude baths (lambda 0.2, gamma 1, T 1,
    one Pade mode) coupled through random Hermitian (or diagonal) V of norm 1;
    random complex bath parameters can make the Redfield generator grow without
    bound, a physical bath does not."""
    from libheom_amd import bath
    rng = np.random.default_rng(seed)

    def herm():
        X = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        X = 0.5 * (X + X.conj().T)
        return X / np.abs(np.linalg.eigvalsh(X)).max()

    nu, c, delta = bath.drude_modes(0.2, 1.0, 1.0, 1, "pade")
    noises = [bath.noise_params(herm() if dense_v else
                                np.diag(rng.normal(size=N)).astype(np.complex128), nu, c, delta)
              for _ in range(2)]
    return {"N": N, "H": herm(), "noises": noises, "max_depth": 0}


def spin_boson_redfield(**kw):
    m = configs.spin_boson(max_depth=0, **kw)
    m["rho0"] = m["rho0"][:4].copy()
    return m


@pytest.mark.parametrize("N,n_noise,dense_v", [(2, 1, True), (5, 3, True), (8, 8, False),
                                               (16, 2, True), (24, 3, False), (32, 2, True)])
def test_lambda_matches_oracle(N, n_noise, dense_v):
    m = configs.random_model(N, n_noise, 2, 0, seed=N, dense_v=dense_v)
    g, o = pair("hilb", m)
    for u in range(n_noise):
        ref, w_o = o.redfield_lambda(u)
        assert rel(g.Lambda(u), ref) <= 1e-12
    assert np.abs(np.sort(g.eigenvalues()) - np.sort(w_o)).max() <= 1e-12 * max(
        1.0, np.abs(w_o).max())


@pytest.mark.parametrize("form", ["hilb", "liou"])
@pytest.mark.parametrize("N,n_noise,dense_v", [(1, 1, True), (2, 1, True), (3, 2, True),
                                               (5, 3, False), (8, 4, True), (16, 3, True),
                                               (16, 16, False), (17, 2, True), (24, 4, False),
                                               (32, 3, True)])
def test_calc_diff_parity(form, N, n_noise, dense_v):
    m = configs.random_model(N, n_noise, 2, 0, seed=100 + N, dense_v=dense_v)
    g, o = pair(form, m)
    rng = np.random.default_rng(N)
    rho = rng.normal(size=N * N) + 1j * rng.normal(size=N * N)
    d0 = rng.normal(size=N * N) + 1j * rng.normal(size=N * N)
    for alpha, beta in [(1.0, 0.0), (0.5 - 0.25j, 0.3 + 0.1j)]:
        ref = o.calc_diff(rho, alpha, beta, drho=d0)
        got = g.calc_diff_impl(None, d0.copy(), rho, alpha, beta)
        assert rel(got, ref) <= TOL, (alpha, beta)


@pytest.mark.parametrize("form", ["hilb", "liou"])
def test_complex64_error(form):
    m = configs.random_model(8, 3, 2, 0, seed=7, dense_v=True)
    g, o = pair(form, m, dtype="complex64")
    rho = m["rho0"]
    err = rel(g.calc_diff_impl(None, np.zeros_like(rho), rho, 1.0, 0.0), o.calc_diff(rho))
    print("complex64 redfield_%s calc_diff relative error: %.3e" % (form, err))
    assert err <= 1e-5


@pytest.mark.parametrize("form", ["hilb", "liou"])
@pytest.mark.parametrize("N", [3, 20])
def test_col_major_is_transposed_storage(form, N):
    m = configs.random_model(N, 2, 2, 0, seed=300 + N, dense_v=True)
    g_r, _ = pair(form, m)
    g_c, _ = pair(form, m, order=H.col_major)
    rng = np.random.default_rng(3)
    rho = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    d_r = g_r.calc_diff_impl(None, np.zeros(N * N, complex), rho.reshape(-1), 1.0, 0.0)
    d_c = g_c.calc_diff_impl(None, np.zeros(N * N, complex), rho.T.copy().reshape(-1), 1.0, 0.0)
    assert rel(d_c.reshape(N, N).T, d_r.reshape(N, N)) <= TOL


@pytest.mark.parametrize("form", ["hilb", "liou"])
@pytest.mark.parametrize("solver", ["rk4", "lsrk4"])
def test_fixed_step_solve_spin_boson(form, solver):
    """Spin-boson Redfield (Drude bath, 2 modes), dt=0.01 x 2000 with rho_S every 10 steps."""
    m = spin_boson_redfield()
    g, o = pair(form, m)
    s = H.qme_solver(eng(), g, getattr(H, solver)())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": m["dt"]})
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], getattr(O, solver.upper()), m["dt"])
    assert s.stats()["n_rhs"] == st_o.n_rhs
    assert np.abs(rs - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL


@pytest.mark.parametrize("form", ["hilb", "liou"])
def test_rkdp_chain7_same_steps(form):
    """7-site chain, 7 Drude baths, rkdp rtol 1e-8 over 200 fs: same accepted and
    rejected steps as the oracle, rho every 10 fs within 1e-10."""
    m = configs.exciton_chain7(max_depth=0, t_final=200.0)
    m["rho0"] = m["rho0"][:49].copy()
    g, o = pair(form, m)
    kw = {"dt": m["dt"], "atol": m["atol"], "rtol": m["rtol"]}
    s = H.qme_solver(eng(), g, H.rkdp())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"], kw["atol"],
                                      kw["rtol"])
    st = s.stats()
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    assert np.abs(rs - rs_o).max() <= TOL
    # a density matrix: unit trace, Hermitian
    assert abs(np.trace(rho.reshape(7, 7)) - 1.0) <= 1e-10


@pytest.mark.parametrize("N,dense_v", [(4, True), (32, True), (32, False)])
def test_rk4_solve_each_hilbert_kernel(N, dense_v):
    """RK4 dt=0.005 to t=1 through the Hilbert-form paths not covered by the
    spin-boson and chain solves (one workgroup, diagonal V): one workgroup with
    dense V, two passes with dense and with diagonal V."""
    m = physical_model(N, dense_v, seed=N)
    t = np.linspace(0.0, 1.0, 5)
    g, o = pair("hilb", m)
    X = np.random.default_rng(1).normal(size=(N, N))
    rho0 = (X @ X.T).astype(np.complex128)
    rho0 = (rho0 / np.trace(rho0)).reshape(-1)
    s = H.qme_solver(eng(), g, H.rk4())
    _, rs = s.solve_rho_s(rho0, t, {"dt": 0.005})
    _, rs_o, _, _, _ = o.solve(rho0, t, O.RK4, 0.005)
    assert rel(rs, rs_o.reshape(rs.shape)) <= TOL


def test_graphs_bitwise():
    m = spin_boson_redfield()
    m["t_list"] = np.linspace(0.0, 3.0, 8)
    g, _ = pair("hilb", m)
    out = []
    for graphs in (False, True):
        s = H.qme_solver(eng(), g, H.rk4())
        s.set_graphs(graphs)
        out.append(s.solve_rho_s(m["rho0"], m["t_list"], {"dt": 0.03}))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_redfield_handles_are_not_batch_members():
    m = spin_boson_redfield()
    g, _ = pair("hilb", m)
    with pytest.raises(H.HeomError):
        H.heom_batch([g])
