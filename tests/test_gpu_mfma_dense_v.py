"""heom_hilb with dense (non-diagonal) coupling operators V for 16 < n_level <= 64
on matrix cores (heom_hilb.h:108-209: i[H,rho], delta[V,[V,rho]], i[V,Phi] +
{V,Psi} as V X + Y V with X = i Phi + Psi + delta [V,rho],
Y = Psi - i Phi - delta [V,rho]), blocks zero-padded to tiles of 32 or 64.

FP64 (f64 16x16x4 MFMA) must meet the FP64 bar against the oracle (1e-10,
relative to the magnitude of the result), for complex and real H and V, for
every block size class (17..32 -> tile 32, 33..64 -> tile 64), and general
alpha / beta; a Dormand-Prince solve must take the oracle's steps.  The FP32
operator (f32 MFMA) states its error.
"""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def rel_err(got, ref):
    return np.abs(got - ref).max() / max(1.0, np.abs(ref).max())


def model(N, complex_params, n_noise=2, L=2, seed=0):
    return configs.random_model(N, n_noise, 2, L, seed=seed + N, dense_v=True,
                                complex_params=complex_params)


@pytest.mark.parametrize("N", [17, 24, 32, 33, 48, 64])
@pytest.mark.parametrize("complex_params", [True, False])
def test_calc_diff_fp64(N, complex_params):
    print("dense V fp64", N, complex_params, flush=True)
    m = model(N, complex_params)
    g = H.heom_hilb(m["max_depth"]).load(m, H.hip(0))
    o = O.Oracle(O.HILB, N, m["max_depth"], n_threads=8).load(m)
    rng = np.random.default_rng(N)
    d0 = rng.normal(size=g.main_size()) + 1j * rng.normal(size=g.main_size())
    for alpha, beta in [(1.0, 0.0), (0.5 - 0.25j, 0.3 + 0.2j)]:
        ref = o.calc_diff(m["rho0"], alpha, beta, d0)
        got = g.calc_diff_impl(None, d0.copy(), m["rho0"], alpha, beta)
        print("relative error %.3e" % rel_err(got, ref))
        assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("N", [20, 64])
def test_calc_diff_off_diagonal_gamma(N):
    """parity_q1 = False keeps the off-diagonal gamma links (heom_hilb.h:128-147)."""
    m = configs.random_model(N, 1, 2, 2, seed=N, dense_v=True, offdiag_gamma=True)
    g = H.heom_hilb(2, parity_q1=False).load(m, H.hip(0))
    o = O.Oracle(O.HILB, N, 2, parity_q1=False).load(m)
    ref = o.calc_diff(m["rho0"], 1.0, 0.0)
    got = g.calc_diff_impl(None, np.zeros_like(ref), m["rho0"], 1.0, 0.0)
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("N", [32, 48])
def test_calc_diff_complex64_error(N):
    m = model(N, True)
    g = H.heom_hilb(m["max_depth"], dtype="complex64").load(m, H.hip(0))
    o = O.Oracle(O.HILB, N, m["max_depth"]).load(m)
    ref = o.calc_diff(m["rho0"], 1.0, 0.0)
    got = g.calc_diff_impl(None, np.zeros_like(ref), m["rho0"], 1.0, 0.0)
    err = rel_err(got, ref)
    print("dense V complex64 N=%d relative error %.3e" % (N, err))
    assert err <= 1e-5


@pytest.mark.parametrize("N", [24, 64])
def test_rkdp_matches_oracle(N):
    m = model(N, True, n_noise=1, L=1, seed=5)
    m["rho0"] = np.zeros_like(m["rho0"])
    m["rho0"][0] = 1.0
    t = np.linspace(0.0, 0.4, 3)
    kw = {"dt": 0.01, "atol": 1e-10, "rtol": 1e-8}
    g = H.heom_hilb(m["max_depth"]).load(m, H.hip(0))
    s = H.qme_solver(H.hip(0), g, H.rkdp())
    rho, rs = s.solve_rho_s(m["rho0"], t, kw)
    st = s.stats()
    o = O.Oracle(O.HILB, N, m["max_depth"], n_threads=8).load(m)
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], t, O.RKDP, kw["dt"], kw["atol"], kw["rtol"])
    print("N=%d |rho_S| max %.3e, rho_S error %.3e, state relative error %.3e"
          % (N, np.abs(rs_o).max(), np.abs(rs - rs_o).max(), rel_err(rho, rho_o)))
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    # the random (non-Hermitian) parameters let the state grow: bars relative to it
    assert rel_err(rs, rs_o) <= TOL
    assert rel_err(rho, rho_o) <= TOL


def test_dense_v_outside_the_matrix_cores_is_refused():
    m = configs.random_model(17, 1, 1, 1, seed=2, dense_v=True)
    g = H.heom_hilb(1, matrix_cores=False).load(m, H.hip(0))
    with pytest.raises(H.HeomError) as e:
        g.calc_diff_impl(None, np.zeros(g.main_size(), np.complex128), m["rho0"], 1.0, 0.0)
    assert e.value.code == 5
    with pytest.raises(H.HeomError) as e:
        H.heom_hilb(1, dtype="complex64_bf16x3").load(m, H.hip(0))
    assert e.value.code == 5
