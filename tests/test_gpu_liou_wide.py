"""heom_liou / heom_ado beyond n_level 16 and in col_major storage, against the
CPU oracle (bar 1e-10, as every FP64 form).

The reference templates take any n_level (heom_liou.h:18-27, heom_ado.h:62-189)
and an `order`: its kron_x_1 / kron_1_x_T (linalg_engine_nil.h.j2:199-294)
place the superoperator entries so that with col_major the Liouvillian acts on
the column-major vector of rho_n -- the same operator on transposed storage.
The col_major results, transposed back per ADO, must therefore equal the
oracle's row-major ones; a mistake in the index map of either layout (a
missing transpose, a sign of the commutator) fails that comparison.
"""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10
FORMS = {"liou": (H.heom_liou, O.LIOU), "ado": (H.heom_ado, O.ADO)}
ENG = None


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def pair(form, m, order=H.row_major, sqrt_norm=False):
    cls, okind = FORMS[form]
    g = cls(m["max_depth"], order=order, sqrt_normalization=sqrt_norm).load(m, eng())
    o = O.Oracle(okind, m["N"], m["max_depth"], n_threads=8, sqrt_norm=sqrt_norm).load(m)
    return g, o


@pytest.mark.parametrize("form", ["liou", "ado"])
@pytest.mark.parametrize("N", [17, 24, 32])
@pytest.mark.parametrize("dense_v", [True, False])
def test_wide_calc_diff_parity(form, N, dense_v):
    print("wide calc_diff", form, N, dense_v, flush=True)
    m = configs.random_model(N, 2, 1, 2, seed=300 + N, dense_v=dense_v, offdiag_gamma=True)
    g, o = pair(form, m)
    assert g.n_hrchy == o.n_hrchy
    rng = np.random.default_rng(N)
    rho = m["rho0"]
    d0 = rng.normal(size=rho.size) + 1j * rng.normal(size=rho.size)
    for alpha, beta in [(1.0, 0.0), (0.3 - 0.2j, 0.7 + 0.1j)]:
        ref = o.calc_diff(rho, alpha, beta, d0)
        got = g.calc_diff_impl(None, d0.copy(), rho, alpha, beta)
        err = np.abs(got - ref).max()
        assert err <= TOL * max(1.0, np.abs(ref).max()), (err, alpha, beta)


@pytest.mark.parametrize("form", ["liou", "ado"])
def test_wide_rkdp_solve(form):
    """Dormand-Prince at n_level 20 with dense V (every stage kernel variant):
    the oracle's accepted / rejected counts and its state at 1e-10."""
    print("wide rkdp", form, flush=True)
    m = configs.random_model(20, 2, 1, 2, seed=7, dense_v=True, complex_params=False)
    m["H"] = m["H"] * 0.2
    m["t_list"] = np.linspace(0.0, 0.5, 6)
    kw = {"dt": 0.01, "atol": 1e-10, "rtol": 1e-8}
    g, o = pair(form, m)
    s = H.qme_solver(eng(), g, H.rkdp())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"], kw["atol"],
                                      kw["rtol"])
    st = s.stats()
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    scale = max(1.0, np.abs(rho_o).max())
    assert np.abs(rs - rs_o).max() <= TOL * scale
    assert np.abs(rho - rho_o).max() <= TOL * scale


@pytest.mark.parametrize("form", ["liou", "ado"])
@pytest.mark.parametrize("N,dense_v,sq", [(3, True, False), (4, False, True), (20, True, False)])
def test_col_major_is_transposed_storage(form, N, dense_v, sq):
    print("col_major", form, N, dense_v, sq, flush=True)
    m = configs.random_model(N, 2, 2, 2, seed=40 + N, dense_v=dense_v, offdiag_gamma=True)
    _, o = pair(form, m, sqrt_norm=sq)
    g_c, _ = pair(form, m, order=H.col_major, sqrt_norm=sq)
    rho = m["rho0"].reshape(-1, N, N)
    rho_c = np.ascontiguousarray(rho.transpose(0, 2, 1)).reshape(-1)
    ref = o.calc_diff(m["rho0"], 1.0, 0.0).reshape(-1, N, N)
    got = g_c.calc_diff_impl(None, np.zeros_like(rho_c), rho_c, 1.0, 0.0).reshape(-1, N, N)
    assert np.abs(got.transpose(0, 2, 1) - ref).max() <= TOL * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("form", ["liou", "ado"])
def test_wide_complex64_calc_diff(form):
    """complex64 state and operator (FP32 arithmetic): error against the FP64
    oracle stated, bar 1e-5 relative as for the other complex64 forms."""
    cls, okind = FORMS[form]
    m = configs.random_model(24, 2, 1, 2, seed=324, dense_v=True, offdiag_gamma=True)
    g = cls(m["max_depth"], dtype="complex64").load(m, eng())
    o = O.Oracle(okind, m["N"], m["max_depth"], n_threads=8).load(m)
    ref = o.calc_diff(m["rho0"], 1.0, 0.0)
    got = g.calc_diff_impl(None, np.zeros_like(ref), m["rho0"], 1.0, 0.0)
    err = np.abs(got - ref).max() / max(1.0, np.abs(ref).max())
    print("wide complex64", form, "relative error vs FP64 oracle %.2e" % err)
    assert err <= 1e-5


def test_liou_beyond_n_level_32_is_refused():
    m = configs.random_model(33, 1, 1, 1, seed=1, dense_v=False)
    with pytest.raises(H.HeomError) as e:
        H.heom_liou(1).load(m, eng())
    assert "n_level > 32" in str(e.value)
