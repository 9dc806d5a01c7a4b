"""Matrix-core commutator of heom_hilb (16 < n_level <= 64, diagonal V; blocks
other than 32 and 64 are zero-padded to the next tile of 32 or 64 rows).

FP64 (f64 16x16x4 MFMA) must meet the FP64 bar against the oracle (1e-10,
relative to the magnitude of the result); the FP32 (f32 MFMA) and split-bf16x3
(bf16 16x16x32 MFMA, hi*hi + hi*lo + lo*hi) variants state their error.  The
vector-ALU kernels (matrix_cores = 0) are checked against the same oracle so
that both paths stay covered.
"""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10
ENG = None


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def model(N, real_h, complex_diag, L=2, seed=0):
    m = configs.random_model(N, 2, 2, L, seed=seed + N, dense_v=False)
    rng = np.random.default_rng(seed + 7 * N)
    if real_h:
        X = rng.normal(size=(N, N))
        m["H"] = (0.5 * (X + X.T)).astype(np.complex128)
    if complex_diag:
        for z in m["noises"]:
            z["V"] = np.diag(rng.normal(size=N) + 1j * rng.normal(size=N))
    return m


def rel_err(got, ref):
    return np.abs(got - ref).max() / max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("N", [17, 24, 32, 33, 48, 63, 64])
@pytest.mark.parametrize("real_h", [True, False])
@pytest.mark.parametrize("complex_diag", [False, True])
@pytest.mark.parametrize("matrix_cores", [True, False])
def test_calc_diff_fp64(N, real_h, complex_diag, matrix_cores):
    print("mfma fp64", N, real_h, complex_diag, matrix_cores, flush=True)
    m = model(N, real_h, complex_diag)
    g = H.heom_hilb(m["max_depth"], matrix_cores=matrix_cores).load(m, eng())
    o = O.Oracle(O.HILB, N, m["max_depth"], n_threads=8).load(m)
    rng = np.random.default_rng(N)
    d0 = rng.normal(size=g.main_size()) + 1j * rng.normal(size=g.main_size())
    for alpha, beta in [(1.0, 0.0), (0.5 - 0.25j, 0.3 + 0.2j)]:
        ref = o.calc_diff(m["rho0"], alpha, beta, d0)
        got = g.calc_diff_impl(None, d0.copy(), m["rho0"], alpha, beta)
        assert rel_err(got, ref) <= TOL


def test_calc_diff_dense64_full_config():
    """N=64, 16 modes, depth 3: all 969 ADOs, real H, projector baths."""
    m = configs.dense64()
    g = H.heom_hilb(m["max_depth"]).load(m, eng())
    assert g.n_hrchy == 969
    o = O.Oracle(O.HILB, 64, m["max_depth"], n_threads=16).load(m)
    rng = np.random.default_rng(64)
    rho = (rng.normal(size=g.main_size()) + 1j * rng.normal(size=g.main_size())) * 1e-2
    ref = o.calc_diff(rho, 1.0, 0.0)
    got = g.calc_diff_impl(None, np.zeros_like(ref), rho, 1.0, 0.0)
    print("dense64 fp64 mfma relative error %.3e" % rel_err(got, ref))
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("N", [20, 32, 48, 64])
@pytest.mark.parametrize("real_h", [True, False])
def test_calc_diff_low_precision_variants(N, real_h):
    m = model(N, real_h, False)
    o = O.Oracle(O.HILB, N, m["max_depth"], n_threads=8).load(m)
    ref = o.calc_diff(m["rho0"], 1.0, 0.0)
    errs = {}
    for dt in ("complex64", "complex64_bf16x3"):
        g = H.heom_hilb(m["max_depth"], dtype=dt).load(m, eng())
        got = g.calc_diff_impl(None, np.zeros_like(ref), m["rho0"], 1.0, 0.0)
        errs[dt] = rel_err(got, ref)
    print("N=%d real_h=%s relative error: fp32 %.3e, bf16x3 %.3e"
          % (N, real_h, errs["complex64"], errs["complex64_bf16x3"]))
    assert errs["complex64"] <= 1e-5
    assert errs["complex64_bf16x3"] <= 1e-3


def test_rkdp_dense64_reduced_depth_parity():
    """N=64, depth 1 (17 ADOs), Dormand-Prince rtol=1e-6 to t=1: same steps, 1e-10."""
    m = configs.dense64(max_depth=1, t_final=1.0, out_every=0.5)
    kw = {"dt": m["dt"], "atol": m["atol"], "rtol": m["rtol"]}
    g = H.heom_hilb(1).load(m, eng())
    s = H.qme_solver(eng(), g, H.rkdp())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    o = O.Oracle(O.HILB, 64, 1, n_threads=16).load(m)
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"], kw["atol"],
                                      kw["rtol"])
    st = s.stats()
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    assert np.abs(rs - rs_o).max() <= TOL
    assert rel_err(rho, rho_o) <= TOL


def test_rkdp_dense64_precision_variants_error():
    """rho_S error of the FP32 and bf16x3 matrix-core variants against FP64."""
    m = configs.dense64(max_depth=2, t_final=2.0, out_every=0.5)
    kw = {"dt": m["dt"], "atol": m["atol"], "rtol": m["rtol"]}
    out = {}
    for dt in ("complex128", "complex64", "complex64_bf16x3"):
        g = H.heom_hilb(2, dtype=dt).load(m, eng())
        s = H.qme_solver(eng(), g, H.rkdp())
        _, out[dt] = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    e32 = np.abs(out["complex64"] - out["complex128"]).max()
    e16 = np.abs(out["complex64_bf16x3"] - out["complex128"]).max()
    print("dense64 L=2 rho_S error vs FP64: fp32 %.3e, bf16x3 %.3e" % (e32, e16))
    assert e32 <= 1e-5
    assert e16 <= 1e-3


def test_rkdp_padded_block_parity():
    """N=48 (padded to the 64-row tile), depth 1, Dormand-Prince to t=1: the
    oracle's accepted / rejected steps and rho_S within 1e-10."""
    m = configs.dense64(N=48, max_depth=1, t_final=1.0, out_every=0.5)
    kw = {"dt": m["dt"], "atol": m["atol"], "rtol": m["rtol"]}
    g = H.heom_hilb(1).load(m, eng())
    s = H.qme_solver(eng(), g, H.rkdp())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    o = O.Oracle(O.HILB, 48, 1, n_threads=16).load(m)
    rho_o, rs_o, st_o, _, _ = o.solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"], kw["atol"],
                                      kw["rtol"])
    st = s.stats()
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    assert np.abs(rs - rs_o).max() <= TOL
    assert rel_err(rho, rho_o) <= TOL


def test_bf16x3_outside_matrix_core_path_is_unsupported():
    m = model(32, True, False)
    with pytest.raises(H.HeomError) as e:
        H.heom_hilb(2, dtype="complex64_bf16x3", matrix_cores=False).load(m, eng())
    assert e.value.code == 5
    m = configs.spin_boson(max_depth=2)
    with pytest.raises(H.HeomError) as e:
        H.heom_hilb(2, dtype="complex64_bf16x3").load(m, eng())
    assert e.value.code == 5
