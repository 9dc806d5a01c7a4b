"""Batches of independent Hilbert-form hierarchies (parameter sweeps).

The batch packs its members into one state; every member must come out as if
it had been propagated alone.  Oracle: the CPU restatement run member by
member.  Bar: FP64 <= 1e-10 max-abs.

DoPri on a batch takes one step size for the whole batch (the RMS error norm of
the concatenated state), so member-by-member parity holds exactly for a batch
of one member and for k identical copies (the RMS of k copies equals the RMS of
one); a batch of different members is checked against each member's own
adaptive solution at the solver tolerance.
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


def member(m, dtype="complex128"):
    return H.heom_hilb(m["max_depth"], dtype=dtype).load(m, eng())


def close(got, ref):
    """FP64 bar, relative to the magnitude of the reference state when above 1."""
    return np.abs(got - ref).max() <= TOL * max(1.0, np.abs(ref).max())


def oracle(m):
    return O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=8).load(m)


def diag_models(seed, Ns, complex_diag=False):
    rng = np.random.default_rng(seed)
    out = []
    for k, N in enumerate(Ns):
        n_noise = int(rng.integers(1, 3))
        ln = int(rng.integers(1, 3))
        L = int(rng.integers(1, 5))
        m = configs.random_model(N, n_noise, ln, L, seed=seed * 100 + k, dense_v=False)
        if complex_diag and k % 2:
            for z in m["noises"]:
                z["V"] = np.diag(rng.normal(size=N) + 1j * rng.normal(size=N))
        out.append(m)
    return out


@pytest.mark.parametrize("complex_diag", [False, True])
def test_batch_calc_diff_mixed_n(complex_diag):
    ms = diag_models(11, [1, 2, 3, 4, 5, 6, 7, 8, 2, 8, 3], complex_diag)
    b = H.heom_batch([member(m) for m in ms])
    assert b.main_size() == sum(m["rho0"].size for m in ms)
    rho = b.pack([m["rho0"] for m in ms])
    d0 = b.pack([np.full(m["rho0"].size, 0.3 - 0.2j) for m in ms])
    alpha, beta = 0.7 - 0.1j, 0.25 + 0.5j
    got = b.calc_diff_impl(None, d0.copy(), rho, alpha, beta)
    for k, (m, g) in enumerate(zip(ms, b.unpack(got))):
        ref = oracle(m).calc_diff(m["rho0"], alpha, beta,
                                  drho=np.full(m["rho0"].size, 0.3 - 0.2j))
        err = np.abs(g - ref).max()
        assert err <= TOL * max(1.0, np.abs(ref).max()), (k, m["name"], err)


def test_batch_calc_diff_full_sweep():
    """All 512 instances of the sweep (N in {2,4,8}, K 2..4, L 4..12) in one launch."""
    ms = configs.batched_sweep()
    b = H.heom_batch([member(m) for m in ms])
    rng = np.random.default_rng(5)
    rho = (rng.normal(size=b.main_size()) + 1j * rng.normal(size=b.main_size())) * 1e-2
    got = b.calc_diff_impl(None, np.zeros_like(rho), rho, 1.0, 0.0)
    worst = 0.0
    for m, r, g in zip(ms, b.unpack(rho), b.unpack(got)):
        ref = oracle(m).calc_diff(r, 1.0, 0.0)
        worst = max(worst, np.abs(g - ref).max() / max(1.0, np.abs(ref).max()))
    print("512-instance calc_diff worst relative error: %.3e" % worst)
    assert worst <= TOL


def test_batch_tile_shapes_agree_and_match_the_oracle():
    """The autotuner times every tile shape (one or two work items per wavefront
    step, one, two or four wavefronts per workgroup) over several grids and
    keeps the fastest; every shape and grid must give the same derivative to
    rounding (the same expression per element; the compiler may contract a
    product and a sum into one fma differently in the kernel variants), equal
    to the oracle's member by member."""
    ms = diag_models(23, [2, 4, 8, 3, 8, 2, 6, 4, 5])
    b = H.heom_batch([member(m) for m in ms])
    chosen, tried = b.tuning()
    print("autotuner:", chosen, tried, flush=True)
    assert {t[0] for t in tried} == {1, 2} and {t[1] for t in tried} == {1, 2, 4}
    assert all(t[3] > 0 for t in tried)
    # the fastest shape, or one item per one-wavefront workgroup when nothing is
    # 3% faster than that
    best = min(t[3] for t in tried)
    plain = [t for t in tried if t[:2] == (1, 1) and t[2] == max(u[2] for u in tried
                                                                 if u[:2] == (1, 1))][0]
    shape = (chosen["items_per_step"], chosen["waves_per_group"], chosen["grid"])
    if best > 0.97 * plain[3]:
        assert shape == plain[:3]
    else:
        assert shape in [t[:3] for t in tried if t[3] == best]
    rng = np.random.default_rng(8)
    rho = rng.normal(size=b.main_size()) + 1j * rng.normal(size=b.main_size())
    outs = []
    for shape in [(1, 1, 1), (1, 1, 7), (2, 1, 1), (2, 1, 5), (1, 2, 3), (2, 4, 1), (1, 4, 2),
                  (2, 2, 4096), (1, 1, 4096)]:
        b.set_tuning(*shape)
        outs.append(b.calc_diff_impl(None, np.zeros_like(rho), rho, 1.0, 0.0))
    scale = np.abs(outs[0]).max()
    for o in outs[1:]:
        assert np.abs(o - outs[0]).max() <= 1e-14 * scale
    for o in outs:
        for m, r, g in zip(ms, b.unpack(rho), b.unpack(o)):
            ref = oracle(m).calc_diff(r, 1.0, 0.0)
            assert close(g, ref), m["name"]
    with pytest.raises(H.HeomError):
        b.set_tuning(3, 1, 16)
    with pytest.raises(H.HeomError):
        b.set_tuning(1, 3, 16)


def test_batch_rkdp_with_several_wavefronts_per_workgroup():
    """Dormand-Prince on 7 copies of the well-conditioned spin-boson of
    test_batch_rkdp_identical_copies_match_single_solve, in every tile shape:
    the error norm's partial sums are one per wavefront; the accepted and
    rejected attempts must be the oracle's and rho_S within the FP64 bar."""
    m = configs.spin_boson(max_depth=6, T=0.5)
    m["t_list"] = np.linspace(0.0, 5.0, 11)
    kw = {"dt": 0.8, "atol": 1e-10, "rtol": 1e-8}
    _, rs_o, st_o, _, _ = oracle(m).solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"], kw["atol"],
                                          kw["rtol"])
    g = member(m)
    for shape in [(1, 1, 3), (2, 2, 1), (1, 4, 1), (2, 4, 64)]:
        b = H.heom_batch([g] * 7)
        b.set_tuning(*shape)
        s = H.qme_solver(eng(), b, H.rkdp())
        _, rs = s.solve_rho_s(b.pack([m["rho0"]] * 7), m["t_list"], kw)
        st = s.stats()
        assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected), shape
        for got_rs in b.split_rho_s(rs):
            assert close(got_rs, rs_o), shape


def test_batch_rk4_sweep_rho_s_every_output():
    """24 sweep instances, RK4 dt=0.005 x 400 steps, rho_S every 100 steps.

    The sweep's parameter ranges put some members outside RK4's stability
    region at dt=0.005 (Pade frequency x depth beyond 2.8/dt): their states
    grow without bound in the oracle as well.  Finite members must agree to
    1e-10 on rho_S and to 1e-10 relative to the state's magnitude on the
    whole hierarchy (deep ADOs of an unnormalised hierarchy reach 1e6 and
    more); divergent members must diverge on the GPU too.
    """
    ms = configs.batched_sweep(n_inst=24, seed=24, n_steps=400)
    b = H.heom_batch([member(m) for m in ms])
    s = H.qme_solver(eng(), b, H.rk4())
    t = ms[0]["t_list"]
    rho, rs = s.solve_rho_s(b.pack([m["rho0"] for m in ms]), t, {"dt": ms[0]["dt"]})
    assert rs.shape == (t.size, b.rho_s_size())
    per = b.split_rho_s(rs)
    n_div = 0
    for m, got_rho, got_rs in zip(ms, b.unpack(rho), per):
        rho_o, rs_o, st_o, _, _ = oracle(m).solve(m["rho0"], t, O.RK4, m["dt"])
        if not np.isfinite(rho_o).all() or np.abs(rho_o).max() > 1e30:
            n_div += 1
            assert not np.isfinite(got_rho).all() or np.abs(got_rho).max() > 1e25, m["name"]
            continue
        scale = max(1.0, np.abs(rho_o).max())
        assert np.abs(got_rs - rs_o).max() <= TOL * max(1.0, np.abs(rs_o).max()), m["name"]
        assert np.abs(got_rho - rho_o).max() <= TOL * scale, m["name"]
    print("sweep members outside RK4 stability: %d of %d" % (n_div, len(ms)))
    assert n_div < len(ms) // 2
    # same step sequence as the oracle, shortened last steps included
    st = s.stats()
    assert st["n_rhs"] == st_o.n_rhs
    assert st["ado_stage_updates"] == st_o.n_rhs * b.n_hrchy


def test_batch_host_solve_and_callback():
    ms = configs.batched_sweep(n_inst=5, seed=3, n_steps=100, out_every=50)
    b = H.heom_batch([member(m) for m in ms])
    s = H.qme_solver(eng(), b, H.rk4())
    rho = b.pack([m["rho0"] for m in ms])
    seen = []
    s.solve(rho, ms[0]["t_list"], lambda t: seen.append(t), {"dt": ms[0]["dt"]})
    assert np.allclose(seen, ms[0]["t_list"])
    for m, got in zip(ms, b.unpack(rho)):
        rho_o, _, _, _, _ = oracle(m).solve(m["rho0"], m["t_list"], O.RK4, m["dt"])
        assert close(got, rho_o), m["name"]


def test_batch_rkdp_single_member_is_exact():
    m = configs.spin_boson(max_depth=6, T=0.5)
    m["t_list"] = np.linspace(0.0, 5.0, 11)
    kw = {"dt": 0.05, "atol": 1e-10, "rtol": 1e-8}
    b = H.heom_batch([member(m)])
    s = H.qme_solver(eng(), b, H.rkdp())
    rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
    rho_o, rs_o, st_o, _, _ = oracle(m).solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"],
                                              kw["atol"], kw["rtol"])
    st = s.stats()
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    assert np.abs(rs.reshape(rs_o.shape) - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL


def test_batch_rkdp_identical_copies_match_single_solve():
    """7 copies of a well-conditioned spin-boson; the large first dt forces rejections.

    The model is chosen so that no error norm comes near the accept threshold
    (nearest 0.25 below it), so the step sequence does not depend on the
    summation order of the norm.
    """
    m = configs.spin_boson(max_depth=6, T=0.5)
    m["t_list"] = np.linspace(0.0, 5.0, 11)
    kw = {"dt": 0.8, "atol": 1e-10, "rtol": 1e-8}
    k = 7
    g = member(m)
    b = H.heom_batch([g] * k)
    s = H.qme_solver(eng(), b, H.rkdp())
    rho, rs = s.solve_rho_s(b.pack([m["rho0"]] * k), m["t_list"], kw)
    rho_o, rs_o, st_o, _, _ = oracle(m).solve(m["rho0"], m["t_list"], O.RKDP, kw["dt"],
                                              kw["atol"], kw["rtol"])
    st = s.stats()
    assert st_o.n_rejected > 0
    assert (st["n_accepted"], st["n_rejected"]) == (st_o.n_accepted, st_o.n_rejected)
    for got_rho, got_rs in zip(b.unpack(rho), b.split_rho_s(rs)):
        assert close(got_rs, rs_o)
        assert close(got_rho, rho_o)


def _mixed_physical_members():
    return [configs.spin_boson(max_depth=6, T=0.5), configs.spin_boson(max_depth=8),
            configs.chain4(1, max_depth=3), configs.chain4(2, max_depth=2),
            configs.chain4(3, max_depth=1, N=8), configs.spin_boson(max_depth=3, n_pade=2)]


def test_batch_rkdp_mixed_members_within_tolerance():
    ms = _mixed_physical_members()
    t = np.linspace(0.0, 2.0, 5)
    kw = {"dt": 0.01, "atol": 1e-12, "rtol": 1e-10}
    b = H.heom_batch([member(m) for m in ms])
    s = H.qme_solver(eng(), b, H.rkdp())
    _, rs = s.solve_rho_s(b.pack([m["rho0"] for m in ms]), t, kw)
    worst = 0.0
    for m, got in zip(ms, b.split_rho_s(rs)):
        _, rs_o, _, _, _ = oracle(m).solve(m["rho0"], t, O.RK4, 1e-3)
        worst = max(worst, np.abs(got - rs_o).max())
    print("mixed batch rkdp vs fine-step rk4 rho_S: %.3e" % worst)
    assert worst <= 1e-7


def test_batch_complex64_error_vs_fp64():
    ms = _mixed_physical_members()
    t = np.linspace(0.0, 2.0, 5)
    b = H.heom_batch([member(m, "complex64") for m in ms])
    s = H.qme_solver(eng(), b, H.rk4())
    _, rs = s.solve_rho_s(b.pack([m["rho0"] for m in ms]), t, {"dt": 0.005})
    worst = 0.0
    for m, got in zip(ms, b.split_rho_s(rs)):
        _, rs_o, _, _, _ = oracle(m).solve(m["rho0"], t, O.RK4, 0.005)
        worst = max(worst, np.abs(got - rs_o).max())
    print("complex64 batch rho_S error vs FP64 oracle: %.3e" % worst)
    assert worst <= 1e-5


def test_batch_rejects_unsupported_members():
    dense = configs.random_model(3, 1, 1, 2, seed=1, dense_v=True)
    with pytest.raises(H.HeomError) as e:
        H.heom_batch([member(dense)])
    assert e.value.code == 5
    big = configs.random_model(9, 1, 1, 1, seed=2, dense_v=False)
    with pytest.raises(H.HeomError) as e:
        H.heom_batch([member(big)])
    assert e.value.code == 5
    small = configs.random_model(2, 1, 1, 2, seed=3, dense_v=False)
    with pytest.raises(H.HeomError) as e:
        H.heom_batch([member(small), member(small, "complex64")])
    assert e.value.code == 1
    unset = H.heom_hilb(2)
    with pytest.raises(H.HeomError) as e:
        H.heom_batch([unset])
    assert e.value.code == 2
