"""Learned truncation closure: network kernels against the CPU restatement
(oracle/mlp_oracle.py), training rows against the CPU oracle's derivatives,
the closure's plumbing into calc_diff / RK4, and what a trained network gains
over plain truncation on held-out models.
"""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import closure as C
from libheom_amd import configs
from oracle import mlp_oracle as MO
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ENG = None


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def small_net(seed=3, n_in=13, n_hid=40, n_out=10):
    net = C.Mlp(n_in, n_hid, n_out, seed=seed)
    rng = np.random.default_rng(seed)
    sc = (rng.normal(size=n_in), rng.uniform(0.5, 2.0, n_in), rng.normal(size=n_out),
          rng.uniform(0.5, 2.0, n_out))
    net.set_scaling(*sc)
    return net, [np.asarray(s, np.float32).astype(np.float64) for s in sc], rng


def test_forward_matches_oracle():
    net, sc, rng = small_net()
    X = rng.normal(size=(77, net.n_in)).astype(np.float32)
    got = net.forward(X)
    ref = MO.forward(net.get_params().astype(np.float64), net.n_in, net.n_hidden, net.n_out,
                     X.astype(np.float64), sc)
    assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("rows", [1, 37, 130])
def test_gradient_matches_oracle(rows):
    net, sc, rng = small_net(seed=rows)
    X = rng.normal(size=(rows, net.n_in)).astype(np.float32)
    Y = rng.normal(size=(rows, net.n_out)).astype(np.float32)
    loss, g = net.loss_grad(X, Y)
    loss_o, g_o = MO.loss_grad(net.get_params().astype(np.float64), net.n_in, net.n_hidden,
                               net.n_out, X.astype(np.float64), Y.astype(np.float64), sc)
    assert abs(loss - loss_o) <= 1e-5 * loss_o
    assert np.abs(g - g_o).max() <= 1e-4 * np.abs(g_o).max()


def test_adam_step_matches_oracle():
    net, sc, rng = small_net(seed=11)
    X = rng.normal(size=(50, net.n_in)).astype(np.float32)
    Y = rng.normal(size=(50, net.n_out)).astype(np.float32)
    p0 = net.get_params().astype(np.float64)
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    for t in (1, 2, 3):
        _, g = net.loss_grad(X, Y)
        net.adam_step(1e-3)
        p0, m, v = MO.adam(p0, g.astype(np.float64), m, v, t, 1e-3)
        assert np.abs(net.get_params() - p0).max() <= 1e-6


def test_fit_scaling_matches_column_statistics():
    net = C.Mlp(5, 8, 3, seed=1)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(300, 5)).astype(np.float32) * [1, 2, 3, 0, 5] + [0, 1, 2, 7, 4]
    Y = rng.normal(size=(300, 3)).astype(np.float32)
    dX, dY = C.DevArray.from_host(X.astype(np.float32)), C.DevArray.from_host(Y)
    net.fit_scaling(dX, dY, 300)
    xm, xs, ym, ys = net.get_scaling()
    mu, sd = MO.column_scaling(X.astype(np.float32).astype(np.float64))
    assert np.allclose(xm, mu, atol=1e-5) and np.allclose(xs, sd, rtol=1e-5)
    assert xs[3] == 1.0  # constant column


def test_training_is_reproducible_and_reduces_loss():
    rng = np.random.default_rng(4)
    X = rng.uniform(-2, 2, size=(2000, 4)).astype(np.float32)
    Y = np.stack([np.sin(X[:, 0]) * X[:, 1], np.cos(X[:, 2] + X[:, 3])], 1).astype(np.float32)
    dX, dY = C.DevArray.from_host(X), C.DevArray.from_host(Y)
    params = []
    for _ in range(2):
        net = C.Mlp(4, 64, 2, seed=7)
        net.fit_scaling(dX, dY, 2000)
        losses = net.train(dX, dY, 2000, batch=100, epochs=30, lr=3e-3, seed=5)
        params.append(net.get_params())
    print("training loss: first %.4f last %.5f" % (losses[0], losses[-1]))
    assert losses[-1] < 0.05 * losses[0]
    assert np.array_equal(params[0], params[1])


@pytest.mark.parametrize("layout", [0, 1])
def test_closure_rows_match_oracle_derivatives(layout):
    seed, L, Ld = 5, 3, 5
    md, ms = configs.chain4(seed, max_depth=Ld), configs.chain4(seed, max_depth=L)
    deep = H.heom_hilb(Ld).load(md, eng())
    sh = H.heom_hilb(L).load(ms, eng())
    rng = np.random.default_rng(seed)
    rho = (rng.normal(size=deep.main_size()) + 1j * rng.normal(size=deep.main_size())) * 0.1
    f = C.chain_features(md)
    nL = configs.hrchy_count(8, L) - configs.hrchy_count(8, L - 1)
    rows, n_in, n_out = C.dims(sh, layout, f.size)
    assert (rows, n_in, n_out) == ((nL, 32 + 8 + f.size, 32) if layout == 0 else
                                   (1, 32 * nL + f.size, 32 * nL))
    dX, dY = C.DevArray(rows * n_in * 4), C.DevArray(rows * n_out * 4)
    drho = C.DevArray.from_host(rho)
    from libheom_amd import _capi
    _capi.check(_capi.load().heom_closure_rows(deep._h, sh._h, drho.ptr, C._p(f), f.size,
                                               layout, dX.ptr, dY.ptr))
    Xr = dX.to_host(np.float32, (rows, n_in))
    Y = dY.to_host(np.float32, (rows, n_out)).reshape(nL, 32)
    od = O.Oracle(O.HILB, 4, Ld, n_threads=8).load(md)
    osh = O.Oracle(O.HILB, 4, L, n_threads=8).load(ms)
    lo, hi = configs.hrchy_count(8, L - 1) * 16, configs.hrchy_count(8, L) * 16
    # calc_diff = -D: the dropped term D_full - D_trunc = -(cd_full - cd_trunc)
    term = -(od.calc_diff(rho)[lo:hi] - osh.calc_diff(rho[:hi])[lo:]).reshape(nL, 16)
    assert np.abs(Y[:, :16] - term.real).max() <= 1e-6 * np.abs(term).max()
    assert np.abs(Y[:, 16:] - term.imag).max() <= 1e-6 * np.abs(term).max()
    n_tab, _, _ = osh.tables()
    r = rho[lo:hi].reshape(nL, 16)
    if layout == 0:
        X = Xr
        assert np.array_equal(X[:, 32:40], n_tab[lo // 16:].astype(np.float32))
        assert np.array_equal(X[:, 40:], np.broadcast_to(f, (nL, f.size)))
    else:
        X = Xr[0, :32 * nL].reshape(nL, 32)
        assert np.array_equal(Xr[0, 32 * nL:], f)
    assert np.array_equal(X[:, :16], r.real.astype(np.float32))
    assert np.array_equal(X[:, 16:32], r.imag.astype(np.float32))


def test_closure_plumbing_with_constant_network():
    """Zero weights and output shift c: the closure adds exactly c to D of the
    last tier; checked on calc_diff and on an RK4 solve against the oracle."""
    m = configs.chain4(2, max_depth=3, t_final=1.0)
    q = H.heom_hilb(3).load(m, eng())
    f = C.chain_features(m)
    net = C.Mlp(32 + 8 + f.size, 16, 32, seed=0)
    net.set_params(np.zeros(net.n_params, np.float32))
    c = (np.arange(32, dtype=np.float32) - 15.5) / 64.0
    net.set_scaling(np.zeros(net.n_in), np.ones(net.n_in), c, np.ones(32))
    rng = np.random.default_rng(0)
    rho = (rng.normal(size=q.main_size()) + 1j * rng.normal(size=q.main_size())) * 0.1
    base = q.calc_diff_impl(None, np.zeros_like(rho), rho, 1.0, 0.0)
    q.set_closure(net, f, layout=0)
    got = q.calc_diff_impl(None, np.zeros_like(rho), rho, 1.0, 0.0)
    lo = configs.hrchy_count(8, 2) * 16
    add = np.zeros_like(rho)
    add[lo:] = np.tile(c[:16] + 1j * c[16:], q.n_hrchy - lo // 16)
    assert np.abs((got - base) + add).max() <= 1e-14
    s = H.qme_solver(eng(), q, H.rk4())
    rho_g, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": m["dt"]})
    o = O.Oracle(O.HILB, 4, 3, n_threads=8).load(m)
    y = m["rho0"].copy()
    f_o = lambda r: o.calc_diff(r) - add
    dt = m["dt"]
    for _ in range(int(round(m["t_list"][-1] / dt))):
        k1 = f_o(y)
        k2 = f_o(y + 0.5 * dt * k1)
        k3 = f_o(y + 0.5 * dt * k2)
        k4 = f_o(y + dt * k3)
        y = y + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert np.abs(rho_g - y).max() <= 1e-10
    q.set_closure(None)
    assert np.abs(q.calc_diff_impl(None, np.zeros_like(rho), rho, 1.0, 0.0) - base).max() == 0.0


def test_bound_closure_matches_network_on_the_same_rows():
    """Inference path (gather, folded normalisation, three GEMMs, add to D)
    against the CPU restatement of the network on the same rows."""
    m = configs.chain4(7, max_depth=3)
    q = H.heom_hilb(3).load(m, eng())
    f = C.chain_features(m)
    for layout in (0, 1):
        rows, n_in, n_out = C.dims(q, layout, f.size)
        net, sc, rng = small_net(seed=21 + layout, n_in=n_in, n_hid=48, n_out=n_out)
        rho = (rng.normal(size=q.main_size()) + 1j * rng.normal(size=q.main_size())) * 0.1
        q.set_closure(None)
        base = q.calc_diff_impl(None, np.zeros_like(rho), rho, 1.0, 0.0)
        q.set_closure(net, f, layout=layout)
        got = q.calc_diff_impl(None, np.zeros_like(rho), rho, 1.0, 0.0)
        lo = configs.hrchy_count(8, 2)
        blk = rho[lo * 16:].reshape(-1, 16)
        n_tab = O.Oracle(O.HILB, 4, 3).load(m).tables()[0][lo:].astype(np.float64)
        if layout == 0:
            X = np.concatenate([blk.real, blk.imag, n_tab, np.broadcast_to(f, (len(blk), f.size))], 1)
        else:
            X = np.concatenate([np.concatenate([blk.real, blk.imag], 1).reshape(1, -1), f[None, :]], 1)
        X = X.astype(np.float32).astype(np.float64)
        y = MO.forward(net.get_params().astype(np.float64), n_in, 48, n_out, X, sc).reshape(-1, 32)
        term = (y[:, :16] + 1j * y[:, 16:]).reshape(-1)
        diff = -(got - base)[lo * 16:]
        assert np.abs(diff - term).max() <= 2e-5 * max(1.0, np.abs(term).max()), layout
        q.set_closure(None)


def test_trained_closure_recovers_the_dropped_term():
    """What training gains: on held-out models' deep states the trained network
    recovers part of the tier-(L+1) term that plain truncation drops (plain
    truncation recovers none of it).  Test scale: 500 training models at depth
    5 (20 states each), 20 epochs; the full-scale study (2000 depth-8 models,
    libheom_amd/closure_study.py) recovers about 0.9."""
    L, Ld, dt = 3, 5, 0.01
    t_list = np.arange(21) * 0.5
    tr, te = list(range(500)), [10000 + s for s in range(20)]
    mk = lambda seeds, d: [configs.chain4(s, max_depth=d) for s in seeds]
    f_tr = [C.chain_features(m) for m in mk(tr, L)]
    f_te = [C.chain_features(m) for m in mk(te, L)]
    d_tr = C.generate(mk(tr, Ld), mk(tr, L), dt, t_list, f_tr, eng(), chunk=250)
    d_te = C.generate(mk(te, Ld), mk(te, L), dt, t_list, f_te, eng(), chunk=250)
    net = C.Mlp(49, 256, 32, seed=1)
    net.fit_scaling(d_tr.X, d_tr.Y, d_tr.n_rows)
    losses = net.train(d_tr.X, d_tr.Y, d_tr.n_rows, batch=4096, epochs=20, lr=1e-3, seed=2)
    X = d_te.X.to_host(np.float32, (d_te.n_rows, 49))
    Y = d_te.Y.to_host(np.float32, (d_te.n_rows, 32)).astype(np.float64)
    P = net.forward(X).astype(np.float64)
    recovered = 1.0 - np.sum((P - Y) ** 2) / np.sum(Y ** 2)
    print("training loss %.4f -> %.4f; held-out fraction of the dropped term recovered: %.3f"
          % (losses[0], losses[-1], recovered))
    assert losses[-1] < 0.5 * losses[0]
    assert recovered > 0.15


def test_rollout_rows_pair_closure_driven_inputs_with_deep_targets():
    """closure.rollout_rows (training on the closure's own runs): the targets
    are the deep run's dropped term (the deep-state rows' targets, to the FP64
    rounding of the deep-minus-truncated derivative difference, whose tier-3
    part now comes from another state), the constant columns the same, and the
    Re/Im columns the last tier of the closure-driven depth-3 run at that time."""
    L, Ld, dt = 3, 5, 0.01
    t_list = np.arange(4) * 0.5
    seeds = list(range(6))
    mk = lambda d: [configs.chain4(s, max_depth=d) for s in seeds]
    f = [C.chain_features(m) for m in mk(L)]
    d = C.generate(mk(Ld), mk(L), dt, t_list, f, eng(), chunk=4)
    net = C.Mlp(49, 256, 32, seed=1)
    net.fit_scaling(d.X, d.Y, d.n_rows)
    xr, yr, nr = C.rollout_rows(mk(Ld), mk(L), dt, t_list, f, eng(), net, chunk=4)
    assert nr == d.n_rows
    Xg = d.X.to_host(np.float32, (nr, 49))
    Yg = d.Y.to_host(np.float32, (nr, 32))
    Xr = xr.to_host(np.float32, (nr, 49))
    Yr = yr.to_host(np.float32, (nr, 32))
    err = float(np.abs(Yr.astype(np.float64) - Yg).max())
    print("rollout targets vs deep-state targets: max abs difference %.2e" % err)
    assert err <= 1e-10
    assert np.array_equal(Xr[:, 32:], Xg[:, 32:])
    assert not np.array_equal(Xr[:, :32], Xg[:, :32])  # the closure moved the run
    # model 0 at t = 0.5: its rows are the first 120; rerun it with the closure
    m = configs.chain4(0, max_depth=L)
    q = H.heom_hilb(L).load(m, eng())
    q.set_closure(net, f[0])
    rho, _ = H.qme_solver(eng(), q, H.rk4()).solve_rho_s(m["rho0"], t_list[:2], {"dt": dt})
    last = rho.reshape(-1, 16)[45:]
    assert last.shape[0] == 120
    assert np.array_equal(Xr[:120, :16], last.real.astype(np.float32))
    assert np.array_equal(Xr[:120, 16:32], last.imag.astype(np.float32))
