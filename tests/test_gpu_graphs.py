"""HIP-graph replay of the stage sequences gives the same bits as launching
every kernel directly: RK4 with shortened last steps, Dormand-Prince with
rejections, a batch, and a bound closure; each graph is replayed many times
with the library's own eager work (outputs, copies) in between."""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import closure as C
from libheom_amd import configs

pytestmark = pytest.mark.gpu
ENG = None


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def both(qme, solver, m, kw):
    out = []
    for graphs in (False, True):
        s = H.qme_solver(eng(), qme, solver())
        s.set_graphs(graphs)
        rho, rs = s.solve_rho_s(m["rho0"], m["t_list"], kw)
        out.append((rho, rs, s.stats()))
    return out


def test_rk4_graphs_bitwise():
    m = configs.spin_boson(max_depth=6)   # 0.1 intervals of dt 0.01: shortened steps
    q = H.heom_hilb(6).load(m, eng())
    (r0, s0, st0), (r1, s1, st1) = both(q, H.rk4, m, {"dt": m["dt"]})
    assert np.array_equal(r0, r1) and np.array_equal(s0, s1)
    assert st0["n_rhs"] == st1["n_rhs"]


def test_rkdp_graphs_bitwise():
    """Depth 25 (1404 elements): above the resident integrator's size, so both
    runs take the stage kernels, replayed from graphs or launched one by one."""
    m = configs.spin_boson(max_depth=25, T=0.5)
    m["t_list"] = np.linspace(0.0, 5.0, 11)
    q = H.heom_hilb(25).load(m, eng())
    (r0, s0, st0), (r1, s1, st1) = both(q, H.rkdp, m, {"dt": 0.8, "atol": 1e-10, "rtol": 1e-8})
    assert st0["n_rejected"] > 0
    assert np.array_equal(r0, r1) and np.array_equal(s0, s1)
    assert (st0["n_accepted"], st0["n_rejected"]) == (st1["n_accepted"], st1["n_rejected"])


def test_batch_graphs_bitwise():
    ms = configs.batched_sweep(n_inst=6, seed=6, n_steps=200, out_every=50)
    b = H.heom_batch([H.heom_hilb(m["max_depth"]).load(m, eng()) for m in ms])
    mm = {"rho0": b.pack([m["rho0"] for m in ms]), "t_list": ms[0]["t_list"]}
    (r0, s0, _), (r1, s1, _) = both(b, H.rk4, mm, {"dt": ms[0]["dt"]})
    assert np.array_equal(r0, r1) and np.array_equal(s0, s1)


def test_closure_graphs_bitwise():
    m = configs.chain4(4, max_depth=3, t_final=2.0)
    q = H.heom_hilb(3).load(m, eng())
    f = C.chain_features(m)
    net = C.Mlp(49, 32, 32, seed=5)
    rng = np.random.default_rng(1)
    net.set_params(rng.normal(size=net.n_params).astype(np.float32) * 0.01)
    q.set_closure(net, f)
    (r0, s0, _), (r1, s1, _) = both(q, H.rk4, m, {"dt": m["dt"]})
    assert np.array_equal(r0, r1) and np.array_equal(s0, s1)
