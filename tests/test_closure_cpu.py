"""CPU checks of the closure network restatement (oracle/mlp_oracle.py).

Central finite differences pin the analytic gradient; Adam's first step is
checked against its closed form (|step| = lr for every parameter with a
non-zero gradient when beta-corrected moments are used).
"""
import numpy as np

from oracle import mlp_oracle as MO


def _setup(seed=0, n_in=6, n_hid=5, n_out=4, rows=7):
    rng = np.random.default_rng(seed)
    n_par = n_hid * n_in + n_hid + n_hid * n_hid + n_hid + n_out * n_hid + n_out
    p = rng.normal(size=n_par) * 0.5
    X = rng.normal(size=(rows, n_in))
    T = rng.normal(size=(rows, n_out))
    sc = (rng.normal(size=n_in), rng.uniform(0.5, 2.0, n_in), rng.normal(size=n_out),
          rng.uniform(0.5, 2.0, n_out))
    return p, (n_in, n_hid, n_out), X, T, sc


def test_gradient_matches_finite_differences():
    p, dims, X, T, sc = _setup()
    loss, g = MO.loss_grad(p, *dims, X, T, sc)
    h = 1e-6
    fd = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (MO.loss_grad(p + e, *dims, X, T, sc)[0] -
                 MO.loss_grad(p - e, *dims, X, T, sc)[0]) / (2 * h)
    assert np.abs(g - fd).max() <= 1e-7 * max(1.0, np.abs(fd).max())


def test_forward_denormalises_and_matches_loss():
    p, dims, X, T, sc = _setup(1)
    Y = MO.forward(p, *dims, X, sc)
    loss, _ = MO.loss_grad(p, *dims, X, T, sc)
    ref = np.mean(((Y - T) / sc[3]) ** 2)
    assert abs(loss - ref) <= 1e-12 * max(1.0, ref)


def test_adam_first_step_is_lr_sign():
    rng = np.random.default_rng(2)
    p = rng.normal(size=9)
    g = rng.normal(size=9)
    p1, m, v = MO.adam(p, g, np.zeros(9), np.zeros(9), 1, lr=1e-2, eps=0.0)
    assert np.allclose(p1 - p, -1e-2 * np.sign(g), atol=1e-15)


def test_column_scaling_constant_column():
    X = np.array([[1.0, 3.0], [2.0, 3.0], [3.0, 3.0]])
    mu, sd = MO.column_scaling(X)
    assert np.allclose(mu, [2.0, 3.0]) and np.allclose(sd, [np.sqrt(2.0 / 3.0), 1.0])


def test_shallow_hierarchy_is_the_prefix_of_the_deep_one():
    """closure.rollout_rows builds a deep-sized state whose first ADOs are the
    closure-driven depth-L state: the depth-L hierarchy must be the first
    C(K+L, L) ADOs of the deep one, in the same order (breadth-first, tier by
    tier, hrchy_space.cc:160-229), and the next ADO must start tier L+1."""
    from libheom_amd import configs
    from oracle import oracle as O
    for K, L, deep in [(8, 3, 8), (2, 4, 9), (5, 1, 4), (3, 3, 5)]:
        m = configs.random_model(2, K, 1, deep, seed=K, dense_v=False)
        n_deep = O.Oracle(O.HILB, 2, deep).load(m).tables()[0]
        m["max_depth"] = L
        n_sh = O.Oracle(O.HILB, 2, L).load(m).tables()[0]
        assert np.array_equal(n_deep[:n_sh.shape[0]], n_sh), (K, L, deep)
        assert n_sh.sum(1).max() == L and n_deep[n_sh.shape[0]].sum() == L + 1
