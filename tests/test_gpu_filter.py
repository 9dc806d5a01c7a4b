"""Filtered hierarchies (hrchy_filter, hrchy_space.h:29-36) on the GPU: the
tables match the oracle's enumeration (which the CPU tests pin to the
reference's own hrchy_space.cc through tests/golden/hrchy_ref.npz), and
calc_diff / RK4 on the pruned hierarchy match the oracle built with the same
filter (FP64, 1e-10)."""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import configs
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10
FORMS = {"hilb": (H.heom_hilb, O.HILB), "liou": (H.heom_liou, O.LIOU),
         "ado": (H.heom_ado, O.ADO)}
FILTERS = {
    "max occupation 2": lambda n, d: n.max() <= 2,
    "prune through an intermediate": lambda n, d: not (n[0] == 1 and n[1:].sum() == 0),
}


def build(form, m, fn):
    cls, okind = FORMS[form]
    g = cls(m["max_depth"])
    g.hrchy_filter = fn
    g.load(m, H.hip(0))
    o = O.Oracle(okind, m["N"], m["max_depth"], n_threads=8).set_filter(fn).load(m)
    return g, o


@pytest.mark.parametrize("name", sorted(FILTERS))
@pytest.mark.parametrize("form", ["hilb", "liou", "ado"])
def test_filtered_calc_diff(form, name):
    m = configs.random_model(3, 2, 2, 4, seed=5, dense_v=(form != "hilb"))
    fn = FILTERS[name]
    g, o = build(form, m, fn)
    assert g.n_hrchy == o.n_hrchy < configs.hrchy_count(4, 4)
    n, p1, m1 = g.hrchy_tables()
    on, op1, om1 = o.tables()
    assert np.array_equal(n, on) and np.array_equal(p1, op1) and np.array_equal(m1, om1)
    rng = np.random.default_rng(3)
    rho = (rng.normal(size=g.main_size()) + 1j * rng.normal(size=g.main_size())) * 0.1
    ref_d = o.calc_diff(rho, 0.5 - 0.25j, 0.0)
    got = g.calc_diff_impl(None, np.zeros_like(rho), rho, 0.5 - 0.25j, 0.0)
    assert np.abs(got - ref_d).max() <= TOL * max(1.0, np.abs(ref_d).max())


def test_filtered_rk4_spin_boson():
    m = configs.spin_boson(max_depth=8)
    fn = lambda n, d: n.max() <= 5
    g, o = build("hilb", m, fn)
    assert g.n_hrchy < configs.hrchy_count(2, 8)
    s = H.qme_solver(H.hip(0), g, H.rk4())
    rho0 = np.zeros(g.main_size(), np.complex128)
    rho0[0] = 1.0
    rho, rs = s.solve_rho_s(rho0, m["t_list"][:51], {"dt": m["dt"]})
    rho_o, rs_o, _, _, _ = o.solve(rho0, m["t_list"][:51], O.RK4, m["dt"])
    assert np.abs(rs - rs_o).max() <= TOL
    assert np.abs(rho - rho_o).max() <= TOL * max(1.0, np.abs(rho_o).max())


def test_filter_exception_is_raised():
    m = configs.spin_boson(max_depth=3)
    g = H.heom_hilb(3)

    def bad(n, d):
        raise ValueError("filter failed")

    g.hrchy_filter = bad
    with pytest.raises(ValueError):
        g.load(m, H.hip(0))
