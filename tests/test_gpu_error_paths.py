"""Error statuses of the entry points that have no reference counterpart or
extend it: Redfield set-up products, the rho_0 projection and the resident
integrator's fall-back, each surfacing as a HeomError with the C status."""
import numpy as np
import pytest

import libheom_amd as H
from libheom_amd import _capi, configs

pytestmark = pytest.mark.gpu
ENG = None


def eng():
    global ENG
    if ENG is None:
        ENG = H.hip(0)
    return ENG


def test_redfield_lambda_on_a_heom_handle():
    m = configs.spin_boson(max_depth=2)
    g = H.heom_hilb(2).load(m, eng())
    lam = np.zeros((2, 2), np.complex128)
    rc = _capi.load().heom_redfield_lambda(g._h, 0, lam.ctypes.data_as(_capi.ctypes.c_void_p),
                                            None)
    assert rc == _capi.HEOM_ERR_ARG


def test_redfield_lambda_noise_index_out_of_range():
    m = configs.spin_boson(max_depth=0)
    m["rho0"] = m["rho0"][:4].copy()
    g = H.redfield_hilb().load(m, eng())
    with pytest.raises(H.HeomError) as e:
        g.Lambda(1)
    assert e.value.code == _capi.HEOM_ERR_ARG


def test_redfield_singular_correlation_solve():
    """gamma - i*omega singular (a zero mode rate at omega = 0) is an argument error."""
    m = configs.spin_boson(max_depth=0)
    z = m["noises"][0]
    z["gamma"] = np.zeros_like(np.atleast_2d(z["gamma"]))
    with pytest.raises(H.HeomError) as e:
        H.redfield_hilb().load(m, eng())
    assert e.value.code == _capi.HEOM_ERR_ARG


def test_projection_beyond_n_level_64():
    m = configs.random_model(65, 1, 1, 1, seed=1, dense_v=False)
    g = H.heom_hilb(1, renorm_rho0=True).load(m, eng())
    with pytest.raises(H.HeomError) as e:
        H.qme_solver(eng(), g, H.rk4())
    assert e.value.code == _capi.HEOM_ERR_UNSUPPORTED


def test_resident_eligible_problem_with_projection_takes_the_stage_kernels():
    """renorm_rho0 turns the resident integrator off; the result still meets the bar."""
    from oracle import oracle as O
    m = configs.spin_boson(max_depth=6)
    g = H.heom_hilb(6, renorm_rho0=True).load(m, eng())
    s = H.qme_solver(eng(), g, H.rk4())
    _, rs = s.solve_rho_s(m["rho0"], m["t_list"], {"dt": m["dt"]})
    o = O.Oracle(O.HILB, 2, 6).load(m)
    _, rs_o, _, _, _ = o.solve(m["rho0"], m["t_list"], O.RK4, m["dt"])
    assert np.abs(rs - rs_o).max() <= 1e-10
