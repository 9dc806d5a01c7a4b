"""The C++17 drop-in headers (include/libheom_amd.hpp) used the way a
libheom caller uses the reference templates (libheom_amd/examples/heom_run.cc)."""
import os
import subprocess

import numpy as np
import pytest

from libheom_amd import configs
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "libheom_amd", "_lib", "heom_run")


def write_model(path, form, m, solver, dt, atol, rtol, t_list, rho0):
    def c(z):
        z = np.asarray(z, np.complex128).reshape(-1)
        return " ".join("%.17g %.17g" % (v.real, v.imag) for v in z)
    lines = ["%d %d %d %d" % (form, m["N"], m["max_depth"], len(m["noises"])), c(m["H"])]
    for z in m["noises"]:
        g = np.atleast_2d(z["gamma"])
        lines += ["%d" % g.shape[0], c(z["V"]), c(g), c(z["phi_0"]), c(z["sigma"]), c(z["S"]),
                  c(z["A"]), c([z["s_delta"]])]
    lines.append("%d %.17g %.17g %.17g %d" % (solver, dt, atol, rtol, len(t_list)))
    lines.append(" ".join("%.17g" % t for t in t_list))
    lines.append(c(rho0))
    open(path, "w").write("\n".join(lines) + "\n")


def test_header_compiles_with_plain_gxx():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "libheom_amd", "examples", "heom_run.cc")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("form,solver", [(0, 0), (0, 1), (1, 0), (2, 1), (0, 2), (3, 0),
                                         (4, 1)])
def test_heom_run_matches_oracle(tmp_path, form, solver):
    m = configs.spin_boson(max_depth=0 if form >= 3 else 6)
    t_list = np.linspace(0.0, 2.0, 11)
    dt, atol, rtol = (0.05, 1e-10, 1e-8) if solver == 1 else (0.01, 0.0, 0.0)
    f = tmp_path / "model.txt"
    write_model(str(f), form, m, solver, dt, atol, rtol, t_list, m["rho0"])
    print("heom_run", form, solver, flush=True)
    r = subprocess.run([EXE, str(f)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().split("\n")])
    assert np.allclose(rows[:, 0], t_list)
    rs = (rows[:, 1::2] + 1j * rows[:, 2::2]).reshape(len(t_list), 2, 2)
    o = O.Oracle(form, 2, m["max_depth"]).load(m)
    _, rs_o, _, _, _ = o.solve(m["rho0"], t_list, [O.RK4, O.RKDP, O.LSRK4][solver], dt, atol,
                               rtol)
    assert np.abs(rs - rs_o).max() <= 1e-10
