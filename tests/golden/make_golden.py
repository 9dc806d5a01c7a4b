"""Writes the golden fixtures of the full-size workloads from the CPU oracle.

TEST INFRASTRUCTURE.  The oracle (oracle/heom_oracle.cc, an FP64 restatement
of heom_hilb.h:70-215, rkdp.h:81-179, rk4.h:40-65 and solver_base.h:64-91) is
too slow to run at these sizes inside a test, so its outputs are stored here and
the GPU tests compare against them:

  chain7_dopri.npz   7-site exciton chain (configs.exciton_chain7, seed 7):
                     N=7, 21 modes, depth 4, 12650 ADOs, Dormand-Prince
                     rtol=1e-8 atol=1e-11, t = 0..1000 fs, rho_S every 10 fs;
                     rho_S at all 101 output times, the accepted / rejected
                     attempt counts of every output interval, the final step
                     size and the final hierarchy's largest magnitude; and the
                     same for the oracle with its coupling sums reversed (its
                     own sensitivity to summation order).
  dense64_dopri.npz  N=64 config (configs.dense64, seed 64): 969 ADOs of
                     64 x 64, Dormand-Prince rtol=1e-6 atol=1e-9, t = 0..2
                     (outputs at 0, 1, 2); rho_S and the counts.
  sweep24_rk4.npz    24 members of the batched sweep (configs.batched_sweep
                     with n_inst=24, seed=24), RK4 dt=0.005 x 4000 steps, rho_S
                     every 100 steps; per member rho_S at all 41 output times
                     and the final hierarchy's largest magnitude (inf when
                     the member leaves RK4's stability region).

Usage: python tests/golden/make_golden.py [chain7|dense64|sweep24 ...]
(all three when no name is given; uses every host core).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from libheom_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402

THREADS = len(os.sched_getaffinity(0))


def _save(name, meta, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
    print("wrote %s (%d bytes)" % (path, os.path.getsize(path)), flush=True)


def _chain7_run(m, reverse):
    """The 7-site job solved interval by interval with dt_save carried between
    the solves (the same step sequence as one solve over all of t_list,
    adaptive_step_size_solver.h:39-55), recording each interval's counts."""
    o = O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=THREADS).load(m)
    o.set_sum_order(reverse)
    rho = m["rho0"].copy()
    dts = None
    t = m["t_list"]
    rs = [rho[:m["N"] ** 2].reshape(m["N"], m["N"]).copy()]
    per = []
    for i in range(len(t) - 1):
        rho, r, st, dts, rc = o.solve(rho, t[i:i + 2], O.RKDP, m["dt"], m["atol"], m["rtol"],
                                      dt_save=dts)
        assert rc == 0
        rs.append(r[-1])
        per.append([st.n_accepted, st.n_rejected])
        print("reversed" if reverse else "plain", i, per[-1], dts, flush=True)
    return rho, np.array(rs), np.array(per, np.int64), dts


def chain7():
    """Plus the same job with the oracle's per-noise coupling sums reversed (the
    same operator with other rounding): the spread between the two runs is the
    reference algorithm's own sensitivity to summation order at this tolerance
    (their step sequences fork after some 270 fs)."""
    m = configs.exciton_chain7()
    t0 = time.time()
    rho, rs, per, dt_save = _chain7_run(m, False)
    _, rs_rev, per_rev, _ = _chain7_run(m, True)
    meta = {"workload": "exciton_chain7(max_depth=4, seed=7, t_final=1000, out_every=10, "
                        "rtol=1e-8, atol=1e-11, dt0=0.1)",
            "n_hrchy": int(len(m["rho0"]) // 49), "n_rhs": int(7 * per.sum()),
            "accepted": int(per[:, 0].sum()), "rejected": int(per[:, 1].sum()),
            "dt_save": dt_save, "final_max_abs": float(np.abs(rho).max()),
            "reversed_accepted": int(per_rev[:, 0].sum()),
            "reversed_rejected": int(per_rev[:, 1].sum()),
            "oracle_seconds": time.time() - t0}
    print(meta, flush=True)
    _save("chain7_dopri.npz", meta, t_list=m["t_list"], rho_s=rs, per_interval=per,
          rho_s_reversed=rs_rev, per_interval_reversed=per_rev)


def dense64():
    m = configs.dense64(t_final=2.0)
    o = O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=THREADS).load(m)
    t0 = time.time()
    rho, rs, st, dt_save, rc = o.solve(m["rho0"], m["t_list"], O.RKDP, m["dt"], m["atol"],
                                       m["rtol"])
    meta = {"workload": "dense64(N=64, max_depth=3, seed=64, t_final=2, out_every=1, "
                        "rtol=1e-6, atol=1e-9, dt0=0.01)",
            "n_hrchy": o.n_hrchy, "n_rhs": st.n_rhs, "accepted": st.n_accepted,
            "rejected": st.n_rejected, "dt_save": dt_save, "rc": rc,
            "final_max_abs": float(np.abs(rho).max()), "oracle_seconds": time.time() - t0}
    print(meta, flush=True)
    _save("dense64_dopri.npz", meta, t_list=m["t_list"], rho_s=rs)


def sweep24():
    ms = configs.batched_sweep(n_inst=24, seed=24, n_steps=4000)
    t_list = ms[0]["t_list"]
    rs_all, scale, names, nrhs = [], [], [], []
    t0 = time.time()
    for m in ms:
        o = O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=THREADS).load(m)
        with np.errstate(all="ignore"):
            rho, rs, st, _, _ = o.solve(m["rho0"], t_list, O.RK4, m["dt"])
            mx = float(np.abs(rho).max()) if np.isfinite(rho).all() else float("inf")
        rs_all.append(rs.reshape(t_list.size, -1))
        scale.append(mx)
        names.append(m["name"])
        nrhs.append(st.n_rhs)
        print(m["name"], st.n_rhs, mx, flush=True)
    meta = {"workload": "batched_sweep(n_inst=24, seed=24, dt=0.005, n_steps=4000, "
                        "out_every=100)", "names": names, "n_rhs": nrhs,
            "oracle_seconds": time.time() - t0}
    sizes = np.array([r.shape[1] for r in rs_all])
    _save("sweep24_rk4.npz", meta, t_list=t_list, rho_s=np.concatenate(rs_all, axis=1),
          n2=sizes, final_max_abs=np.array(scale))


if __name__ == "__main__":
    todo = sys.argv[1:] or ["chain7", "dense64", "sweep24"]
    for name in todo:
        {"chain7": chain7, "dense64": dense64, "sweep24": sweep24}[name]()
