"""Benchmark: HEOM propagation of the 7-site exciton chain on one MI355X.

Workload (one "step" = one full propagation, the time-to-rho_S(t) job):
  7-site chain, N = 7, per-site Drude bath (1 Drude + 2 Pade modes) -> K = 21,
  depth L = 4 -> 12650 ADOs (main_size 619850 complex128), Dormand-Prince
  rtol = 1e-8, atol = 1e-11, t = 0 .. 1000 fs, rho_S recorded every 10 fs,
  plain FP64 (complex128) Hilbert form.  Inputs are synthetic (seeded).
The state is resident in HBM when the timed region starts.

Metric: ADO-stage updates/s = n_hrchy * (RHS evaluations performed) / s.
Also reported: the roofline of the dominant kernel (the fused Hilbert-form
stage pass; algorithmic bytes per launch / average launch duration measured
with HIP events on the library stream) and the CPU oracle timed on the host
cores on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "hilb_fast_kernel_7site_traffic.json")


def summarize_pmc(fetch_csv, write_csv, state_bytes):
    """HBM bytes per dispatch of the stage kernel from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch).  Each counter is scaled by what
    it reads for the largest rocclr copy in the same run, a device-to-device
    copy of one state buffer (state_bytes read, state_bytes written)."""
    import csv
    import statistics

    def load(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        copies = [r for r in rows if "copyBuffer" in r["Kernel_Name"]]
        big = max(int(r["Grid_Size"]) for r in copies)
        cal = statistics.median(float(r["Counter_Value"]) for r in copies
                                if int(r["Grid_Size"]) == big)
        stage = [float(r["Counter_Value"]) for r in rows if "hilb_fast_kernel" in r["Kernel_Name"]]
        return statistics.mean(stage), cal, len(stage)

    f, fcal, n1 = load(fetch_csv, "FETCH_SIZE")
    w, wcal, n2 = load(write_csv, "WRITE_SIZE")
    fs, ws = state_bytes / (fcal * 1024.0), state_bytes / (wcal * 1024.0)
    return {"fetch_bytes": f * 1024.0 * fs, "write_bytes": w * 1024.0 * ws,
            "traffic": f * 1024.0 * fs + w * 1024.0 * ws,
            "fetch_scale": fs, "write_scale": ws, "dispatches": [n1, n2],
            "fetch_kib_raw": f, "write_kib_raw": w}


def cpu_baseline(m, seconds=12.0):
    """Oracle (FP64 CPU restatement) on a bounded sample: Dormand-Prince
    attempts of the same workload from rho(0), timed on the host cores."""
    from oracle import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    o = O.Oracle(O.HILB, m["N"], m["max_depth"], n_threads=cores).load(m)
    rho = m["rho0"].copy()
    t, dt = 0.0, m["dt"]
    n_rhs = 0
    t0 = time.perf_counter()
    attempts = 0
    while time.perf_counter() - t0 < seconds and attempts < 200:
        rho, t, dt, st, rc = o.rkdp_step(rho, t, m["t_list"][-1], dt, m["atol"], m["rtol"])
        n_rhs += st.n_rhs
        attempts += st.n_accepted + st.n_rejected
    el = time.perf_counter() - t0
    return {
        "value": o.n_hrchy * n_rhs / el, "unit": "ADO-stage updates/s", "cores": cores,
        "kind": "port",
        "sample": "%d Dormand-Prince attempts (%d RHS evaluations) of the same 12650-ADO "
                  "workload from rho(0), %.1f s, OpenMP over ADOs" % (attempts, n_rhs, el),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--t-final", type=float, default=1000.0,
                    help="propagation length in fs (profiling runs use a shorter one)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write the HBM traffic of the stage kernel from two rocprofv3 "
                         "--pmc counter_collection.csv files and exit")
    args = ap.parse_args()
    if args.summarize_pmc:
        from libheom_amd import configs as _c
        m0 = _c.exciton_chain7(t_final=10.0)
        t = summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], m0["rho0"].nbytes)
        json.dump(t, open(TRAFFIC_JSON, "w"), indent=1)
        print(json.dumps(t))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not provided "
              "(see DESIGN.md for how the hierarchy would shard)", file=sys.stderr)
        sys.exit(2)

    import libheom_amd as H
    from libheom_amd import _capi, configs

    m = configs.exciton_chain7(t_final=args.t_final)
    eng = H.hip(0)
    qme = H.heom_hilb(m["max_depth"]).load(m, eng)
    solver = H.qme_solver(eng, qme, H.rkdp())
    qme32 = H.heom_hilb(m["max_depth"], dtype="complex64").load(m, eng)
    solver32 = H.qme_solver(eng, qme32, H.rkdp())
    L = _capi.load()
    kw = (m["dt"], m["atol"], m["rtol"])
    t_list = np.ascontiguousarray(m["t_list"])
    nb = m["rho0"].nbytes
    d = ctypes.c_void_p()
    _capi.check(L.heom_dev_alloc(ctypes.byref(d), nb))
    rs = np.zeros((t_list.size, 7, 7), np.complex128)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)

    def one_step(sv=solver, out_rs=rs):
        sv.reset()
        _capi.check(L.heom_memcpy_h2d(d, p(m["rho0"]), nb))
        _capi.check(L.heom_device_sync())
        t0 = time.perf_counter()
        _capi.check(L.heom_solve_dev(sv._s, d, p(t_list), t_list.size, kw[0], kw[1],
                                     kw[2], p(out_rs)))
        _capi.check(L.heom_device_sync())
        return time.perf_counter() - t0, sv.stats()

    for _ in range(args.warmup):
        one_step()
    times, updates, rhs = [], 0, 0
    for _ in range(args.steps):
        el, st = one_step()
        times.append(el)
        updates += st["ado_stage_updates"]
        rhs += st["n_rhs"]
    total = sum(times)
    value = updates / total

    # complex64 variant: FP32 operator and stage inputs, FP64 state, stage
    # combination and error norm; its rho_S error is taken against the FP64 run
    rs32 = np.zeros_like(rs)
    for _ in range(args.warmup):
        one_step(solver32, rs32)
    t32, u32 = 0.0, 0
    for _ in range(args.steps):
        el, st = one_step(solver32, rs32)
        t32 += el
        u32 += st["ado_stage_updates"]
    variant32 = {
        "dtype": "f32 operator / f64 combination", "value": u32 / t32,
        "ms_per_step": 1e3 * t32 / args.steps,
        "rho_s_max_abs_err_vs_fp64": float(np.abs(rs32 - rs).max()),
        "speedup_time_to_rho_s": total / t32,
    }

    # dominant kernel: per-stage device time (HIP events) and algorithmic bytes
    ns = ctypes.c_int()
    ms = np.zeros(8)
    by = np.zeros(8)
    _capi.check(L.heom_solver_time_stages(solver._s, args.reps, p(ms), ctypes.byref(ns)))
    _capi.check(L.heom_solver_stage_bytes(solver._s, p(by), ctypes.byref(ns)))
    n = ns.value
    avg_bytes = float(by[:n].mean())
    avg_ms = float(ms[:n].mean())
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    L.heom_dev_free(d)
    # measured HBM bytes per stage dispatch (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    # calibrated on a copy of known size), recorded by --summarize-pmc
    traffic = None
    if os.path.exists(TRAFFIC_JSON):
        traffic = json.load(open(TRAFFIC_JSON)).get("traffic")

    out = {
        "metric": "ADO-stage updates/s + time-to-ρ_S(t) at err≤1e-6 vs plain FP64 on 1 GPU",
        "value": value, "unit": "ADO-stage updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded 7-site chain, default_rng(7))",
        "config": {
            "workload": "7-site exciton chain: N=7, 7 Drude baths x (1 Drude + 2 Pade) = 21 "
                        "modes, depth 4, 12650 ADOs, heom_hilb complex128, Dormand-Prince "
                        "rtol=1e-8 atol=1e-11, t=0..%g fs, rho_S every 10 fs" % args.t_final,
            "n_hrchy": qme.n_hrchy, "main_size": qme.main_size(),
            "rhs_evals_per_step": rhs // args.steps,
            "time_to_rho_s_ms": 1e3 * total / args.steps,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "hilb_fast_kernel<7,double> (fused Dormand-Prince stage pass)",
            "bytes_per_launch": avg_bytes, "avg_launch_ms": avg_ms,
            "stage_ms": [float(x) for x in ms[:n]],
        },
        "cpu_baseline": None,
        "variants": {"complex64": variant32},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
