"""Learned truncation closure on the 4-site chain: the full pipeline.

    python -m libheom_amd.closure_study [--train 2000] [--held-out 100] [--epochs 40]

Workload: 4-site chain (site energies ~N(0, 0.5), couplings ~U(-1, 1)), one
Drude bath per site (1 Drude + 1 Pade mode, lambda ~U(0.1, 1), gamma = 1,
T ~U(0.5, 2)), 8 modes.  Training models are seeds 0 .. train-1, held-out
models seeds 10000 .. 10000+held_out-1.

1. Plain depth-8 runs (RK4, dt = 0.01, t = 0 .. 10, one batch per 250 models)
   give 20 states per model; each state gives 120 closure rows (the depth-3
   last tier): input (Re, Im of rho_n, n, model features), target the
   tier-4 term that depth-3 truncation drops.
2. A 3-layer width-256 tanh network is trained on them (Adam, batch 4096).
2b. Rollout rounds (--rollout-rounds, default 2): the first --rollout-models
   training models are run at depth 3 with the network as their closure; the
   rows of those closure-driven states (inputs from the closure-driven last
   tier, targets from the deep run at the same t) join the training rows and
   the network is trained on the union for --rollout-epochs more epochs.
3. Gain: the network's output scaled by g in {0, 0.25, 0.5, 0.75, 1} on 20
   validation models (seeds 20000 ..), keeping the g with the lowest median
   rho_S error against their depth-8 runs (g = 0 is plain truncation).
4. Held-out models: the fraction of the dropped term the network recovers on
   their deep states, and the rho_S(t) error against the depth-8 run of plain
   depth 3, depth 3 with the closure (gain 1 and the validated gain), and
   plain depths 4 and 5, with the time per run.

Prints one JSON object; everything runs on the GPU in the library's kernels.
"""
import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--held-out", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--deep", type=int, default=8)
    ap.add_argument("--rollout-rounds", type=int, default=2)
    ap.add_argument("--rollout-models", type=int, default=500)
    ap.add_argument("--rollout-epochs", type=int, default=10)
    a = ap.parse_args(argv)

    import libheom_amd as H
    from libheom_amd import closure as C
    from libheom_amd import configs

    eng = H.hip(0)
    dt, L = 0.01, 3
    t_list = np.arange(21) * 0.5
    t0 = time.time()

    def log(msg):
        print("[%.0f s] %s" % (time.time() - t0, msg), file=sys.stderr, flush=True)

    def data(seeds, rows=True):
        md = [configs.chain4(s, max_depth=a.deep) for s in seeds]
        ms = [configs.chain4(s, max_depth=L) for s in seeds]
        f = [C.chain_features(m) for m in ms]
        return C.generate(md, ms, dt, t_list, f, eng, chunk=250, rows=rows,
                          progress=lambda i, n: log("deep runs %d / %d" % (i, n))), f

    te_seeds = [10000 + s for s in range(a.held_out)]
    tg = time.time()
    d_tr, _ = data(range(a.train))
    t_gen = time.time() - tg
    d_te, f_te = data(te_seeds)
    net = C.Mlp(49, 256, 32, seed=1)
    net.fit_scaling(d_tr.X, d_tr.Y, d_tr.n_rows)
    tt = time.time()
    losses = []
    for e0 in range(0, a.epochs, 5):
        losses += net.train(d_tr.X, d_tr.Y, d_tr.n_rows, batch=4096, epochs=min(5, a.epochs - e0),
                            lr=1e-3, seed=2 + e0).tolist()
        log("epoch %d loss %.4f" % (len(losses), losses[-1]))
    t_train = time.time() - tt
    # rollout rounds: rows of the closure-driven states join the training set
    parts = [(d_tr.X, d_tr.Y, d_tr.n_rows)]
    ro = {"rounds": a.rollout_rounds, "models": a.rollout_models, "rows": [], "loss": []}
    tr = time.time()
    ro_seeds = list(range(min(a.rollout_models, a.train)))
    for r in range(a.rollout_rounds):
        md = [configs.chain4(s, max_depth=a.deep) for s in ro_seeds]
        ms = [configs.chain4(s, max_depth=L) for s in ro_seeds]
        fr = [C.chain_features(m) for m in ms]
        xr, yr, nr = C.rollout_rows(md, ms, dt, t_list, fr, eng, net, chunk=250,
                                    progress=lambda i, n: log("rollouts %d / %d" % (i, n)))
        parts.append((xr, yr, nr))
        ro["rows"].append(nr)
        Xu, Yu, nu = C.concat_rows(parts, 49, 32)
        lr_ = net.train(Xu, Yu, nu, batch=4096, epochs=a.rollout_epochs, lr=3e-4,
                        seed=100 + r).tolist()
        ro["loss"].append([lr_[0], lr_[-1]])
        log("rollout round %d: %d rows, loss %.4f" % (r + 1, nr, lr_[-1]))
        del Xu, Yu
    ro["seconds"] = time.time() - tr
    X = d_te.X.to_host(np.float32, (d_te.n_rows, 49))
    Y = d_te.Y.to_host(np.float32, (d_te.n_rows, 32)).astype(np.float64)
    P = net.forward(X).astype(np.float64)
    rec = float(1.0 - np.sum((P - Y) ** 2) / np.sum(Y ** 2))

    xm, xs, ym, ys = net.get_scaling()

    def solve(seed, depth, closure=None, gain=1.0):
        m = configs.chain4(seed, max_depth=depth)
        q = H.heom_hilb(depth).load(m, eng)
        if closure is not None and gain > 0.0:  # gain 0: plain truncation
            # the bound network is a snapshot: its output scaled by the gain
            net.set_scaling(xm, xs, gain * ym, gain * ys)
            q.set_closure(net, closure)
            net.set_scaling(xm, xs, ym, ys)
        s = H.qme_solver(eng, q, H.rk4())
        t = time.perf_counter()
        _, rs = s.solve_rho_s(m["rho0"], t_list, {"dt": dt})
        return rs, time.perf_counter() - t

    va_seeds = [20000 + s for s in range(20)]
    d_va, f_va = data(va_seeds, rows=False)
    gains = [0.0, 0.25, 0.5, 0.75, 1.0]
    val = {}
    for g in gains:
        e = [float(np.abs(solve(s, L, f_va[i], g)[0] - d_va.rho_s[i]).max())
             for i, s in enumerate(va_seeds)]
        val[g] = float(np.median(e))
        log("validation gain %.2f: median rho_S error %.3e" % (g, val[g]))
    g_best = min(gains, key=lambda g: val[g])

    runs = {"plain depth 3": (L, None), "depth 3 + closure": (L, 1.0),
            "depth 3 + closure, validated gain": (L, g_best),
            "plain depth 4": (4, None), "plain depth 5": (5, None),
            "plain depth %d" % a.deep: (a.deep, None)}
    err = {k: [] for k in runs}
    tms = {k: [] for k in runs}
    for i, s in enumerate(te_seeds):
        for k, (depth, g) in runs.items():
            rs, el = solve(s, depth, f_te[i] if g is not None else None, 1.0 if g is None else g)
            err[k].append(float(np.abs(rs - d_te.rho_s[i]).max()))
            tms[k].append(el)
        if i % 10 == 9:
            log("held-out %d / %d" % (i + 1, len(te_seeds)))
    plain = np.array(err["plain depth 3"])
    clo = np.array(err["depth 3 + closure"])
    clo_g = np.array(err["depth 3 + closure, validated gain"])
    out = {
        "workload": "4-site chain, 8 modes; closure at depth 3 trained on %d plain depth-%d "
                    "runs (RK4 dt=0.01, t=0..10, 20 states each)" % (a.train, a.deep),
        "training_rows": d_tr.n_rows, "deep_runs_s": t_gen, "training_s": t_train,
        "epochs": a.epochs, "loss_first_last": [losses[0], losses[-1]],
        "rollout_training": ro,
        "held_out_models": a.held_out,
        "held_out_fraction_of_dropped_term_recovered": rec,
        "rho_s_error_vs_deep": {k: {"median": float(np.median(v)), "mean": float(np.mean(v)),
                                    "max": float(np.max(v))} for k, v in err.items()},
        "ms_per_run": {k: 1e3 * float(np.mean(v)) for k, v in tms.items()},
        "closure_better_than_plain_depth3_fraction": float(np.mean(clo < plain)),
        "validation_median_rho_s_error_by_gain": {"%.2f" % g: v for g, v in val.items()},
        "validated_gain": g_best,
        "validated_gain_closure_better_than_plain_depth3_fraction":
            float(np.mean(clo_g < plain)),
    }
    print(json.dumps(out, indent=1))
    return out


if __name__ == "__main__":
    main()
