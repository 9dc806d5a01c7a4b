"""Learned hierarchy-truncation closure over the C ABI.

A depth-L heom_hilb drops every ADO beyond tier L, so each tier-L ADO loses
its couplings to tier L+1.  The closure network predicts that dropped term
from the tier-L ADO itself (Re/Im of rho_n), its index vector n and a few
model features, and the operator kernels add the prediction to D_n at every
derivative evaluation.  Training rows come from deeper hierarchies of the
same models (heom_closure_rows: the exact dropped term is the difference of
the deep and truncated derivatives on the deep state); training and inference
run in the library's own HIP kernels.

    net = Mlp(n_in, 256, n_out)
    data = generate(models_deep, models_shallow, ...)   # rows in HBM
    net.fit_scaling(data.X, data.Y, data.n_rows)
    losses = net.train(data.X, data.Y, data.n_rows, batch, epochs, lr)
    qme.set_closure(net, features)                      # depth-L heom_hilb
"""
import ctypes

import numpy as np

from . import _capi

_p = lambda a: a.ctypes.data_as(ctypes.c_void_p)


class DevArray:
    """Device memory from heom_dev_alloc, freed with the object."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _capi.check(_capi.load().heom_dev_alloc(ctypes.byref(self.ptr), max(self.nbytes, 1)))

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            _capi.load().heom_dev_free(self.ptr)
            self.ptr = None

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _capi.check(_capi.load().heom_memcpy_h2d(d.ptr, _p(a), a.nbytes))
        return d

    def addr(self, byte_offset=0):
        return ctypes.c_void_p(self.ptr.value + int(byte_offset))

    def to_host(self, dtype, shape, byte_offset=0):
        out = np.empty(shape, dtype)
        _capi.check(_capi.load().heom_memcpy_d2h(_p(out), self.addr(byte_offset), out.nbytes))
        return out


class Mlp:
    """3-layer tanh network (n_in -> n_hidden -> n_hidden -> n_out), FP32."""

    def __init__(self, n_in, n_hidden, n_out, seed=0, device=0):
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        h = ctypes.c_void_p()
        _capi.check(_capi.load().heom_mlp_create(n_in, n_hidden, n_out, seed, device,
                                                 ctypes.byref(h)))
        self._h = h
        n = ctypes.c_longlong()
        _capi.check(_capi.load().heom_mlp_n_params(h, ctypes.byref(n)))
        self.n_params = n.value

    def __del__(self):
        if getattr(self, "_h", None):
            _capi.load().heom_mlp_destroy(self._h)
            self._h = None

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _capi.check(_capi.load().heom_mlp_get_params(self._h, _p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _capi.check(_capi.load().heom_mlp_set_params(self._h, _p(p)))

    def set_scaling(self, xm, xs, ym, ys):
        a = [np.ascontiguousarray(v, np.float32) for v in (xm, xs, ym, ys)]
        _capi.check(_capi.load().heom_mlp_set_scaling(self._h, *[_p(v) for v in a]))

    def get_scaling(self):
        a = [np.zeros(n, np.float32) for n in (self.n_in, self.n_in, self.n_out, self.n_out)]
        _capi.check(_capi.load().heom_mlp_get_scaling(self._h, *[_p(v) for v in a]))
        return a

    def fit_scaling(self, X, Y, n_rows):
        _capi.check(_capi.load().heom_mlp_fit_scaling(self._h, X.ptr, Y.ptr, n_rows))

    def forward(self, X):
        """Host rows in, host rows out (through device buffers)."""
        X = np.ascontiguousarray(X, np.float32)
        dX = DevArray.from_host(X)
        dY = DevArray(X.shape[0] * self.n_out * 4)
        _capi.check(_capi.load().heom_mlp_forward(self._h, dX.ptr, dY.ptr, X.shape[0]))
        return dY.to_host(np.float32, (X.shape[0], self.n_out))

    def loss_grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        dX, dY = DevArray.from_host(X), DevArray.from_host(Y)
        loss = ctypes.c_double()
        g = np.zeros(self.n_params, np.float32)
        _capi.check(_capi.load().heom_mlp_loss_grad(self._h, dX.ptr, dY.ptr, X.shape[0],
                                                    ctypes.byref(loss), _p(g)))
        return loss.value, g

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _capi.check(_capi.load().heom_mlp_adam_step(self._h, lr, beta1, beta2, eps))

    def train(self, X, Y, n_rows, batch, epochs, lr, seed=0):
        losses = np.zeros(max(epochs, 1))
        _capi.check(_capi.load().heom_mlp_train(self._h, X.ptr, Y.ptr, n_rows, batch, epochs,
                                                lr, seed, _p(losses)))
        return losses[:epochs]


def chain_features(m):
    """Model features of a chain config: lambda, T, site energies, couplings."""
    Hr = np.real(m["H"])
    return np.array([m["lam"], m["T"], *np.diag(Hr), *np.diag(Hr, 1)], np.float32)


class Rows:
    """Training rows in HBM plus the deep runs' rho_S (n_models, n_t, N, N)."""

    def __init__(self, X, Y, n_rows, rho_s, t_list):
        self.X, self.Y, self.n_rows, self.rho_s, self.t_list = X, Y, n_rows, rho_s, t_list


def dims(qme, layout, n_features):
    """(rows per state, n_in, n_out) of a closure layout for a depth-L heom_hilb."""
    r, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _capi.check(_capi.load().heom_closure_dims(qme._h, layout, n_features, ctypes.byref(r),
                                               ctypes.byref(a), ctypes.byref(b)))
    return r.value, a.value, b.value


def generate(deep_models, shallow_models, dt, t_list, features, engine, chunk=256,
             rows=True, layout=_capi.HEOM_CLOSURE_PER_ADO, progress=None):
    """Propagate the deep models (RK4, as one batch per chunk) and, at every
    t in t_list[1:], build the closure rows of each model from its deep state.

    Returns Rows with X (n_rows x n_in) and Y (n_rows x n_out) on the device
    (None when rows is False) and the deep rho_S at every t in t_list."""
    from . import heom_batch, heom_hilb, qme_solver, rk4
    L = _capi.load()
    n_models = len(deep_models)
    N = deep_models[0]["N"]
    t_list = np.asarray(t_list, np.float64)
    per_model = None
    X = Y = None
    row = 0
    rho_s = np.zeros((n_models, t_list.size, N, N), np.complex128)
    for c0 in range(0, n_models, chunk):
        idx = list(range(c0, min(n_models, c0 + chunk)))
        deep = [heom_hilb(deep_models[i]["max_depth"]).load(deep_models[i], engine) for i in idx]
        shallow = [heom_hilb(shallow_models[i]["max_depth"]).load(shallow_models[i], engine)
                   for i in idx]
        if per_model is None:
            per_model, n_in, n_out = dims(shallow[0], layout, len(features[0]))
            total = n_models * (t_list.size - 1) * per_model if rows else 0
            if rows:
                X = DevArray(total * n_in * 4)
                Y = DevArray(total * n_out * 4)
        b = heom_batch(deep)
        sv = qme_solver(engine, b, rk4())
        state = DevArray.from_host(b.pack([deep_models[i]["rho0"] for i in idx]))
        blk = np.zeros((2, b.rho_s_size()), np.complex128)
        for j in range(1, t_list.size):
            tt = np.ascontiguousarray(t_list[j - 1:j + 1])
            _capi.check(L.heom_solve_dev(sv._s, state.ptr, _p(tt), 2, dt, 0.0, 0.0, _p(blk)))
            if j == 1:
                for k, rs in enumerate(b.split_rho_s(blk[:1])):
                    rho_s[idx[k], 0] = rs[0]
            for k, rs in enumerate(b.split_rho_s(blk[1:])):
                rho_s[idx[k], j] = rs[0]
            if rows:
                for k, i in enumerate(idx):
                    f = np.ascontiguousarray(features[i], np.float32)
                    off = int(b.elem_base[k]) * 16
                    _capi.check(L.heom_closure_rows(
                        deep[k]._h, shallow[k]._h, state.addr(off), _p(f), f.size, layout,
                        X.addr(row * n_in * 4), Y.addr(row * n_out * 4)))
                    row += per_model
        del sv, b, state
        if progress:
            progress(idx[-1] + 1, n_models)
    return Rows(X, Y, row, rho_s, t_list)


def rollout_rows(deep_models, shallow_models, dt, t_list, features, engine, net, chunk=250,
                 layout=_capi.HEOM_CLOSURE_PER_ADO, bound=1e6, progress=None):
    """Closure rows along closure-driven runs (data seen by the closure itself).

    Each shallow model is propagated (RK4) with `net` bound as its closure,
    beside its deep run.  At every t in t_list[1:] its rows take the inputs
    from the closure-driven run's last tier and the targets (the tier-(L+1)
    term truncation drops) from the deep state at the same t.  The shallow
    hierarchy is the first ADOs of the deep one (both enumerated breadth-first,
    tier by tier), so the rows are built from the deep state with its first
    ADOs replaced by the closure-driven state.  A run whose state leaves
    |rho| <= bound (or turns non-finite) gives no further rows.

    Returns (X, Y, n_rows) with X, Y on the device."""
    from . import heom_batch, heom_hilb, qme_solver, rk4
    L = _capi.load()
    n_models = len(deep_models)
    t_list = np.asarray(t_list, np.float64)
    per_model, n_in, n_out = None, 0, 0
    X = Y = None
    row = 0
    for c0 in range(0, n_models, chunk):
        idx = list(range(c0, min(n_models, c0 + chunk)))
        deep = [heom_hilb(deep_models[i]["max_depth"]).load(deep_models[i], engine) for i in idx]
        shallow = [heom_hilb(shallow_models[i]["max_depth"]).load(shallow_models[i], engine)
                   for i in idx]
        if per_model is None:
            per_model, n_in, n_out = dims(shallow[0], layout, len(features[0]))
            total = n_models * (t_list.size - 1) * per_model
            X = DevArray(total * n_in * 4)
            Y = DevArray(total * n_out * 4)
        # closure-driven runs: own QMEs with the network bound, device states
        runs = []
        for k, i in enumerate(idx):
            q = heom_hilb(shallow_models[i]["max_depth"]).load(shallow_models[i], engine)
            q.set_closure(net, features[i], layout)
            sv = qme_solver(engine, q, rk4())
            st = DevArray.from_host(np.ascontiguousarray(shallow_models[i]["rho0"],
                                                         np.complex128))
            runs.append([q, sv, st, True])
        b = heom_batch(deep)
        sv_d = qme_solver(engine, b, rk4())
        state = DevArray.from_host(b.pack([deep_models[i]["rho0"] for i in idx]))
        n_deep = max(q.main_size() for q in deep)
        mixed = DevArray(n_deep * 16)
        for j in range(1, t_list.size):
            tt = np.ascontiguousarray(t_list[j - 1:j + 1])
            _capi.check(L.heom_solve_dev(sv_d._s, state.ptr, _p(tt), 2, dt, 0.0, 0.0, None))
            for k, i in enumerate(idx):
                q, sv, st, live = runs[k]
                if not live:
                    continue
                _capi.check(L.heom_solve_dev(sv._s, st.ptr, _p(tt), 2, dt, 0.0, 0.0, None))
                ns = q.main_size()
                rs = st.to_host(np.complex128, (ns,))
                if not np.isfinite(rs).all() or np.abs(rs).max() > bound:
                    runs[k][3] = False
                    continue
                nd = deep[k].main_size()
                full = state.to_host(np.complex128, (nd,), int(b.elem_base[k]) * 16)
                full[:ns] = rs
                _capi.check(L.heom_memcpy_h2d(mixed.ptr, _p(full), full.nbytes))
                f = np.ascontiguousarray(features[i], np.float32)
                _capi.check(L.heom_closure_rows(
                    deep[k]._h, shallow[k]._h, mixed.ptr, _p(f), f.size, layout,
                    X.addr(row * n_in * 4), Y.addr(row * n_out * 4)))
                row += per_model
        del sv_d, b, state, runs
        if progress:
            progress(idx[-1] + 1, n_models)
    return X, Y, row


def concat_rows(parts, n_in, n_out):
    """One device row set from several (X, Y, n_rows) parts (through the host)."""
    n = sum(p[2] for p in parts)
    X, Y = DevArray(n * n_in * 4), DevArray(n * n_out * 4)
    r = 0
    for px, py, m in parts:
        if m == 0:
            continue
        hx = px.to_host(np.float32, (m, n_in))
        hy = py.to_host(np.float32, (m, n_out))
        _capi.check(_capi.load().heom_memcpy_h2d(X.addr(r * n_in * 4), _p(hx), hx.nbytes))
        _capi.check(_capi.load().heom_memcpy_h2d(Y.addr(r * n_out * 4), _p(hy), hy.nbytes))
        r += m
    return X, Y, n
