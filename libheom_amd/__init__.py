"""libheom_amd -- MI355X-native HEOM propagation behind libheom's qme API.

Host-side mirror of the reference's C++ template interface (qme_base.h,
heom_hilb.h, heom_liou.h, heom_ado.h, redfield_hilb.h, redfield_liou.h,
solver/rk4.h, solver/rkdp.h, solver/lsrk4.h, qme_solver.h) over the C ABI of
``include/libheom_amd.h``:

    qme = heom_hilb(max_depth, n_inner_threads, n_outer_threads)
    qme.n_level = N; qme.n_level_2 = N*N
    qme.H = H                       # N x N
    qme.alloc_noises(n_noise)
    qme.V[u] = V; qme.len_gamma[u] = K; qme.gamma[u] = g; qme.phi_0[u] = ...
    qme.sigma[u] = ...; qme.S[u] = ...; qme.A[u] = ...; qme.s_delta[u] = ...
    qme.set_param(engine)           # engine = hip()
    solver = qme_solver(engine, qme, rkdp())
    solver.solve(rho, t_list, callback, {"dt": ..., "atol": ..., "rtol": ...})

``rho`` is the host state of main_size = N*N*n_hrchy complex128 values
(ADO-major, each ADO an N x N row-major block), propagated in place; the
callback receives each output time t before propagation (solver_base.h:73-78).
Missing kwargs raise KeyError (the reference's kwargs.at throws
std::out_of_range).  Every computation runs in the HIP library; there is no
CPU fallback.
"""
import ctypes

import numpy as np

from . import _capi
from ._capi import HeomError

__all__ = ["hip", "heom_hilb", "heom_liou", "heom_ado", "heom_batch", "redfield_hilb",
           "redfield_liou", "rk4", "rkdp", "lsrk4",
           "qme_solver",
           "HeomError", "row_major", "col_major"]

row_major = True
col_major = False


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c128(a, shape):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    return a.reshape(shape)


class hip:
    """The HIP linalg engine (engine tag replacing eigen/mkl/cuda, linalg_engine.h:100-134)."""

    def __init__(self, device=0):
        self.device = device
        _capi.load()


class _heom:
    _kind = None
    # every new C handle gets a fresh generation number, so that a solver can
    # tell a replaced handle from its own even when the allocator reuses the address
    _generations = iter(range(1, 1 << 62))

    def __init__(self, max_depth, n_inner_threads=1, n_outer_threads=1, dtype="complex128",
                 order=row_major, parity_q1=True, sqrt_normalization=False, matrix_cores=True,
                 renorm_rho0=False):
        # heom.h:44-50; thread counts are accepted for API parity, the device replaces them
        self.max_depth = int(max_depth)
        self.n_inner_threads = n_inner_threads
        self.n_outer_threads = n_outer_threads
        # "complex64_bf16x3": complex64 state, commutator products on bf16 matrix
        # cores with split (hi + lo) operands
        self.bf16x3 = dtype == "complex64_bf16x3"
        if self.bf16x3:
            dtype = "complex64"
        if dtype not in ("complex128", "complex64", np.complex128, np.complex64):
            raise ValueError("dtype must be complex128, complex64 or complex64_bf16x3")
        self.dtype = np.dtype(dtype)
        self.matrix_cores = matrix_cores
        # project rho_0 on Hermitian matrices of the initial trace after every step
        self.renorm_rho0 = renorm_rho0
        self.order = order
        self.parity_q1 = parity_q1
        self.sqrt_normalization = sqrt_normalization
        self.n_level = 0
        self.n_level_2 = 0
        self.H = None
        self.n_noise = 0
        # hrchy_filter(index: ndarray of n_dim ints, depth: int) -> bool, applied
        # where the breadth-first enumeration adds an index (hrchy_space.cc:183-187)
        self.hrchy_filter = None
        self._h = None
        self._ready = False

    def __del__(self):
        if getattr(self, "_h", None):
            _capi.load().heom_destroy(self._h)
            self._h = None

    # qme_base.h:54-67
    def alloc_noises(self, n_noise):
        self.n_noise = int(n_noise)
        self.V = [None] * n_noise
        self.len_gamma = [0] * n_noise
        self.gamma = [None] * n_noise
        self.phi_0 = [None] * n_noise
        self.sigma = [None] * n_noise
        self.S = [None] * n_noise
        self.A = [None] * n_noise
        self.s_delta = [0.0] * n_noise

    def set_param(self, engine=None):
        """heom::set_param + form set_param: hierarchy, tables, upload to HBM."""
        L = _capi.load()
        dev = engine.device if engine is not None else 0
        if self.n_level <= 0:
            raise HeomError(_capi.HEOM_ERR_ARG, "n_level not set")
        N = self.n_level
        opt = _capi.heom_options(int(self.parity_q1), int(self.sqrt_normalization),
                                 0 if self.order == row_major else 1, dev,
                                 int(bool(self.matrix_cores)), int(bool(self.renorm_rho0)))
        dt = _capi.HEOM_C64 if self.dtype == np.complex64 else _capi.HEOM_C128
        if self.bf16x3:
            dt = _capi.HEOM_C64_BF16X3
        key = (self._kind, dt, N, self.max_depth) + tuple(getattr(opt, f) for f, _ in
                                                          opt._fields_)
        if self._h and key != getattr(self, "_hkey", None):
            # create-time options changed: a new handle (solvers bound to the old
            # one keep it alive and rebind on their next solve)
            L.heom_destroy(self._h)
            self._h = None
        if not self._h:
            h = ctypes.c_void_p()
            _capi.check(L.heom_create(self._kind, dt, N, self.max_depth, ctypes.byref(opt),
                                      ctypes.byref(h)))
            self._h = h
            self._hkey = key
            self._hgen = next(_heom._generations)
        # the same handle is set up again in place, so solvers created on this
        # qme see the new parameters (qme_solver holds the qme, qme_solver.h:18-53)
        h = self._h
        self._ready = False
        if self.H is None:
            raise HeomError(_capi.HEOM_ERR_ARG, "H not set")
        H = _c128(self.H, (N, N))
        _capi.check(L.heom_set_H(h, _ptr(H)))
        if self.n_noise < 1:
            raise HeomError(_capi.HEOM_ERR_ARG, "alloc_noises not called")
        _capi.check(L.heom_alloc_noises(h, self.n_noise))
        for u in range(self.n_noise):
            ln = int(self.len_gamma[u]) if self.len_gamma[u] else \
                np.atleast_2d(self.gamma[u]).shape[0]
            V = _c128(self.V[u], (N, N))
            g = _c128(self.gamma[u], (ln, ln))
            p0 = _c128(self.phi_0[u], (ln,))
            sg = _c128(self.sigma[u], (ln,))
            S = _c128(self.S[u], (ln, ln))
            A = _c128(self.A[u], (ln, ln))
            sd = complex(self.s_delta[u])
            _capi.check(L.heom_set_noise(h, u, _ptr(V), ln, _ptr(g), _ptr(p0), _ptr(sg),
                                         _ptr(S), _ptr(A), sd.real, sd.imag))
        errors = []
        if self.hrchy_filter is not None:
            user_filter = self.hrchy_filter

            def _filter(idx, n_dim, depth, user):
                try:
                    return 1 if user_filter(np.ctypeslib.as_array(idx, shape=(n_dim,)).copy(),
                                            depth) else 0
                except Exception as e:  # never unwind through the C frame
                    errors.append(e)
                    return 0

            self._filter_cb = _capi.FILTER(_filter)
            _capi.check(L.heom_set_hrchy_filter(h, self._filter_cb, None))
        else:
            _capi.check(L.heom_set_hrchy_filter(h, _capi.FILTER(), None))
            self._filter_cb = None
        _capi.check(L.heom_set_param(h))
        if errors:
            raise errors[0]
        n = ctypes.c_int()
        _capi.check(L.heom_get_n_hrchy(h, ctypes.byref(n)))
        self.n_hrchy = n.value
        _capi.check(L.heom_get_n_dim(h, ctypes.byref(n)))
        self.n_dim = n.value
        self._ready = True

    def load(self, model, engine=None):
        """Fill the public members from a libheom_amd.configs model dict and set_param."""
        self.n_level = model["N"]
        self.n_level_2 = model["N"] ** 2
        self.H = model["H"]
        self.alloc_noises(len(model["noises"]))
        for u, z in enumerate(model["noises"]):
            self.V[u] = z["V"]
            self.gamma[u] = np.atleast_2d(z["gamma"])
            self.len_gamma[u] = self.gamma[u].shape[0]
            self.phi_0[u] = z["phi_0"]
            self.sigma[u] = z["sigma"]
            self.S[u] = z["S"]
            self.A[u] = z["A"]
            self.s_delta[u] = z["s_delta"]
        self.set_param(engine)
        return self

    def get_n_hrchy(self):  # heom.h:52-55
        return self.n_hrchy

    def main_size(self):  # heom_hilb.h:45-49
        return self.n_level * self.n_level * self.n_hrchy

    def temp_size(self):  # heom_hilb.h:51-55 (no host temporaries are needed here)
        return 0

    def hrchy_tables(self):
        K, nh = self.n_dim, self.n_hrchy
        n = np.zeros((nh, K), np.int32)
        p1 = np.zeros((nh + 1, K), np.int32)
        m1 = np.zeros((nh + 1, K), np.int32)
        _capi.check(_capi.load().heom_get_hrchy(self._h, _ptr(n), _ptr(p1), _ptr(m1)))
        return n, p1, m1

    def calc_diff_impl(self, engine, drho_dt, rho, alpha, beta, temp=None):
        """drho_dt <- beta*drho_dt - alpha*D(rho) on host arrays (qme_base.h:74-79)."""
        if not self._ready:
            raise HeomError(_capi.HEOM_ERR_STATE, "set_param has not been called")
        rho = np.ascontiguousarray(rho, dtype=np.complex128).reshape(-1)
        out = np.ascontiguousarray(drho_dt, dtype=np.complex128).reshape(-1)
        a, b = complex(alpha), complex(beta)
        _capi.check(_capi.load().heom_calc_diff_host(self._h, _ptr(out), _ptr(rho), a.real,
                                                     a.imag, b.real, b.imag))
        if isinstance(drho_dt, np.ndarray) and drho_dt.dtype == np.complex128 and \
                drho_dt.flags.c_contiguous and out is not drho_dt:
            drho_dt.reshape(-1)[:] = out
        return out

    def set_closure(self, mlp, features=(), layout=_capi.HEOM_CLOSURE_PER_ADO):
        """Bind a snapshot of a trained closure network (libheom_amd.closure.Mlp)
        to this depth-L heom_hilb; None removes it."""
        f = np.ascontiguousarray(features, np.float32)
        _capi.check(_capi.load().heom_set_closure(self._h, mlp._h if mlp is not None else None,
                                                  _ptr(f) if f.size else None, f.size, layout))

    def calc_diff_dev(self, drho_dev, rho_dev, alpha, beta, stream=None):
        """Device-pointer variant (heom_calc_diff)."""
        a, b = complex(alpha), complex(beta)
        _capi.check(_capi.load().heom_calc_diff(self._h, drho_dev, rho_dev, a.real, a.imag,
                                                b.real, b.imag, stream))


class heom_hilb(_heom):
    """Hilbert-space form, per-ADO commutators (heom_hilb.h:26)."""
    _kind = _capi.HEOM_HILB


class heom_liou(_heom):
    """Liouville-space form, per-ADO superoperators (heom_liou.h:27)."""
    _kind = _capi.HEOM_LIOU


class heom_ado(_heom):
    """Global sparse form, one matrix-vector product (heom_ado.h:27)."""
    _kind = _capi.HEOM_ADO


class _redfield(_heom):
    """Redfield QME (redfield.h:19-174): the qme_base parameter members; set_param
    diagonalises H and builds Lambda_u from the bath correlations.  The state is
    one N x N density matrix (main_size = N^2)."""

    def __init__(self, dtype="complex128", order=row_major, renorm_rho0=False):
        super().__init__(0, dtype=dtype, order=order, renorm_rho0=renorm_rho0)

    def Lambda(self, u):
        """Lambda_u after set_param (redfield.h:30, 166), N x N."""
        return self._lambda_eig(u)[0]

    def eigenvalues(self):
        """Eigenvalues of H from set_param (in the eigensolver's order)."""
        return self._lambda_eig(0)[1]

    def _lambda_eig(self, u):
        if not self._ready:
            raise HeomError(_capi.HEOM_ERR_STATE, "set_param has not been called")
        lam = np.zeros((self.n_level, self.n_level), np.complex128)
        w = np.zeros(self.n_level)
        _capi.check(_capi.load().heom_redfield_lambda(self._h, u, _ptr(lam), _ptr(w)))
        return lam, w


class redfield_hilb(_redfield):
    """Hilbert-space Redfield form (redfield_hilb.h:24)."""
    _kind = _capi.HEOM_REDFIELD_HILB


class redfield_liou(_redfield):
    """Liouville-space Redfield form, one sparse superoperator (redfield_liou.h:25)."""
    _kind = _capi.HEOM_REDFIELD_LIOU


class heom_batch:
    """Independent heom_hilb hierarchies advanced together (no reference counterpart).

    Members need diagonal coupling operators and n_level <= 8.  The batch state
    is the concatenation of the members' states; ``pack`` / ``unpack`` convert
    between the two.  One operator launch per Runge-Kutta stage covers every
    member, so a parameter sweep of small systems fills the device.
    """

    def __init__(self, members):
        members = list(members)
        for q in members:
            if not getattr(q, "_ready", False):
                raise HeomError(_capi.HEOM_ERR_STATE, "member without set_param")
        L = _capi.load()
        arr = (ctypes.c_void_p * len(members))(*[q._h.value for q in members])
        h = ctypes.c_void_p()
        _capi.check(L.heom_batch_create(arr, len(members), ctypes.byref(h)))
        self._h = h
        self._ready = True
        self.members = members
        n = len(members)
        self.elem_base = np.zeros(n, np.int64)
        self.rho_s_offset = np.zeros(n, np.int32)
        self.n_levels = np.zeros(n, np.int32)
        _capi.check(L.heom_batch_layout(h, _ptr(self.elem_base), _ptr(self.rho_s_offset),
                                        _ptr(self.n_levels)))
        m = ctypes.c_longlong()
        _capi.check(L.heom_main_size(h, ctypes.byref(m)))
        self._main_size = m.value
        self.n_level = int(self.n_levels[0])
        self.dtype = members[0].dtype
        nh = ctypes.c_int()
        _capi.check(L.heom_get_n_hrchy(h, ctypes.byref(nh)))
        self.n_hrchy = nh.value

    def __del__(self):
        if getattr(self, "_h", None):
            _capi.load().heom_destroy(self._h)
            self._h = None

    def main_size(self):
        return self._main_size

    def tuning(self):
        """The autotuner's choice and every candidate it timed:
        ({"items_per_step": k, "waves_per_group": w, "grid": g},
         [(items_per_step, waves_per_group, grid, us_per_pass), ...])."""
        L = _capi.load()
        cap = 64
        shape = np.zeros(3, np.int32)
        tried = np.zeros(3 * cap, np.int32)
        us = np.zeros(cap, np.float32)
        n = ctypes.c_int()
        _capi.check(L.heom_batch_tuning(self._h, _ptr(shape), _ptr(tried), _ptr(us), cap,
                                        ctypes.byref(n)))
        k = min(n.value, cap)
        return ({"items_per_step": int(shape[0]), "waves_per_group": int(shape[1]),
                 "grid": int(shape[2])},
                [(int(tried[3 * i]), int(tried[3 * i + 1]), int(tried[3 * i + 2]), float(us[i]))
                 for i in range(k)])

    def set_tuning(self, items_per_step, waves_per_group, grid):
        """Override the autotuner's tile shape and grid (results change only by rounding)."""
        _capi.check(_capi.load().heom_batch_set_tuning(self._h, int(items_per_step),
                                                       int(waves_per_group), int(grid)))

    def rho_s_size(self):
        return int(np.sum(self.n_levels.astype(np.int64) ** 2))

    def pack(self, rhos):
        out = np.zeros(self._main_size, np.complex128)
        for m, r in enumerate(rhos):
            r = np.asarray(r, np.complex128).reshape(-1)
            out[self.elem_base[m]:self.elem_base[m] + r.size] = r
        return out

    def unpack(self, rho):
        ends = list(self.elem_base[1:]) + [self._main_size]
        return [np.asarray(rho)[b:e] for b, e in zip(self.elem_base, ends)]

    def split_rho_s(self, block):
        """(n_t, rho_s_size) gathered rho_S -> list of (n_t, N_m, N_m) per member."""
        block = np.asarray(block).reshape(block.shape[0], -1)
        res = []
        for o, n in zip(self.rho_s_offset, self.n_levels):
            res.append(block[:, o:o + n * n].reshape(-1, n, n))
        return res

    def calc_diff_impl(self, engine, drho_dt, rho, alpha, beta, temp=None):
        return _heom.calc_diff_impl(self, engine, drho_dt, rho, alpha, beta, temp)

    def calc_diff_dev(self, drho_dev, rho_dev, alpha, beta, stream=None):
        return _heom.calc_diff_dev(self, drho_dev, rho_dev, alpha, beta, stream)


class rk4:
    """Classic RK4 (solver/rk4.h); kwargs: dt."""
    _kind = _capi.HEOM_RK4


class rkdp:
    """Dormand-Prince 5(4) (solver/rkdp.h); kwargs: dt (initial), atol, rtol."""
    _kind = _capi.HEOM_RKDP


class lsrk4:
    """Low-storage four-stage RK (solver/lsrk4.h); kwargs: dt."""
    _kind = _capi.HEOM_LSRK4


class qme_solver:
    """Binds engine + qme + solver (qme_solver.h:18-53)."""

    def __init__(self, engine, qme, solver):
        if not qme._ready:
            raise HeomError(_capi.HEOM_ERR_STATE, "set_param has not been called")
        self.engine, self.qme, self.solver = engine, qme, solver
        self._graphs = True
        self._bind()

    def _bind(self):
        s = ctypes.c_void_p()
        _capi.check(_capi.load().heom_solver_create(self.qme._h, self.solver._kind,
                                                    ctypes.byref(s)))
        if getattr(self, "_s", None):
            _capi.load().heom_solver_destroy(self._s)
        self._s = s
        self._qgen = getattr(self.qme, "_hgen", 0)
        if not self._graphs:
            _capi.check(_capi.load().heom_solver_set_graphs(self._s, 0))

    def _sync(self):
        """Rebind when the qme's set_param replaced its handle (changed dtype,
        order, device or options); a qme set up again in place keeps the solver,
        whose buffers must still fit its main_size (else HEOM_ERR_STATE)."""
        if not self.qme._ready:
            raise HeomError(_capi.HEOM_ERR_STATE, "set_param has not been called")
        if getattr(self.qme, "_hgen", 0) != self._qgen:
            self._bind()

    def __del__(self):
        if getattr(self, "_s", None):
            _capi.load().heom_solver_destroy(self._s)
            self._s = None

    def calc_diff(self, drho_dt, rho):
        """One derivative evaluation, drho_dt = -D(rho).

        The reference forwards to an empty stub (solver_base.h:55-62); this
        computes the evident intent, calc_diff_impl(alpha=1, beta=0).
        """
        return self.qme.calc_diff_impl(self.engine, drho_dt, rho, 1.0, 0.0)

    def _kw(self, kwargs):
        dt = float(kwargs["dt"])
        if self.solver._kind == _capi.HEOM_RKDP:
            return dt, float(kwargs["atol"]), float(kwargs["rtol"])
        return dt, 0.0, 0.0

    def solve(self, rho, t_list, callback=None, kwargs=None, full_copy=True):
        """solver_base::solve (solver_base.h:64-79): rho is propagated in place."""
        self._sync()
        dt, atol, rtol = self._kw(kwargs or {})
        if not (isinstance(rho, np.ndarray) and rho.dtype == np.complex128 and
                rho.flags.c_contiguous):
            raise TypeError("rho must be a C-contiguous complex128 numpy array")
        t = np.ascontiguousarray(np.asarray(t_list, dtype=np.float64))
        N2 = self.qme.n_level ** 2
        errors = []

        def _cb(tt, rs, rf, user):
            if callback is None:
                return
            try:
                callback(tt)
            except Exception as e:  # never unwind through the C frame
                errors.append(e)

        cfun = _capi.CALLBACK(_cb)
        rc = _capi.load().heom_solve(self._s, _ptr(rho), _ptr(t), t.size, cfun, None, dt,
                                     atol, rtol, int(bool(full_copy)))
        if errors:
            raise errors[0]
        _capi.check(rc)
        return rho

    def solve_rho_s(self, rho, t_list, kwargs):
        """Propagate a copy of rho; return (final state, rho_S at every t).

        Device-resident path: rho_S is gathered in HBM and copied once at the end.
        """
        self._sync()
        dt, atol, rtol = self._kw(kwargs)
        L = _capi.load()
        rho = np.array(rho, dtype=np.complex128, copy=True).reshape(-1)
        t = np.ascontiguousarray(np.asarray(t_list, dtype=np.float64))
        if isinstance(self.qme, heom_batch):
            out = np.zeros((t.size, self.qme.rho_s_size()), np.complex128)
        else:
            N = self.qme.n_level
            out = np.zeros((t.size, N, N), np.complex128)
        d = ctypes.c_void_p()
        nb = rho.nbytes
        _capi.check(L.heom_dev_alloc(ctypes.byref(d), nb))
        try:
            _capi.check(L.heom_memcpy_h2d(d, _ptr(rho), nb))
            rc = L.heom_solve_dev(self._s, d, _ptr(t), t.size, dt, atol, rtol, _ptr(out))
            _capi.check(L.heom_memcpy_d2h(_ptr(rho), d, nb))
            _capi.check(rc)
        finally:
            L.heom_dev_free(d)
        return rho, out

    def stats(self):
        st = _capi.heom_stats()
        _capi.check(_capi.load().heom_solver_stats(self._s, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def reset(self):
        _capi.check(_capi.load().heom_solver_reset(self._s))

    def set_graphs(self, enable):
        """Replay fixed stage sequences from HIP graphs (default) or launch directly."""
        self._graphs = bool(enable)
        _capi.check(_capi.load().heom_solver_set_graphs(self._s, int(bool(enable))))
