"""ctypes front end of the CPU oracle (oracle/heom_oracle.cc).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module, and only as the checker;
the product library (``libheom_amd``) never loads it.

The oracle is an FP64 restatement of libheom's propagation path:
  heom::set_param (heom.h:57-139), alloc_hrchy_space (hrchy_space.cc:137-231),
  heom_hilb/heom_liou/heom_ado calc_diff_impl (heom_hilb.h:70-215,
  heom_liou.h:121-258, heom_ado.h:191-202), rk4 (rk4.h:40-65),
  rkdp (rkdp.h:81-179), lsrk4 (lsrk4.h:33-48), solver_base::solve
  (solver_base.h:64-79), and the Redfield QMEs (redfield.h:33-174,
  redfield_hilb.h:66-89, redfield_liou.h:56-102).
"""
import ctypes
import os

import numpy as np

# int filter(const int* index, int n_dim, int depth, void* user)
FILTER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                          ctypes.c_void_p)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libheom_oracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libhrchy_ref.so")

HILB, LIOU, ADO = 0, 1, 2
REDFIELD_HILB, REDFIELD_LIOU = 3, 4
RK4, RKDP, LSRK4 = 0, 1, 2

_lib = None


class OrcStats(ctypes.Structure):
    _fields_ = [("n_rhs", ctypes.c_long), ("n_steps", ctypes.c_long),
                ("n_accepted", ctypes.c_long), ("n_rejected", ctypes.c_long),
                ("gave_up", ctypes.c_int)]


def build():
    """Compile the oracle (and oracle/_ref when the reference is present)."""
    import subprocess
    subprocess.run(["make", "-C", _HERE, "all"], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ip, ip, ip, ip]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_options.argtypes = [vp, ip, ip]
        L.orc_set_filter.argtypes = [vp, FILTER, vp]
        L.orc_set_threads.argtypes = [vp, ip]
        L.orc_set_sum_order.argtypes = [vp, ip]
        L.orc_set_H.argtypes = [vp, vp]
        L.orc_alloc_noises.argtypes = [vp, ip]
        L.orc_set_noise.argtypes = [vp, ip, vp, ip, vp, vp, vp, vp, vp, dp, dp]
        L.orc_set_param.restype = ip
        L.orc_set_param.argtypes = [vp]
        L.orc_hrchy_count.restype = ctypes.c_longlong
        L.orc_hrchy_count.argtypes = [ip, ip]
        L.orc_n_hrchy.restype = ip
        L.orc_n_hrchy.argtypes = [vp]
        L.orc_n_dim.restype = ip
        L.orc_n_dim.argtypes = [vp]
        L.orc_main_size.restype = ctypes.c_long
        L.orc_main_size.argtypes = [vp]
        L.orc_ado_nnz.restype = ctypes.c_long
        L.orc_ado_nnz.argtypes = [vp]
        L.orc_hrchy_tables.argtypes = [vp, vp, vp, vp]
        L.orc_redfield_lambda.restype = ip
        L.orc_redfield_lambda.argtypes = [vp, ip, vp, vp]
        L.orc_redfield_nnz.restype = ctypes.c_long
        L.orc_redfield_nnz.argtypes = [vp]
        L.orc_ngamma_diag.argtypes = [vp, vp]
        L.orc_calc_diff.argtypes = [vp, vp, vp, dp, dp, dp, dp]
        L.orc_errnrm1.restype = dp
        L.orc_errnrm1.argtypes = [vp, vp, vp, dp, dp, ctypes.c_long]
        L.orc_rk4_step.argtypes = [vp, vp, dp]
        L.orc_lsrk4_step.argtypes = [vp, vp, dp]
        L.orc_rkdp_step.restype = ip
        L.orc_rkdp_step.argtypes = [vp, vp, vp, dp, vp, dp, dp, ctypes.POINTER(OrcStats)]
        L.orc_solve.restype = ip
        L.orc_solve.argtypes = [vp, ip, vp, vp, ip, dp, dp, dp, vp, vp,
                                ctypes.POINTER(OrcStats)]
        _lib = L
    return _lib


def _c128(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def hrchy_count(level, dim):
    return int(lib().orc_hrchy_count(level, dim))


def errnrm1(e, x1, x2, atol, rtol):
    e, x1, x2 = _c128(e), _c128(x1), _c128(x2)
    return lib().orc_errnrm1(_ptr(e), _ptr(x1), _ptr(x2), atol, rtol, e.size)


class Oracle:
    """One HEOM model (hilb / liou / ado) on the CPU oracle.

    Parameters follow qme_base (qme_base.h:32-48): ``H`` (N x N), and per noise
    ``V`` (N x N), ``gamma`` (len x len), ``phi_0`` (len), ``sigma`` (len),
    ``S``/``A`` (len x len), ``s_delta`` (scalar).
    """

    def __init__(self, kind, n_level, max_depth, n_threads=1, parity_q1=True,
                 sqrt_norm=False):
        self._L = lib()
        self.kind, self.N, self.max_depth = kind, n_level, max_depth
        self.h = self._L.orc_create(kind, n_level, max_depth, n_threads)
        self._L.orc_set_options(self.h, int(parity_q1), int(sqrt_norm))
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_destroy(self.h)
            self.h = None

    def set_threads(self, n):
        self._L.orc_set_threads(self.h, n)

    def set_sum_order(self, reverse):
        """Sensitivity experiments: sum the heom_hilb coupling terms of the
        noises in reverse order (the same operator, other rounding)."""
        self._L.orc_set_sum_order(self.h, int(bool(reverse)))
        return self

    def set_filter(self, fn):
        """hrchy_filter(index ndarray, depth) -> bool (hrchy_space.h:29-36); before load()."""
        cb = FILTER(lambda idx, n, depth, user:
                    1 if fn(np.ctypeslib.as_array(idx, shape=(n,)).copy(), depth) else 0)
        self._keep.append(cb)
        self._L.orc_set_filter(self.h, cb, None)
        return self

    def set_H(self, H):
        H = _c128(H, (self.N, self.N))
        self._L.orc_set_H(self.h, _ptr(H))

    def alloc_noises(self, n):
        self._L.orc_alloc_noises(self.h, n)

    def set_noise(self, u, V, gamma, phi_0, sigma, S, A, s_delta):
        V = _c128(V, (self.N, self.N))
        gamma = np.atleast_2d(_c128(gamma))
        ln = gamma.shape[0]
        phi_0 = _c128(phi_0, (ln,))
        sigma = _c128(sigma, (ln,))
        S = _c128(S, (ln, ln))
        A = _c128(A, (ln, ln))
        sd = complex(s_delta)
        self._L.orc_set_noise(self.h, u, _ptr(V), ln, _ptr(gamma), _ptr(phi_0), _ptr(sigma),
                              _ptr(S), _ptr(A), sd.real, sd.imag)

    def set_param(self):
        if self._L.orc_set_param(self.h) != 0:
            raise RuntimeError("oracle set_param failed")
        self.n_hrchy = self._L.orc_n_hrchy(self.h)
        self.n_dim = self._L.orc_n_dim(self.h)
        self.main_size = self._L.orc_main_size(self.h)

    def load(self, model):
        """Fill from a libheom_amd.configs model description (dict)."""
        self.set_H(model["H"])
        self.alloc_noises(len(model["noises"]))
        for u, z in enumerate(model["noises"]):
            self.set_noise(u, z["V"], z["gamma"], z["phi_0"], z["sigma"], z["S"], z["A"],
                           z["s_delta"])
        self.set_param()
        return self

    def redfield_lambda(self, u):
        """(Lambda_u, eigenvalues of H) of a Redfield model (redfield.h:144-168)."""
        lam = np.zeros((self.N, self.N), np.complex128)
        w = np.zeros(self.N)
        if self._L.orc_redfield_lambda(self.h, u, _ptr(lam), _ptr(w)) != 0:
            raise ValueError("not a Redfield model or noise index out of range")
        return lam, w

    def redfield_nnz(self):
        return int(self._L.orc_redfield_nnz(self.h))

    def tables(self):
        K, nh = self.n_dim, self.n_hrchy
        n = np.zeros((nh, K), np.int32)
        p1 = np.zeros((nh + 1, K), np.int32)
        m1 = np.zeros((nh + 1, K), np.int32)
        self._L.orc_hrchy_tables(self.h, _ptr(n), _ptr(p1), _ptr(m1))
        return n, p1, m1

    def ngamma_diag(self):
        out = np.zeros(self.n_hrchy, np.complex128)
        self._L.orc_ngamma_diag(self.h, _ptr(out))
        return out

    def calc_diff(self, rho, alpha=1.0, beta=0.0, drho=None):
        """drho <- beta*drho - alpha*D(rho) (heom_hilb.h:210-211)."""
        rho = _c128(rho).reshape(-1)
        if drho is None:
            drho = np.zeros_like(rho)
        else:
            drho = _c128(drho).reshape(-1).copy()
        a, b = complex(alpha), complex(beta)
        self._L.orc_calc_diff(self.h, _ptr(drho), _ptr(rho), a.real, a.imag, b.real, b.imag)
        return drho

    def rk4_step(self, rho, dt):
        rho = _c128(rho).reshape(-1).copy()
        self._L.orc_rk4_step(self.h, _ptr(rho), dt)
        return rho

    def lsrk4_step(self, rho, dt):
        rho = _c128(rho).reshape(-1).copy()
        self._L.orc_lsrk4_step(self.h, _ptr(rho), dt)
        return rho

    def rkdp_step(self, rho, t, t_bound, dt, atol, rtol):
        rho = _c128(rho).reshape(-1).copy()
        tt = ctypes.c_double(t)
        dd = ctypes.c_double(dt)
        st = OrcStats()
        rc = self._L.orc_rkdp_step(self.h, _ptr(rho), ctypes.byref(tt), t_bound,
                                   ctypes.byref(dd), atol, rtol, ctypes.byref(st))
        return rho, tt.value, dd.value, st, rc

    def solve(self, rho, t_list, solver, dt, atol=0.0, rtol=0.0, dt_save=None):
        """solver_base::solve; returns (rho_final, rho_S at each t, stats, dt_save)."""
        rho = _c128(rho).reshape(-1).copy()
        t_list = np.ascontiguousarray(np.asarray(t_list, dtype=np.float64))
        n_t = t_list.size
        rs = np.zeros((n_t, self.N, self.N), np.complex128)
        dts = ctypes.c_double(-1.0 if dt_save is None else dt_save)
        st = OrcStats()
        rc = self._L.orc_solve(self.h, solver, _ptr(rho), _ptr(t_list), n_t, dt, atol, rtol,
                               _ptr(rs), ctypes.byref(dts), ctypes.byref(st))
        return rho, rs, st, dts.value, rc


def ref_hrchy_filtered(n_dim, max_depth, fn):
    """Reference src/hrchy_space.cc enumeration with an hrchy_filter."""
    if not os.path.exists(_REF_PATH):
        return None
    L = ctypes.CDLL(_REF_PATH)
    L.ref_hrchy_filtered.restype = ctypes.c_int
    L.ref_hrchy_filtered.argtypes = [ctypes.c_int, ctypes.c_int, FILTER, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int]
    cb = FILTER(lambda idx, n, depth, user:
                1 if fn(np.ctypeslib.as_array(idx, shape=(n,)).copy(), depth) else 0)
    nh = L.ref_hrchy_filtered(n_dim, max_depth, cb, None, None, None, None, 0)
    n = np.zeros((nh, n_dim), np.int32)
    p1 = np.zeros((nh + 1, n_dim), np.int32)
    m1 = np.zeros((nh + 1, n_dim), np.int32)
    L.ref_hrchy_filtered(n_dim, max_depth, cb, None, _ptr(n), _ptr(p1), _ptr(m1), nh)
    return n, p1, m1


def ref_hrchy(n_dim, max_depth):
    """Reference src/hrchy_space.cc enumeration (compiled into oracle/_ref)."""
    if not os.path.exists(_REF_PATH):
        return None
    L = ctypes.CDLL(_REF_PATH)
    L.ref_hrchy.restype = ctypes.c_int
    L.ref_hrchy.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_void_p, ctypes.c_int]
    nh = L.ref_hrchy(n_dim, max_depth, None, None, None, 0)
    n = np.zeros((nh, n_dim), np.int32)
    p1 = np.zeros((nh + 1, n_dim), np.int32)
    m1 = np.zeros((nh + 1, n_dim), np.int32)
    L.ref_hrchy(n_dim, max_depth, _ptr(n), _ptr(p1), _ptr(m1), nh)
    return n, p1, m1
