// HIP kernels of the libheom_amd engine (gfx950).
//
// One kernel pass = one HEOM derivative evaluation fused with the Runge-Kutta
// stage work that consumes it:
//   D_n   = the HEOM operator of heom_hilb.h:108-209 (or its Liouville /
//           global-sparse forms, heom_liou.h:151-255, heom_ado.h:191-202)
//   k     = -dt*D  (calc_diff_impl with alpha = dt, beta = 0; rk4.h:48)
//   y_i+1 = y + sum_j a_ij k_j  (the next stage input, written in the same pass)
//   e     = sum_j E_j k_j and the partial sums of the error norm (rkdp.h:142-153)
// The last workgroup of the final Dormand-Prince stage reduces the partials
// in a fixed order and runs the step control on the device (rkdp.h:97-177).
//
// Threads map to matrix elements: thread (a, e) of a workgroup handles element
// e = i*N + j of the a-th ADO of the group, so every ADO block (row-major
// N x N, the reference layout) is read and written with contiguous lanes.
#include <hip/hip_runtime.h>

#include "../../include/libheom_amd.h"
#include "common.h"
#include "kernels.h"

namespace hamd {

namespace {

__device__ __forceinline__ void* slot_ptr(const Slots& S, int s, int cur) {
  if (s < 4) return S.p[(s & 2) + ((s & 1) ^ cur)];
  return S.p[s];
}

__device__ __forceinline__ z128 ld(const void* p, int type, long long i) {
  if (type == ST_C128) return static_cast<const z128*>(p)[i];
  z64 v = static_cast<const z64*>(p)[i];
  return z128{static_cast<double>(v.re), static_cast<double>(v.im)};
}

__device__ __forceinline__ void st(void* p, int type, long long i, z128 v) {
  if (type == ST_C128)
    static_cast<z128*>(p)[i] = v;
  else
    static_cast<z64*>(p)[i] = z64{static_cast<float>(v.re), static_cast<float>(v.im)};
}

__device__ __forceinline__ double zabs(z128 a) { return sqrt(a.re * a.re + a.im * a.im); }

// Epilogue shared by all operator kernels.  D is the operator applied to the
// stage input at flat element gi; returns w^2 of the error norm (or 0).
template <typename R>
__device__ __forceinline__ double epilogue(const StageArgs& A, const Slots& S, int cur,
                                           long long gi, cx<R> Dr, z128 alpha) {
  z128 D = conv<double>(Dr);
  // -alpha*D
  z128 k = z128{-(alpha.re * D.re - alpha.im * D.im), -(alpha.re * D.im + alpha.im * D.re)};
  if (A.mode == MODE_DIFF) {
    const int o = A.out_slot;
    void* op = slot_ptr(S, o, cur);
    z128 v = k;
    if (A.b_re != 0.0 || A.b_im != 0.0) {
      z128 old = ld(op, S.type[o], gi);
      v = z128{A.b_re * old.re - A.b_im * old.im + k.re, A.b_re * old.im + A.b_im * old.re + k.im};
    }
    st(op, S.type[o], gi, v);
    return 0.0;
  }
  if (A.k_slot >= 0) st(slot_ptr(S, A.k_slot, cur), S.type[A.k_slot], gi, k);
  for (int c = 0; c < A.ncombo; ++c) {
    const Combo& cb = A.combo[c];
    z128 v{0.0, 0.0};
    for (int s = 0; s < cb.nsrc; ++s) {
      z128 x = ld(slot_ptr(S, cb.src[s], cur), S.type[cb.src[s]], gi);
      if (s == 0) {
        v = cb.c[0] * x;
      } else {
        v.re += cb.c[s] * x.re;
        v.im += cb.c[s] * x.im;
      }
    }
    v.re += cb.ck * k.re;
    v.im += cb.ck * k.im;
    st(slot_ptr(S, cb.dst, cur), S.type[cb.dst], gi, v);
    if (cb.dst2 >= 0) st(slot_ptr(S, cb.dst2, cur), S.type[cb.dst2], gi, v);
  }
  if (A.do_err) {
    z128 e{0.0, 0.0};
    for (int s = 0; s < A.ne; ++s) {
      z128 x = ld(slot_ptr(S, A.esrc[s], cur), S.type[A.esrc[s]], gi);
      if (s == 0) {
        e = A.ec[0] * x;
      } else {
        e.re += A.ec[s] * x.re;
        e.im += A.ec[s] * x.im;
      }
    }
    e.re += A.ek * k.re;
    e.im += A.ek * k.im;
    double y1 = zabs(ld(slot_ptr(S, A.ynew_slot, cur), S.type[A.ynew_slot], gi));
    double y0 = zabs(ld(slot_ptr(S, A.yold_slot, cur), S.type[A.yold_slot], gi));
    double w = zabs(e) / (A.atol + A.rtol * (y1 > y0 ? y1 : y0));
    return w * w;
  }
  return 0.0;
}

// Step control of rkdp::solve_adaptive_step (rkdp.h:97-177) for the attempt
// just finished, followed by the prologue of the next attempt.  The
// std::min/std::max argument orders of the reference are kept so that a NaN
// error norm is handled the same way (factor 10 / 0.2).
__device__ void dopri_control(Ctrl* c, double err) {
  const double safety = 0.9, expo = -1.0 / 5.0, dt_min = 1e-16;
  c->err = err;
  c->attempts += 1;
  double dt = c->dt;
  double f = pow(err, expo);
  if (err < 1.0) {
    c->t += dt;
    dt *= safety * ((f < 10.0) ? f : 10.0);
    c->cur ^= 1;
    c->accepted += 1;
    c->count = 0;
    if (c->t >= c->t_bound) {
      c->dt = dt;
      c->done = 1;
      return;
    }
  } else {
    dt *= safety * ((0.2 < f) ? f : 0.2);
    if (dt < dt_min) dt = dt_min;
    c->rejected += 1;
  }
  c->count += 1;
  if (c->count > 1000) {
    c->t = c->t_bound;
    c->done = 1;
    c->gave_up = 1;
    c->dt = dt;
    return;
  }
  if (c->t + dt > c->t_bound) dt = c->t_bound - c->t;
  c->dt = dt;
}

// Block reduction of the error-norm partials; the last block to finish sums
// all partials in block order (deterministic) and runs the control.
__device__ void finish_error(const StageArgs& A, double v) {
  __shared__ double s_red[16];
  __shared__ int s_last;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
  if (lane == 0) s_red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += s_red[w];
    __hip_atomic_store(&A.partials[blockIdx.x], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned ticket = atomicAdd(A.counter, 1u);
    s_last = (ticket == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  double x = 0.0;
  for (int b = threadIdx.x; b < static_cast<int>(gridDim.x); b += blockDim.x)
    x += __hip_atomic_load(&A.partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  __syncthreads();
  if (lane == 0) s_red[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += s_red[w];
    double err = sqrt(t * A.inv_n);
    if (A.control)
      dopri_control(A.ctrl, err);
    else if (A.ctrl)
      A.ctrl->err = err;
    __hip_atomic_store(A.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int N>
struct HilbCfg {
  static constexpr int N2 = N * N;
  static constexpr int BS = N2 <= 64 ? 256 : ((N2 + 63) / 64) * 64;
  static constexpr int APB = BS / N2;  // ADOs per workgroup
};

// ---------------------------------------------------------------------------
// heom_hilb: per-ADO commutator form (heom_hilb.h:96-212), one element per lane
// ---------------------------------------------------------------------------
template <int N, typename R, bool DIAGV>
__global__ void __launch_bounds__(HilbCfg<N>::BS)
    hilb_kernel(const FormDev d, const Slots S, const StageArgs A) {
  using C = cx<R>;
  constexpr int N2 = N * N, APB = HilbCfg<N>::APB;
  __shared__ C s_rho[APB * N2];
  __shared__ C s_X[DIAGV ? 1 : APB * N2];
  __shared__ C s_Y[DIAGV ? 1 : APB * N2];

  int cur = 0;
  z128 alpha{A.a_re, A.a_im};
  if (A.ctrl) {
    if (A.ctrl->done) return;
    cur = A.ctrl->cur;
    if (A.mode == MODE_STAGE) alpha = z128{A.ctrl->dt, 0.0};
  }
  const int tid = threadIdx.x;
  const int la = tid / N2;
  const int e = tid - la * N2;
  const int i = e / N, j = e - (e / N) * N;
  const bool lane_ok = la < APB;

  C Hr[N], Hc[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    Hr[k] = lane_ok ? conv<R>(d.H[i * N + k]) : C{0, 0};
    Hc[k] = lane_ok ? conv<R>(d.H[k * N + j]) : C{0, 0};
  }
  const C dterm = (DIAGV && lane_ok) ? conv<R>(d.dterm[e]) : C{0, 0};
  const C* in = static_cast<const C*>(slot_ptr(S, A.in_slot, cur));

  double err_acc = 0.0;
  const int n_groups = (d.n_hrchy + APB - 1) / APB;
  for (int g = blockIdx.x; g < n_groups; g += gridDim.x) {
    const int n = g * APB + la;
    const bool ok = lane_ok && n < d.n_hrchy;
    const long long gi = static_cast<long long>(n) * N2 + e;
    const C r = ok ? in[gi] : C{0, 0};
    if (lane_ok) s_rho[tid] = r;
    __syncthreads();
    C acc{0, 0};
    if (lane_ok) {
      const C* b = s_rho + la * N2;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        cmac(acc, Hr[k], b[k * N + j]);
        cmac(acc, C{-b[i * N + k].re, -b[i * N + k].im}, Hc[k]);
      }
    }
    C D = times_i(acc);
    int l0 = 0, l1 = 0;
    if (ok) {
      const C gm = conv<R>(d.gam[n]);
      cmac(D, gm + dterm, r);
      l0 = d.link_ptr[n];
      l1 = d.link_ptr[n + 1];
    }
    if (DIAGV) {
      for (int l = l0; l < l1; ++l) {
        const Link lk = d.links[l];
        const Coef c = d.coef[lk.cid];
        const C x = in[static_cast<long long>(lk.nbr) * N2 + e];
        C cf = conv<R>(c.dir);
        if (c.u < d.n_noise) {
          const C vi = conv<R>(d.Vd[c.u * N + i]);
          const C vj = conv<R>(d.Vd[c.u * N + j]);
          cf = cf + times_i(vi - vj) * conv<R>(c.phi) + (vi + vj) * conv<R>(c.psi);
        }
        cmac(D, cf, x);
      }
    } else {
      int l = l0;
      for (int u = 0; u < d.n_noise; ++u) {
        C phi{0, 0}, psi{0, 0};
        while (l < l1) {
          const Link lk = d.links[l];
          const Coef c = d.coef[lk.cid];
          if (c.u != u) break;
          const C x = in[static_cast<long long>(lk.nbr) * N2 + e];
          cmac(phi, conv<R>(c.phi), x);
          cmac(psi, conv<R>(c.psi), x);
          ++l;
        }
        const z128* V = d.V + static_cast<long long>(u) * N2;
        C cc{0, 0};
        if (lane_ok) {
          const C* b = s_rho + la * N2;
#pragma unroll
          for (int k = 0; k < N; ++k) {
            cmac(cc, conv<R>(V[i * N + k]), b[k * N + j]);
            const C t = b[i * N + k];
            cmac(cc, C{-t.re, -t.im}, conv<R>(V[k * N + j]));
          }
        }
        const C sd = conv<R>(d.sdelta[u]);
        const C dcc = sd * cc;
        const C iphi = times_i(phi);
        // i[V,Phi] + {V,Psi} + delta[V,[V,rho]] = V X + Y V
        const C X = iphi + psi + dcc;
        const C Y = psi - iphi - dcc;
        if (lane_ok) {
          s_X[tid] = X;
          s_Y[tid] = Y;
        }
        __syncthreads();
        if (lane_ok) {
          const C* bx = s_X + la * N2;
          const C* by = s_Y + la * N2;
#pragma unroll
          for (int k = 0; k < N; ++k) {
            cmac(D, conv<R>(V[i * N + k]), bx[k * N + j]);
            cmac(D, by[i * N + k], conv<R>(V[k * N + j]));
          }
        }
        __syncthreads();
      }
      for (; l < l1; ++l) {
        const Link lk = d.links[l];
        const Coef c = d.coef[lk.cid];
        cmac(D, conv<R>(c.dir), in[static_cast<long long>(lk.nbr) * N2 + e]);
      }
    }
    if (ok) err_acc += epilogue<R>(A, S, cur, gi, D, alpha);
    __syncthreads();
  }
  if (A.do_err) finish_error(A, err_acc);
}

// ---------------------------------------------------------------------------
// heom_liou: per-ADO superoperator form (heom_liou.h:151-255), N <= 16
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256)
    liou_kernel(const FormDev d, const Slots S, const StageArgs A) {
  using C = cx<R>;
  __shared__ C s_rho[256];
  __shared__ C s_phi[256];
  __shared__ C s_psi[256];
  int cur = 0;
  z128 alpha{A.a_re, A.a_im};
  if (A.ctrl) {
    if (A.ctrl->done) return;
    cur = A.ctrl->cur;
    if (A.mode == MODE_STAGE) alpha = z128{A.ctrl->dt, 0.0};
  }
  const int N2 = d.N2;
  const int APB = 256 / N2;
  const int tid = threadIdx.x;
  const int la = tid / N2;
  const int e = tid - la * N2;
  const bool lane_ok = la < APB;
  const C* in = static_cast<const C*>(slot_ptr(S, A.in_slot, cur));
  const int stride = N2 + 1;
  double err_acc = 0.0;
  const int n_groups = (d.n_hrchy + APB - 1) / APB;
  for (int g = blockIdx.x; g < n_groups; g += gridDim.x) {
    const int n = g * APB + la;
    const bool ok = lane_ok && n < d.n_hrchy;
    const long long gi = static_cast<long long>(n) * N2 + e;
    const C r = ok ? in[gi] : C{0, 0};
    if (lane_ok) s_rho[tid] = r;
    __syncthreads();
    C D{0, 0};
    int l0 = 0, l1 = 0;
    if (ok) {
      for (int p = d.sop_ptr[e]; p < d.sop_ptr[e + 1]; ++p)
        cmac(D, conv<R>(d.sop_val[p]), s_rho[la * N2 + d.sop_col[p]]);
      cmac(D, conv<R>(d.gam[n]), r);
      l0 = d.link_ptr[n];
      l1 = d.link_ptr[n + 1];
    }
    int l = l0;
    for (int u = 0; u < d.n_noise; ++u) {
      C phi{0, 0}, psi{0, 0};
      while (l < l1) {
        const Link lk = d.links[l];
        const Coef c = d.coef[lk.cid];
        if (c.u != u) break;
        const C x = in[static_cast<long long>(lk.nbr) * N2 + e];
        cmac(phi, conv<R>(c.phi), x);
        cmac(psi, conv<R>(c.psi), x);
        ++l;
      }
      if (lane_ok) {
        s_phi[tid] = phi;
        s_psi[tid] = psi;
      }
      __syncthreads();
      if (ok) {
        const int* pp = d.sop_ptr + (1 + 2 * u) * stride;
        for (int p = pp[e]; p < pp[e + 1]; ++p)
          cmac(D, conv<R>(d.sop_val[p]), s_phi[la * N2 + d.sop_col[p]]);
        const int* pq = d.sop_ptr + (2 + 2 * u) * stride;
        for (int p = pq[e]; p < pq[e + 1]; ++p)
          cmac(D, conv<R>(d.sop_val[p]), s_psi[la * N2 + d.sop_col[p]]);
      }
      __syncthreads();
    }
    for (; l < l1; ++l) {
      const Link lk = d.links[l];
      const Coef c = d.coef[lk.cid];
      cmac(D, conv<R>(c.dir), in[static_cast<long long>(lk.nbr) * N2 + e]);
    }
    if (ok) err_acc += epilogue<R>(A, S, cur, gi, D, alpha);
    __syncthreads();
  }
  if (A.do_err) finish_error(A, err_acc);
}

// ---------------------------------------------------------------------------
// heom_ado: one global CSR matrix-vector product (heom_ado.h:191-202)
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256)
    ado_kernel(const FormDev d, const Slots S, const StageArgs A) {
  using C = cx<R>;
  int cur = 0;
  z128 alpha{A.a_re, A.a_im};
  if (A.ctrl) {
    if (A.ctrl->done) return;
    cur = A.ctrl->cur;
    if (A.mode == MODE_STAGE) alpha = z128{A.ctrl->dt, 0.0};
  }
  const C* in = static_cast<const C*>(slot_ptr(S, A.in_slot, cur));
  double err_acc = 0.0;
  const long long rows = d.main_size;
  for (long long r = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; r < rows;
       r += static_cast<long long>(gridDim.x) * blockDim.x) {
    C D{0, 0};
    for (long long p = d.R_ptr[r]; p < d.R_ptr[r + 1]; ++p)
      cmac(D, conv<R>(d.R_val[p]), in[d.R_col[p]]);
    err_acc += epilogue<R>(A, S, cur, r, D, alpha);
  }
  if (A.do_err) finish_error(A, err_acc);
}

__global__ void ctrl_init_kernel(Ctrl* c, double t0, double t_bound, double dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (dt < 0.0) dt = c->dt;  // keep dt_save (adaptive_step_size_solver.h:39-41,55)
  c->t = t0;
  c->t_bound = t_bound;
  c->done = 0;
  c->count = 1;
  c->gave_up = 0;
  if (t0 + dt > t_bound) dt = t_bound - t0;
  c->dt = dt;
}

__global__ void narrow_kernel(z64* dst, const z128* src, long long n) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    dst[i] = z64{static_cast<float>(src[i].re), static_cast<float>(src[i].im)};
}

template <typename T>
__global__ void copy_kernel(T* dst, const T* src, long long n) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

template <int N, typename R>
int launch_hilb_n(const FormDev& d, const Slots& S, const StageArgs& A, hipStream_t st) {
  constexpr int APB = HilbCfg<N>::APB, BS = HilbCfg<N>::BS;
  int groups = (d.n_hrchy + APB - 1) / APB;
  int grid = groups < kMaxGrid ? groups : kMaxGrid;
  if (grid < 1) grid = 1;
  if (d.diag_v)
    hipLaunchKernelGGL((hilb_kernel<N, R, true>), dim3(grid), dim3(BS), 0, st, d, S, A);
  else
    hipLaunchKernelGGL((hilb_kernel<N, R, false>), dim3(grid), dim3(BS), 0, st, d, S, A);
  return grid;
}

template <typename R>
int launch_hilb(const FormDev& d, const Slots& S, const StageArgs& A, hipStream_t st) {
  switch (d.N) {
#define HAMD_CASE(n) \
  case n:            \
    return launch_hilb_n<n, R>(d, S, A, st);
    HAMD_CASE(1) HAMD_CASE(2) HAMD_CASE(3) HAMD_CASE(4) HAMD_CASE(5) HAMD_CASE(6) HAMD_CASE(7)
    HAMD_CASE(8) HAMD_CASE(9) HAMD_CASE(10) HAMD_CASE(11) HAMD_CASE(12) HAMD_CASE(13)
    HAMD_CASE(14) HAMD_CASE(15) HAMD_CASE(16)
#undef HAMD_CASE
    default:
      throw HeomError{HEOM_ERR_UNSUPPORTED, "heom_hilb: n_level > 16 not provided"};
  }
}

template <typename R>
int launch_form(const FormDev& d, const Slots& S, const StageArgs& A, hipStream_t st) {
  if (d.kind == 0) return launch_hilb<R>(d, S, A, st);
  if (d.kind == 1) {
    const int APB = 256 / d.N2;
    int groups = (d.n_hrchy + APB - 1) / APB;
    int grid = groups < kMaxGrid ? (groups < 1 ? 1 : groups) : kMaxGrid;
    hipLaunchKernelGGL((liou_kernel<R>), dim3(grid), dim3(256), 0, st, d, S, A);
    return grid;
  }
  long long blocks = (d.main_size + 255) / 256;
  int grid = blocks < kMaxGrid ? static_cast<int>(blocks < 1 ? 1 : blocks) : kMaxGrid;
  hipLaunchKernelGGL((ado_kernel<R>), dim3(grid), dim3(256), 0, st, d, S, A);
  return grid;
}

}  // namespace

int launch_operator(const FormDev& d, int real_bytes, const Slots& S, const StageArgs& A,
                    hipStream_t stream) {
  int g = real_bytes == 4 ? launch_form<float>(d, S, A, stream)
                          : launch_form<double>(d, S, A, stream);
  HAMD_HIP(hipGetLastError());
  return g;
}

void launch_ctrl_init(Ctrl* c, double t0, double t_bound, double dt, hipStream_t stream) {
  hipLaunchKernelGGL(ctrl_init_kernel, dim3(1), dim3(64), 0, stream, c, t0, t_bound, dt);
  HAMD_HIP(hipGetLastError());
}

void launch_narrow(void* dst, const void* src, long long n, hipStream_t stream) {
  long long b = (n + 255) / 256;
  int grid = b < 4096 ? static_cast<int>(b < 1 ? 1 : b) : 4096;
  hipLaunchKernelGGL(narrow_kernel, dim3(grid), dim3(256), 0, stream, static_cast<z64*>(dst),
                     static_cast<const z128*>(src), n);
  HAMD_HIP(hipGetLastError());
}

void launch_copy(void* dst, const void* src, long long n, int type, hipStream_t stream) {
  long long b = (n + 255) / 256;
  int grid = b < 4096 ? static_cast<int>(b < 1 ? 1 : b) : 4096;
  if (type == ST_C128)
    hipLaunchKernelGGL((copy_kernel<z128>), dim3(grid), dim3(256), 0, stream,
                       static_cast<z128*>(dst), static_cast<const z128*>(src), n);
  else
    hipLaunchKernelGGL((copy_kernel<z64>), dim3(grid), dim3(256), 0, stream,
                       static_cast<z64*>(dst), static_cast<const z64*>(src), n);
  HAMD_HIP(hipGetLastError());
}

}  // namespace hamd
