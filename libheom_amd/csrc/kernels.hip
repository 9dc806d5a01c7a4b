// HIP kernels of the libheom_amd engine (gfx950).
//
// One kernel pass = one HEOM derivative evaluation fused with the Runge-Kutta
// stage work that consumes it:
//   D_n   = the HEOM operator of heom_hilb.h:108-209 (or its Liouville /
//           global-sparse forms, heom_liou.h:151-255, heom_ado.h:191-202)
//   k     = -dt*D  (calc_diff_impl with alpha = dt, beta = 0; rk4.h:48)
//   y_i+1 = y + sum_j a_ij k_j  (the next stage input, written in the same pass)
//   e     = sum_j E_j k_j and per-workgroup partial sums of the error norm
//           (rkdp.h:142-153); err_finalize_kernel reduces them in a fixed
//           order and runs the step control on the device (rkdp.h:97-177).
//
// Threads map to matrix elements: thread e of an ADO's lane group handles
// element e = i*N + j, so every ADO block (row-major N x N, the reference
// layout) is read and written by contiguous lanes.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/libheom_amd.h"
#include "common.h"
#include "kernels.h"

namespace hamd {

namespace {

// read-only tables through the constant address space: uniform indices then
// compile to scalar loads
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* cst(const T* p) {
  return (const __attribute__((address_space(4))) T*)(p);
}

template <typename T>
using cptr = const __attribute__((address_space(4))) T*;

__device__ __forceinline__ LinkFlat ld_flat(cptr<LinkFlat> p) {
  LinkFlat f;
  f.nbr = p->nbr;
  f.u = p->u;
  f.iphi = z128{p->iphi.re, p->iphi.im};
  f.psi = z128{p->psi.re, p->psi.im};
  f.dir = z128{p->dir.re, p->dir.im};
  return f;
}

// Link l from the 8-byte record (neighbour, coefficient index) and its entry in
// the hierarchy's deduplicated coefficient table (one entry per mode,
// occupation and direction: a few KB, held by the scalar cache); i*phi is
// formed exactly (a swap and a negation), as the host does for LinkFlat.
__device__ __forceinline__ LinkFlat ld_link_coef(cptr<Link> lk, cptr<Coef> cf, int l) {
  LinkFlat f;
  const int c = lk[l].cid;
  f.nbr = lk[l].nbr;
  f.u = cf[c].u;
  f.iphi = z128{-cf[c].phi.im, cf[c].phi.re};
  f.psi = z128{cf[c].psi.re, cf[c].psi.im};
  f.dir = z128{cf[c].dir.re, cf[c].dir.im};
  return f;
}

__device__ __forceinline__ z128 ldz(const void* p, bool lo, long long i) {
  if (!lo) return static_cast<const z128*>(p)[i];
  const z64 v = static_cast<const z64*>(p)[i];
  return z128{static_cast<double>(v.re), static_cast<double>(v.im)};
}

// ldz with a non-temporal load (streamed operands that later loads must not evict)
__device__ __forceinline__ z128 ldz_nt(const void* p, bool lo, long long i) {
  typedef double d2v __attribute__((ext_vector_type(2)));
  typedef float f2v __attribute__((ext_vector_type(2)));
  if (!lo) {
    const d2v w = __builtin_nontemporal_load(static_cast<const d2v*>(p) + i);
    return z128{w.x, w.y};
  }
  const f2v w = __builtin_nontemporal_load(static_cast<const f2v*>(p) + i);
  return z128{static_cast<double>(w.x), static_cast<double>(w.y)};
}

__device__ __forceinline__ void stz(void* p, bool lo, long long i, z128 v) {
  if (!lo)
    static_cast<z128*>(p)[i] = v;
  else
    static_cast<z64*>(p)[i] = z64{static_cast<float>(v.re), static_cast<float>(v.im)};
}

__device__ __forceinline__ double zabs(z128 a) { return sqrt(a.re * a.re + a.im * a.im); }

// parity select without indexing kernel-argument arrays by a runtime value
// (which would copy the arguments to scratch memory)
template <typename T>
__device__ __forceinline__ T pick(const T (&a)[2], int cur) {
  return cur ? a[1] : a[0];
}

// Per-element epilogue operands, loaded before the operator work so that
// their memory latency overlaps it.  F64: every operand is complex128.
template <int KP>
struct PreT {
  z128 v[KP];
};
using Pre = PreT<kMaxPre>;

template <bool F64, int KP, bool NT = false>
__device__ __forceinline__ void prefetch(PreT<KP>& p, const StageArgs& A, int cur, long long gi,
                                         bool ok) {
#pragma unroll
  for (int s = 0; s < KP; ++s) {
    const void* ptr = cur ? A.pre[1][s] : A.pre[0][s];
    if (ok && s < A.npre && ptr != nullptr)
      p.v[s] = NT ? ldz_nt(ptr, F64 ? false : ((A.pre_lo >> s) & 1u), gi)
                  : ldz(ptr, F64 ? false : ((A.pre_lo >> s) & 1u), gi);
    else
      p.v[s] = z128{0.0, 0.0};
  }
}

// Linear combination sum_s c[s]*p[s] + ck*k, accumulated in s order.  Terms
// with a zero coefficient add an exact zero, so the result equals the
// sequential sum over the non-zero terms (reference order).
template <int KP>
__device__ __forceinline__ z128 combine(const PreT<KP>& p, const double* c, int npre, double ck,
                                        z128 k) {
  z128 v{0.0, 0.0};
#pragma unroll
  for (int s = 0; s < KP; ++s) {
    if (s < npre) {
      v.re = fma(c[s], p.v[s].re, v.re);
      v.im = fma(c[s], p.v[s].im, v.im);
    }
  }
  v.re = fma(ck, k.re, v.re);
  v.im = fma(ck, k.im, v.im);
  return v;
}

// Epilogue shared by all operator kernels.  D is the operator applied to the
// stage input at flat element gi; returns w^2 of the error norm (or 0).
template <bool F64, typename R, int KP>
__device__ __forceinline__ double epilogue(const PreT<KP>& p, const StageArgs& A, int cur,
                                           long long gi, cx<R> Dr, z128 alpha) {
  const z128 D = conv<double>(Dr);
  const z128 k{-(alpha.re * D.re - alpha.im * D.im), -(alpha.re * D.im + alpha.im * D.re)};
  if (A.mode == MODE_DIFF) {
    z128 v = k;
    if (A.npre > 0) {
      const z128 o = p.v[0];
      v = z128{A.b_re * o.re - A.b_im * o.im + k.re, A.b_re * o.im + A.b_im * o.re + k.im};
    }
    stz(pick(A.dst[0], cur), F64 ? false : A.dst_lo[0], gi, v);
    return 0.0;
  }
  if (void* kp = pick(A.k, cur)) stz(kp, F64 ? false : A.k_lo, gi, k);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (c < A.ncombo) {
      const z128 v = combine(p, A.cc[c], A.npre, A.ck[c], k);
      stz(pick(A.dst[c], cur), F64 ? false : A.dst_lo[c], gi, v);
      if (void* d2 = pick(A.dst2[c], cur)) stz(d2, F64 ? false : A.dst2_lo[c], gi, v);
    }
  }
  if (A.do_err) {
    // the last two operands hold y_new / y_old (zero weights in ec)
    const int np = A.npre;
    const z128 e = combine(p, A.ec, np - 2, A.ek, k);
    z128 yn{0.0, 0.0}, yo{0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KP; ++s) {
      if (s == np - 2) yn = p.v[s];
      if (s == np - 1) yo = p.v[s];
    }
    const double a1 = zabs(yn), a0 = zabs(yo);
    const double w = zabs(e) / (A.atol + A.rtol * (a1 > a0 ? a1 : a0));
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
  // a step that is not positive would never reach t_bound: give up instead
  // (unreachable from a positive step, which the rule above keeps positive)
  if (!(dt > 0.0)) {
    c->t = c->t_bound;
    c->done = 1;
    c->gave_up = 1;
  }
}

// Block reduction of the error-norm partial sums: one value per workgroup,
// written without atomics; err_finalize_kernel combines them.
__device__ void write_partial(const StageArgs& A, double v) {
  __shared__ double s_red[16];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
  if (lane == 0) s_red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += s_red[w];
    A.partials[blockIdx.x] = t;
  }
}

// Sums the partials of the last stage pass in a fixed order (deterministic)
// and runs the step control (one workgroup of 1024 threads; every thread sums
// a strided slice with four independent accumulators).
__global__ void __launch_bounds__(1024) err_finalize_kernel(const StageArgs A, int nparts) {
  __shared__ double s_red[16];
  if (A.ctrl && A.ctrl->done) return;
  double x0 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0;
  int b = threadIdx.x;
  for (; b + 3 * 1024 < nparts; b += 4 * 1024) {
    x0 += A.partials[b];
    x1 += A.partials[b + 1024];
    x2 += A.partials[b + 2048];
    x3 += A.partials[b + 3072];
  }
  for (; b < nparts; b += 1024) x0 += A.partials[b];
  double x = (x0 + x1) + (x2 + x3);
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += s_red[w];
    double err = sqrt(t * A.inv_n);
    if (A.control)
      dopri_control(A.ctrl, err);
    else if (A.ctrl)
      A.ctrl->err = err;
  }
}

// Kernel prologue: leave early when the control block says the interval is
// done, pick the ping-pong parity and the step.
struct Prologue {
  int cur;
  z128 alpha;
  bool skip;
};

__device__ __forceinline__ Prologue prologue(const StageArgs& A) {
  Prologue p{0, z128{A.a_re, A.a_im}, false};
  if (A.ctrl) {
    if (A.ctrl->done) {
      p.skip = true;
      return p;
    }
    p.cur = A.ctrl->cur;
    if (A.mode == MODE_STAGE) p.alpha = z128{A.ctrl->dt, 0.0};
  }
  return p;
}

// Learned truncation closure: the network's prediction of the tier-(L+1)
// coupling term of a last-tier ADO, added to D before the epilogue.
template <typename R>
__device__ __forceinline__ void add_closure(const FormDev& d, long long gi, cx<R>& D) {
  if (d.clos == nullptr) return;
  const long long a = gi / d.N2;
  const int r = static_cast<int>(a) - d.clos_lo;
  if (r < 0 || r >= d.clos_n) return;
  const int e = static_cast<int>(gi - a * d.N2);
  const float* y = d.clos + static_cast<long long>(r) * 2 * d.N2;
  D.re += static_cast<R>(y[e]);
  D.im += static_cast<R>(y[d.N2 + e]);
}

// ---------------------------------------------------------------------------
// heom_hilb, fast path: one ADO per workgroup (N = 6..16), diagonal V
// (heom_hilb.h:96-212 with [V,X]_ij = (v_i - v_j) X_ij, {V,X}_ij = (v_i + v_j) X_ij)
//
// Workgroup b takes the ADO of work unit bmap[b] (its link range in the same
// record), read beside the control block.  H and the diagonals of every V_u
// are staged in LDS once per ADO (no per-link coefficient tables), each link is
// an 8-byte scalar record, its coefficient-table entry and one neighbour-element
// load (the 64-byte LinkFlat records cost 8% more time per stage).
// REAL: H and every V_u real (the commutator and the link coefficients in real
// arithmetic on the real parts; a complex H or V takes REAL = false).
// ---------------------------------------------------------------------------
template <int N>
struct WaveCfg {
  static constexpr int N2 = N * N;
  static constexpr int BS = ((N2 + 63) / 64) * 64;
};

// LDS bytes of the diagonals of every V_u (plus one zero row for direct links)
inline size_t fast_vd_lds(const FormDev& d, bool real, int real_bytes) {
  const size_t rows = static_cast<size_t>(d.n_noise + 1) * d.N;
  return rows * (real ? 1 : 2) * static_cast<size_t>(real_bytes);
}

// Partial sum of the error norm of one wavefront (no workgroup barrier),
// written to slot `slot`.
__device__ void write_partial_wave(const StageArgs& A, double v, int slot) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) A.partials[slot] = v;
}

// LDS written by a wavefront, then read by other lanes of the same wavefront:
// the LDS serves one wavefront's operations in order, so only the compiler
// must keep them in order.
// The scheduling barrier keeps the code on either side apart, as a workgroup
// barrier would (without it the compiler hoists later loads above the LDS
// phase and holds their registers across it).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  __builtin_amdgcn_sched_barrier(0);
}

template <int N, typename R, bool REAL, bool F64, int KP>
__global__ void __launch_bounds__(WaveCfg<N>::BS, 1)
    hilb_fast_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  using HT = typename std::conditional<REAL, R, C>::type;
  constexpr int N2 = N * N, BS = WaveCfg<N>::BS;
  __shared__ C s_rho[N2];
  __shared__ HT s_H[N2];
  extern __shared__ __align__(16) unsigned char s_dyn[];
  R* s_vr = reinterpret_cast<R*>(s_dyn);  // real V: v_u,i at u*N + i
  C* s_vc = reinterpret_cast<C*>(s_dyn);  // complex V

  const auto bmp = cst(d.bmap) + blockIdx.x;
  const BlockMap bm{bmp->g, bmp->l0, bmp->l1, 0};
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const int e = threadIdx.x;
  const bool lane_ok = e < N2;
  const int ee = lane_ok ? e : 0;
  const int i = ee / N, j = ee - (ee / N) * N;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const auto lks = cst(d.links);
  const auto cfs = cst(d.coef);

  // idle lanes (e >= N2) read element 0 and store nothing
  const int g = bm.g;
  const long long gi = static_cast<long long>(g) * N2 + ee;
  const int nvd = (d.n_noise + 1) * N, nvd0 = d.n_noise * N;
  const z128 hv = d.H[ee];  // BS >= N2: one element of H per lane
  const z128 vdv = e < nvd0 ? d.Vd[e] : z128{0.0, 0.0};
  const C dterm = conv<R>(d.dterm[ee]);
  const C r = in[gi];
  const int l0 = bm.l0, l1 = bm.l1;
  // the wide-operand stages issue their first four links before the barrier
  // (measured: faster there, slower for stages of at most four operands)
  const int nb0 = KP > 4 ? ((l1 - l0 < 4) ? l1 - l0 : 4) : 0;
  LinkFlat q[4];
  C x[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (t < nb0) {
      q[t] = ld_link_coef(lks, cfs, l0 + t);
      x[t] = in[static_cast<long long>(q[t].nbr) * N2 + ee];
    }
  }
  PreT<KP> pv;
  // stages of at most four epilogue operands load them here, ahead of the
  // operator work; the wider stages load them after the links (76 instead of 87
  // VGPRs in FP64: six waves per SIMD instead of five; measured faster there)
  if constexpr (KP <= 4) prefetch<F64, KP>(pv, A, cur, gi, true);
  const C gm = conv<R>(z128{cst(d.gam)[g].re, cst(d.gam)[g].im});
  // LDS: H, the diagonals of V (zero row for direct links last), rho_n
  if (lane_ok) {
    if constexpr (REAL)
      s_H[e] = static_cast<R>(hv.re);
    else
      s_H[e] = conv<R>(hv);
  }
  for (int xx = e; xx < nvd; xx += BS) {
    const z128 v = xx == e ? vdv : (xx < nvd0 ? d.Vd[xx] : z128{0.0, 0.0});
    if (REAL)
      s_vr[xx] = static_cast<R>(v.re);
    else
      s_vc[xx] = conv<R>(v);
  }
  if (lane_ok) s_rho[e] = r;
  __syncthreads();

  C acc{0, 0};
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const C b = s_rho[k * N + j], t = s_rho[i * N + k];
    if constexpr (REAL) {
      const R hik = s_H[i * N + k], hkj = s_H[k * N + j];
      acc.re += hik * b.re;
      acc.im += hik * b.im;
      acc.re -= t.re * hkj;
      acc.im -= t.im * hkj;
    } else {
      cmac(acc, s_H[i * N + k], b);
      cmac(acc, C{-t.re, -t.im}, s_H[k * N + j]);
    }
  }
  C D = times_i(acc);
  cmac(D, gm + dterm, r);
  // link l adds ((v_i - v_j) i*phi + (v_i + v_j) psi + dir) X_nbr, in link order
  auto apply = [&](const LinkFlat& lq, const C& xv) {
    C w;
    if constexpr (REAL) {
      const R vi = s_vr[lq.u * N + i], vj = s_vr[lq.u * N + j];
      const R c0 = vi - vj, a0 = vi + vj;
      w.re = c0 * static_cast<R>(lq.iphi.re) + a0 * static_cast<R>(lq.psi.re) +
             static_cast<R>(lq.dir.re);
      w.im = c0 * static_cast<R>(lq.iphi.im) + a0 * static_cast<R>(lq.psi.im) +
             static_cast<R>(lq.dir.im);
    } else {
      const C vi = s_vc[lq.u * N + i], vj = s_vc[lq.u * N + j];
      w = (vi - vj) * conv<R>(lq.iphi) + (vi + vj) * conv<R>(lq.psi) + conv<R>(lq.dir);
    }
    cmac(D, w, xv);
  };
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (t < nb0) apply(q[t], x[t]);
  for (int l = l0 + nb0; l < l1; l += 4) {
    const int nb = (l1 - l < 4) ? l1 - l : 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t < nb) {
        q[t] = ld_link_coef(lks, cfs, l + t);
        x[t] = in[static_cast<long long>(q[t].nbr) * N2 + ee];
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if (t < nb) apply(q[t], x[t]);
  }
  if constexpr (KP > 4) prefetch<F64, KP>(pv, A, cur, gi, true);
  double w2 = 0.0;
  if (lane_ok) {
    add_closure(d, gi, D);
    w2 = epilogue<F64, R, KP>(pv, A, cur, gi, D, pg.alpha);
  }
  if (A.do_err) write_partial(A, w2);
}

// ---------------------------------------------------------------------------
// heom_hilb, general path: any V, N <= 16; N <= 8 packs 64/N^2 ADOs per
// wavefront (heom_hilb.h:96-212)
// ---------------------------------------------------------------------------
template <int N>
struct HilbCfg {
  static constexpr int N2 = N * N;
  static constexpr int BS = N2 <= 64 ? 64 : ((N2 + 63) / 64) * 64;
  static constexpr int APB = BS / N2;  // ADOs per workgroup
};

template <int N, typename R, bool DIAGV, bool F64, int KP>
__global__ void __launch_bounds__(HilbCfg<N>::BS)
    hilb_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  constexpr int N2 = N * N, APB = HilbCfg<N>::APB, BS = HilbCfg<N>::BS;
  __shared__ C s_rho[APB * N2];
  __shared__ C s_H[N2];
  __shared__ C s_X[DIAGV ? 1 : APB * N2];
  __shared__ C s_Y[DIAGV ? 1 : APB * N2];
  extern __shared__ __align__(16) unsigned char s_dyn[];
  C* s_vd = reinterpret_cast<C*>(s_dyn);

  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const int tid = threadIdx.x;
  const int la = tid / N2;
  const int e = tid - la * N2;
  const int i = e / N, j = e - (e / N) * N;
  const bool lane_ok = la < APB;
  for (int x = tid; x < N2; x += BS) s_H[x] = conv<R>(d.H[x]);
  if (DIAGV) {
    // diagonal of every V_u, then one zero row for direct (off-diagonal gamma) links
    for (int x = tid; x < (d.n_noise + 1) * N; x += BS)
      s_vd[x] = x < d.n_noise * N ? conv<R>(d.Vd[x]) : C{0, 0};
  }
  const C dterm = (DIAGV && lane_ok) ? conv<R>(d.dterm[e]) : C{0, 0};
  const C* in = static_cast<const C*>(pick(A.in, cur));

  double err_acc = 0.0;
  const int n_groups = (d.n_hrchy + APB - 1) / APB;
  for (int g = blockIdx.x; g < n_groups; g += gridDim.x) {
    const int n = g * APB + la;
    const bool ok = lane_ok && n < d.n_hrchy;
    const long long gi = static_cast<long long>(n) * N2 + e;
    const C r = ok ? in[gi] : C{0, 0};
    PreT<KP> pv;
    prefetch<F64, KP>(pv, A, cur, gi, ok);
    const int l0 = ok ? d.link_ptr[n] : 0;
    const int l1 = ok ? d.link_ptr[n + 1] : 0;
    if (lane_ok) s_rho[tid] = r;
    __syncthreads();
    C acc{0, 0};
    if (lane_ok) {
      const C* b = s_rho + la * N2;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        cmac(acc, s_H[i * N + k], b[k * N + j]);
        const C t = b[i * N + k];
        cmac(acc, C{-t.re, -t.im}, s_H[k * N + j]);
      }
    }
    C D = times_i(acc);
    if (ok) cmac(D, conv<R>(d.gam[n]) + dterm, r);
    if (DIAGV) {
      int l = l0;
      for (; l + 2 <= l1; l += 2) {
        const LinkFlat q0 = d.flat[l], q1 = d.flat[l + 1];
        const C x0 = in[static_cast<long long>(q0.nbr) * N2 + e];
        const C x1 = in[static_cast<long long>(q1.nbr) * N2 + e];
        const C vi0 = s_vd[q0.u * N + i], vj0 = s_vd[q0.u * N + j];
        const C vi1 = s_vd[q1.u * N + i], vj1 = s_vd[q1.u * N + j];
        cmac(D, (vi0 - vj0) * conv<R>(q0.iphi) + (vi0 + vj0) * conv<R>(q0.psi) + conv<R>(q0.dir),
             x0);
        cmac(D, (vi1 - vj1) * conv<R>(q1.iphi) + (vi1 + vj1) * conv<R>(q1.psi) + conv<R>(q1.dir),
             x1);
      }
      for (; l < l1; ++l) {
        const LinkFlat q0 = d.flat[l];
        const C x0 = in[static_cast<long long>(q0.nbr) * N2 + e];
        const C vi0 = s_vd[q0.u * N + i], vj0 = s_vd[q0.u * N + j];
        cmac(D, (vi0 - vj0) * conv<R>(q0.iphi) + (vi0 + vj0) * conv<R>(q0.psi) + conv<R>(q0.dir),
             x0);
      }
    } else {
      // dense V: Phi_u, Psi_u per noise, then V_u X + Y V_u (heom_hilb.h:113-208)
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
    if (ok) {
      add_closure(d, gi, D);
      err_acc += epilogue<F64, R, KP>(pv, A, cur, gi, D, pg.alpha);
    }
    __syncthreads();
  }
  if (A.do_err) write_partial(A, err_acc);
}

// ---------------------------------------------------------------------------
// heom_hilb, large N (N > 16), diagonal V: one ADO per workgroup of 256 lanes,
// each lane owns a 4x4 tile of elements.  The block sits in LDS (and H too
// when both fit in 64 KB); the commutator is a register-tiled complex matrix
// product, the neighbour terms and the stage epilogue run per owned element.
// ---------------------------------------------------------------------------
constexpr int kBigBS = 256;
constexpr int kBigTile = 4;

template <typename R, bool HLDS, bool F64, int KP>
__global__ void __launch_bounds__(kBigBS) hilb_big_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  extern __shared__ __align__(16) unsigned char s_dyn[];
  const int N = d.N, N2 = d.N2;
  C* s_rho = reinterpret_cast<C*>(s_dyn);
  C* s_H = s_rho + N2;
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const auto link_ptr = cst(d.link_ptr);
  const auto flat = cst(d.flat);
  const auto gam = cst(d.gam);
  const int nt = (N + kBigTile - 1) / kBigTile;  // tiles per row
  if (HLDS)
    for (int x = threadIdx.x; x < N2; x += kBigBS) s_H[x] = conv<R>(d.H[x]);
  double err_acc = 0.0;
  for (int g = blockIdx.x; g < d.n_hrchy; g += gridDim.x) {
    const long long base = static_cast<long long>(g) * N2;
    for (int x = threadIdx.x; x < N2; x += kBigBS) s_rho[x] = in[base + x];
    __syncthreads();
    const int l0 = link_ptr[g], l1 = link_ptr[g + 1];
    const C gm = conv<R>(z128{gam[g].re, gam[g].im});
    for (int t = threadIdx.x; t < nt * nt; t += kBigBS) {
      const int r0 = (t / nt) * kBigTile, c0 = (t % nt) * kBigTile;
      C acc[kBigTile][kBigTile];
#pragma unroll
      for (int a = 0; a < kBigTile; ++a)
#pragma unroll
        for (int b = 0; b < kBigTile; ++b) acc[a][b] = C{0, 0};
      for (int k = 0; k < N; ++k) {
        C hr[kBigTile], rr[kBigTile], rc[kBigTile], hc[kBigTile];
#pragma unroll
        for (int a = 0; a < kBigTile; ++a) {
          const int i = r0 + a < N ? r0 + a : N - 1;
          const int j = c0 + a < N ? c0 + a : N - 1;
          hr[a] = HLDS ? s_H[i * N + k] : conv<R>(d.H[i * N + k]);
          rc[a] = s_rho[i * N + k];
          rr[a] = s_rho[k * N + j];
          hc[a] = HLDS ? s_H[k * N + j] : conv<R>(d.H[k * N + j]);
        }
#pragma unroll
        for (int a = 0; a < kBigTile; ++a)
#pragma unroll
          for (int b = 0; b < kBigTile; ++b) {
            cmac(acc[a][b], hr[a], rr[b]);
            cmac(acc[a][b], C{-rc[a].re, -rc[a].im}, hc[b]);
          }
      }
#pragma unroll
      for (int a = 0; a < kBigTile; ++a)
#pragma unroll
        for (int b = 0; b < kBigTile; ++b) {
          const int i = r0 + a, j = c0 + b;
          if (i >= N || j >= N) continue;
          const int e = i * N + j;
          const C r = s_rho[e];
          C D = times_i(acc[a][b]);
          cmac(D, gm + conv<R>(d.dterm[e]), r);
          for (int l = l0; l < l1; ++l) {
            const LinkFlat q = ld_flat(flat + l);
            const C x = in[static_cast<long long>(q.nbr) * N2 + e];
            const long long tb = static_cast<long long>(q.u) * N2 + e;
            C cf;
            if (d.real_v) {
              const z128 p = d.cmac_tab[tb];
              cf.re = static_cast<R>(p.re) * static_cast<R>(q.iphi.re) +
                      static_cast<R>(p.im) * static_cast<R>(q.psi.re) + static_cast<R>(q.dir.re);
              cf.im = static_cast<R>(p.re) * static_cast<R>(q.iphi.im) +
                      static_cast<R>(p.im) * static_cast<R>(q.psi.im) + static_cast<R>(q.dir.im);
            } else {
              cf = conv<R>(d.cmac_tab[2 * tb]) * conv<R>(q.iphi) +
                   conv<R>(d.cmac_tab[2 * tb + 1]) * conv<R>(q.psi) + conv<R>(q.dir);
            }
            cmac(D, cf, x);
          }
          PreT<KP> pv;
          prefetch<F64, KP>(pv, A, cur, base + e, true);
          add_closure(d, base + e, D);
          err_acc += epilogue<F64, R, KP>(pv, A, cur, base + e, D, pg.alpha);
        }
    }
    __syncthreads();
  }
  if (A.do_err) write_partial(A, err_acc);
}

// ---------------------------------------------------------------------------
// heom_hilb, 16 < N <= 64, diagonal V: the commutator H*rho - rho*H on the
// matrix cores (heom_hilb.h:108-125 computes it with two gemm calls per ADO).
// One ADO per 256-thread workgroup.  rho_n is staged in LDS as separate real
// and imaginary planes with row stride N+1, so that the 16 rows of an MFMA
// fragment fall in distinct banks.  H (shared by every ADO) is read from L2
// as MFMA operands.  Wave w owns NB*NB/4 of the
// (N/16)^2 output tiles, all in one 16-row band, and accumulates H*rho and
// -rho*H into the same accumulators: 4 real products per k step when H is
// real, 8 otherwise.  The accumulators then feed the element-wise part of
// the operator and the fused Runge-Kutta epilogue.
//
// MFMA 16x16xK lane layout (checked against host products): lane l holds
// A[l%16][k0 + l/16] and B[k0 + l/16][l%16] (K = 4), or the 8 consecutive k
// from k0 + 8*(l/16) (bf16, K = 32); accumulator element r of lane l is
// C[4*(l/16) + r][l%16] for the f32 accumulators and C[l/16 + 4*r][l%16]
// for f64.
// ---------------------------------------------------------------------------
typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 mf_b8 __attribute__((ext_vector_type(8)));

template <typename R>
struct Mfma;
template <>
struct Mfma<double> {
  using acc = mf_d4;
  static __device__ __forceinline__ acc mac(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
  static __device__ __forceinline__ const double* re(const FormDev& d) { return d.Hre; }
  static __device__ __forceinline__ const double* im(const FormDev& d) { return d.Him; }
};
template <>
struct Mfma<float> {
  using acc = mf_f4;
  static __device__ __forceinline__ acc mac(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
  static __device__ __forceinline__ const float* re(const FormDev& d) { return d.Hre32; }
  static __device__ __forceinline__ const float* im(const FormDev& d) { return d.Him32; }
};

// x = hi + lo with hi = bf16(x) and lo = bf16(x - hi): 16 significant bits
struct Split8 {
  mf_b8 hi, lo;
};

__device__ __forceinline__ Split8 split8(const float (&x)[8], bool neg) {
  Split8 s;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const float v = neg ? -x[q] : x[q];
    const __bf16 h = static_cast<__bf16>(v);
    s.hi[q] = h;
    s.lo[q] = static_cast<__bf16>(v - static_cast<float>(h));
  }
  return s;
}

// a*b ~ hi*hi + hi*lo + lo*hi (the lo*lo term, 2^-16 relative, is dropped);
// the small terms are accumulated first
__device__ __forceinline__ mf_f4 mac3(const Split8& a, const Split8& b, mf_f4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, c, 0, 0, 0);
}

constexpr int kMfmaBS = 256;
constexpr size_t kMfmaLdsCap = 150 * 1024;  // dynamic LDS the matrix-core kernel may use

// Coefficient of a neighbour element at (i, j) with the diagonals of every V_u
// in LDS (sv: n_noise + 1 rows of n, the last row zero; real V as R, complex V
// as cx<R>): (v_i - v_j)*iphi + (v_i + v_j)*psi + dir.
template <typename R>
__device__ __forceinline__ cx<R> link_coef_lds(const FormDev& d, const LinkFlat& q, const R* sv,
                                               int n, int i, int j) {
  cx<R> cf;
  if (d.real_v) {
    const R vi = sv[q.u * n + i], vj = sv[q.u * n + j];
    const R c0 = vi - vj, a0 = vi + vj;
    cf.re = c0 * static_cast<R>(q.iphi.re) + a0 * static_cast<R>(q.psi.re) +
            static_cast<R>(q.dir.re);
    cf.im = c0 * static_cast<R>(q.iphi.im) + a0 * static_cast<R>(q.psi.im) +
            static_cast<R>(q.dir.im);
  } else {
    const cx<R>* sc = reinterpret_cast<const cx<R>*>(sv);
    const cx<R> vi = sc[q.u * n + i], vj = sc[q.u * n + j];
    cf = (vi - vj) * conv<R>(q.iphi) + (vi + vj) * conv<R>(q.psi) + conv<R>(q.dir);
  }
  return cf;
}

template <int N>
struct MfmaCfg {
  static constexpr int N2 = N * N;
  static constexpr int LD = N + 1;                // LDS row stride
  static constexpr int NB = N / 16;               // 16-row bands
  static constexpr int TPW = NB * NB / 4;         // output tiles per wave
  static constexpr int KS = N / 4;                // k steps of the 16x16x4 products
  static constexpr int KB = N / 32;               // k steps of the bf16 products
  static_assert(N == 32 || N == 64, "MFMA path: tiles of 32 or 64 rows");
};

// H*rho - rho*H for the wave's tiles, operands in the state precision R
template <int N, typename R, bool REALH, int TPW>
__device__ __forceinline__ void mfma_commutator(const FormDev& d, const R* s_re, const R* s_im,
                                                int band, int t0, int lr, int lk,
                                                typename Mfma<R>::acc (&cre)[TPW],
                                                typename Mfma<R>::acc (&cim)[TPW]) {
  using M = Mfma<R>;
  using Cfg = MfmaCfg<N>;
  const R* Hr = M::re(d);
  const R* Hi = M::im(d);
  const int row = band * 16 + lr;
  // fewer tiles per wave go with a smaller register budget (see hilb_mfma_kernel)
  constexpr int KU = TPW > 2 ? Cfg::KS : 4;
#pragma unroll KU
  for (int ks = 0; ks < Cfg::KS; ++ks) {
    const int kk = 4 * ks + lk;
    const R arr = -s_re[row * Cfg::LD + kk], ari = -s_im[row * Cfg::LD + kk];
    const R ahr = Hr[row * N + kk], ahi = REALH ? R(0) : Hi[row * N + kk];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int col = ((t0 + t) % Cfg::NB) * 16 + lr;
      const R brr = s_re[kk * Cfg::LD + col], bri = s_im[kk * Cfg::LD + col];
      const R bhr = Hr[kk * N + col];
      cre[t] = M::mac(ahr, brr, cre[t]);   //  Hr rho_r
      cim[t] = M::mac(ahr, bri, cim[t]);   //  Hr rho_i
      cre[t] = M::mac(arr, bhr, cre[t]);   // -rho_r Hr
      cim[t] = M::mac(ari, bhr, cim[t]);   // -rho_i Hr
      if (!REALH) {
        const R bhi = Hi[kk * N + col];
        cre[t] = M::mac(-ahi, bri, cre[t]);  // -Hi rho_i
        cim[t] = M::mac(ahi, brr, cim[t]);   //  Hi rho_r
        cre[t] = M::mac(-ari, bhi, cre[t]);  //  rho_i Hi
        cim[t] = M::mac(arr, bhi, cim[t]);   // -rho_r Hi
      }
    }
  }
}

// the same products on bf16 matrix cores with split operands (FP32 state)
template <int N, bool REALH, int TPW>
__device__ __forceinline__ void mfma_commutator_bf16x3(const FormDev& d, const float* s_re,
                                                       const float* s_im, int band, int t0,
                                                       int lr, int lk, mf_f4 (&cre)[TPW],
                                                       mf_f4 (&cim)[TPW]) {
  using Cfg = MfmaCfg<N>;
  const float* Hr = d.Hre32;
  const float* Hi = d.Him32;
  const int row = band * 16 + lr;
  for (int kb = 0; kb < Cfg::KB; ++kb) {
    const int k0 = 32 * kb + 8 * lk;
    float x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = Hr[row * N + k0 + q];
    const Split8 ahr = split8(x, false);
    Split8 ahi{}, nahi{};
    if (!REALH) {
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q] = Hi[row * N + k0 + q];
      ahi = split8(x, false);
      nahi = split8(x, true);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = s_re[row * Cfg::LD + k0 + q];
    const Split8 arr = split8(x, true);   // -rho_r
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = s_im[row * Cfg::LD + k0 + q];
    const Split8 ari = split8(x, true);   // -rho_i
    const Split8 pari = split8(x, false);
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int col = ((t0 + t) % Cfg::NB) * 16 + lr;
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q] = s_re[(k0 + q) * Cfg::LD + col];
      const Split8 brr = split8(x, false);
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q] = s_im[(k0 + q) * Cfg::LD + col];
      const Split8 bri = split8(x, false);
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q] = Hr[(k0 + q) * N + col];
      const Split8 bhr = split8(x, false);
      cre[t] = mac3(ahr, brr, cre[t]);
      cim[t] = mac3(ahr, bri, cim[t]);
      cre[t] = mac3(arr, bhr, cre[t]);
      cim[t] = mac3(ari, bhr, cim[t]);
      if (!REALH) {
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = Hi[(k0 + q) * N + col];
        const Split8 bhi = split8(x, false);
        cre[t] = mac3(nahi, bri, cre[t]);
        cim[t] = mac3(ahi, brr, cim[t]);
        cre[t] = mac3(pari, bhi, cre[t]);
        cim[t] = mac3(arr, bhi, cim[t]);
      }
    }
  }
}

// BS threads (BS / 64 waves) per ADO; every form asks for two workgroups per
// CU (LDS holds two FP64 workgroups per CU), i.e. BS / 128 waves per SIMD:
// 256 threads leave 256 registers per lane, 512 threads 128.
// BF16X3: R = float, products via mfma_commutator_bf16x3.  N is the tile size
// (32 or 64); the ADO block is n x n with n = d.N in (N/2, N] for N = 64, in
// (16, 32] for N = 32, zero-padded in LDS (and H in its planes), so that the
// padded rows and columns of the products are exactly zero.
template <int N, typename R, bool REALH, bool F64, int KP, bool BF16X3, bool PADDED, int BS>
__global__ void __launch_bounds__(BS, BS / 128)
    hilb_mfma_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  using M = Mfma<R>;
  using Cfg = MfmaCfg<N>;
  constexpr int N2 = Cfg::N2, LD = Cfg::LD, NB = Cfg::NB;
  constexpr int TPW = NB * NB / (BS / 64);  // output tiles per wave
  static_assert(TPW >= 1 && TPW * (BS / 64) == NB * NB, "tiles must divide among the waves");
  extern __shared__ __align__(16) unsigned char s_dyn[];
  R* s_re = reinterpret_cast<R*>(s_dyn);
  R* s_im = s_re + N * LD;
  R* s_v = s_im + N * LD;  // diagonals of every V_u (and a zero row), see link_coef_lds
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  // the ADO block (n <= N); unpadded blocks keep the index arithmetic constant
  const int n = PADDED ? d.N : N, n2 = PADDED ? d.N2 : N * N;
  if (PADDED) {  // the padding stays zero: only rows and columns < n are ever written
    for (int x = threadIdx.x; x < 2 * N * LD; x += BS) s_re[x] = R(0);
  }
  {
    const int nvd = (d.n_noise + 1) * d.N, nvd0 = d.n_noise * d.N;
    for (int x = threadIdx.x; x < nvd; x += BS) {
      const z128 v = x < nvd0 ? d.Vd[x] : z128{0.0, 0.0};
      if (d.real_v) {
        s_v[x] = static_cast<R>(v.re);
      } else {
        s_v[2 * x] = static_cast<R>(v.re);
        s_v[2 * x + 1] = static_cast<R>(v.im);
      }
    }
  }
  __syncthreads();
  const int cur = pg.cur;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const auto link_ptr = cst(d.link_ptr);
  const auto flat = cst(d.flat);
  const auto gam = cst(d.gam);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int t0 = wave * TPW;
  const int band = t0 / NB;
  double err_acc = 0.0;
  for (int w = blockIdx.x; w < d.n_hrchy; w += gridDim.x) {
    // work unit w: the XCD-aware block map when present
    BlockMap bm{w, 0, 0, 0};
    if (d.bmap) {
      const auto p = cst(d.bmap) + w;
      bm = BlockMap{p->g, p->l0, p->l1, 0};
    } else {
      bm.l0 = link_ptr[w];
      bm.l1 = link_ptr[w + 1];
    }
    const int g = bm.g;
    const long long base = static_cast<long long>(g) * n2;
    {
      // all loads of rho_n in flight before the LDS stores
      constexpr int PER = N2 / BS;
      C v[PER];
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const int x = threadIdx.x + c * BS;
        v[c] = x < n2 ? in[base + x] : C{0, 0};
      }
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const int x = threadIdx.x + c * BS;
        if (x < n2) {
          const int i = x / n, j = x - (x / n) * n;
          s_re[i * LD + j] = v[c].re;
          s_im[i * LD + j] = v[c].im;
        }
      }
    }
    __syncthreads();
    typename M::acc cre[TPW], cim[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      cre[t] = typename M::acc{0, 0, 0, 0};
      cim[t] = typename M::acc{0, 0, 0, 0};
    }
    if constexpr (BF16X3)
      mfma_commutator_bf16x3<N, REALH, TPW>(d, s_re, s_im, band, t0, lr, lk, cre, cim);
    else
      mfma_commutator<N, R, REALH, TPW>(d, s_re, s_im, band, t0, lr, lk, cre, cim);
    // every wave has read rho_n as MFMA operands: each lane now replaces its
    // own elements by i[H, rho] + (Gamma_n + delta term) rho (one owner each)
    __syncthreads();
    const C gm = conv<R>(z128{gam[g].re, gam[g].im});
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int col = ((t0 + t) % NB) * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = band * 16 + M::row(lk, r);
        if (i < n && col < n) {
          C D = times_i(C{cre[t][r], cim[t][r]});
          const C rv{s_re[i * LD + col], s_im[i * LD + col]};
          cmac(D, gm + conv<R>(d.dterm[i * n + col]), rv);
          s_re[i * LD + col] = D.re;
          s_im[i * LD + col] = D.im;
        }
      }
    }
    __syncthreads();
    // Neighbour couplings and the Runge-Kutta epilogue, element-parallel and
    // coalesced.  Two schedules of the same per-element arithmetic (D, then
    // the links in order, then the epilogue):
    //  - links outer (FP64 and BF16X3 forms, and every form with eight waves
    //    per ADO): for each link the loads of LCH of the thread's elements of
    //    the neighbour block are in flight together; each thread updates only
    //    its own elements (x = tid + c*BS) in LDS and reads them back in the
    //    epilogue, so the phases need no barrier;
    //  - links inside the epilogue passes (complex64 form with four waves per
    //    ADO, whose register budget keeps three waves per SIMD that way).
    // Each epilogue pass takes CH elements per thread and issues every load
    // before its stores (the stores may alias later loads); elements past the
    // block read element 0 and store nothing.
    constexpr bool kLinksOuter = sizeof(R) == 8 || BF16X3 || BS > 256;
    constexpr int EPT = N2 / BS;              // elements per thread (tile)
    constexpr int LMAX = BS > 256 ? 4 : 8;    // register budget: 256 / 128 per lane
    constexpr int LCH = EPT < LMAX ? EPT : LMAX;  // elements per load group
    constexpr int CH = (KP <= 4 && BS <= 256) ? 2 : 1;
    const int l0 = bm.l0, l1 = bm.l1;
    if constexpr (kLinksOuter) {
      for (int l = l0; l < l1; ++l) {
        const LinkFlat q = ld_flat(flat + l);
        const C* nb = in + static_cast<long long>(q.nbr) * n2;
#pragma unroll 1
        for (int c0 = 0; c0 < EPT; c0 += LCH) {
          C xv[LCH];
#pragma unroll
          for (int c = 0; c < LCH; ++c) {
            const int x = threadIdx.x + (c0 + c) * BS;
            const int xs = x < n2 ? x : 0;
            xv[c] = nb[xs];
          }
#pragma unroll
          for (int c = 0; c < LCH; ++c) {
            const int x = threadIdx.x + (c0 + c) * BS;
            if (x < n2) {
              const int i = x / n, j = x - (x / n) * n;
              C Dv{s_re[i * LD + j], s_im[i * LD + j]};
              cmac(Dv, link_coef_lds<R>(d, q, s_v, d.N, i, j), xv[c]);
              s_re[i * LD + j] = Dv.re;
              s_im[i * LD + j] = Dv.im;
            }
          }
        }
      }
    }
    for (int x0 = threadIdx.x; x0 < n2; x0 += BS * CH) {
      C Dc[CH];
      PreT<KP> pv[CH];
      int xs[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int x = x0 + c * BS;
        xs[c] = x < n2 ? x : 0;
        const int i = xs[c] / n, j = xs[c] - (xs[c] / n) * n;
        Dc[c] = C{s_re[i * LD + j], s_im[i * LD + j]};
        // FP64 at N = 64: the streamed epilogue operands are loaded non-temporally,
        // which keeps more of the stage input (the neighbour blocks) cached
        prefetch<F64, KP, sizeof(R) == 8 && N == 64>(pv[c], A, cur, base + xs[c], true);
      }
      if constexpr (!kLinksOuter) {
        for (int l = l0; l < l1; ++l) {
          const LinkFlat q = ld_flat(flat + l);
          C xv[CH];
#pragma unroll
          for (int c = 0; c < CH; ++c) xv[c] = in[static_cast<long long>(q.nbr) * n2 + xs[c]];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            const int i = xs[c] / n, j = xs[c] - (xs[c] / n) * n;
            cmac(Dc[c], link_coef_lds<R>(d, q, s_v, d.N, i, j), xv[c]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int x = x0 + c * BS;
        if (x < n2) {
          add_closure(d, base + x, Dc[c]);
          err_acc += epilogue<F64, R, KP>(pv[c], A, cur, base + x, Dc[c], pg.alpha);
        }
      }
    }
    __syncthreads();
  }
  if (A.do_err) write_partial(A, err_acc);
}

// ---------------------------------------------------------------------------
// heom_hilb on matrix cores with dense (non-diagonal) V, 16 < N <= 64
// (heom_hilb.h:108-209): per ADO
//   D = i[H, rho] + Gamma_n rho + sum_u (V_u X_u + Y_u V_u) + sum_direct dir rho_l
//   X_u = i Phi_u + Psi_u + delta_u [V_u, rho],  Y_u = Psi_u - i Phi_u - delta_u [V_u, rho]
// (i[V,Phi] + {V,Psi} + delta [V,[V,rho]] = V X + Y V).  D stays in the MFMA
// accumulators from i[H, rho] + Gamma rho on; per noise [V_u, rho] goes to a
// second accumulator set, X_u and then Y_u are built in an LDS workspace
// (delta C at the accumulator positions, then the links of noise u added by
// coalesced element-parallel passes) and multiplied on the matrix cores.
// ---------------------------------------------------------------------------
// acc += sgn * (A B) for the wave's tiles.  A and B are complex; an operand in
// LDS has leading dimension LD, a global one (a V_u plane pair) N, and a global
// operand is real when GREAL (its imaginary plane is not read).
template <int N, typename R, bool A_LDS, bool B_LDS, bool GREAL, int TPW>
__device__ __forceinline__ void mfma_cprod(const R* ar, const R* ai, const R* br, const R* bi,
                                           R sgn, int band, int t0, int lr, int lk,
                                           typename Mfma<R>::acc (&cre)[TPW],
                                           typename Mfma<R>::acc (&cim)[TPW]) {
  using M = Mfma<R>;
  using Cfg = MfmaCfg<N>;
  constexpr int LDA = A_LDS ? Cfg::LD : N, LDB = B_LDS ? Cfg::LD : N;
  constexpr bool AR = !A_LDS && GREAL, BR = !B_LDS && GREAL;
  const int row = band * 16 + lr;
#pragma unroll 2
  for (int ks = 0; ks < Cfg::KS; ++ks) {
    const int kk = 4 * ks + lk;
    const R a_r = sgn * ar[row * LDA + kk];
    const R a_i = AR ? R(0) : sgn * ai[row * LDA + kk];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int col = ((t0 + t) % Cfg::NB) * 16 + lr;
      const R b_r = br[kk * LDB + col];
      cre[t] = M::mac(a_r, b_r, cre[t]);
      if (!BR) {
        const R b_i = bi[kk * LDB + col];
        cim[t] = M::mac(a_r, b_i, cim[t]);
        if (!AR) cre[t] = M::mac(-a_i, b_i, cre[t]);
      }
      if (!AR) cim[t] = M::mac(a_i, b_r, cim[t]);
    }
  }
}

template <typename R>
struct MfmaV;
template <>
struct MfmaV<double> {
  static __device__ __forceinline__ const double* re(const FormDev& d) { return d.Vre; }
  static __device__ __forceinline__ const double* im(const FormDev& d) { return d.Vim; }
};
template <>
struct MfmaV<float> {
  static __device__ __forceinline__ const float* re(const FormDev& d) { return d.Vre32; }
  static __device__ __forceinline__ const float* im(const FormDev& d) { return d.Vim32; }
};

constexpr int kDenseCh = 4;  // elements per thread whose loads are in flight together

template <int N, typename R, bool REALV, bool REALH, bool F64, int KP, bool PADDED, int BS>
__global__ void __launch_bounds__(BS, BS / 256)
    hilb_mfma_dense_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  using M = Mfma<R>;
  using Cfg = MfmaCfg<N>;
  using acc_t = typename M::acc;
  constexpr int N2 = Cfg::N2, LD = Cfg::LD, NB = Cfg::NB;
  constexpr int TPW = NB * NB / (BS / 64);  // output tiles per wave
  constexpr int EPT = N2 / BS;  // elements per thread (tile)
  extern __shared__ __align__(16) unsigned char s_dyn[];
  R* s_re = reinterpret_cast<R*>(s_dyn);  // rho_n, at the end D
  R* s_im = s_re + N * LD;
  R* w_re = s_im + N * LD;  // workspace: X_u, then Y_u
  R* w_im = w_re + N * LD;
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int n = PADDED ? d.N : N, n2 = PADDED ? d.N2 : N * N;
  if (PADDED) {  // the padding stays zero: only rows and columns < n are ever written
    for (int x = threadIdx.x; x < 4 * N * LD; x += BS) s_re[x] = R(0);
    __syncthreads();
  }
  const int cur = pg.cur;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const auto link_ptr = cst(d.link_ptr);
  const auto flat = cst(d.flat);
  const auto gam = cst(d.gam);
  const auto sdl = cst(d.sdelta);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int t0 = wave * TPW;
  const int band = t0 / NB;
  const R* Vr0 = MfmaV<R>::re(d);
  const R* Vi0 = MfmaV<R>::im(d);
  double err_acc = 0.0;
  for (int w = blockIdx.x; w < d.n_hrchy; w += gridDim.x) {
    BlockMap bm{w, 0, 0, 0};
    if (d.bmap) {
      const auto p = cst(d.bmap) + w;
      bm = BlockMap{p->g, p->l0, p->l1, 0};
    } else {
      bm.l0 = link_ptr[w];
      bm.l1 = link_ptr[w + 1];
    }
    const int g = bm.g;
    const long long base = static_cast<long long>(g) * n2;
#pragma unroll 1
    for (int c0 = 0; c0 < EPT; c0 += kDenseCh) {
      C v[kDenseCh];
#pragma unroll
      for (int c = 0; c < kDenseCh; ++c) {
        const int x = threadIdx.x + (c0 + c) * BS;
        v[c] = x < n2 ? in[base + x] : C{0, 0};
      }
#pragma unroll
      for (int c = 0; c < kDenseCh; ++c) {
        const int x = threadIdx.x + (c0 + c) * BS;
        if (x < n2) {
          const int i = x / n, j = x - (x / n) * n;
          s_re[i * LD + j] = v[c].re;
          s_im[i * LD + j] = v[c].im;
        }
      }
    }
    __syncthreads();
    acc_t cre[TPW], cim[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      cre[t] = acc_t{0, 0, 0, 0};
      cim[t] = acc_t{0, 0, 0, 0};
    }
    mfma_commutator<N, R, REALH, TPW>(d, s_re, s_im, band, t0, lr, lk, cre, cim);
    // D = i[H, rho] + Gamma_n rho at the accumulator positions
    const C gm = conv<R>(z128{gam[g].re, gam[g].im});
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int col = ((t0 + t) % NB) * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = band * 16 + M::row(lk, r);
        const C rv{s_re[i * LD + col], s_im[i * LD + col]};
        C Dv = times_i(C{cre[t][r], cim[t][r]});
        cmac(Dv, gm, rv);
        cre[t][r] = Dv.re;
        cim[t][r] = Dv.im;
      }
    }
    const int nn = d.n_noise;
    int l = bm.l0;
    for (int u = 0; u < nn; ++u) {
      const R* Vr = Vr0 + static_cast<size_t>(u) * N2;
      const R* Vi = Vi0 + static_cast<size_t>(u) * N2;
      acc_t ccr[TPW], cci[TPW];
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        ccr[t] = acc_t{0, 0, 0, 0};
        cci[t] = acc_t{0, 0, 0, 0};
      }
      // C = V rho - rho V
      mfma_cprod<N, R, false, true, REALV, TPW>(Vr, Vi, s_re, s_im, R(1), band, t0, lr, lk, ccr,
                                                cci);
      mfma_cprod<N, R, true, false, REALV, TPW>(s_re, s_im, Vr, Vi, R(-1), band, t0, lr, lk, ccr,
                                                cci);
      const C dl = conv<R>(z128{sdl[u].re, sdl[u].im});
      // links of noise u: [lu0, lu1) (the links of an ADO are sorted by noise)
      const int lu0 = l;
      while (l < bm.l1 && flat[l].u == u) ++l;
      const int lu1 = l;
#pragma unroll 1
      for (int pass = 0; pass < 2; ++pass) {
        // workspace = +-delta C at the accumulator positions
        const R sg = pass == 0 ? R(1) : R(-1);
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
          const int col = ((t0 + t) % NB) * 16 + lr;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = band * 16 + M::row(lk, r);
            if (i < n && col < n) {
              const C cv{sg * ccr[t][r], sg * cci[t][r]};
              const C x = dl * cv;
              w_re[i * LD + col] = x.re;
              w_im[i * LD + col] = x.im;
            }
          }
        }
        __syncthreads();
        // + sum_l coefficient * rho_l, element-parallel: X: iphi + psi, Y: psi - iphi
        for (int lq = lu0; lq < lu1; ++lq) {
          const LinkFlat q = ld_flat(flat + lq);
          const C ip = conv<R>(q.iphi), ps = conv<R>(q.psi);
          const C cf = pass == 0 ? ip + ps : ps - ip;
          const C* nb = in + static_cast<long long>(q.nbr) * n2;
#pragma unroll 1
          for (int c0 = 0; c0 < EPT; c0 += kDenseCh) {
            C xv[kDenseCh];
#pragma unroll
            for (int c = 0; c < kDenseCh; ++c) {
              const int x = threadIdx.x + (c0 + c) * BS;
              xv[c] = nb[x < n2 ? x : 0];
            }
#pragma unroll
            for (int c = 0; c < kDenseCh; ++c) {
              const int x = threadIdx.x + (c0 + c) * BS;
              if (x < n2) {
                const int i = x / n, j = x - (x / n) * n;
                C wv{w_re[i * LD + j], w_im[i * LD + j]};
                cmac(wv, cf, xv[c]);
                w_re[i * LD + j] = wv.re;
                w_im[i * LD + j] = wv.im;
              }
            }
          }
        }
        __syncthreads();
        if (pass == 0)  // D += V X
          mfma_cprod<N, R, false, true, REALV, TPW>(Vr, Vi, w_re, w_im, R(1), band, t0, lr, lk,
                                                    cre, cim);
        else  // D += Y V
          mfma_cprod<N, R, true, false, REALV, TPW>(w_re, w_im, Vr, Vi, R(1), band, t0, lr, lk,
                                                    cre, cim);
        __syncthreads();
      }
    }
    // D to LDS (every wave has finished reading rho and the workspace)
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int col = ((t0 + t) % NB) * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = band * 16 + M::row(lk, r);
        if (i < n && col < n) {
          s_re[i * LD + col] = cre[t][r];
          s_im[i * LD + col] = cim[t][r];
        }
      }
    }
    __syncthreads();
    // direct (off-diagonal gamma) links and the Runge-Kutta epilogue, coalesced
    const int ld0 = l, ld1 = bm.l1;
    for (int x0 = threadIdx.x; x0 < n2; x0 += BS) {
      const int i = x0 / n, j = x0 - (x0 / n) * n;
      C Dc{s_re[i * LD + j], s_im[i * LD + j]};
      PreT<KP> pv;
      prefetch<F64, KP>(pv, A, cur, base + x0, true);
      for (int lq = ld0; lq < ld1; ++lq) {
        const LinkFlat q = ld_flat(flat + lq);
        cmac(Dc, conv<R>(q.dir), in[static_cast<long long>(q.nbr) * n2 + x0]);
      }
      add_closure(d, base + x0, Dc);
      err_acc += epilogue<F64, R, KP>(pv, A, cur, base + x0, Dc, pg.alpha);
    }
    __syncthreads();
  }
  if (A.do_err) write_partial(A, err_acc);
}

// ---------------------------------------------------------------------------
// batch of Hilbert-form hierarchies (diagonal V, N <= 8): one persistent
// launch advances every member; a 64-lane wavefront takes IPW consecutive
// work items per step (each 64/N^2 ADOs of one member), loading all of them
// before it computes any (IPW is a tile shape the batch autotuner picks)
// ---------------------------------------------------------------------------
// The lane's share of one work item.
template <typename R, int KP>
struct BatchLane {
  BatchInst I;
  int ado, e, i, j;
  bool ok;
  long long gi;
  cx<R> r;
  PreT<KP> pv;
};

// Load phase of work item `it`: the lane's element, its epilogue operands,
// and the item's H and ADO blocks in LDS.
template <typename R, bool F64, int KP>
__device__ __forceinline__ void batch_load(const FormDev& d, const StageArgs& A, const cx<R>* in,
                                           int it, int cur, int lane, cx<R>* s_rho, cx<R>* s_H,
                                           BatchLane<R, KP>& b) {
  using C = cx<R>;
  const auto items = cst(d.items);
  const auto insts = cst(d.inst);
  const bool live = it < d.n_items;
  const int m = live ? items[it].inst : 0, ado0 = live ? items[it].ado : 0;
  BatchInst& I = b.I;
  I.N = insts[m].N;
  I.N2 = insts[m].N2;
  I.n_hrchy = insts[m].n_hrchy;
  I.apw = insts[m].apw;
  I.ado_base = insts[m].ado_base;
  I.h_off = insts[m].h_off;
  I.tab_off = insts[m].tab_off;
  I.dterm_off = insts[m].dterm_off;
  I.real_v = insts[m].real_v;
  I.coef_off = insts[m].coef_off;
  I.elem_base = insts[m].elem_base;
  const int N = I.N, N2 = I.N2;
  const int la = lane / N2;
  b.e = lane - la * N2;
  b.i = b.e / N;
  b.j = b.e - b.i * N;
  b.ado = ado0 + la;
  b.ok = live && la < I.apw && b.ado < I.n_hrchy;
  b.gi = I.elem_base + static_cast<long long>(b.ok ? b.ado : ado0) * N2 + (b.ok ? b.e : 0);
  b.r = in[b.gi];
  prefetch<F64, KP>(b.pv, A, cur, b.gi, b.ok);
  s_H[lane] = lane < N2 ? conv<R>(d.H[I.h_off + lane]) : C{0, 0};
  s_rho[lane] = b.ok ? b.r : C{0, 0};
}

// Compute phase: D of the lane's element and the Runge-Kutta epilogue.
template <typename R, bool F64, int KP>
__device__ __forceinline__ void batch_compute(const FormDev& d, const StageArgs& A,
                                              const cx<R>* in, int cur, int lane, z128 alpha,
                                              const cx<R>* s_rho, const cx<R>* s_H,
                                              const BatchLane<R, KP>& b, double& err_acc) {
  using C = cx<R>;
  if (!b.ok) return;
  const BatchInst& I = b.I;
  const int N = I.N, N2 = I.N2, i = b.i, j = b.j, e = b.e;
  C acc{0, 0};
  const C* bl = s_rho + (lane / N2) * N2;
  // members have N <= 8: unrolled to 8 with an exit (measured faster than the
  // rolled loop in FP64)
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (k >= N) break;
    cmac(acc, s_H[i * N + k], bl[k * N + j]);
    const C t = bl[i * N + k];
    cmac(acc, C{-t.re, -t.im}, s_H[k * N + j]);
  }
  C D = times_i(acc);
  cmac(D, conv<R>(d.gam[I.ado_base + b.ado]) + conv<R>(d.dterm[I.dterm_off + e]), b.r);
  const int l0 = d.link_ptr[I.ado_base + b.ado], l1 = d.link_ptr[I.ado_base + b.ado + 1];
  // the next link record is loaded with the current link's coefficient,
  // neighbour and table loads (one memory round trip per link instead of two)
  BatchLink qn = l0 < l1 ? d.blink[l0] : BatchLink{0, 0, 0, 0};
  for (int l = l0; l < l1; ++l) {
    const BatchLink ql = qn;
    if (l + 1 < l1) qn = d.blink[l + 1];
    const BatchCoef q = d.bcoef[I.coef_off + ql.cid];
    const C x = in[I.elem_base + static_cast<long long>(ql.nbr) * N2 + e];
    const int tb = ql.u * N2 + e;
    C cf;
    if (I.real_v) {
      const z128 p = d.cmac_tab[I.tab_off + tb];
      cf.re = static_cast<R>(p.re) * static_cast<R>(q.iphi.re) +
              static_cast<R>(p.im) * static_cast<R>(q.psi.re) + static_cast<R>(q.dir.re);
      cf.im = static_cast<R>(p.re) * static_cast<R>(q.iphi.im) +
              static_cast<R>(p.im) * static_cast<R>(q.psi.im) + static_cast<R>(q.dir.im);
    } else {
      cf = conv<R>(d.cmac_tab[I.tab_off + 2 * tb]) * conv<R>(q.iphi) +
           conv<R>(d.cmac_tab[I.tab_off + 2 * tb + 1]) * conv<R>(q.psi) + conv<R>(q.dir);
    }
    cmac(D, cf, x);
  }
  err_acc += epilogue<F64, R, KP>(b.pv, A, cur, b.gi, D, alpha);
}

// KP = 4 for the stages with at most four epilogue operands (every RK4 stage):
// the smaller operand set keeps 7 waves per SIMD (FP32) instead of 4-5.
// WPB wavefronts per workgroup, each with its own items and LDS region and
// none waiting for another (whether that pays is the autotuner's call); the
// error-norm partial sums are one per wavefront.
template <typename R, bool F64, int KP, int IPW, int WPB>
__global__ void __launch_bounds__(64 * WPB) batch_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  __shared__ C s_rho[WPB][IPW][64];
  __shared__ C s_H[WPB][IPW][64];
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const int w = WPB > 1 ? __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / 64) : 0;
  const int lane = static_cast<int>(threadIdx.x) - 64 * w;
  double err_acc = 0.0;
  for (int it = (static_cast<int>(blockIdx.x) * WPB + w) * IPW; it < d.n_items;
       it += gridDim.x * WPB * IPW) {
    BatchLane<R, KP> b[IPW];
#pragma unroll
    for (int q = 0; q < IPW; ++q)
      batch_load<R, F64, KP>(d, A, in, it + q, cur, lane, s_rho[w][q], s_H[w][q], b[q]);
    if constexpr (WPB > 1)
      wave_lds_sync();
    else
      __syncthreads();
#pragma unroll
    for (int q = 0; q < IPW; ++q)
      batch_compute<R, F64, KP>(d, A, in, cur, lane, pg.alpha, s_rho[w][q], s_H[w][q], b[q],
                                err_acc);
    if constexpr (WPB > 1)
      wave_lds_sync();
    else
      __syncthreads();
  }
  if (A.do_err) {
    if constexpr (WPB > 1)
      write_partial_wave(A, err_acc, static_cast<int>(blockIdx.x) * WPB + w);
    else
      write_partial(A, err_acc);
  }
}

// closure network input (Re, Im of the last-tier ADOs of the stage input)
__global__ void closure_gather_kernel(const FormDev d, const StageArgs A, float* x,
                                      int ado_stride) {
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const void* in = pick(A.in, pg.cur);
  const long long n = static_cast<long long>(d.clos_n) * d.N2;
  const long long base = static_cast<long long>(d.clos_lo) * d.N2;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const z128 v = ldz(in, A.in_lo != 0, base + i);
    const long long r = i / d.N2;
    const int e = static_cast<int>(i - r * d.N2);
    x[r * ado_stride + e] = static_cast<float>(v.re);
    x[r * ado_stride + d.N2 + e] = static_cast<float>(v.im);
  }
}

// ---------------------------------------------------------------------------
// Learned closure, one launch per operator pass (per-ADO layout): each
// workgroup takes 16 last-tier ADOs of the stage input, gathers their network
// rows (Re, Im of rho_n, then the constant columns written at binding), and
// runs the three layers on the matrix cores (f32 16x16x4, FP32 accumulation):
// h1 = tanh(x W1^T + b1), h2 = tanh(h1 W2^T + b2), y = h2 W3^T + b3 (the
// normalisation folded into W1, b1, W3, b3).  Rows stay in LDS between layers;
// the weights (row-major, one row per output unit) are read from L2.
// ---------------------------------------------------------------------------
constexpr int kClosRows = 16;   // network rows per workgroup
constexpr int kClosInMax = 64;  // widest input the fused kernel takes
constexpr int kClosHidMax = 256;

// acc[t] += (rows of `a` in LDS, leading dimension lda) x W^T for output
// columns [c0 + 16 t, c0 + 16 t + 16), k over [0, K) (K a multiple of 4;
// columns of W at or past kw read as zero).  Eight k steps at a time: their
// operands are loaded before their products, so the weight loads of a group
// are in flight together.
template <int T>
__device__ __forceinline__ void clos_layer(const float* a, int lda, const float* W, int kw, int K,
                                           int c0, int lr, int lk, mf_f4 (&acc)[T]) {
  constexpr int G = 8;
  for (int k0 = 0; k0 < K; k0 += 4 * G) {
    float av[G], bv[G][T];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int k = k0 + 4 * g + lk;
      const bool in = k0 + 4 * g < K;
      av[g] = in ? a[lr * lda + k] : 0.f;
#pragma unroll
      for (int t = 0; t < T; ++t) {
        const int n = c0 + 16 * t + lr;
        bv[g][t] = (in && k < kw) ? W[static_cast<long long>(n) * kw + k] : 0.f;
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int t = 0; t < T; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[g], bv[g][t], acc[t], 0, 0, 0);
  }
}

__global__ void __launch_bounds__(256) closure_mlp_kernel(const FormDev d, const StageArgs A,
                                                          const float* xc, int n_in, int n_hid,
                                                          int n_out, const float* W1,
                                                          const float* b1, const float* W2,
                                                          const float* b2, const float* W3,
                                                          const float* b3, float* y) {
  constexpr int LDX = kClosInMax + 4, LDH = kClosHidMax + 4;
  __shared__ float s_x[kClosRows * LDX];
  __shared__ float s_h1[kClosRows * LDH];
  __shared__ float s_h2[kClosRows * LDH];
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const void* in = pick(A.in, pg.cur);
  const int r0 = blockIdx.x * kClosRows;
  const int rows = d.clos_n, N2 = d.N2;
  const int kin = (n_in + 3) & ~3;
  // gather: rho of ADO lo + r, then the constant columns (zero past n_in and rows)
  for (int x = threadIdx.x; x < kClosRows * kin; x += blockDim.x) {
    const int r = x / kin, c = x - (x / kin) * kin;
    float v = 0.f;
    if (r0 + r < rows && c < n_in) {
      if (c < 2 * N2) {
        const long long gi = static_cast<long long>(d.clos_lo + r0 + r) * N2 + (c % N2);
        const z128 z = ldz(in, A.in_lo != 0, gi);
        v = static_cast<float>(c < N2 ? z.re : z.im);
      } else {
        v = xc[static_cast<long long>(r0 + r) * n_in + c];
      }
    }
    s_x[r * LDX + c] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  // hidden layers: wave w takes columns [64 w, 64 w + 64) (n_hid <= 256)
  const float* hin = s_x;
  int lda = LDX, kw = n_in, K = kin;
  const float* Ws[2] = {W1, W2};
  const float* bs[2] = {b1, b2};
  float* hout[2] = {s_h1, s_h2};
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const int c0 = 64 * wave;
    if (c0 < n_hid) {
      mf_f4 acc[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = mf_f4{0.f, 0.f, 0.f, 0.f};
      clos_layer<4>(hin, lda, Ws[layer], kw, K, c0, lr, lk, acc);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int n = c0 + 16 * t + lr;
        if (n < n_hid) {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            hout[layer][(4 * lk + q) * LDH + n] = tanhf(acc[t][q] + bs[layer][n]);
        }
      }
    }
    __syncthreads();
    hin = hout[layer];
    lda = LDH;
    kw = n_hid;
    K = n_hid;
  }
  // output layer: wave w takes columns [16 w, 16 w + 16) (n_out <= 64)
  const int c0 = 16 * wave;
  if (c0 < n_out) {
    mf_f4 acc[1] = {mf_f4{0.f, 0.f, 0.f, 0.f}};
    clos_layer<1>(s_h2, LDH, W3, n_hid, n_hid, c0, lr, lk, acc);
    const int n = c0 + lr;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = 4 * lk + q;
      if (n < n_out && r0 + r < rows)
        y[static_cast<long long>(r0 + r) * n_out + n] = acc[0][q] + b3[n];
    }
  }
}

// rho_S of every batch member, gathered into one block (member order)
__global__ void gather_rho_s_kernel(z128* dst, const z128* src, const BatchInst* inst) {
  const BatchInst I = inst[blockIdx.x];
  for (int e = threadIdx.x; e < I.N2; e += blockDim.x) dst[I.rs_off + e] = src[I.elem_base + e];
}

// ---------------------------------------------------------------------------
// heom_liou: per-ADO superoperator form (heom_liou.h:151-255), N <= 16
// ---------------------------------------------------------------------------
template <typename R, bool F64, int KP>
__global__ void __launch_bounds__(256) liou_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  __shared__ C s_rho[256];
  __shared__ C s_phi[256];
  __shared__ C s_psi[256];
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const int N2 = d.N2;
  const int APB = 256 / N2;
  const int tid = threadIdx.x;
  const int la = tid / N2;
  const int e = tid - la * N2;
  const bool lane_ok = la < APB;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const int stride = N2 + 1;
  double err_acc = 0.0;
  const int n_groups = (d.n_hrchy + APB - 1) / APB;
  for (int g = blockIdx.x; g < n_groups; g += gridDim.x) {
    const int n = g * APB + la;
    const bool ok = lane_ok && n < d.n_hrchy;
    const long long gi = static_cast<long long>(n) * N2 + e;
    const C r = ok ? in[gi] : C{0, 0};
    PreT<KP> pv;
    prefetch<F64, KP>(pv, A, cur, gi, ok);
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
    if (ok) err_acc += epilogue<F64, R, KP>(pv, A, cur, gi, D, pg.alpha);
    __syncthreads();
  }
  if (A.do_err) write_partial(A, err_acc);
}

// heom_liou for 16 < N <= 32 (N^2 up to 1024): one ADO per 256-lane workgroup,
// each lane up to four elements e = tid + 256 c, rho_n and the per-noise
// neighbour sums phi, psi staged in LDS (3 N^2 complex values, dynamic);
// the same per-element operation order as liou_kernel.
template <typename R, bool F64, int KP>
__global__ void __launch_bounds__(256) liou_wide_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  constexpr int NC = 4;
  extern __shared__ __align__(16) unsigned char s_dyn[];
  const int N2 = d.N2;
  C* s_rho = reinterpret_cast<C*>(s_dyn);
  C* s_phi = s_rho + N2;
  C* s_psi = s_phi + N2;
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const int tid = threadIdx.x;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const int stride = N2 + 1;
  double err_acc = 0.0;
  for (int n = blockIdx.x; n < d.n_hrchy; n += gridDim.x) {
    const long long base = static_cast<long long>(n) * N2;
    for (int e = tid; e < N2; e += 256) s_rho[e] = in[base + e];
    __syncthreads();
    C D[NC];
    const C gm = conv<R>(d.gam[n]);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int e = tid + 256 * c;
      D[c] = C{0, 0};
      if (e < N2) {
        for (int p = d.sop_ptr[e]; p < d.sop_ptr[e + 1]; ++p)
          cmac(D[c], conv<R>(d.sop_val[p]), s_rho[d.sop_col[p]]);
        cmac(D[c], gm, s_rho[e]);
      }
    }
    const int l1 = d.link_ptr[n + 1];
    int l = d.link_ptr[n];
    for (int u = 0; u < d.n_noise; ++u) {
      C phi[NC], psi[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) phi[c] = psi[c] = C{0, 0};
      while (l < l1) {
        const Link lk = d.links[l];
        const Coef cf = d.coef[lk.cid];
        if (cf.u != u) break;
        const C* nb = in + static_cast<long long>(lk.nbr) * N2;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int e = tid + 256 * c;
          if (e < N2) {
            const C x = nb[e];
            cmac(phi[c], conv<R>(cf.phi), x);
            cmac(psi[c], conv<R>(cf.psi), x);
          }
        }
        ++l;
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int e = tid + 256 * c;
        if (e < N2) {
          s_phi[e] = phi[c];
          s_psi[e] = psi[c];
        }
      }
      __syncthreads();
      const int* pp = d.sop_ptr + (1 + 2 * u) * stride;
      const int* pq = d.sop_ptr + (2 + 2 * u) * stride;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int e = tid + 256 * c;
        if (e < N2) {
          for (int p = pp[e]; p < pp[e + 1]; ++p)
            cmac(D[c], conv<R>(d.sop_val[p]), s_phi[d.sop_col[p]]);
          for (int p = pq[e]; p < pq[e + 1]; ++p)
            cmac(D[c], conv<R>(d.sop_val[p]), s_psi[d.sop_col[p]]);
        }
      }
      __syncthreads();
    }
    for (; l < l1; ++l) {
      const Link lk = d.links[l];
      const Coef cf = d.coef[lk.cid];
      const C* nb = in + static_cast<long long>(lk.nbr) * N2;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int e = tid + 256 * c;
        if (e < N2) cmac(D[c], conv<R>(cf.dir), nb[e]);
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int e = tid + 256 * c;
      if (e < N2) {
        PreT<KP> pv;
        prefetch<F64, KP>(pv, A, cur, base + e, true);
        err_acc += epilogue<F64, R, KP>(pv, A, cur, base + e, D[c], pg.alpha);
      }
    }
    __syncthreads();
  }
  if (A.do_err) write_partial(A, err_acc);
}

// ---------------------------------------------------------------------------
// heom_ado: one global CSR matrix-vector product (heom_ado.h:191-202)
// ---------------------------------------------------------------------------
template <typename R, bool F64, int KP>
__global__ void __launch_bounds__(256) ado_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  double err_acc = 0.0;
  const long long rows = d.main_size;
  for (long long r = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; r < rows;
       r += static_cast<long long>(gridDim.x) * blockDim.x) {
    PreT<KP> pv;
    prefetch<F64, KP>(pv, A, cur, r, true);
    C D{0, 0};
    for (long long p = d.R_ptr[r]; p < d.R_ptr[r + 1]; ++p)
      cmac(D, conv<R>(d.R_val[p]), in[d.R_col[p]]);
    err_acc += epilogue<F64, R, KP>(pv, A, cur, r, D, pg.alpha);
  }
  if (A.do_err) write_partial(A, err_acc);
}

// ---------------------------------------------------------------------------
// Redfield QMEs (redfield_hilb.h:66-89, redfield_liou.h:91-102):
//   D = i[H, rho] + sum_u [V_u, T_u],  T_u = Lambda_u rho - rho Lambda_u^+
// ---------------------------------------------------------------------------
// One workgroup holds the whole problem (n_level <= 16 and every table within
// 60 KB of LDS), one lane per element.  Every table is staged into LDS at the
// start, so that all of their loads are in flight together; then
//   diagonal V (DIAG): with v_u the diagonal of V_u,
//     D = HL rho + rho HR - sum_u ( v_u[i] (rho Lambda_u^+)[i][j] + (Lambda_u rho)[i][j] v_u[j] ),
//     HL = iH + sum_u V_u Lambda_u, HR = -iH + sum_u Lambda_u^+ V_u (formed at set-up),
//     read from LDS with no further barrier;
//   dense V: D = i[H, rho], then per noise T_u = Lambda_u rho - rho Lambda_u^+
//     into LDS, a barrier, and D += [V_u, T_u].
// LDS (cx<R> elements): rho, H or HL, HR (DIAG), per noise Lambda_u, Lambda_u^+
// and V_u (dense), then v_u (DIAG) or T (dense).
template <typename R, bool F64, int KP, bool DIAG>
__global__ void __launch_bounds__(256) redfield_small_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  extern __shared__ __align__(16) unsigned char rf_lds[];
  C* sm = reinterpret_cast<C*>(rf_lds);
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const int N = d.N, N2 = d.N2, nn = d.n_noise;
  constexpr int per = DIAG ? 2 : 3;  // tables per noise
  C* s_rho = sm;
  C* s_H = sm + N2;  // H, or HL
  C* s_HR = sm + 2 * N2;
  C* s_tab = sm + 3 * N2;
  C* s_x = s_tab + static_cast<size_t>(per) * nn * N2;  // v_u (DIAG) or T (dense)
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const int e = threadIdx.x;
  const bool ok = e < N2;
  const int i = ok ? e / N : 0, j = ok ? e - (e / N) * N : 0;
  PreT<KP> pv;
  prefetch<F64, KP>(pv, A, cur, e, ok);
  for (int x = threadIdx.x; x < N2; x += blockDim.x) {
    s_rho[x] = in[x];
    s_H[x] = conv<R>(DIAG ? d.rf_HL[x] : d.H[x]);
    if (DIAG) s_HR[x] = conv<R>(d.rf_HR[x]);
  }
  for (int x = threadIdx.x; x < nn * N2; x += blockDim.x) {
    const int u = x / N2, y = x - u * N2;
    C* t = s_tab + static_cast<size_t>(u) * per * N2;
    t[y] = conv<R>(d.rf_L[x]);
    t[N2 + y] = conv<R>(d.rf_Ld[x]);
    if (!DIAG) t[2 * N2 + y] = conv<R>(d.V[x]);
  }
  if (DIAG)
    for (int x = threadIdx.x; x < nn * N; x += blockDim.x) s_x[x] = conv<R>(d.Vd[x]);
  __syncthreads();
  C D{0, 0};
  if (DIAG) {
    if (ok) {
      for (int k = 0; k < N; ++k) {
        cmac(D, s_H[i * N + k], s_rho[k * N + j]);
        cmac(D, s_rho[i * N + k], s_HR[k * N + j]);
      }
      for (int u = 0; u < nn; ++u) {
        const C* L = s_tab + static_cast<size_t>(u) * 2 * N2;
        const C* Ld = L + N2;
        C x{0, 0}, y{0, 0};
        for (int k = 0; k < N; ++k) {
          cmac(x, s_rho[i * N + k], Ld[k * N + j]);
          cmac(y, L[i * N + k], s_rho[k * N + j]);
        }
        const C vi = s_x[u * N + i], vj = s_x[u * N + j];
        cmac(D, C{-vi.re, -vi.im}, x);
        cmac(D, C{-vj.re, -vj.im}, y);
      }
    }
  } else {
    if (ok) {
      C c{0, 0};
      for (int k = 0; k < N; ++k) {
        cmac(c, s_H[i * N + k], s_rho[k * N + j]);
        cmac(c, s_rho[i * N + k], C{-s_H[k * N + j].re, -s_H[k * N + j].im});
      }
      D = times_i(c);
    }
    for (int u = 0; u < nn; ++u) {
      const C* L = s_tab + static_cast<size_t>(u) * 3 * N2;
      const C* Ld = L + N2;
      const C* V = L + 2 * N2;
      if (ok) {
        C t{0, 0};
        for (int k = 0; k < N; ++k) {
          cmac(t, L[i * N + k], s_rho[k * N + j]);
          cmac(t, s_rho[i * N + k], C{-Ld[k * N + j].re, -Ld[k * N + j].im});
        }
        s_x[e] = t;
      }
      __syncthreads();
      if (ok) {
        for (int k = 0; k < N; ++k) {
          cmac(D, V[i * N + k], s_x[k * N + j]);
          cmac(D, s_x[i * N + k], C{-V[k * N + j].re, -V[k * N + j].im});
        }
      }
      __syncthreads();
    }
  }
  double err = 0.0;
  if (ok) err = epilogue<F64, R, KP>(pv, A, cur, e, D, pg.alpha);
  if (A.do_err) write_partial(A, err);
}

// Larger problems, first pass: T_u = Lambda_u rho - rho Lambda_u^+ for every
// noise, one lane per (u, element); operands come through L1/L2.
template <typename R>
__global__ void __launch_bounds__(256) redfield_t_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const C* in = static_cast<const C*>(pick(A.in, pg.cur));
  C* T = static_cast<C*>(d.rf_T);
  const int N = d.N, N2 = d.N2;
  const long long total = static_cast<long long>(d.n_noise) * N2;
  for (long long x = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; x < total;
       x += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long u = x / N2;
    const int e = static_cast<int>(x - u * N2);
    const int i = e / N, j = e - i * N;
    const z128* L = d.rf_L + u * N2;
    const z128* Ld = d.rf_Ld + u * N2;
    C t{0, 0};
    for (int k = 0; k < N; ++k) {
      cmac(t, conv<R>(L[i * N + k]), in[k * N + j]);
      const C ld = conv<R>(Ld[k * N + j]);
      cmac(t, in[i * N + k], C{-ld.re, -ld.im});
    }
    T[x] = t;
  }
}

// Larger problems, second pass: D = i[H, rho] + sum_u [V_u, T_u] (diagonal V:
// (v_i - v_j) T_u[i][j]) and the stage epilogue, one lane per element.
template <typename R, bool F64, int KP>
__global__ void __launch_bounds__(256) redfield_d_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const C* T = static_cast<const C*>(d.rf_T);
  const int N = d.N, N2 = d.N2;
  double err_acc = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N2; e += gridDim.x * blockDim.x) {
    PreT<KP> pv;
    prefetch<F64, KP>(pv, A, cur, e, true);
    const int i = e / N, j = e - i * N;
    C c{0, 0};
    for (int k = 0; k < N; ++k) {
      cmac(c, conv<R>(d.H[i * N + k]), in[k * N + j]);
      const C h = conv<R>(d.H[k * N + j]);
      cmac(c, in[i * N + k], C{-h.re, -h.im});
    }
    C D = times_i(c);
    for (int u = 0; u < d.n_noise; ++u) {
      const long long b = static_cast<long long>(u) * N2;
      if (d.diag_v) {
        cmac(D, conv<R>(d.Vd[u * N + i]) - conv<R>(d.Vd[u * N + j]), T[b + e]);
      } else {
        for (int k = 0; k < N; ++k) {
          cmac(D, conv<R>(d.V[b + i * N + k]), T[b + k * N + j]);
          const C v = conv<R>(d.V[b + k * N + j]);
          cmac(D, T[b + i * N + k], C{-v.re, -v.im});
        }
      }
    }
    err_acc += epilogue<F64, R, KP>(pv, A, cur, e, D, pg.alpha);
  }
  if (A.do_err) write_partial(A, err_acc);
}

// Sparse superoperator times the state, one 64-lane wavefront per row: the
// lanes stride over the row's entries (coalesced) and reduce with shuffles;
// lane 0 runs the epilogue.  Used for redfield_liou, whose rows hold up to
// N^2 entries (R = L - sum_u Phi_u Theta_u).
template <typename R, bool F64, int KP>
__global__ void __launch_bounds__(256) csr_wave_kernel(const FormDev d, const StageArgs A) {
  using C = cx<R>;
  const Prologue pg = prologue(A);
  if (pg.skip) return;
  const int cur = pg.cur;
  const C* in = static_cast<const C*>(pick(A.in, cur));
  const int lane = threadIdx.x & 63;
  const long long wave = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const long long n_waves = (static_cast<long long>(gridDim.x) * blockDim.x) >> 6;
  double err_acc = 0.0;
  for (long long r = wave; r < d.main_size; r += n_waves) {
    PreT<KP> pv;
    prefetch<F64, KP>(pv, A, cur, r, lane == 0);
    C D{0, 0};
    for (long long p = d.R_ptr[r] + lane; p < d.R_ptr[r + 1]; p += 64)
      cmac(D, conv<R>(d.R_val[p]), in[d.R_col[p]]);
    for (int off = 32; off > 0; off >>= 1) {
      D.re += __shfl_down(D.re, off, 64);
      D.im += __shfl_down(D.im, off, 64);
    }
    if (lane == 0) err_acc += epilogue<F64, R, KP>(pv, A, cur, r, D, pg.alpha);
  }
  if (A.do_err) write_partial(A, err_acc);
}

// rho_0 renormalisation (optional): Tr rho_0 of the initial state, then after
// every step rho_0 <- (rho_0 + rho_0^+)/2 * tr0 / Tr, in FP64.  One workgroup of
// 1024 lanes, up to four elements each (n_level <= 64); every element and its
// transpose are read before any lane writes.  y0/y1 (and the FP32 copies f0/f1,
// may be null) are the two parities of the state; ctrl == null: parity 0.
__global__ void __launch_bounds__(1024) rho0_trace_kernel(const z128* y, int N, double* tr0) {
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < N; ++i) t += y[i * N + i].re;
    *tr0 = t;
  }
}

__global__ void __launch_bounds__(1024) rho0_renorm_kernel(z128* y0, z128* y1, z64* f0, z64* f1,
                                                          Ctrl* ctrl, int N,
                                                          const double* tr0) {
  __shared__ double s_factor;
  // with a control block: project once after each accepted attempt, never after
  // a rejected one or a no-op attempt of a finished interval
  if (ctrl && ctrl->accepted == ctrl->projected) return;
  const int cur = ctrl ? ctrl->cur : 0;
  z128* y = cur ? y1 : y0;
  z64* f = cur ? f1 : f0;
  const int N2 = N * N;
  z128 v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = threadIdx.x + r * 1024;
    if (e < N2) {
      const int i = e / N, j = e - i * N;
      const z128 a = y[e], b = y[j * N + i];
      v[r] = z128{0.5 * (a.re + b.re), 0.5 * (a.im - b.im)};
    }
  }
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < N; ++i) t += y[i * N + i].re;
    const double target = *tr0;
    s_factor = (t != 0.0 && isfinite(t) && isfinite(target)) ? target / t : 1.0;
  }
  __syncthreads();
  const double g = s_factor;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = threadIdx.x + r * 1024;
    if (e < N2) {
      const z128 w{g * v[r].re, g * v[r].im};
      y[e] = w;
      if (f) f[e] = z64{static_cast<float>(w.re), static_cast<float>(w.im)};
    }
  }
  if (ctrl && threadIdx.x == 0) ctrl->projected = ctrl->accepted;
}

// ---------------------------------------------------------------------------
// Resident fixed-step integrator: heom_hilb with diagonal V, FP64, at most
// kResidentMax elements.  One workgroup keeps the state and every stage buffer
// in LDS and runs all steps of an output interval; one lane per element.  The
// per-element link coefficients (neighbour element, c = cm*i*phi + ac*psi +
// dir, the same expression the fast kernel forms per stage) and Gamma_n +
// delta term are computed once into registers.  Stage arithmetic follows the
// epilogue of the stage kernels (k = -dt*D, sum of c_s*operand_s then + ck*k).
// ---------------------------------------------------------------------------
// One lane's element of a resident problem: the link coefficients and
// neighbour elements (computed once) and the operator D at that element.
template <int MAXL>
struct ResidentElem {
  int g, i, j, nl;
  z128 diag;
  int nbr[MAXL];
  z128 cf[MAXL];

  __device__ __forceinline__ void init(const FormDev& d, int e, bool ok) {
    const int N = d.N, N2 = d.N2;
    g = ok ? e / N2 : 0;
    const int ee = ok ? e - g * N2 : 0;
    i = ee / N;
    j = ee - (ee / N) * N;
    const int l0 = d.link_ptr[g], l1 = d.link_ptr[g + 1];
    nl = ok ? l1 - l0 : 0;
    diag = d.gam[g] + d.dterm[ee];
#pragma unroll
    for (int l = 0; l < MAXL; ++l) {
      nbr[l] = 0;
      cf[l] = z128{0.0, 0.0};
      if (l < nl) {
        const LinkFlat q = d.flat[l0 + l];
        nbr[l] = q.nbr * N2 + ee;
        const long long b = static_cast<long long>(q.u) * N2 + ee;
        if (d.real_v) {
          const z128 p = d.cmac_tab[b];
          cf[l].re = p.re * q.iphi.re + p.im * q.psi.re + q.dir.re;
          cf[l].im = p.re * q.iphi.im + p.im * q.psi.im + q.dir.im;
        } else {
          cf[l] = d.cmac_tab[2 * b] * q.iphi + d.cmac_tab[2 * b + 1] * q.psi + q.dir;
        }
      }
    }
  }

  // LDS table: H
  static __host__ __device__ size_t tab_elems(int N, int n_noise) { return static_cast<size_t>(N) * N; }
  static __device__ __forceinline__ void stage(const FormDev& d, z128* tab) {
    for (int x = threadIdx.x; x < d.N2; x += blockDim.x) tab[x] = d.H[x];
  }

  // D at element e of the state X (same expression order as hilb_fast_kernel)
  __device__ __forceinline__ z128 op(const z128* X, const z128* s_H, int e, int N, int N2) const {
    const z128* Xg = X + static_cast<size_t>(g) * N2;
    z128 acc{0.0, 0.0};
    for (int k = 0; k < N; ++k) {
      cmac(acc, s_H[i * N + k], Xg[k * N + j]);
      const z128 t = Xg[i * N + k];
      cmac(acc, z128{-t.re, -t.im}, s_H[k * N + j]);
    }
    z128 D = times_i(acc);
    cmac(D, diag, X[e]);
#pragma unroll
    for (int l = 0; l < MAXL; ++l)
      if (l < nl) cmac(D, cf[l], X[nbr[l]]);
    return D;
  }
};

// One element of a resident redfield_hilb problem with diagonal V: the
// operator of redfield_small_kernel<..., true> (D = HL rho + rho HR -
// sum_u (v_u[i] (rho Lambda_u^+)[i][j] + (Lambda_u rho)[i][j] v_u[j])), same
// expression order, with its tables in LDS.
struct RedfieldElem {
  int i, j, nn;
  // LDS table: HL, HR, per noise Lambda_u and Lambda_u^+, then v_u (n_noise x N)
  static __host__ __device__ size_t tab_elems(int N, int n_noise) {
    return static_cast<size_t>(2 + 2 * n_noise) * N * N + static_cast<size_t>(n_noise) * N;
  }
  static __device__ __forceinline__ void stage(const FormDev& d, z128* tab) {
    const int N2 = d.N2, nn = d.n_noise;
    for (int x = threadIdx.x; x < N2; x += blockDim.x) {
      tab[x] = d.rf_HL[x];
      tab[N2 + x] = d.rf_HR[x];
    }
    for (int x = threadIdx.x; x < nn * N2; x += blockDim.x) {
      const int u = x / N2, y = x - u * N2;
      tab[static_cast<size_t>(2 + 2 * u) * N2 + y] = d.rf_L[x];
      tab[static_cast<size_t>(3 + 2 * u) * N2 + y] = d.rf_Ld[x];
    }
    for (int x = threadIdx.x; x < nn * d.N; x += blockDim.x)
      tab[static_cast<size_t>(2 + 2 * nn) * N2 + x] = d.Vd[x];
  }
  __device__ __forceinline__ void init(const FormDev& d, int e, bool ok) {
    const int ee = ok ? e : 0;
    i = ee / d.N;
    j = ee - i * d.N;
    nn = d.n_noise;
  }
  __device__ __forceinline__ z128 op(const z128* X, const z128* tab, int e, int N, int N2) const {
    const z128* HL = tab;
    const z128* HR = tab + N2;
    z128 D{0.0, 0.0};
    for (int k = 0; k < N; ++k) {
      cmac(D, HL[i * N + k], X[k * N + j]);
      cmac(D, X[i * N + k], HR[k * N + j]);
    }
    const z128* v = tab + static_cast<size_t>(2 + 2 * nn) * N2;
    for (int u = 0; u < nn; ++u) {
      const z128* L = tab + static_cast<size_t>(2 + 2 * u) * N2;
      const z128* Ld = L + N2;
      z128 x{0.0, 0.0}, y{0.0, 0.0};
      for (int k = 0; k < N; ++k) {
        cmac(x, X[i * N + k], Ld[k * N + j]);
        cmac(y, L[i * N + k], X[k * N + j]);
      }
      const z128 vi = v[u * N + i], vj = v[u * N + j];
      cmac(D, z128{-vi.re, -vi.im}, x);
      cmac(D, z128{-vj.re, -vj.im}, y);
    }
    return D;
  }
};

template <class E, int SCHEME>
__global__ void __launch_bounds__(1024) resident_kernel(const FormDev d, const ResidentArgs a) {
  extern __shared__ __align__(16) unsigned char res_lds[];
  z128* buf = reinterpret_cast<z128*>(res_lds);
  // the stage program: uniform reads through a read-only pointer (scalar loads)
  const ResidentStage* __restrict__ prog = a.prog;
  const int ms = static_cast<int>(d.main_size);
  const int N = d.N, N2 = d.N2;
  z128* s_H = buf + static_cast<size_t>(a.n_buf) * ms;  // the element type's tables
  const int e = threadIdx.x;
  const bool ok = e < ms;
  for (int x = e; x < a.n_buf * ms; x += blockDim.x)
    buf[x] = x < ms ? a.y[x] : z128{0.0, 0.0};
  E::stage(d, s_H);
  E el;
  el.init(d, e, ok);
  __syncthreads();
  const int n_int = a.n_int > 0 ? a.n_int : 1;
  for (int iv = 0; iv < n_int; ++iv) {
    const long long n_full = a.n_int > 0 ? a.n_full_i[iv] : a.n_full;
    const double dt_last = a.n_int > 0 ? a.dt_last_i[iv] : a.dt_last;
    if (a.n_int > 0 && e < N2) a.rho_s[static_cast<size_t>(iv) * N2 + e] = buf[e];
    const long long n_steps = n_full + 1;
    for (long long step = 0; step < n_steps; ++step) {
      const double dt = step < n_full ? a.dt : dt_last;
      auto stage = [&](const ResidentStage& st) {
        if (ok) {
          const z128 D = el.op(buf + static_cast<size_t>(st.in) * ms, s_H, e, N, N2);
          const z128 k{-(dt * D.re - 0.0 * D.im), -(dt * D.im + 0.0 * D.re)};
          z128 pv[kResidentPre];
#pragma unroll
          for (int p = 0; p < kResidentPre; ++p)
            pv[p] = p < st.npre ? buf[static_cast<size_t>(st.pre[p]) * ms + e] : z128{0.0, 0.0};
          for (int c = 0; c < st.ncombo; ++c) {
            z128 v{0.0, 0.0};
#pragma unroll
            for (int p = 0; p < kResidentPre; ++p) {
              if (p < st.npre) {
                v.re = fma(st.cc[c][p], pv[p].re, v.re);
                v.im = fma(st.cc[c][p], pv[p].im, v.im);
              }
            }
            v.re = fma(st.ck[c], k.re, v.re);
            v.im = fma(st.ck[c], k.im, v.im);
            buf[static_cast<size_t>(st.dst[c]) * ms + e] = v;
          }
        }
        __syncthreads();
      };
      if constexpr (SCHEME == 0) {
        for (int s = 0; s < a.n_stages; ++s) stage(prog[s]);
      } else {
        constexpr const ResidentProg& P = SCHEME == 1 ? kResidentRk4 : kResidentLsrk4;
#pragma unroll
        for (int s = 0; s < P.n_stages; ++s) stage(P.st[s]);
      }
    }
  }
  if (a.n_int > 0 && e < N2) a.rho_s[static_cast<size_t>(a.n_int) * N2 + e] = buf[e];
  if (ok) a.y[e] = buf[e];
}

// Resident Dormand-Prince: the whole adaptive solve of a small heom_hilb
// problem (diagonal V, FP64) in one workgroup.  Stage arithmetic follows the
// stage kernels (k = -dt*D; y + sum a_ij k_j as one fma chain in operand order;
// the error combination and weights of rkdp.h:136-153), the step control is
// dopri_control on a control block in LDS, and each output interval starts as
// ctrl_init_kernel does.  The error norm is summed in a fixed order (wavefront
// shuffles, then the waves in order).  LDS buffers: 0/1 the y / y_new pair
// (swapped by Ctrl::cur on acceptance), 2..7 k1..k6, 8/9 the stage inputs.
template <class E>
__global__ void __launch_bounds__(1024)
    resident_dopri_kernel(const FormDev d, const ResidentDopriArgs a) {
  extern __shared__ __align__(16) unsigned char res_lds[];
  __shared__ Ctrl s_c;
  __shared__ double s_red[16];
  __shared__ int s_gave_up;
  z128* buf = reinterpret_cast<z128*>(res_lds);
  const int ms = static_cast<int>(d.main_size);
  const int N = d.N, N2 = d.N2;
  z128* s_H = buf + static_cast<size_t>(kResidentDopriBuf) * ms;
  const int e = threadIdx.x;
  const bool ok = e < ms;
  for (int x = e; x < kResidentDopriBuf * ms; x += blockDim.x)
    buf[x] = x < ms ? a.y[x] : z128{0.0, 0.0};
  E::stage(d, s_H);
  E el;
  el.init(d, e, ok);
  if (e == 0) {
    s_c = Ctrl{};
    s_c.dt = a.dt_first;
    s_gave_up = 0;
  }
  __syncthreads();
  if (e < N2) a.rho_s[e] = buf[e];
  z128* K = buf + 2 * static_cast<size_t>(ms);
  z128* SA = buf + 8 * static_cast<size_t>(ms);
  z128* SB = buf + 9 * static_cast<size_t>(ms);
  for (int iv = 0; iv < a.n_int; ++iv) {
    if (e == 0) {  // ctrl_init_kernel: dt carries over from the last attempt
      const double t0 = a.t_list[iv], t1 = a.t_list[iv + 1];
      double dt = s_c.dt;
      s_c.t = t0;
      s_c.t_bound = t1;
      s_c.count = 1;
      s_c.gave_up = 0;
      // a zero-length interval takes no attempt and keeps the step size
      s_c.done = (t1 > t0) ? 0 : 1;
      if (!s_c.done) {
        if (t0 + dt > t1) dt = t1 - t0;
        s_c.dt = dt;
      }
    }
    __syncthreads();
    while (!s_c.done) {
      const double dt = s_c.dt;
      const int cur = s_c.cur;
      const z128* Y = buf + (cur ? ms : 0);
      z128* YN = buf + (cur ? 0 : ms);
#pragma unroll
      for (int st = 0; st < 6; ++st) {
        const z128* in = st == 0 ? Y : ((st & 1) ? SA : SB);
        z128* out = st == 5 ? YN : ((st & 1) ? SB : SA);
        if (ok) {
          const z128 D = el.op(in, s_H, e, N, N2);
          const z128 k{-(dt * D.re - 0.0 * D.im), -(dt * D.im + 0.0 * D.re)};
          K[static_cast<size_t>(st) * ms + e] = k;
          const z128 y = Y[e];
          z128 v{fma(1.0, y.re, 0.0), fma(1.0, y.im, 0.0)};
          if (st < 5) {
#pragma unroll
            for (int j = 0; j < st; ++j) {
              const z128 kj = K[static_cast<size_t>(j) * ms + e];
              v.re = fma(kDpA[st + 1][j], kj.re, v.re);
              v.im = fma(kDpA[st + 1][j], kj.im, v.im);
            }
            v.re = fma(kDpA[st + 1][st], k.re, v.re);
            v.im = fma(kDpA[st + 1][st], k.im, v.im);
          } else {
#pragma unroll
            for (int j = 0; j < 5; ++j) {
              if (kDpB[j] != 0.0) {
                const z128 kj = K[static_cast<size_t>(j) * ms + e];
                v.re = fma(kDpB[j], kj.re, v.re);
                v.im = fma(kDpB[j], kj.im, v.im);
              }
            }
            v.re = fma(kDpB[5], k.re, v.re);
            v.im = fma(kDpB[5], k.im, v.im);
          }
          out[e] = v;
        }
        __syncthreads();
      }
      double w2 = 0.0;
      if (ok) {
        const z128 D = el.op(YN, s_H, e, N, N2);
        const z128 k{-(dt * D.re - 0.0 * D.im), -(dt * D.im + 0.0 * D.re)};
        z128 er{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          if (kDpE[j] != 0.0) {
            const z128 kj = K[static_cast<size_t>(j) * ms + e];
            er.re = fma(kDpE[j], kj.re, er.re);
            er.im = fma(kDpE[j], kj.im, er.im);
          }
        }
        // the stage kernels' padded sixth operand adds 0*0
        er.re = fma(0.0, 0.0, er.re);
        er.im = fma(0.0, 0.0, er.im);
        er.re = fma(kDpE[6], k.re, er.re);
        er.im = fma(kDpE[6], k.im, er.im);
        const double a1 = zabs(YN[e]), a0 = zabs(Y[e]);
        const double w = zabs(er) / (a.atol + a.rtol * (a1 > a0 ? a1 : a0));
        w2 = w * w;
      }
      for (int off = 32; off > 0; off >>= 1) w2 += __shfl_down(w2, off, 64);
      if ((e & 63) == 0) s_red[e >> 6] = w2;
      __syncthreads();
      if (e == 0) {
        double t = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += s_red[w];
        dopri_control(&s_c, sqrt(t * a.inv_n));
        if (s_c.gave_up) s_gave_up = 1;
      }
      __syncthreads();
    }
    // every wave has read done = 1 before thread 0 rewrites the control block
    // for the next interval
    __syncthreads();
    if (e < N2) a.rho_s[static_cast<size_t>(iv + 1) * N2 + e] = buf[(s_c.cur ? ms : 0) + e];
  }
  if (ok) a.y[e] = buf[(s_c.cur ? ms : 0) + e];
  if (e == 0) {
    Ctrl c = s_c;
    c.gave_up = s_gave_up;
    *a.ctrl = c;
  }
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

inline int cap_grid(long long n) {
  if (n < 1) return 1;
  return n < kMaxGrid ? static_cast<int>(n) : kMaxGrid;
}

template <int N, typename R, bool F64>
int launch_hilb_n(const FormDev& d, const StageArgs& A, hipStream_t st) {
  const bool real = d.real_v && d.real_h;
  if (N >= 6 && d.diag_v && d.bmap != nullptr && d.n_hrchy <= kMaxGrid &&
      fast_vd_lds(d, real, sizeof(R)) <= 32 * 1024) {
    constexpr int BS = WaveCfg<N>::BS;
    const int grid = d.n_hrchy;
    const size_t lds = fast_vd_lds(d, real, sizeof(R));
    // four prefetched operands cover every stage except the last Dormand-Prince one
    const bool small = A.npre <= 4;
    const bool mid = A.npre <= 6;  // Dormand-Prince stages 5 and 6
    if (real) {
      if (small)
        hipLaunchKernelGGL((hilb_fast_kernel<N, R, true, F64, 4>), dim3(grid), dim3(BS), lds, st,
                           d, A);
      else if (mid)
        hipLaunchKernelGGL((hilb_fast_kernel<N, R, true, F64, 6>), dim3(grid), dim3(BS), lds, st,
                           d, A);
      else
        hipLaunchKernelGGL((hilb_fast_kernel<N, R, true, F64, kMaxPre>), dim3(grid), dim3(BS),
                           lds, st, d, A);
      return grid;
    } else {
      if (small)
        hipLaunchKernelGGL((hilb_fast_kernel<N, R, false, F64, 4>), dim3(grid), dim3(BS), lds,
                           st, d, A);
      else
        hipLaunchKernelGGL((hilb_fast_kernel<N, R, false, F64, kMaxPre>), dim3(grid), dim3(BS),
                           lds, st, d, A);
      return grid;
    }
  }
  constexpr int APB = HilbCfg<N>::APB, BS = HilbCfg<N>::BS;
  const int grid = cap_grid((d.n_hrchy + APB - 1) / APB);
  const bool small = A.npre <= 4;
  if (d.diag_v) {
    const size_t lds = sizeof(cx<R>) * (d.n_noise + 1) * N;
    if (lds > 48 * 1024)
      throw HeomError{HEOM_ERR_UNSUPPORTED, "heom_hilb: too many noises for the LDS tables"};
    if (small)
      hipLaunchKernelGGL((hilb_kernel<N, R, true, F64, 4>), dim3(grid), dim3(BS), lds, st, d, A);
    else
      hipLaunchKernelGGL((hilb_kernel<N, R, true, F64, kMaxPre>), dim3(grid), dim3(BS), lds, st,
                         d, A);
  } else {
    if (small)
      hipLaunchKernelGGL((hilb_kernel<N, R, false, F64, 4>), dim3(grid), dim3(BS), 0, st, d, A);
    else
      hipLaunchKernelGGL((hilb_kernel<N, R, false, F64, kMaxPre>), dim3(grid), dim3(BS), 0, st, d,
                         A);
  }
  return grid;
}

// dynamic LDS above 64 KB needs the per-kernel limit raised once
template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    HAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 static_cast<int>(bytes)));
}

template <int N, typename R, bool F64, bool BF16X3>
int launch_hilb_mfma_n(const FormDev& d, const StageArgs& A, hipStream_t st) {
  // rho_n planes, then the diagonals of every V_u and a zero row
  const size_t lds = 2 * sizeof(R) * N * MfmaCfg<N>::LD +
                     sizeof(R) * (d.real_v ? 1 : 2) * static_cast<size_t>(d.n_noise + 1) * d.N;
  if (lds > kMfmaLdsCap)
    throw HeomError{HEOM_ERR_UNSUPPORTED, "matrix-core path: too many noises for LDS"};
  const int grid = cap_grid(d.n_hrchy);
  const bool small = A.npre <= 4;
  // N = 64: eight waves per ADO, each with half the output tiles (two
  // workgroups per CU either way, so twice the waves work on the same ADOs;
  // measured at 969 ADOs per Dormand-Prince stage: FP64 135 against 158 us,
  // split-bf16 99 against 138 us, complex64 with the links outer 112 against
  // 126 us)
  constexpr int BS = N == 64 ? 512 : kMfmaBS;
#define HAMD_MFMA(REALH, KP, PAD)                                                \
  {                                                                              \
    auto k = hilb_mfma_kernel<N, R, REALH, F64, KP, BF16X3, PAD, BS>;            \
    static bool once = (allow_lds(k, kMfmaLdsCap), true);                        \
    (void)once;                                                                  \
    hipLaunchKernelGGL(k, dim3(grid), dim3(BS), lds, st, d, A);                  \
  }
#define HAMD_MFMA_P(REALH, KP)                          \
  {                                                     \
    if (d.N == N) HAMD_MFMA(REALH, KP, false) else HAMD_MFMA(REALH, KP, true) \
  }
  if (d.real_h) {
    if (small) HAMD_MFMA_P(true, 4) else HAMD_MFMA_P(true, kMaxPre)
  } else {
    if (small) HAMD_MFMA_P(false, 4) else HAMD_MFMA_P(false, kMaxPre)
  }
#undef HAMD_MFMA_P
#undef HAMD_MFMA
  return grid;
}

template <typename R, bool F64>
int launch_hilb_mfma(const FormDev& d, const StageArgs& A, hipStream_t st) {
  if (d.mfma == 2) {
    if constexpr (std::is_same<R, float>::value) {
      return d.N <= 32 ? launch_hilb_mfma_n<32, float, false, true>(d, A, st)
                       : launch_hilb_mfma_n<64, float, false, true>(d, A, st);
    } else {
      throw HeomError{HEOM_ERR_UNSUPPORTED, "bf16x3 products need the FP32 state"};
    }
  }
  return d.N <= 32 ? launch_hilb_mfma_n<32, R, F64, false>(d, A, st)
                   : launch_hilb_mfma_n<64, R, F64, false>(d, A, st);
}

template <int N, typename R, bool F64>
int launch_hilb_mfma_dense_n(const FormDev& d, const StageArgs& A, hipStream_t st) {
  // rho_n planes and the workspace planes
  const size_t lds = 4 * sizeof(R) * N * MfmaCfg<N>::LD;
  if (lds > kMfmaLdsCap)
    throw HeomError{HEOM_ERR_UNSUPPORTED, "matrix-core path: the blocks exceed LDS"};
  const int grid = cap_grid(d.n_hrchy);
  const bool small = A.npre <= 4;
  // N = 64: eight waves per ADO (measured at 969 ADOs: 497 against 715 us per
  // FP64 stage, 309 against 364 us complex64)
  constexpr int BS = N == 64 ? 512 : kMfmaBS;
#define HAMD_MD(REALV, REALH, KP, PAD)                                                 \
  {                                                                                    \
    auto k = hilb_mfma_dense_kernel<N, R, REALV, REALH, F64, KP, PAD, BS>;             \
    static bool once = (allow_lds(k, kMfmaLdsCap), true);                              \
    (void)once;                                                                        \
    hipLaunchKernelGGL(k, dim3(grid), dim3(BS), lds, st, d, A);                        \
  }
#define HAMD_MD_P(REALV, REALH, KP) \
  {                                 \
    if (d.N == N) HAMD_MD(REALV, REALH, KP, false) else HAMD_MD(REALV, REALH, KP, true) \
  }
#define HAMD_MD_K(REALV, REALH) \
  {                             \
    if (small) HAMD_MD_P(REALV, REALH, 4) else HAMD_MD_P(REALV, REALH, kMaxPre) \
  }
  if (d.real_vd) {
    if (d.real_h) HAMD_MD_K(true, true) else HAMD_MD_K(true, false)
  } else {
    if (d.real_h) HAMD_MD_K(false, true) else HAMD_MD_K(false, false)
  }
#undef HAMD_MD_K
#undef HAMD_MD_P
#undef HAMD_MD
  return grid;
}

template <typename R, bool F64>
int launch_hilb_big(const FormDev& d, const StageArgs& A, hipStream_t st) {
  if (d.mfma == 3)  // 16 < N <= 64, dense V
    return d.N <= 32 ? launch_hilb_mfma_dense_n<32, R, F64>(d, A, st)
                     : launch_hilb_mfma_dense_n<64, R, F64>(d, A, st);
  if (d.mfma) return launch_hilb_mfma<R, F64>(d, A, st);  // 16 < N <= 64, diagonal V
  if (!d.diag_v)
    throw HeomError{HEOM_ERR_UNSUPPORTED,
                    "heom_hilb: n_level > 16 is provided for diagonal coupling operators V"};
  const size_t one = sizeof(cx<R>) * static_cast<size_t>(d.N2);
  const int grid = cap_grid(d.n_hrchy);
  if (2 * one <= 64 * 1024) {
    if (A.npre <= 4)
      hipLaunchKernelGGL((hilb_big_kernel<R, true, F64, 4>), dim3(grid), dim3(kBigBS), 2 * one,
                         st, d, A);
    else
      hipLaunchKernelGGL((hilb_big_kernel<R, true, F64, kMaxPre>), dim3(grid), dim3(kBigBS),
                         2 * one, st, d, A);
  } else if (one <= 64 * 1024) {
    if (A.npre <= 4)
      hipLaunchKernelGGL((hilb_big_kernel<R, false, F64, 4>), dim3(grid), dim3(kBigBS), one, st,
                         d, A);
    else
      hipLaunchKernelGGL((hilb_big_kernel<R, false, F64, kMaxPre>), dim3(grid), dim3(kBigBS), one,
                         st, d, A);
  } else {
    throw HeomError{HEOM_ERR_UNSUPPORTED, "heom_hilb: one ADO block exceeds 64 KB of LDS"};
  }
  return grid;
}

template <typename R, bool F64>
int launch_hilb(const FormDev& d, const StageArgs& A, hipStream_t st) {
  switch (d.N) {
#define HAMD_CASE(n) \
  case n:            \
    return launch_hilb_n<n, R, F64>(d, A, st);
    HAMD_CASE(1) HAMD_CASE(2) HAMD_CASE(3) HAMD_CASE(4) HAMD_CASE(5) HAMD_CASE(6) HAMD_CASE(7)
    HAMD_CASE(8) HAMD_CASE(9) HAMD_CASE(10) HAMD_CASE(11) HAMD_CASE(12) HAMD_CASE(13)
    HAMD_CASE(14) HAMD_CASE(15) HAMD_CASE(16)
#undef HAMD_CASE
    default:
      return launch_hilb_big<R, F64>(d, A, st);
  }
}

template <typename R, bool F64>
int launch_redfield(const FormDev& d, const StageArgs& A, hipStream_t st) {
  const bool small = A.npre <= 4;
  if (d.kind == HEOM_REDFIELD_LIOU) {
    const int grid = cap_grid((d.main_size + 3) / 4);  // four rows per workgroup
    if (small)
      hipLaunchKernelGGL((csr_wave_kernel<R, F64, 4>), dim3(grid), dim3(256), 0, st, d, A);
    else
      hipLaunchKernelGGL((csr_wave_kernel<R, F64, kMaxPre>), dim3(grid), dim3(256), 0, st, d, A);
    return grid;
  }
  const bool diag = d.rf_HL != nullptr;
  const size_t lds = redfield_lds_elems(d.N2, d.N, d.n_noise, diag) * sizeof(cx<R>);
  if (d.N <= 16 && lds <= kRedfieldLds) {
#define HAMD_RF(DG, KPV)                                                                    \
  hipLaunchKernelGGL((redfield_small_kernel<R, F64, KPV, DG>), dim3(1), dim3(256), lds, st, d, \
                     A)
    if (diag) {
      if (small) HAMD_RF(true, 4); else HAMD_RF(true, kMaxPre);
    } else {
      if (small) HAMD_RF(false, 4); else HAMD_RF(false, kMaxPre);
    }
#undef HAMD_RF
    return 1;
  }
  if (d.rf_T == nullptr) throw HeomError{HEOM_ERR_STATE, "redfield_hilb: scratch missing"};
  const long long nt = static_cast<long long>(d.n_noise) * d.N2;
  hipLaunchKernelGGL((redfield_t_kernel<R>), dim3(cap_grid((nt + 255) / 256)), dim3(256), 0, st,
                     d, A);
  const int grid = cap_grid((d.N2 + 255) / 256);
  if (small)
    hipLaunchKernelGGL((redfield_d_kernel<R, F64, 4>), dim3(grid), dim3(256), 0, st, d, A);
  else
    hipLaunchKernelGGL((redfield_d_kernel<R, F64, kMaxPre>), dim3(grid), dim3(256), 0, st, d, A);
  return grid;
}

template <typename R, bool F64>
int launch_form(const FormDev& d, const StageArgs& A, hipStream_t st) {
  if (d.kind == 0) return launch_hilb<R, F64>(d, A, st);
  if (d.kind == HEOM_REDFIELD_HILB || d.kind == HEOM_REDFIELD_LIOU)
    return launch_redfield<R, F64>(d, A, st);
  if (d.kind == kBatch) {
    const int wpb = d.batch_wpb > 0 ? d.batch_wpb : 1;
    // the persistent loop covers every item at any grid; keep one partial slot per wavefront
    const int grid = std::max(1, std::min(d.batch_grid, kMaxGrid / wpb));
    const bool small = A.npre <= 4;
#define HAMD_BATCH(IPW, WPB)                                                               \
  {                                                                                        \
    if (small)                                                                             \
      hipLaunchKernelGGL((batch_kernel<R, F64, 4, IPW, WPB>), dim3(grid), dim3(64 * WPB), 0, \
                         st, d, A);                                                        \
    else                                                                                   \
      hipLaunchKernelGGL((batch_kernel<R, F64, kMaxPre, IPW, WPB>), dim3(grid),            \
                         dim3(64 * WPB), 0, st, d, A);                                     \
  }
#define HAMD_BATCH_W(IPW)                   \
  {                                         \
    if (wpb == 4) HAMD_BATCH(IPW, 4)        \
    else if (wpb == 2) HAMD_BATCH(IPW, 2)   \
    else HAMD_BATCH(IPW, 1)                 \
  }
    if (d.batch_ipw == 2) HAMD_BATCH_W(2) else HAMD_BATCH_W(1)
#undef HAMD_BATCH_W
#undef HAMD_BATCH
    return grid * wpb;  // error-norm partials: one per wavefront
  }
  if (d.kind == 1 && d.N2 > 256) {
    // 16 < N <= 32: one ADO per workgroup (set_param refuses larger N)
    const int grid = cap_grid(d.n_hrchy);
    const size_t lds = 3 * sizeof(cx<R>) * static_cast<size_t>(d.N2);
    if (A.npre <= 4)
      hipLaunchKernelGGL((liou_wide_kernel<R, F64, 4>), dim3(grid), dim3(256), lds, st, d, A);
    else
      hipLaunchKernelGGL((liou_wide_kernel<R, F64, kMaxPre>), dim3(grid), dim3(256), lds, st, d,
                         A);
    return grid;
  }
  if (d.kind == 1) {
    const int APB = 256 / d.N2;
    const int grid = cap_grid((d.n_hrchy + APB - 1) / APB);
    if (A.npre <= 4)
      hipLaunchKernelGGL((liou_kernel<R, F64, 4>), dim3(grid), dim3(256), 0, st, d, A);
    else
      hipLaunchKernelGGL((liou_kernel<R, F64, kMaxPre>), dim3(grid), dim3(256), 0, st, d, A);
    return grid;
  }
  const int grid = cap_grid((d.main_size + 255) / 256);
  if (A.npre <= 4)
    hipLaunchKernelGGL((ado_kernel<R, F64, 4>), dim3(grid), dim3(256), 0, st, d, A);
  else
    hipLaunchKernelGGL((ado_kernel<R, F64, kMaxPre>), dim3(grid), dim3(256), 0, st, d, A);
  return grid;
}

}  // namespace

int launch_operator(const FormDev& d, int real_bytes, const StageArgs& A, hipStream_t stream) {
  // every operand complex128: the epilogue needs no per-operand type checks
  bool all64 = real_bytes == 8 && !A.in_lo && A.pre_lo == 0 && !A.k_lo;
  for (int c = 0; c < 2; ++c) all64 = all64 && !A.dst_lo[c] && !A.dst2_lo[c];
  int g;
  if (real_bytes == 4)
    g = launch_form<float, false>(d, A, stream);
  else if (all64)
    g = launch_form<double, true>(d, A, stream);
  else
    g = launch_form<double, false>(d, A, stream);
  HAMD_HIP(hipGetLastError());
  return g;
}

void launch_err_finalize(const StageArgs& A, int nparts, hipStream_t stream) {
  hipLaunchKernelGGL(err_finalize_kernel, dim3(1), dim3(1024), 0, stream, A, nparts);
  HAMD_HIP(hipGetLastError());
}

void launch_ctrl_init(Ctrl* c, double t0, double t_bound, double dt, hipStream_t stream) {
  hipLaunchKernelGGL(ctrl_init_kernel, dim3(1), dim3(64), 0, stream, c, t0, t_bound, dt);
  HAMD_HIP(hipGetLastError());
}

void launch_closure_gather(const FormDev& d, const StageArgs& A, float* x, int ado_stride,
                           hipStream_t stream) {
  const long long n = static_cast<long long>(d.clos_n) * d.N2;
  hipLaunchKernelGGL(closure_gather_kernel, dim3(cap_grid((n + 255) / 256)), dim3(256), 0, stream,
                     d, A, x, ado_stride);
  HAMD_HIP(hipGetLastError());
}

bool closure_mlp_fits(int n_in, int n_hid, int n_out) {
  return n_in <= kClosInMax && n_hid % 16 == 0 && n_hid <= kClosHidMax && n_out % 16 == 0 &&
         n_out <= 64;
}

void launch_closure_mlp(const FormDev& d, const StageArgs& A, const float* xc, int n_in,
                        int n_hid, int n_out, const float* W1, const float* b1, const float* W2,
                        const float* b2, const float* W3, const float* b3, float* y,
                        hipStream_t stream) {
  if (!closure_mlp_fits(n_in, n_hid, n_out))
    throw HeomError{HEOM_ERR_UNSUPPORTED, "closure: network too wide for the fused kernel"};
  const int grid = (d.clos_n + kClosRows - 1) / kClosRows;
  hipLaunchKernelGGL(closure_mlp_kernel, dim3(grid), dim3(256), 0, stream, d, A, xc, n_in, n_hid,
                     n_out, W1, b1, W2, b2, W3, b3, y);
  HAMD_HIP(hipGetLastError());
}

void launch_gather_rho_s(const FormDev& d, void* dst, const void* src, hipStream_t stream) {
  hipLaunchKernelGGL(gather_rho_s_kernel, dim3(d.n_inst), dim3(64), 0, stream,
                     static_cast<z128*>(dst), static_cast<const z128*>(src), d.inst);
  HAMD_HIP(hipGetLastError());
}

void launch_rho0_trace(const void* y, int N, double* tr0, hipStream_t stream) {
  hipLaunchKernelGGL(rho0_trace_kernel, dim3(1), dim3(64), 0, stream, static_cast<const z128*>(y),
                     N, tr0);
  HAMD_HIP(hipGetLastError());
}

void launch_rho0_renorm(void* y0, void* y1, void* f0, void* f1, Ctrl* ctrl, int N,
                        const double* tr0, hipStream_t stream) {
  if (N > 64) throw HeomError{HEOM_ERR_UNSUPPORTED, "renorm_rho0: n_level > 64"};
  hipLaunchKernelGGL(rho0_renorm_kernel, dim3(1), dim3(1024), 0, stream, static_cast<z128*>(y0),
                     static_cast<z128*>(y1), static_cast<z64*>(f0), static_cast<z64*>(f1), ctrl,
                     N, tr0);
  HAMD_HIP(hipGetLastError());
}

bool resident_ok(const FormDev& d, int max_links) {
  if (d.main_size > kResidentMax || d.clos != nullptr) return false;
  if (d.kind == HEOM_HILB) return d.diag_v && max_links <= 8;
  return d.kind == HEOM_REDFIELD_HILB && d.rf_HL != nullptr;  // diagonal V
}

namespace {
size_t resident_tab_elems(const FormDev& d) {
  return d.kind == HEOM_REDFIELD_HILB ? RedfieldElem::tab_elems(d.N, d.n_noise)
                                      : ResidentElem<4>::tab_elems(d.N, d.n_noise);
}
}  // namespace

size_t resident_lds_bytes(const FormDev& d, int n_buf) {
  return sizeof(z128) * (static_cast<size_t>(n_buf) * d.main_size + resident_tab_elems(d));
}

void launch_resident(const FormDev& d, int max_links, const ResidentArgs& a, hipStream_t stream) {
  const int ms = static_cast<int>(d.main_size);
  if (!resident_ok(d, max_links) || a.n_stages < 1 || a.n_stages > 4 || a.n_buf < 1 || !a.prog)
    throw HeomError{HEOM_ERR_STATE, "resident integrator: unsupported problem"};
  const size_t lds = resident_lds_bytes(d, a.n_buf);
  if (lds > kResidentLds) throw HeomError{HEOM_ERR_STATE, "resident integrator: LDS exceeded"};
  const int bs = ((ms + 63) / 64) * 64;
#define HAMD_RES(E, S)                                                                \
  do {                                                                                \
    static bool once = (allow_lds(resident_kernel<E, S>, kResidentLds), true);        \
    (void)once;                                                                       \
    hipLaunchKernelGGL((resident_kernel<E, S>), dim3(1), dim3(bs), lds, stream, d, a); \
  } while (0)
#define HAMD_RES_S(E)       \
  do {                      \
    if (a.scheme == 1)      \
      HAMD_RES(E, 1);       \
    else if (a.scheme == 2) \
      HAMD_RES(E, 2);       \
    else                    \
      HAMD_RES(E, 0);       \
  } while (0)
  if (d.kind == HEOM_REDFIELD_HILB)
    HAMD_RES_S(RedfieldElem);
  else if (max_links <= 4)
    HAMD_RES_S(ResidentElem<4>);
  else
    HAMD_RES_S(ResidentElem<8>);
#undef HAMD_RES_S
#undef HAMD_RES
  HAMD_HIP(hipGetLastError());
}

bool resident_dopri_ok(const FormDev& d, int max_links) {
  return resident_ok(d, max_links) && d.main_size <= kResidentDopriMax &&
         resident_lds_bytes(d, kResidentDopriBuf) <= kResidentLds;
}

void launch_resident_dopri(const FormDev& d, int max_links, const ResidentDopriArgs& a,
                           hipStream_t stream) {
  if (!resident_dopri_ok(d, max_links) || a.n_int < 1)
    throw HeomError{HEOM_ERR_STATE, "resident Dormand-Prince: unsupported problem"};
  const int ms = static_cast<int>(d.main_size);
  const size_t lds = resident_lds_bytes(d, kResidentDopriBuf);
  if (lds > kResidentLds) throw HeomError{HEOM_ERR_STATE, "resident Dormand-Prince: LDS exceeded"};
  const int bs = ((ms + 63) / 64) * 64;
#define HAMD_RDP(E)                                                                      \
  do {                                                                                   \
    static bool once = (allow_lds(resident_dopri_kernel<E>, kResidentLds), true);        \
    (void)once;                                                                          \
    hipLaunchKernelGGL((resident_dopri_kernel<E>), dim3(1), dim3(bs), lds, stream, d, a); \
  } while (0)
  if (d.kind == HEOM_REDFIELD_HILB)
    HAMD_RDP(RedfieldElem);
  else if (max_links <= 4)
    HAMD_RDP(ResidentElem<4>);
  else
    HAMD_RDP(ResidentElem<8>);
#undef HAMD_RDP
  HAMD_HIP(hipGetLastError());
}

void launch_narrow(void* dst, const void* src, long long n, hipStream_t stream) {
  hipLaunchKernelGGL(narrow_kernel, dim3(cap_grid((n + 255) / 256)), dim3(256), 0, stream,
                     static_cast<z64*>(dst), static_cast<const z128*>(src), n);
  HAMD_HIP(hipGetLastError());
}

void launch_copy_grid(void* dst, const void* src, long long n, int grid, hipStream_t stream) {
  hipLaunchKernelGGL((copy_kernel<z128>), dim3(grid), dim3(256), 0, stream, static_cast<z128*>(dst),
                     static_cast<const z128*>(src), n);
  HAMD_HIP(hipGetLastError());
}

void launch_copy(void* dst, const void* src, long long n, int type, hipStream_t stream) {
  const int grid = cap_grid((n + 255) / 256);
  if (type == ST_C128)
    hipLaunchKernelGGL((copy_kernel<z128>), dim3(grid), dim3(256), 0, stream,
                       static_cast<z128*>(dst), static_cast<const z128*>(src), n);
  else
    hipLaunchKernelGGL((copy_kernel<z64>), dim3(grid), dim3(256), 0, stream,
                       static_cast<z64*>(dst), static_cast<const z64*>(src), n);
  HAMD_HIP(hipGetLastError());
}

}  // namespace hamd
