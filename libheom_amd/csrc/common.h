// Shared host/device definitions of the libheom_amd engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace hamd {

// ---------------------------------------------------------------------------
// complex arithmetic on plain structs (host + device)
// ---------------------------------------------------------------------------
template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R re, im;
};
using z128 = cx<double>;
using z64 = cx<float>;

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) {
  return cx<R>{a, b};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) {
  return cx<R>{a.re + b.re, a.im + b.im};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) {
  return cx<R>{a.re - b.re, a.im - b.im};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return cx<R>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(R s, cx<R> b) {
  return cx<R>{s * b.re, s * b.im};
}
// a += b*c
template <typename R>
__host__ __device__ __forceinline__ void cmac(cx<R>& a, cx<R> b, cx<R> c) {
  a.re += b.re * c.re - b.im * c.im;
  a.im += b.re * c.im + b.im * c.re;
}
// i*a
template <typename R>
__host__ __device__ __forceinline__ cx<R> times_i(cx<R> a) {
  return cx<R>{-a.im, a.re};
}
template <typename R, typename S>
__host__ __device__ __forceinline__ cx<R> conv(cx<S> a) {
  return cx<R>{static_cast<R>(a.re), static_cast<R>(a.im)};
}

// ---------------------------------------------------------------------------
// device-side tables
// ---------------------------------------------------------------------------
// Coefficient of one neighbour link (type of link x mode x occupation).
// Contribution of neighbour block X to D_n (heom_hilb.h:128-208):
//   D_n += c_dir * X                         (off-diagonal gamma term)
//   Phi_u += c_phi * X,  Psi_u += c_psi * X  (then i[V_u,Phi_u] + {V_u,Psi_u})
struct Coef {
  z128 phi;
  z128 psi;
  z128 dir;
  int u;  // noise index; == n_noise for pure direct links
  int pad;
};

struct Link {
  int nbr;  // neighbour ADO index
  int cid;  // index into the Coef table
};

// Dormand-Prince control block, resident in device memory.  The last block of
// the final stage kernel of an attempt updates it (rkdp.h:97-177) so that the
// host never reads back inside a step.
struct Ctrl {
  double t, t_bound, dt, err;
  int cur;      // which buffer of each ping-pong pair holds the current state
  int done;     // interval finished (t >= t_bound) or gave up
  int count;    // attempts of the current step (rkdp.h:94-102)
  int gave_up;
  long long attempts, accepted, rejected;
};

enum SlotType { ST_C128 = 0, ST_C64 = 1 };
constexpr int kMaxSlots = 16;
// Slots 0..3 are ping-pong pairs resolved through Ctrl::cur when a Ctrl is
// attached: 0 = state (FP64), 1 = new state (FP64), 2 = state (stage-input
// precision), 3 = new state (stage-input precision).
struct Slots {
  void* p[kMaxSlots];
  int type[kMaxSlots];
};

// dst (and dst2) <- sum_j c[j]*src[j] + ck*k   (sequential, FP64)
struct Combo {
  int nsrc;
  int src[7];
  double c[7];
  double ck;
  int dst;
  int dst2;  // -1: none
};

enum Mode { MODE_DIFF = 0, MODE_STAGE = 1 };

struct StageArgs {
  int mode;
  int in_slot;   // operator input (neighbour reads)
  // MODE_DIFF: out = beta*out - alpha*D
  int out_slot;
  double a_re, a_im, b_re, b_im;
  // MODE_STAGE: k = -dt*D (dt from ctrl when attached, else a_re)
  int k_slot;  // -1: k not stored
  int ncombo;
  Combo combo[2];
  // error norm (rkdp.h:142-153): e = sum ec[j]*k_src[j] + ek*k,
  // w = |e| / (atol + rtol*max(|y_new|, |y_old|)), partial sums of w^2
  int do_err;
  int ne;
  int esrc[6];
  double ec[6];
  double ek;
  int ynew_slot, yold_slot;
  double atol, rtol;
  int control;  // 1: last block runs the step control
  double inv_n;
  Ctrl* ctrl;
  double* partials;
  unsigned* counter;
  int max_partials;
};

// Operator tables of one QME in device memory.
struct FormDev {
  int kind;  // 0 hilb, 1 liou, 2 ado
  int N, N2, n_hrchy, n_noise;
  long long main_size;
  const int* link_ptr;   // n_hrchy+1
  const Link* links;     // sorted by (ADO, noise)
  const Coef* coef;
  const z128* gam;       // Gamma_n (n_hrchy)
  const z128* H;         // N x N
  const z128* V;         // n_noise x N x N
  const z128* Vd;        // n_noise x N (diagonal V)
  const z128* dterm;     // N2: sum_u delta_u (V_ii - V_jj)^2 (diagonal V)
  const z128* sdelta;    // n_noise
  int diag_v;
  // liou: CSR superoperators (N2 x N2): R0, then Phi_u, Psi_u
  const int* sop_ptr;    // (1 + 2*n_noise) * (N2 + 1)
  const int* sop_col;
  const z128* sop_val;
  // ado: global CSR
  const long long* R_ptr;
  const int* R_col;
  const z128* R_val;
};

// ---------------------------------------------------------------------------
// host-side error helpers
// ---------------------------------------------------------------------------
struct HeomError {
  int code;
  std::string msg;
};

void set_error(int code, const std::string& msg);

#define HAMD_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::hamd::HeomError{3, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

}  // namespace hamd
