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
  long long projected;  // accepted attempts whose rho_0 the projection has handled
};

enum SlotType { ST_C128 = 0, ST_C64 = 1 };

enum Mode { MODE_DIFF = 0, MODE_STAGE = 1 };

constexpr int kMaxPre = 8;  // state-sized operands an epilogue may read per element

// Epilogue of one operator pass, applied per element after D is known.  All
// buffer pointers are resolved on the host, once for each value of Ctrl::cur
// (which buffer of each ping-pong pair holds the current state); without a
// Ctrl both entries are equal.  "lo" bits mark complex64 operands.
//   pv[s] = pre[cur][s] at this element (loaded before the operator work)
// MODE_DIFF : dst[0] = beta*pv[0] - alpha*D      (pv[0] = old drho when beta != 0)
// MODE_STAGE: k = -dt*D (dt from the control block when attached, else a_re)
//             k[cur] <- k (optional)
//             dst[c] (and dst2[c]) <- sum_s cc[c][s]*pv[s] + ck[c]*k, summed in s order
//             with zero coefficients skipped (so the order follows the reference)
//   error norm (rkdp.h:142-153): e = sum_s ec[s]*pv[s] + ek*k,
//             w = |e| / (atol + rtol*max(|pv[ynew]|, |pv[yold]|)); sum of w^2 per workgroup
struct StageArgs {
  int mode;
  const void* in[2];  // operator input (neighbour reads)
  int in_lo;
  double a_re, a_im, b_re, b_im;
  int npre;
  const void* pre[2][kMaxPre];
  unsigned pre_lo;
  void* k[2];  // null: k not stored
  int k_lo;
  int ncombo;
  double cc[2][kMaxPre];
  double ck[2];
  void* dst[2][2];   // [combo][cur]
  void* dst2[2][2];  // null: none
  int dst_lo[2], dst2_lo[2];
  int do_err;
  double ec[kMaxPre];
  double ek;
  int ynew_pre, yold_pre;
  double atol, rtol;
  int control;  // 1: the finalize pass runs the step control
  double inv_n;
  Ctrl* ctrl;
  double* partials;
  int max_partials;
};

// One neighbour link with its coefficients inlined (fast Hilbert kernel):
// contribution cm_u(e)*iphi + ac_u(e)*psi + dir times the neighbour element,
// with cm = v_i - v_j, ac = v_i + v_j for diagonal V_u (u == n_noise: none).
struct alignas(16) LinkFlat {
  int nbr, u, pad0, pad1;
  z128 iphi, psi, dir;
};

// Work unit of the one-ADO-per-workgroup operator kernels: workgroup b
// computes ADO g, whose links are links[l0 .. l1).  Workgroups are dealt to the
// eight XCDs round-robin (b mod 8), so the map places ADOs that share
// neighbours on one XCD (one L2) -- see Qme::build_block_map.
struct alignas(16) BlockMap {
  int g, l0, l1, pad;
};

// One member of a batch of independent Hilbert-form hierarchies packed into
// one set of device tables (Qme::set_param_batch).
struct BatchInst {
  int N, N2, n_hrchy;
  int apw;        // ADOs per 64-lane wavefront (64 / N^2)
  int ado_base;   // first index in the packed link_ptr / gam arrays
  int h_off;      // H in the packed H array (z128 units)
  int tab_off;    // (v_i - v_j, v_i + v_j) table (z128 units)
  int dterm_off;  // delta table (z128 units)
  int real_v;
  int rs_off;     // offset of the member's rho_S in a gathered all-members block
  int coef_off;   // first entry of the member's link coefficients (BatchCoef)
  long long elem_base;  // first element of the member's state
};

// Link of the batch kernel: the neighbour, the noise index and the index of
// its coefficients in the member's table (a member has few distinct ones: one
// per mode, occupation and direction), 16 bytes instead of the 64 of LinkFlat.
struct alignas(16) BatchLink {
  int nbr, u, cid, pad;
};
struct alignas(16) BatchCoef {
  z128 iphi, psi, dir;
};

// one wavefront of work: `apw` ADOs of member `inst` starting at local ADO `ado`
struct BatchItem {
  int inst, ado;
};

constexpr int kBatch = 16;  // Qme / FormDev kind of a packed batch (not a public kind)

// Operator tables of one QME in device memory.
struct FormDev {
  int kind;  // public kind (HEOM_HILB .. HEOM_REDFIELD_LIOU) or kBatch
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
  int real_v;            // diagonal V with real entries
  int max_links;         // most links of any single ADO
  const LinkFlat* flat;  // links with inlined coefficients (same order as links)
  const BlockMap* bmap;  // n_hrchy work units, one per workgroup (single hierarchies)
  // diagonal V: per noise u (plus one zero row) and element e = i*N+j the pair
  // (v_i - v_j, v_i + v_j); real V stores the pair as (re, im) of one z128,
  // complex V as two z128
  const z128* cmac_tab;
  // liou: CSR superoperators (N2 x N2): R0, then Phi_u, Psi_u
  const int* sop_ptr;    // (1 + 2*n_noise) * (N2 + 1)
  const int* sop_col;
  const z128* sop_val;
  // ado: global CSR
  const long long* R_ptr;
  const int* R_col;
  const z128* R_val;
  // matrix-core path (heom_hilb, 16 < N <= 64): H as separate real / imaginary
  // planes in FP64 and FP32, leading dimension 32 or 64 (the MFMA tile,
  // zero-padded); mfma = 0 off, 1 on for diagonal V (operands in the state
  // precision), 2 split-bf16x3 products with FP32 accumulation (diagonal V),
  // 3 dense V (operands in the state precision)
  int real_h;
  int mfma;
  const double* Hre;
  const double* Him;
  const float* Hre32;
  const float* Him32;
  // mfma == 3 (dense V, 16 < N <= 64): every V_u as real / imaginary planes
  // of the same zero-padded tile (n_noise planes each), FP64 and FP32;
  // real_vd: every V_u real
  const double* Vre;
  const double* Vim;
  const float* Vre32;
  const float* Vim32;
  int real_vd;
  // learned truncation closure: rows clos_lo .. clos_lo + clos_n - 1 of the
  // ADOs (the last tier) get clos[(a - clos_lo) * 2 N^2 + e] (real part) and
  // [... + N^2 + e] (imaginary part) added to D; null when no closure is bound
  const float* clos;
  int clos_lo, clos_n;
  // Redfield: Lambda_u and Lambda_u^+ (n_noise x N x N); rf_T: n_noise x N x N
  // scratch of the two-pass Hilbert kernel (N > 16), in the operator precision
  const z128* rf_L;
  const z128* rf_Ld;
  void* rf_T;
  // diagonal V: HL = iH + sum_u V_u Lambda_u, HR = -iH + sum_u Lambda_u^+ V_u
  const z128* rf_HL;
  const z128* rf_HR;
  // batch
  int n_inst, n_items, batch_grid;
  int batch_ipw;  // work items per wavefront step (1 or 2), picked by the autotuner
  int batch_wpb;  // wavefronts per workgroup (1, 2 or 4), picked by the autotuner
  const BatchInst* inst;
  const BatchItem* items;
  const BatchLink* blink;  // links of all members (member order, as link_ptr)
  const BatchCoef* bcoef;  // the members' coefficient tables
};

// Dormand-Prince 5(4) tableau (rkdp.h:21-57): A (stage inputs), B (5th-order
// solution), E (B minus the embedded 4th-order weights, for the error)
constexpr double kDpA[6][6] = {
    {0, 0, 0, 0, 0, 0},
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0}};
constexpr double kDpB[6] = {35.0 / 384, 0.0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784,
                            11.0 / 84};
constexpr double kDpE[7] = {-71.0 / 57600, 0.0, 71.0 / 16695, -71.0 / 1920,
                            17253.0 / 339200, -22.0 / 525, 1.0 / 40};

// Resident fixed-step integrator (hierarchies of at most kResidentMax
// elements): the stage program of a fixed-step scheme with the solver's slots
// renumbered to LDS buffers 0..n_buf-1 (buffer 0 holds the state), run for a
// whole output interval by one workgroup.
constexpr int kResidentMax = 1024;
constexpr size_t kResidentLds = 156 * 1024;  // dynamic LDS (the stage program is static)
constexpr int kResidentPre = 4;
struct ResidentStage {
  int in;  // buffer read by the operator
  int npre;
  int pre[kResidentPre];
  int ncombo;
  double cc[2][kResidentPre];
  double ck[2];
  int dst[2];
};
// The stage programs of rk4 and lsrk4 as the solver builds them (slots
// renumbered by first use: rk4 y, sa, acc, sb; lsrk4 y, y_new, buff), compiled
// into the resident kernel; the solver checks its own program against them.
struct ResidentProg {
  int n_stages, n_buf;
  ResidentStage st[4];
};
constexpr ResidentProg kResidentRk4{
    4, 4,
    {{0, 1, {0, 0, 0, 0}, 2, {{1.0, 0, 0, 0}, {1.0, 0, 0, 0}}, {0.5, 1.0 / 6.0}, {1, 2}},
     {1, 2, {0, 2, 0, 0}, 2, {{1.0, 0.0, 0, 0}, {0.0, 1.0, 0, 0}}, {0.5, 2.0 / 6.0}, {3, 2}},
     {3, 2, {0, 2, 0, 0}, 2, {{1.0, 0.0, 0, 0}, {0.0, 1.0, 0, 0}}, {1.0, 2.0 / 6.0}, {1, 2}},
     {1, 1, {2, 0, 0, 0}, 1, {{1.0, 0, 0, 0}, {0, 0, 0, 0}}, {1.0 / 6.0, 0.0}, {0, 0}}}};
constexpr ResidentProg kResidentLsrk4{
    4, 3,
    {{0, 1, {0, 0, 0, 0}, 2, {{1.0, 0, 0, 0}, {0.0, 0, 0, 0}}, {1.0 / 3.0, 1.0}, {1, 2}},
     {1, 2, {1, 2, 0, 0}, 2, {{1.0, 3.0 / 4.0 * -1.0, 0, 0}, {0.0, -1.0, 0, 0}}, {3.0 / 4.0, 1.0}, {0, 2}},
     {0, 2, {0, 2, 0, 0}, 2, {{1.0, 2.0 / 3.0 * -1.0, 0, 0}, {0.0, -1.0, 0, 0}}, {2.0 / 3.0, 1.0}, {1, 2}},
     {1, 2, {1, 2, 0, 0}, 2, {{1.0, 1.0 / 4.0 * -1.0, 0, 0}, {0.0, -1.0, 0, 0}}, {1.0 / 4.0, 1.0}, {0, 2}}}};

struct ResidentArgs {
  int n_stages, n_buf;
  ResidentStage st[4];
  const ResidentStage* prog;  // the same program in device memory (read by the kernel)
  z128* y;          // the FP64 state in HBM (read at the start, written at the end)
  long long n_full; // steps of dt, then one step of dt_last
  double dt, dt_last;
  // whole solve in one launch (n_int > 0): interval i takes n_full_i steps of dt
  // and one of dt_last_i; rho_S (the first N^2 elements) is written to
  // rho_s + i N^2 before interval i and after the last one
  int n_int;
  const long long* n_full_i;
  const double* dt_last_i;
  z128* rho_s;
  int scheme;  // 0: the program at prog; 1: kResidentRk4; 2: kResidentLsrk4
};

// Resident Dormand-Prince: the whole solve of a small problem in one launch
// (same eligibility as the fixed-step integrator, at most kResidentDopriMax
// elements: LDS holds y, y_new, k1..k6 and two stage inputs).
constexpr int kResidentDopriBuf = 10;
constexpr int kResidentDopriMax = 960;
struct ResidentDopriArgs {
  z128* y;               // FP64 state in HBM (read at the start, written at the end)
  int n_int;             // output intervals
  const double* t_list;  // n_int + 1 output times, device memory
  double dt_first;       // dt of the first attempt (dt_save, else the caller's dt)
  double atol, rtol, inv_n;
  z128* rho_s;           // rho_S at the n_int + 1 output times
  Ctrl* ctrl;            // the final control block (counters, dt_save, any give-up)
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
