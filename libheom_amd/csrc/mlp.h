// Small fully connected network (3 layers, tanh) for the learned hierarchy
// truncation closure: device parameters, training (minibatch Adam on the MSE
// of normalised outputs) and batched inference, all in hand-written HIP
// kernels (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.h"
#include "qme.h"

namespace hamd {

// C[M x N] (ldc) = op(A)[M x K] * op(B)[K x N] (+ bias[N]) (tanh if act);
// op(A)(m, k) = ta ? A[k*lda + m] : A[m*lda + k], op(B)(k, n) likewise
void launch_sgemm(int M, int N, int K, const float* A, int lda, bool ta, const float* B, int ldb,
                  bool tb, float* C, int ldc, const float* bias, bool act, hipStream_t st);

class Mlp {
 public:
  Mlp(int n_in, int n_hidden, int n_out, uint64_t seed, int device);
  ~Mlp();
  int n_in, n_hid, n_out, device;
  long long n_params() const { return n_par; }
  // parameter order: W1 (hid x in), b1, W2 (hid x hid), b2, W3 (out x hid), b3
  void get_params(float* host) const;
  void set_params(const float* host);
  void set_scaling(const float* xm, const float* xs, const float* ym, const float* ys);
  void get_scaling(float* xm, float* xs, float* ym, float* ys) const;
  void fit_scaling(const float* X, const float* Y, long long n_rows);
  // raw inputs -> raw outputs (normalisation applied inside)
  void forward(const float* X, float* Y, int n_rows);
  // MSE of normalised outputs; gradient left in the device gradient buffer
  double loss_grad(const float* X, const float* Y, int n_rows, const int* perm_dev);
  void get_grad(float* host) const;
  void adam_step(double lr, double beta1, double beta2, double eps);
  void train(const float* X, const float* Y, long long n_rows, int batch, int epochs, double lr,
             uint64_t seed, double* epoch_loss);
  hipStream_t stream = nullptr;

 private:
  long long n_par;
  DevBuf params, grads, adam_m, adam_v;
  DevBuf xm_, xs_, ym_, ys_;  // normalisation: x' = (x - xm)/xs, y = y'*ys + ym
  DevBuf xb, tb, h1, h2, yb, dh1, dh2, dyb, partials, perm;
  int cap_rows = 0;
  long long t_adam = 0;
  void ensure_rows(int rows);
  float* W1() const;
  float* b1() const;
  float* W2() const;
  float* b2() const;
  float* W3() const;
  float* b3() const;
  void forward_normalised(int rows);  // xb -> h1, h2, yb
  friend struct Closure;
};

// Row layouts of the closure network (HEOM_CLOSURE_PER_ADO / _GLOBAL):
//   per ADO: one row per last-tier ADO n = (Re rho_n, Im rho_n, the index
//            vector n, features) -> (Re, Im of n's dropped term);
//   global:  one row per state = (Re, Im of every last-tier ADO in order,
//            features) -> the dropped terms of all of them.
// Both write the prediction as n_ado blocks of 2 N^2 (Re, Im).
struct ClosureDims {
  int lo, n_ado;      // last tier: first ADO, count
  int rows;           // network rows per state
  int n_in, n_out;    // network widths
  int ado_stride;     // distance between ADO blocks in an input row buffer
  int cst_rows, n_cst, cst_off;  // constant input columns (n vectors, features)
};
ClosureDims closure_dims(const Qme& q, int layout, int n_features);

// Closure bound to one heom_hilb QME of depth L: before every operator pass
// the network maps the last-tier ADOs of the stage input to their dropped
// tier-(L+1) coupling terms, which the operator kernels add to D_n.  The
// network is copied at binding time with its normalisation folded into the
// first and last layers.
struct Closure {
  ClosureDims dm{};
  int n_hid = 0;
  DevBuf W1, b1, W2, b2, W3, b3;  // folded copies
  DevBuf x, h1, h2, y;            // y: n_ado blocks of 2 N^2, added by the operator kernels
  bool fused = false;             // one fused launch (launch_closure_mlp) per operator pass
  void bind(Qme& q, const Mlp& m, const float* features, int n_features, int layout);
  void forward(const Qme& q, const StageArgs& A, hipStream_t st);
};

// first ADO of the last tier (tier == max_depth) of q and the number of them
void last_tier(const Qme& q, int& first, int& count);

// training rows of the closure from a state of a deeper hierarchy (see
// heom_closure_rows in include/libheom_amd.h)
void closure_rows(Qme& deep, Qme& shallow, const void* rho_deep, const float* features,
                  int n_features, int layout, float* x, float* y);

}  // namespace hamd
