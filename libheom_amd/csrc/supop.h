// Host-side sparse superoperators (row -> (col -> value)) used to set up the
// Liouville-space forms: the engine's LIL kron / gemm / axpy / optimize
// semantics (linalg_engine_nil.h.j2, lil_matrix.h.j2).
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <vector>

#include "common.h"

namespace hamd {
namespace supop {

inline z128 zmul(z128 a, z128 b) { return a * b; }
inline z128 zscale(double s, z128 a) { return z128{s * a.re, s * a.im}; }
inline bool zzero(z128 a) { return a.re == 0.0 && a.im == 0.0; }
inline double zabs(z128 a) { return std::hypot(a.re, a.im); }
const z128 kI{0.0, 1.0};
const z128 kOne{1.0, 0.0};

using SupOp = std::vector<std::map<int, z128>>;  // row -> (col -> value)

// row-major kron(X, 1): op[(i*n+k),(j*n+k)] += X_ij*alpha
// (linalg_engine_nil.h.j2:197-248)
inline void kron_x_1(SupOp& op, z128 alpha, const std::vector<z128>& X, int n) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      z128 v = X[i * n + j];
      if (zzero(v)) continue;
      for (int k = 0; k < n; ++k) {
        z128& d = op[i * n + k][j * n + k];
        d = d + zmul(v, alpha);
      }
    }
}

// row-major kron(1, X^T): op[(i*n+k),(i*n+j)] += X_jk*alpha
// (linalg_engine_nil.h.j2:250-299)
inline void kron_1_x_T(SupOp& op, z128 alpha, const std::vector<z128>& X, int n) {
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      z128 v = X[j * n + k];
      if (zzero(v)) continue;
      for (int i = 0; i < n; ++i) {
        z128& d = op[i * n + k][i * n + j];
        d = d + zmul(v, alpha);
      }
    }
}

// lil_matrix::optimize (lil_matrix.h.j2:79-124)
inline void optimize(SupOp& op) {
  double mx = 0.0;
  for (auto& row : op)
    for (auto& kv : row) mx = std::max(mx, zabs(kv.second));
  if (mx == 0) mx = 1;
  const double tol = std::numeric_limits<double>::epsilon() * mx;
  for (auto& row : op)
    for (auto it = row.begin(); it != row.end();) {
      if (zabs(it->second) <= tol)
        it = row.erase(it);
      else
        ++it;
    }
}

// C = alpha*A*B, (i,k,j) order, then optimize (linalg_engine_nil.h.j2:59-96)
inline SupOp supop_gemm(z128 alpha, const SupOp& A, const SupOp& B) {
  SupOp C(A.size());
  for (size_t i = 0; i < A.size(); ++i)
    for (auto& ak : A[i])
      for (auto& bj : B[ak.first]) {
        z128& d = C[i][bj.first];
        d = d + zmul(zmul(alpha, ak.second), bj.second);
      }
  optimize(C);
  return C;
}

// y += alpha*x (linalg_engine_nil.h.j2:38-56)
inline void supop_axpy(SupOp& y, z128 alpha, const SupOp& x) {
  for (size_t i = 0; i < x.size(); ++i)
    for (auto& kv : x[i]) {
      z128& d = y[i][kv.first];
      d = d + zmul(alpha, kv.second);
    }
}

}  // namespace supop
}  // namespace hamd
