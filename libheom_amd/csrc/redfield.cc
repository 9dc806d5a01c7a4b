// Set-up of the Redfield QMEs (redfield.h:43-174, redfield_hilb.h:51-64,
// redfield_liou.h:56-89): eigendecomposition of H, the one-sided Fourier
// transform of every bath correlation at the transition frequencies, the
// relaxation operators Lambda_u, and the tables of the two forms in HBM.
//
// Lambda_u = U (Vt_u o Ct_u) U^+ with Vt_u = U^+ V_u U, H = U diag(w) U^+ and
//   Ct_u[r][c] = sigma^T (S + iA) (gamma - i*omega)^-1 phi_0 + s_delta,
//   omega = w_c - w_r                                   (redfield.h:33-41, 144-168)
// D(rho) = i[H, rho] + sum_u [V_u, Lambda_u rho - rho Lambda_u^+]
// (redfield_hilb.h:81-88 with drho = beta*drho - alpha*D), and the
// Liouville form R = L - sum_u Phi_u Theta_u (redfield_liou.h:61-87).
#include "kernels.h"
#include "qme.h"
#include "supop.h"

#include <algorithm>
#include <cmath>
#include <string>

namespace hamd {

using namespace supop;

namespace {

inline z128 zconj(z128 a) { return z128{a.re, -a.im}; }

// Eigendecomposition of a Hermitian n x n matrix (row-major) by cyclic
// Jacobi rotations: A = U diag(w) U^+, U unitary (columns are eigenvectors).
// Each rotation first turns a_pq real with a diagonal phase, then applies the
// real symmetric Jacobi rotation that zeroes it (Golub & Van Loan, sec. 8.5).
void hermitian_eig(std::vector<z128> A, int n, std::vector<double>& w, std::vector<z128>& U) {
  U.assign(static_cast<size_t>(n) * n, z128{0.0, 0.0});
  for (int i = 0; i < n; ++i) U[i * n + i] = z128{1.0, 0.0};
  auto at = [&](int i, int j) -> z128& { return A[static_cast<size_t>(i) * n + j]; };
  double scale = 0.0;
  for (auto& x : A) scale = std::max(scale, zabs(x));
  if (scale == 0.0) scale = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += zabs(at(p, q)) * zabs(at(p, q));
    if (std::sqrt(off) <= 1e-17 * scale * n) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double b = zabs(at(p, q));
        if (b <= 1e-300) continue;
        const z128 ph{at(p, q).re / b, at(p, q).im / b};  // e^{i phi}
        const z128 phc = zconj(ph);                        // e^{-i phi}
        const double alpha = at(p, p).re, beta = at(q, q).re;
        const double tau = (beta - alpha) / (2.0 * b);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        // G = [[c, s], [-s e^{-i phi}, c e^{-i phi}]] on (p, q); A <- G^+ A G, U <- U G
        for (int k = 0; k < n; ++k) {
          const z128 akp = at(k, p), akq = at(k, q);
          at(k, p) = zscale(c, akp) - zscale(s, zmul(phc, akq));
          at(k, q) = zscale(s, akp) + zscale(c, zmul(phc, akq));
        }
        for (int k = 0; k < n; ++k) {
          const z128 apk = at(p, k), aqk = at(q, k);
          at(p, k) = zscale(c, apk) - zscale(s, zmul(ph, aqk));
          at(q, k) = zscale(s, apk) + zscale(c, zmul(ph, aqk));
        }
        for (int k = 0; k < n; ++k) {
          z128& ukp = U[static_cast<size_t>(k) * n + p];
          z128& ukq = U[static_cast<size_t>(k) * n + q];
          const z128 a0 = ukp, a1 = ukq;
          ukp = zscale(c, a0) - zscale(s, zmul(phc, a1));
          ukq = zscale(s, a0) + zscale(c, zmul(phc, a1));
        }
        at(p, q) = at(q, p) = z128{0.0, 0.0};
        at(p, p).im = at(q, q).im = 0.0;
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = at(i, i).re;
}

// x = M^-1 b by Gaussian elimination with partial pivoting (the role of
// lu_solve, linalg_engine_nil.h.j2:483-533; any pivot order gives the same
// solution of a non-singular system up to rounding).
std::vector<z128> lu_solve(std::vector<z128> M, std::vector<z128> b, int n) {
  auto at = [&](int i, int j) -> z128& { return M[static_cast<size_t>(i) * n + j]; };
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (zabs(at(i, k)) > zabs(at(piv, k))) piv = i;
    if (zabs(at(piv, k)) == 0.0)
      throw HeomError{HEOM_ERR_ARG, "redfield: gamma - i*omega is singular"};
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(at(k, j), at(piv, j));
      std::swap(b[k], b[piv]);
    }
    const z128 d = at(k, k);
    const double dd = d.re * d.re + d.im * d.im;
    const z128 inv{d.re / dd, -d.im / dd};
    for (int i = k + 1; i < n; ++i) {
      const z128 f = zmul(at(i, k), inv);
      if (zzero(f)) continue;
      for (int j = k; j < n; ++j) at(i, j) = at(i, j) - zmul(f, at(k, j));
      b[i] = b[i] - zmul(f, b[k]);
    }
  }
  std::vector<z128> x(n);
  for (int i = n - 1; i >= 0; --i) {
    z128 s = b[i];
    for (int j = i + 1; j < n; ++j) s = s - zmul(at(i, j), x[j]);
    const z128 d = at(i, i);
    const double dd = d.re * d.re + d.im * d.im;
    x[i] = zmul(s, z128{d.re / dd, -d.im / dd});
  }
  return x;
}

// C = A * B (n x n, row-major); adj: A^+ instead of A; adjB: B^+ instead of B
std::vector<z128> matmul(const std::vector<z128>& A, const std::vector<z128>& B, int n,
                         bool adjA, bool adjB) {
  std::vector<z128> C(static_cast<size_t>(n) * n, z128{0.0, 0.0});
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const z128 a = adjA ? zconj(A[static_cast<size_t>(k) * n + i]) : A[static_cast<size_t>(i) * n + k];
      if (zzero(a)) continue;
      for (int j = 0; j < n; ++j) {
        const z128 b = adjB ? zconj(B[static_cast<size_t>(j) * n + k]) : B[static_cast<size_t>(k) * n + j];
        z128& c = C[static_cast<size_t>(i) * n + j];
        c = c + zmul(a, b);
      }
    }
  return C;
}

std::vector<z128> adjoint(const std::vector<z128>& A, int n) {
  std::vector<z128> B(A.size());
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) B[static_cast<size_t>(j) * n + i] = zconj(A[static_cast<size_t>(i) * n + j]);
  return B;
}

}  // namespace

// One-sided Fourier transform of noise u's correlation at omega
// (redfield::correlation, redfield.h:33-41).
static z128 correlation(const NoiseParams& z, const std::vector<z128>& sTC, double omega) {
  const int n = z.len;
  if (n == 0) return z.s_delta;
  std::vector<z128> M = z.gamma;
  for (int k = 0; k < n; ++k) M[k * n + k] = M[k * n + k] - z128{0.0, omega};  // gamma - i*omega
  const std::vector<z128> x = lu_solve(M, z.phi0, n);
  z128 c{0.0, 0.0};
  for (int k = 0; k < n; ++k) c = c + zmul(sTC[k], x[k]);
  return c + z.s_delta;
}

void Qme::set_param_redfield() {
  const int nn = static_cast<int>(noise.size());
  // Hermitian eigenproblem from the upper triangle (the 'U' argument of the
  // heev calls, linalg_engine_mkl.h.j2:365-387)
  std::vector<z128> Hh(N2);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      Hh[i * N + j] = i <= j ? H[i * N + j] : zconj(H[j * N + i]);
  for (int i = 0; i < N; ++i) Hh[i * N + i].im = 0.0;
  std::vector<double> w;
  std::vector<z128> U;
  hermitian_eig(Hh, N, w, U);
  rf_eig = w;

  rf_lambda.assign(nn, {});
  for (int u = 0; u < nn; ++u) {
    const NoiseParams& z = noise[u];
    // sigma^T C with C = S + iA (redfield.h:62-75)
    std::vector<z128> sTC(z.len, z128{0.0, 0.0});
    for (int i = 0; i < z.len; ++i)
      for (int j = 0; j < z.len; ++j) {
        const z128 c = z.S[i * z.len + j] + zmul(kI, z.A[i * z.len + j]);
        sTC[j] = sTC[j] + zmul(z.sigma[i], c);
      }
    // V in the eigenbasis, weighted by the correlation at w_c - w_r, and back
    // (redfield.h:144-168)
    std::vector<z128> Vt = matmul(U, matmul(z.V, U, N, false, false), N, true, false);
    for (int r = 0; r < N; ++r)
      for (int c = 0; c < N; ++c)
        Vt[r * N + c] = zmul(Vt[r * N + c], correlation(z, sTC, w[c] - w[r]));
    rf_lambda[u] = matmul(U, matmul(Vt, U, N, false, true), N, false, false);
  }

  // col_major storage holds rho^T, which follows the row-major equation with
  // H -> -H^T, V -> -V^T and Lambda -> -conj(Lambda)
  std::vector<z128> Hw = H;
  std::vector<NoiseParams> nz = noise;
  std::vector<std::vector<z128>> Lw = rf_lambda;
  if (opt.order == 1) {
    auto neg_t = [&](const std::vector<z128>& X) {
      std::vector<z128> Y(N2);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) Y[i * N + j] = zscale(-1.0, X[j * N + i]);
      return Y;
    };
    Hw = neg_t(H);
    for (int u = 0; u < nn; ++u) {
      nz[u].V = neg_t(noise[u].V);
      for (auto& x : Lw[u]) x = zscale(-1.0, zconj(x));
    }
  }

  // no hierarchy: one "ADO", no links
  hs = Hierarchy{};
  hs.n_hrchy = 1;
  lk.assign(nn, {});
  ngamma.assign(1, z128{0.0, 0.0});
  s.assign(nn, {});
  a.assign(nn, {});
  link_ptr.assign(2, 0);
  links.clear();
  coef.clear();
  diag_v = true;
  for (int u = 0; u < nn && diag_v; ++u)
    for (int i = 0; i < N && diag_v; ++i)
      for (int j = 0; j < N; ++j)
        if (i != j && !zzero(nz[u].V[i * N + j])) {
          diag_v = false;
          break;
        }
  real_v = false;
  noise_work_ = nz;
  H_work_ = Hw;
  sop_ptr_.clear();
  sop_col_.clear();
  sop_val_.clear();

  std::vector<z128> L(static_cast<size_t>(nn) * N2), Ld(static_cast<size_t>(nn) * N2);
  for (int u = 0; u < nn; ++u) {
    const std::vector<z128> ad = adjoint(Lw[u], N);
    std::copy(Lw[u].begin(), Lw[u].end(), L.begin() + static_cast<size_t>(u) * N2);
    std::copy(ad.begin(), ad.end(), Ld.begin() + static_cast<size_t>(u) * N2);
  }

  if (kind == HEOM_REDFIELD_LIOU) {
    // R = L - sum_u Phi_u Theta_u, then optimize (redfield_liou.h:61-87)
    SupOp Lop(N2);
    kron_x_1(Lop, kI, Hw, N);
    kron_1_x_T(Lop, zscale(-1.0, kI), Hw, N);
    SupOp R(N2);
    for (int u = 0; u < nn; ++u) {
      SupOp Phi(N2), Theta(N2);
      kron_x_1(Phi, kI, nz[u].V, N);
      kron_1_x_T(Phi, zscale(-1.0, kI), nz[u].V, N);
      std::vector<z128> Lu(L.begin() + static_cast<size_t>(u) * N2,
                           L.begin() + static_cast<size_t>(u + 1) * N2);
      std::vector<z128> Ldu(Ld.begin() + static_cast<size_t>(u) * N2,
                            Ld.begin() + static_cast<size_t>(u + 1) * N2);
      kron_x_1(Theta, kI, Lu, N);
      kron_1_x_T(Theta, zscale(-1.0, kI), Ldu, N);
      supop_axpy(R, kOne, supop_gemm(zscale(-1.0, kOne), Phi, Theta));
    }
    supop_axpy(R, kOne, Lop);
    optimize(R);
    R_ptr_.assign(1, 0);
    R_col_.clear();
    R_val_.clear();
    for (int r = 0; r < N2; ++r) {
      for (auto& kv : R[r]) {
        R_col_.push_back(kv.first);
        R_val_.push_back(kv.second);
      }
      R_ptr_.push_back(static_cast<long long>(R_col_.size()));
    }
    ado_nnz = static_cast<long long>(R_col_.size());
  }

  upload();
  d_rf_L.upload(L);
  d_rf_Ld.upload(Ld);
  dev.rf_L = d_rf_L.as<z128>();
  dev.rf_Ld = d_rf_Ld.as<z128>();
  dev.rf_HL = dev.rf_HR = nullptr;
  if (kind == HEOM_REDFIELD_HILB && diag_v) {
    // [V, Lambda rho - rho Lambda^+] = V Lambda rho + rho Lambda^+ V
    //                                  - V rho Lambda^+ - Lambda rho V
    // the first two terms of every noise fold into HL / HR
    std::vector<z128> HL(N2), HR(N2);
    for (int e = 0; e < N2; ++e) {
      HL[e] = zmul(kI, Hw[e]);
      HR[e] = zscale(-1.0, HL[e]);
    }
    for (int u = 0; u < nn; ++u) {
      const std::vector<z128>& Vu = nz[u].V;
      const z128* Lu = L.data() + static_cast<size_t>(u) * N2;
      const z128* Ldu = Ld.data() + static_cast<size_t>(u) * N2;
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          HL[i * N + j] = HL[i * N + j] + zmul(Vu[i * N + i], Lu[i * N + j]);
          HR[i * N + j] = HR[i * N + j] + zmul(Ldu[i * N + j], Vu[j * N + j]);
        }
    }
    d_rf_HL.upload(HL);
    d_rf_HR.upload(HR);
    dev.rf_HL = d_rf_HL.as<z128>();
    dev.rf_HR = d_rf_HR.as<z128>();
  }
  // the two-pass Hilbert kernels keep Lambda_u rho - rho Lambda_u^+ here
  if (kind == HEOM_REDFIELD_HILB &&
      (N > 16 || redfield_lds_elems(N2, N, nn, diag_v) * (lo() ? sizeof(z64) : sizeof(z128)) >
                     kRedfieldLds)) {
    d_rf_T.alloc(static_cast<size_t>(nn) * N2 * sizeof(z128));
    dev.rf_T = d_rf_T.p;
  } else {
    dev.rf_T = nullptr;
  }
  ready = true;
}

}  // namespace hamd
