// HEOM CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// This file is a plain FP64 C++17 restatement of the libheom time-propagation
// path, written to check the HIP product against.  Only tests/, the smoke()
// entry and bench.py's CPU-baseline leg may load it; the product library never
// links or calls it.
//
// Each function cites the reference file:line it restates.  The operation
// order follows the reference (per-ADO gemm/axpy sequence, RK stage order,
// Dormand-Prince step control) so that the results agree with the reference to
// rounding.  Dense naive loops stand in for the reference's Eigen/MKL/cuBLAS
// primitives (their semantics: C = alpha*A*B + beta*C, y = y + a*x, ...).
//
// Parity pinning (what checks each part; see tests/test_oracle_*.py):
//   hierarchy (BFS order, +-1 tables)   -> compiled reference hrchy_space.cc
//                                          (oracle/_ref), closed-form count,
//                                          table invariants
//   heom_hilb / heom_liou / heom_ado    -> brute-force dense Liouvillian built
//                                          with numpy Kronecker products,
//                                          cross-form equivalence, closed forms
//                                          (no bath, pure dephasing), trace /
//                                          hermiticity invariants
//   rk4 / rkdp single steps             -> matrix polynomial / scipy RK45
//                                          tableau on the dense Liouvillian
//   rkdp step control + solve loop      -> hand-computed control cases,
//                                          scipy solve_ivp end-to-end
//   errnrm1                             -> textbook RMS formula (numpy)
//   sqrt normalisation variant          -> brute force with sqrt factors
//   off-diagonal gamma (non-parity)     -> brute force with off-diag gamma
//
// Threading mirrors the reference's n_outer_threads OpenMP loop over ADOs.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;
using std::vector;

const cplx I1(0.0, 1.0);

// Dense N x N row-major helpers standing in for the engine primitives
// (linalg_engine.h:136-363).  gemm: C = alpha*A*B + beta*C (beta == 0 means
// C is overwritten, BLAS semantics).
void gemm(int n, cplx alpha, const cplx* A, const cplx* B, cplx beta, cplx* C) {
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      cplx s = 0.0;
      for (int k = 0; k < n; ++k) s += A[i * n + k] * B[k * n + j];
      C[i * n + j] = (beta == cplx(0.0) ? cplx(0.0) : beta * C[i * n + j]) + alpha * s;
    }
  }
}

void axpy(int n, cplx a, const cplx* x, cplx* y) {
  for (int i = 0; i < n; ++i) y[i] += a * x[i];
}

void scal(int n, cplx a, cplx* y) {
  if (a == cplx(0.0)) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
  } else {
    for (int i = 0; i < n; ++i) y[i] *= a;
  }
}

void nullify(int n, cplx* y) {
  for (int i = 0; i < n; ++i) y[i] = 0.0;
}

// errnrm1 (linalg_engine_mkl.h.j2:187-212, linalg_engine_cuda.cu:102-201):
// sqrt( sum_i (|e_i| / (atol + rtol*max(|x1_i|,|x2_i|)))^2 / n ).
double errnrm1(const cplx* e, const cplx* x1, const cplx* x2, double atol, double rtol,
               long n) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) {
    double w = std::abs(e[i]) / (atol + rtol * std::max(std::abs(x1[i]), std::abs(x2[i])));
    s += w * w;
  }
  return std::sqrt(s / static_cast<double>(n));
}

// Sparse N2 x N2 superoperator kept as a dense array with an "exists" mask so
// that lil_matrix::optimize (lil_matrix.h.j2:79-124) can be restated exactly.
struct SupOp {
  int n = 0;
  vector<cplx> v;
  vector<char> nz;
  void shape(int n_) {
    n = n_;
    v.assign(static_cast<size_t>(n) * n, 0.0);
    nz.assign(static_cast<size_t>(n) * n, 0);
  }
  void add(int a, int b, cplx x) {
    v[static_cast<size_t>(a) * n + b] += x;
    nz[static_cast<size_t>(a) * n + b] = 1;
  }
  // lil_matrix.h.j2:79-124: drop entries with |v| <= eps*max|v|.
  void optimize() {
    double mx = 0.0;
    for (size_t i = 0; i < v.size(); ++i)
      if (nz[i]) mx = std::max(mx, std::abs(v[i]));
    if (mx == 0) mx = 1;
    const double tol = std::numeric_limits<double>::epsilon();
    for (size_t i = 0; i < v.size(); ++i)
      if (nz[i] && std::abs(v[i]) <= tol * mx) {
        nz[i] = 0;
        v[i] = 0.0;
      }
  }
};

// kron_x_1 for row-major operands (linalg_engine_nil.h.j2:197-248):
// y[(i*n+k),(j*n+k)] += alpha*x[i][j]  for every stored x[i][j] != 0.
void kron_x_1(cplx alpha, int n, const cplx* x, const char* xnz, SupOp& y) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      if (!xnz[i * n + j]) continue;
      cplx val = x[i * n + j];
      if (val == cplx(0.0)) continue;
      for (int k = 0; k < n; ++k) y.add(i * n + k, j * n + k, val * alpha);
    }
}

// kron_1_x_T for row-major operands (linalg_engine_nil.h.j2:250-299):
// y[(i*n+k),(i*n+j)] += alpha*x[j][k].
void kron_1_x_T(cplx alpha, int n, const cplx* x, const char* xnz, SupOp& y) {
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      if (!xnz[j * n + k]) continue;
      cplx val = x[j * n + k];
      if (val == cplx(0.0)) continue;
      for (int i = 0; i < n; ++i) y.add(i * n + k, i * n + j, val * alpha);
    }
}

// nil gemm for row-major LIL operands (linalg_engine_nil.h.j2:59-96):
// C = beta*C; C[i][j] += alpha*A[i][k]*B[k][j] in (i,k,j) order; optimize.
void supop_gemm(cplx alpha, const SupOp& A, const SupOp& B, SupOp& C) {
  const int n = A.n;
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      if (!A.nz[static_cast<size_t>(i) * n + k]) continue;
      cplx av = A.v[static_cast<size_t>(i) * n + k];
      for (int j = 0; j < n; ++j) {
        if (!B.nz[static_cast<size_t>(k) * n + j]) continue;
        C.add(i, j, alpha * av * B.v[static_cast<size_t>(k) * n + j]);
      }
    }
  C.optimize();
}

// axpym (linalg_engine_nil.h.j2:38-56): y[i][j] += alpha*x[i][j].
void supop_axpy(cplx alpha, const SupOp& x, SupOp& y) {
  for (size_t i = 0; i < x.v.size(); ++i)
    if (x.nz[i]) {
      y.v[i] += alpha * x.v[i];
      y.nz[i] = 1;
    }
}

// gemv with a superoperator (dense engine semantics): y = alpha*A*x + beta*y.
void supop_gemv(cplx alpha, const SupOp& A, const cplx* x, cplx beta, cplx* y) {
  const int n = A.n;
  for (int a = 0; a < n; ++a) {
    cplx s = 0.0;
    for (int b = 0; b < n; ++b) s += A.v[static_cast<size_t>(a) * n + b] * x[b];
    y[a] = (beta == cplx(0.0) ? cplx(0.0) : beta * y[a]) + alpha * s;
  }
}

struct Noise {
  vector<cplx> V;        // N x N
  vector<char> Vnz;      // stored-entry mask (caller fills only nonzeros)
  int len = 0;
  vector<cplx> gamma;    // len x len
  vector<cplx> phi0;     // len
  vector<cplx> sigma;    // len
  vector<cplx> S;        // len x len
  vector<cplx> A;        // len x len
  cplx s_delta = 0.0;
};

enum Kind { HILB = 0, LIOU = 1, ADO = 2, REDFIELD_HILB = 3, REDFIELD_LIOU = 4 };

struct Model {
  int kind = HILB;
  int N = 0, N2 = 0;
  int max_depth = 0;
  int n_threads = 1;
  bool parity_q1 = true;   // reference quirk: off-diagonal gamma dropped
  bool sqrt_norm = false;  // LIBHEOM_SQRT_NORMALIZATION variant
  // sensitivity experiments only: accumulate the per-noise coupling terms of
  // heom_hilb in reverse noise order (the same operator with other rounding)
  bool reverse_noise_order = false;
  vector<cplx> H;
  vector<char> Hnz;
  vector<Noise> noise;

  // heom::set_param products (heom.h:57-139)
  int n_dim = 0, n_hrchy = 0, ptr_void = 0;
  vector<vector<int>> lk;
  vector<vector<int>> n;       // n_hrchy x n_dim
  vector<vector<int>> ptr_p1;  // (n_hrchy+1) x n_dim
  vector<vector<int>> ptr_m1;
  vector<cplx> ngamma_diag;
  vector<vector<std::tuple<int, int, cplx>>> gamma_offdiag;  // per u: (j,k,v)
  vector<vector<cplx>> s, a;

  // heom_liou products (heom_liou.h:66-107)
  SupOp L, R0;
  vector<SupOp> Phi, Psi, Xi;
  // heom_ado products (heom_ado.h:62-189), CSR
  vector<long> R_ptr;
  vector<int> R_col;
  vector<cplx> R_val;

  vector<cplx> temp;  // 3*N2 per thread
  // redfield::set_param products (redfield.h:43-174)
  vector<double> rf_w;                         // eigenvalues of H
  vector<vector<cplx>> rf_Lambda, rf_Lambda_dgr;  // per noise, N x N
  SupOp rf_R;                                  // redfield_liou R
  bool ready = false;
  // hrchy_filter of alloc_hrchy_space (hrchy_space.h:29-36); null: none
  int (*filter)(const int*, int, int, void*) = nullptr;
  void* filter_user = nullptr;
};

// calc_multicombination / calc_hrchy_element_count (hrchy_space.cc:36-71).
long long multicombination(long long nn, long long r) {
  long long num = 1, den = 1;
  for (long long i = 1; i <= r; ++i) {
    num *= nn + i - 1;
    den *= i;
    long long g = std::gcd(num, den);
    num /= g;
    den /= g;
  }
  return num / den;
}

// alloc_hrchy_space, breadth-first order (hrchy_space.cc:137-231).
void alloc_hrchy(Model& m) {
  const int K = m.n_dim;
  vector<int> index(K, 0);
  std::map<vector<int>, int> book;
  std::queue<vector<int>> next;
  std::queue<int> k_last;
  next.push(index);
  k_last.push(0);
  int lidx = 0;
  m.n.clear();
  while (!next.empty()) {
    index = next.front();
    next.pop();
    int last = k_last.front();
    k_last.pop();
    book[index] = lidx;
    m.n.push_back(index);
    ++lidx;
    for (int k = last; k < K; ++k) {
      ++index[k];
      int depth = std::accumulate(index.begin(), index.end(), 0);
      // hrchy_space.cc:183-184: depth <= max_depth && (!filter_flag || hrchy_filter(...))
      const bool pass =
          depth <= m.max_depth && (!m.filter || m.filter(index.data(), K, depth, m.filter_user));
      if (pass) {
        next.push(index);
        k_last.push(k);
      }
      --index[k];
    }
  }
  m.n_hrchy = lidx;
  m.ptr_void = lidx;
  m.ptr_p1.assign(lidx + 1, vector<int>(K, lidx));
  m.ptr_m1.assign(lidx + 1, vector<int>(K, lidx));
  for (int l = 0; l < lidx; ++l) {
    index = m.n[l];
    for (int k = 0; k < K; ++k) {
      ++index[k];
      auto it = book.find(index);
      m.ptr_p1[l][k] = (it == book.end()) ? lidx : it->second;
      index[k] -= 2;
      it = book.find(index);
      m.ptr_m1[l][k] = (it == book.end()) ? lidx : it->second;
      ++index[k];
    }
  }
}

// heom::set_param (heom.h:57-139) plus the form-specific set_param.
// ---------------------------------------------------------------------------
// Redfield QMEs (redfield.h, redfield_hilb.h, redfield_liou.h)
// ---------------------------------------------------------------------------

// Hermitian eigendecomposition standing in for the engine's eig (heev with
// the upper triangle, linalg_engine_mkl.h.j2:365-387): cyclic Jacobi sweeps on
// A = U diag(w) U^+, U's columns the eigenvectors.  Each rotation zeroes one
// off-diagonal pair: with a_pq = |a_pq| e^{i phi}, the unitary
//   G = diag(1, e^{-i phi}) * [[c, s], [-s, c]]
// (t = tan(theta) the smaller root of t^2 + 2 tau t - 1 = 0,
// tau = (a_qq - a_pp) / (2 |a_pq|)) diagonalises the 2 x 2 block.
void jacobi_eig(int n, vector<cplx> A, vector<double>& w, vector<cplx>& U) {
  U.assign(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) U[static_cast<size_t>(i) * n + i] = 1.0;
  auto a = [&](int i, int j) -> cplx& { return A[static_cast<size_t>(i) * n + j]; };
  double nrm = 0.0;
  for (const cplx& x : A) nrm += std::norm(x);
  nrm = std::sqrt(nrm);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q)
        if (p != q) off += std::norm(a(p, q));
    if (std::sqrt(off) <= 1e-16 * (nrm > 0 ? nrm : 1.0)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double r = std::abs(a(p, q));
        if (r == 0.0) continue;
        const cplx e = a(p, q) / r;  // e^{i phi}
        const double tau = (a(q, q).real() - a(p, p).real()) / (2.0 * r);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::hypot(1.0, tau));
        const double c = 1.0 / std::hypot(1.0, t), sn = t * c;
        const cplx gpp = c, gpq = sn, gqp = -sn * std::conj(e), gqq = c * std::conj(e);
        for (int k = 0; k < n; ++k) {  // A <- A G
          const cplx x = a(k, p), y = a(k, q);
          a(k, p) = x * gpp + y * gqp;
          a(k, q) = x * gpq + y * gqq;
        }
        for (int k = 0; k < n; ++k) {  // A <- G^+ A
          const cplx x = a(p, k), y = a(q, k);
          a(p, k) = std::conj(gpp) * x + std::conj(gqp) * y;
          a(q, k) = std::conj(gpq) * x + std::conj(gqq) * y;
        }
        for (int k = 0; k < n; ++k) {  // U <- U G
          cplx& x = U[static_cast<size_t>(k) * n + p];
          cplx& y = U[static_cast<size_t>(k) * n + q];
          const cplx x0 = x, y0 = y;
          x = x0 * gpp + y0 * gqp;
          y = x0 * gpq + y0 * gqq;
        }
        a(p, q) = a(q, p) = 0.0;
        a(p, p) = a(p, p).real();
        a(q, q) = a(q, q).real();
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = a(i, i).real();
}

// lu / lu_solve of the nil engine for row-major operands
// (linalg_engine_nil.h.j2:323-389 Doolittle with row pivoting, 483-533
// substitutions).  As in the reference, the pivot search compares against the
// original diagonal magnitude (the running maximum it declares inside the
// loop shadows the outer one), so the last row with a larger entry is taken,
// and a column whose original diagonal is zero is not eliminated.
vector<cplx> nil_lu_solve(int n, const vector<cplx>& A, const vector<cplx>& b) {
  vector<vector<cplx>> U(n, vector<cplx>(n, 0.0));
  vector<vector<char>> Unz(n, vector<char>(n, 0));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      U[i][j] = A[static_cast<size_t>(i) * n + j];
      Unz[i][j] = U[i][j] != cplx(0.0);
    }
  vector<int> pivot(n);
  std::iota(pivot.begin(), pivot.end(), 0);
  for (int i = 0; i < n - 1; ++i) {
    const double mx = std::abs(U[i][i]);
    int swap = -1;
    for (int j = i + 1; j < n; ++j)
      if (Unz[j][i] && std::abs(U[j][i]) > mx) swap = j;
    if (swap != -1) {
      std::swap(pivot[i], pivot[swap]);
      std::swap(U[i], U[swap]);
      std::swap(Unz[i], Unz[swap]);
    }
    if (mx == 0.0) continue;
    for (int j = i + 1; j < n; ++j) {
      if (!Unz[j][i]) continue;
      U[j][i] /= U[i][i];
      for (int k = i + 1; k < n; ++k)
        if (Unz[i][k]) {
          U[j][k] -= U[j][i] * U[i][k];
          Unz[j][k] = 1;
        }
    }
  }
  vector<cplx> y(n), x(n);
  for (int i = 0; i < n; ++i) {
    y[i] = b[pivot[i]];
    for (int j = 0; j < i; ++j)
      if (Unz[i][j]) y[i] -= U[i][j] * y[j];
  }
  for (int i = n - 1; i >= 0; --i) {
    x[i] = y[i];
    for (int j = i + 1; j < n; ++j)
      if (Unz[i][j]) x[i] -= U[i][j] * x[j];
    x[i] /= U[i][i];
  }
  return x;
}

// redfield::correlation (redfield.h:33-41):
// sigma^T C (gamma - i*omega)^-1 phi_0 + s_delta with C = S + iA.
cplx rf_correlation(const Noise& z, const vector<cplx>& sTC, double omega) {
  if (z.len == 0) return z.s_delta;
  vector<cplx> M = z.gamma;
  for (int k = 0; k < z.len; ++k) M[static_cast<size_t>(k) * z.len + k] -= I1 * omega;
  const vector<cplx> x = nil_lu_solve(z.len, M, z.phi0);
  cplx c = 0.0;
  for (int k = 0; k < z.len; ++k) c += sTC[k] * x[k];
  return c + z.s_delta;
}

// redfield::set_param (redfield.h:43-174) and the form set-ups
// (redfield_hilb.h:51-64, redfield_liou.h:56-89).
void set_param_redfield(Model& m) {
  const int N = m.N, N2 = m.N2, nn = static_cast<int>(m.noise.size());
  m.n_dim = 0;
  m.n_hrchy = 1;
  vector<cplx> Hh(N2);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      Hh[i * N + j] = i <= j ? m.H[i * N + j] : std::conj(m.H[j * N + i]);
  vector<cplx> U;
  jacobi_eig(N, Hh, m.rf_w, U);
  vector<cplx> Ud(N2);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) Ud[j * N + i] = std::conj(U[i * N + j]);
  m.rf_Lambda.assign(nn, vector<cplx>(N2));
  m.rf_Lambda_dgr.assign(nn, vector<cplx>(N2));
  vector<cplx> tmp(N2), Vv(N2);
  for (int s = 0; s < nn; ++s) {
    const Noise& z = m.noise[s];
    // sigma_T_C = sigma^T C by gevm (redfield.h:68-75, linalg_engine_nil.h.j2:169-192)
    vector<cplx> sTC(z.len, 0.0);
    for (int i = 0; i < z.len; ++i)
      for (int j = 0; j < z.len; ++j)
        sTC[j] += z.sigma[i] *
                  (z.S[static_cast<size_t>(i) * z.len + j] + I1 * z.A[static_cast<size_t>(i) * z.len + j]);
    // utf: V_v = U^+ (V U) (linalg_engine_mkl.h.j2:389-419)
    gemm(N, 1.0, z.V.data(), U.data(), 0.0, tmp.data());
    gemm(N, 1.0, Ud.data(), tmp.data(), 0.0, Vv.data());
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j)
        Vv[i * N + j] *= rf_correlation(z, sTC, m.rf_w[j] - m.rf_w[i]);  // omega_ji
    // utb: Lambda = U (Lambda_v U^+) (linalg_engine_mkl.h.j2:421-450)
    gemm(N, 1.0, Vv.data(), Ud.data(), 0.0, tmp.data());
    gemm(N, 1.0, U.data(), tmp.data(), 0.0, m.rf_Lambda[s].data());
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) m.rf_Lambda_dgr[s][j * N + i] = std::conj(m.rf_Lambda[s][i * N + j]);
  }
  if (m.kind == REDFIELD_LIOU) {
    auto mask = [](const vector<cplx>& x) {
      vector<char> nz(x.size());
      for (size_t i = 0; i < x.size(); ++i) nz[i] = x[i] != cplx(0.0);
      return nz;
    };
    SupOp L;
    L.shape(N2);
    kron_x_1(I1, N, m.H.data(), m.Hnz.data(), L);
    kron_1_x_T(-I1, N, m.H.data(), m.Hnz.data(), L);
    m.rf_R.shape(N2);
    for (int s = 0; s < nn; ++s) {
      SupOp Phi, Theta;
      Phi.shape(N2);
      Theta.shape(N2);
      kron_x_1(I1, N, m.noise[s].V.data(), m.noise[s].Vnz.data(), Phi);
      kron_1_x_T(-I1, N, m.noise[s].V.data(), m.noise[s].Vnz.data(), Phi);
      const vector<char> lnz = mask(m.rf_Lambda[s]), dnz = mask(m.rf_Lambda_dgr[s]);
      kron_x_1(I1, N, m.rf_Lambda[s].data(), lnz.data(), Theta);
      kron_1_x_T(-I1, N, m.rf_Lambda_dgr[s].data(), dnz.data(), Theta);
      supop_gemm(-1.0, Phi, Theta, m.rf_R);  // R = -Phi Theta + R
    }
    supop_axpy(1.0, L, m.rf_R);
    m.rf_R.optimize();
  }
  m.ready = true;
}

// redfield_hilb::calc_diff_impl (redfield_hilb.h:66-89), same gemm sequence.
void calc_diff_redfield_hilb(Model& m, cplx* drho, const cplx* rho, cplx alpha, cplx beta) {
  const int N = m.N;
  vector<cplx> temp(m.N2);
  gemm(N, -I1 * alpha, m.H.data(), rho, beta, drho);
  gemm(N, I1 * alpha, rho, m.H.data(), 1.0, drho);
  for (size_t s = 0; s < m.noise.size(); ++s) {
    gemm(N, I1, m.rf_Lambda[s].data(), rho, 0.0, temp.data());
    gemm(N, -I1, rho, m.rf_Lambda_dgr[s].data(), 1.0, temp.data());
    gemm(N, I1 * alpha, m.noise[s].V.data(), temp.data(), 1.0, drho);
    gemm(N, -I1 * alpha, temp.data(), m.noise[s].V.data(), 1.0, drho);
  }
}

void set_param(Model& m) {
  const int nn = static_cast<int>(m.noise.size());
  if (m.kind == REDFIELD_HILB || m.kind == REDFIELD_LIOU) {
    m.N2 = m.N * m.N;
    set_param_redfield(m);
    return;
  }
  m.N2 = m.N * m.N;
  m.n_dim = 0;
  for (auto& z : m.noise) m.n_dim += z.len;
  m.lk.assign(nn, {});
  int ctr = 0;
  for (int u = 0; u < nn; ++u)
    for (int k = 0; k < m.noise[u].len; ++k) m.lk[u].push_back(ctr++);

  alloc_hrchy(m);

  // ngamma_diag (heom.h:93-104)
  m.ngamma_diag.assign(m.n_hrchy, 0.0);
  for (int l = 0; l < m.n_hrchy; ++l) {
    cplx g = 0.0;
    for (int u = 0; u < nn; ++u)
      for (int k = 0; k < m.noise[u].len; ++k)
        g += static_cast<double>(m.n[l][m.lk[u][k]]) *
             m.noise[u].gamma[static_cast<size_t>(k) * m.noise[u].len + k];
    m.ngamma_diag[l] = g;
  }

  // gamma_offdiag (heom.h:106-121; set_shape at 110).  In the reference, set_shape clears
  // gamma[u] before the copy loop, so gamma_offdiag is always empty; the
  // parity flag reproduces that, the non-parity mode keeps the intended term.
  m.gamma_offdiag.assign(nn, {});
  if (!m.parity_q1) {
    for (int u = 0; u < nn; ++u) {
      const int len = m.noise[u].len;
      for (int j = 0; j < len; ++j)
        for (int k = 0; k < len; ++k) {
          cplx v = m.noise[u].gamma[static_cast<size_t>(j) * len + k];
          if (j != k && v != cplx(0.0)) m.gamma_offdiag[u].emplace_back(j, k, v);
        }
    }
  }

  // s = S phi_0, a = A phi_0 (heom.h:123-138; nil gemv, linalg_engine_nil.h.j2:137-164)
  m.s.assign(nn, {});
  m.a.assign(nn, {});
  for (int u = 0; u < nn; ++u) {
    const int len = m.noise[u].len;
    m.s[u].assign(len, 0.0);
    m.a[u].assign(len, 0.0);
    for (int i = 0; i < len; ++i)
      for (int j = 0; j < len; ++j) {
        m.s[u][i] += m.noise[u].S[static_cast<size_t>(i) * len + j] * m.noise[u].phi0[j];
        m.a[u][i] += m.noise[u].A[static_cast<size_t>(i) * len + j] * m.noise[u].phi0[j];
      }
  }

  if (m.kind == LIOU || m.kind == ADO) {
    // heom_liou::set_param (heom_liou.h:66-107)
    const int N = m.N, N2 = m.N2;
    m.L.shape(N2);
    kron_x_1(+I1, N, m.H.data(), m.Hnz.data(), m.L);
    kron_1_x_T(-I1, N, m.H.data(), m.Hnz.data(), m.L);
    m.Phi.assign(nn, {});
    m.Psi.assign(nn, {});
    m.Xi.assign(nn, {});
    for (int u = 0; u < nn; ++u) {
      m.Phi[u].shape(N2);
      kron_x_1(+I1, N, m.noise[u].V.data(), m.noise[u].Vnz.data(), m.Phi[u]);
      kron_1_x_T(-I1, N, m.noise[u].V.data(), m.noise[u].Vnz.data(), m.Phi[u]);
      m.Phi[u].optimize();
      m.Psi[u].shape(N2);
      kron_x_1(1.0, N, m.noise[u].V.data(), m.noise[u].Vnz.data(), m.Psi[u]);
      kron_1_x_T(1.0, N, m.noise[u].V.data(), m.noise[u].Vnz.data(), m.Psi[u]);
      m.Psi[u].optimize();
      m.Xi[u].shape(N2);
      supop_gemm(-m.noise[u].s_delta, m.Phi[u], m.Phi[u], m.Xi[u]);
    }
    m.R0.shape(N2);
    supop_axpy(1.0, m.L, m.R0);
    for (int u = 0; u < nn; ++u) supop_axpy(1.0, m.Xi[u], m.R0);
  }

  if (m.kind == ADO) {
    // heom_ado::set_param (heom_ado.h:62-189): assemble R row by row with
    // push-accumulate semantics, then compress to CSR.
    const int N2 = m.N2;
    vector<vector<SupOp>> Theta(nn);
    for (int u = 0; u < nn; ++u) {
      Theta[u].assign(m.noise[u].len, {});
      for (int k = 0; k < m.noise[u].len; ++k) {
        Theta[u][k].shape(N2);
        supop_axpy(+m.s[u][k], m.Phi[u], Theta[u][k]);
        supop_axpy(-m.a[u][k], m.Psi[u], Theta[u][k]);
        Theta[u][k].optimize();
      }
    }
    const long nrow = static_cast<long>(m.n_hrchy) * N2;
    m.R_ptr.assign(nrow + 1, 0);
    m.R_col.clear();
    m.R_val.clear();
    std::map<int, cplx> row;
    for (int l = 0; l < m.n_hrchy; ++l) {
      for (int a_ = 0; a_ < N2; ++a_) {
        row.clear();
        auto push = [&](long col, cplx v) { row[static_cast<int>(col)] += v; };
        // -1 terms
        for (int u = 0; u < nn; ++u)
          for (int k = 0; k < m.noise[u].len; ++k) {
            int lm1 = m.ptr_m1[l][m.lk[u][k]];
            if (lm1 == m.ptr_void) continue;
            double nk = static_cast<double>(m.n[l][m.lk[u][k]]);
            for (int b = 0; b < N2; ++b) {
              if (!Theta[u][k].nz[static_cast<size_t>(a_) * N2 + b]) continue;
              cplx v = Theta[u][k].v[static_cast<size_t>(a_) * N2 + b];
              v *= m.sqrt_norm ? std::sqrt(nk) : nk;
              if (v != cplx(0.0)) push(static_cast<long>(lm1) * N2 + b, v);
            }
          }
        // 0 terms
        push(static_cast<long>(l) * N2 + a_, m.ngamma_diag[l]);
        for (int u = 0; u < nn; ++u)
          for (auto& t : m.gamma_offdiag[u]) {
            int j = std::get<0>(t), k = std::get<1>(t);
            cplx v = std::get<2>(t);
            int lm1j = m.ptr_m1[l][m.lk[u][j]];
            if (lm1j == m.ptr_void) continue;
            int lm1jp1k = m.ptr_p1[lm1j][m.lk[u][k]];
            if (lm1jp1k == m.ptr_void) continue;
            double nj = static_cast<double>(m.n[l][m.lk[u][j]]);
            double nk = static_cast<double>(m.n[l][m.lk[u][k]]);
            push(static_cast<long>(lm1jp1k) * N2 + a_,
                 v * (m.sqrt_norm ? std::sqrt(nj * (nk + 1)) : nj));
          }
        for (int b = 0; b < N2; ++b)
          if (m.R0.nz[static_cast<size_t>(a_) * N2 + b])
            push(static_cast<long>(l) * N2 + b, m.R0.v[static_cast<size_t>(a_) * N2 + b]);
        // +1 terms
        for (int u = 0; u < nn; ++u)
          for (int k = 0; k < m.noise[u].len; ++k) {
            int lp1 = m.ptr_p1[l][m.lk[u][k]];
            if (lp1 == m.ptr_void) continue;
            double nk = static_cast<double>(m.n[l][m.lk[u][k]]);
            for (int b = 0; b < N2; ++b) {
              if (!m.Phi[u].nz[static_cast<size_t>(a_) * N2 + b]) continue;
              cplx v = m.Phi[u].v[static_cast<size_t>(a_) * N2 + b];
              if (m.sqrt_norm) v *= std::sqrt(nk + 1);
              v *= m.noise[u].sigma[k];
              if (v != cplx(0.0)) push(static_cast<long>(lp1) * N2 + b, v);
            }
          }
        for (auto& kv : row) {
          m.R_col.push_back(kv.first);
          m.R_val.push_back(kv.second);
        }
        m.R_ptr[static_cast<long>(l) * N2 + a_ + 1] = static_cast<long>(m.R_col.size());
      }
    }
  }

  m.temp.assign(static_cast<size_t>(3) * m.N2 * std::max(1, m.n_threads), 0.0);
  m.ready = true;
}

// heom_hilb::calc_diff_impl (heom_hilb.h:70-215).
void calc_diff_hilb(Model& m, cplx* drho, const cplx* rho, cplx alpha, cplx beta) {
  const int N = m.N, N2 = m.N2, nn = static_cast<int>(m.noise.size());
  const int pv = m.ptr_void;
#pragma omp parallel for num_threads(m.n_threads) schedule(static)
  for (int l = 0; l < m.n_hrchy; ++l) {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    const cplx* rho_n = rho + static_cast<size_t>(l) * N2;
    cplx* d = &m.temp[static_cast<size_t>(3 * tid + 0) * N2];
    cplx* tPhi = &m.temp[static_cast<size_t>(3 * tid + 1) * N2];
    cplx* tPsi = &m.temp[static_cast<size_t>(3 * tid + 2) * N2];
    // 0 terms (109-118)
    gemm(N, I1, m.H.data(), rho_n, 0.0, d);
    gemm(N, -I1, rho_n, m.H.data(), 1.0, d);
    axpy(N2, m.ngamma_diag[l], rho_n, d);
    for (int u = 0; u < nn; ++u) {
      const cplx* V = m.noise[u].V.data();
      gemm(N, 1.0, V, rho_n, 0.0, tPhi);
      gemm(N, -1.0, rho_n, V, 1.0, tPhi);
      gemm(N, m.noise[u].s_delta, V, tPhi, 1.0, d);
      gemm(N, -m.noise[u].s_delta, tPhi, V, 1.0, d);
    }
    for (int uu = 0; uu < nn; ++uu) {
      const int u = m.reverse_noise_order ? nn - 1 - uu : uu;
      const auto& lk_u = m.lk[u];
      const int len = m.noise[u].len;
      const cplx* V = m.noise[u].V.data();
      // off-diagonal gamma, row-major branch (128-147)
      for (auto& t : m.gamma_offdiag[u]) {
        int j = std::get<0>(t), k = std::get<1>(t);
        cplx v = std::get<2>(t);
        int lm1j = m.ptr_m1[l][lk_u[j]];
        int lm1jp1k = m.ptr_p1[lm1j][lk_u[k]];
        if (lm1j != pv && lm1jp1k != pv) {
          double nj = static_cast<double>(m.n[l][lk_u[j]]);
          double nk = static_cast<double>(m.n[l][lk_u[k]]);
          double f = m.sqrt_norm ? std::sqrt(nj * (nk + 1)) : nj;
          axpy(N2, f * v, rho + static_cast<size_t>(lm1jp1k) * N2, d);
        }
      }
      nullify(N2, tPhi);
      nullify(N2, tPsi);
      // +1 terms (170-182)
      for (int k = 0; k < len; ++k) {
        int lp1 = m.ptr_p1[l][lk_u[k]];
        if (lp1 != pv) {
          double nf = static_cast<double>(m.n[l][lk_u[k]]);
          cplx c = m.sqrt_norm ? m.noise[u].sigma[k] * std::sqrt(nf + 1) : m.noise[u].sigma[k];
          axpy(N2, c, rho + static_cast<size_t>(lp1) * N2, tPhi);
        }
      }
      // -1 terms (184-198)
      for (int k = 0; k < len; ++k) {
        int lm1 = m.ptr_m1[l][lk_u[k]];
        if (lm1 != pv) {
          double nf = static_cast<double>(m.n[l][lk_u[k]]);
          double f = m.sqrt_norm ? std::sqrt(nf) : nf;
          const cplx* rm = rho + static_cast<size_t>(lm1) * N2;
          axpy(N2, m.s[u][k] * f, rm, tPhi);
          axpy(N2, -m.a[u][k] * f, rm, tPsi);
        }
      }
      // (200-208)
      gemm(N, I1, V, tPhi, 1.0, d);
      gemm(N, -I1, tPhi, V, 1.0, d);
      gemm(N, 1.0, V, tPsi, 1.0, d);
      gemm(N, 1.0, tPsi, V, 1.0, d);
    }
    // (210-211)
    cplx* out = drho + static_cast<size_t>(l) * N2;
    scal(N2, beta, out);
    axpy(N2, -alpha, d, out);
  }
}

// heom_liou::calc_diff_impl (heom_liou.h:121-258).
void calc_diff_liou(Model& m, cplx* drho, const cplx* rho, cplx alpha, cplx beta) {
  const int N2 = m.N2, nn = static_cast<int>(m.noise.size());
  const int pv = m.ptr_void;
#pragma omp parallel for num_threads(m.n_threads) schedule(static)
  for (int l = 0; l < m.n_hrchy; ++l) {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    const cplx* rho_n = rho + static_cast<size_t>(l) * N2;
    cplx* d = &m.temp[static_cast<size_t>(3 * tid + 0) * N2];
    cplx* tPhi = &m.temp[static_cast<size_t>(3 * tid + 1) * N2];
    cplx* tPsi = &m.temp[static_cast<size_t>(3 * tid + 2) * N2];
    supop_gemv(1.0, m.R0, rho_n, 0.0, d);
    axpy(N2, m.ngamma_diag[l], rho_n, d);
    for (int uu = 0; uu < nn; ++uu) {
      const int u = m.reverse_noise_order ? nn - 1 - uu : uu;
      const auto& lk_u = m.lk[u];
      const int len = m.noise[u].len;
      for (auto& t : m.gamma_offdiag[u]) {
        int j = std::get<0>(t), k = std::get<1>(t);
        cplx v = std::get<2>(t);
        int lm1j = m.ptr_m1[l][lk_u[j]];
        int lm1jp1k = m.ptr_p1[lm1j][lk_u[k]];
        if (lm1j != pv && lm1jp1k != pv) {
          double nj = static_cast<double>(m.n[l][lk_u[j]]);
          double nk = static_cast<double>(m.n[l][lk_u[k]]);
          double f = m.sqrt_norm ? std::sqrt(nj * (nk + 1)) : nj;
          axpy(N2, f * v, rho + static_cast<size_t>(lm1jp1k) * N2, d);
        }
      }
      nullify(N2, tPhi);
      nullify(N2, tPsi);
      for (int k = 0; k < len; ++k) {
        int lp1 = m.ptr_p1[l][lk_u[k]];
        if (lp1 != pv) {
          double nf = static_cast<double>(m.n[l][lk_u[k]]);
          cplx c = m.sqrt_norm ? m.noise[u].sigma[k] * std::sqrt(nf + 1) : m.noise[u].sigma[k];
          axpy(N2, c, rho + static_cast<size_t>(lp1) * N2, tPhi);
        }
      }
      for (int k = 0; k < len; ++k) {
        int lm1 = m.ptr_m1[l][lk_u[k]];
        if (lm1 != pv) {
          double nf = static_cast<double>(m.n[l][lk_u[k]]);
          double f = m.sqrt_norm ? std::sqrt(nf) : nf;
          const cplx* rm = rho + static_cast<size_t>(lm1) * N2;
          axpy(N2, m.s[u][k] * f, rm, tPhi);
          axpy(N2, -m.a[u][k] * f, rm, tPsi);
        }
      }
      supop_gemv(1.0, m.Phi[u], tPhi, 1.0, d);
      supop_gemv(1.0, m.Psi[u], tPsi, 1.0, d);
    }
    cplx* out = drho + static_cast<size_t>(l) * N2;
    scal(N2, beta, out);
    axpy(N2, -alpha, d, out);
  }
}

// heom_ado::calc_diff_impl (heom_ado.h:191-202): drho = beta*drho - alpha*R*rho.
void calc_diff_ado(Model& m, cplx* drho, const cplx* rho, cplx alpha, cplx beta) {
  const long nrow = static_cast<long>(m.n_hrchy) * m.N2;
#pragma omp parallel for num_threads(m.n_threads) schedule(static)
  for (long r = 0; r < nrow; ++r) {
    cplx s = 0.0;
    for (long p = m.R_ptr[r]; p < m.R_ptr[r + 1]; ++p) s += m.R_val[p] * rho[m.R_col[p]];
    drho[r] = (beta == cplx(0.0) ? cplx(0.0) : beta * drho[r]) - alpha * s;
  }
}

void calc_diff(Model& m, cplx* drho, const cplx* rho, cplx alpha, cplx beta) {
  if (m.kind == REDFIELD_HILB)
    calc_diff_redfield_hilb(m, drho, rho, alpha, beta);
  else if (m.kind == REDFIELD_LIOU)
    supop_gemv(-alpha, m.rf_R, rho, beta, drho);  // redfield_liou.h:91-102
  else if (m.kind == HILB)
    calc_diff_hilb(m, drho, rho, alpha, beta);
  else if (m.kind == LIOU)
    calc_diff_liou(m, drho, rho, alpha, beta);
  else
    calc_diff_ado(m, drho, rho, alpha, beta);
}

// Vector primitives on main_size buffers (copy/axpy/scal of linalg_engine.h).
void vcopy(long n, const cplx* x, cplx* y) { std::memcpy(y, x, sizeof(cplx) * n); }
void vaxpy(long n, double a, const cplx* x, cplx* y) {
  for (long i = 0; i < n; ++i) y[i] += a * x[i];
}
void vscal(long n, double a, cplx* y) {
  for (long i = 0; i < n; ++i) y[i] *= a;
}

struct Stats {
  long n_rhs = 0;
  long n_steps = 0;
  long n_accepted = 0;
  long n_rejected = 0;
  int gave_up = 0;
};

// rk4::solve_fixed_step (rk4.h:40-65).
void rk4_step(Model& m, cplx* rho, double dt, vector<cplx>& buf, Stats& st) {
  const long ms = static_cast<long>(m.n_hrchy) * m.N2;
  cplx* rho_n = &buf[0];
  cplx* kh1 = &buf[ms];
  cplx* kh2 = &buf[2 * ms];
  cplx* kh3 = &buf[3 * ms];
  vcopy(ms, rho, rho_n);
  calc_diff(m, kh1, rho_n, dt, 0.0);
  vaxpy(ms, 0.5, kh1, rho_n);
  calc_diff(m, kh2, rho_n, dt, 0.0);
  vcopy(ms, rho, rho_n);
  vaxpy(ms, 0.5, kh2, rho_n);
  calc_diff(m, kh3, rho_n, dt, 0.0);
  vcopy(ms, rho, rho_n);
  vaxpy(ms, 1.0, kh3, rho_n);
  vaxpy(ms, 1.0, kh3, kh2);
  calc_diff(m, kh3, rho_n, dt, 0.0);
  vaxpy(ms, 1.0 / 6.0, kh1, rho);
  vaxpy(ms, 2.0 / 6.0, kh2, rho);
  vaxpy(ms, 1.0 / 6.0, kh3, rho);
  st.n_rhs += 4;
  st.n_steps += 1;
}

// lsrk4::solve_fixed_step (lsrk4.h:33-48): buff = beta*buff - dt*D(rho) with
// beta = 0, -1, -1, -1, each followed by rho += c*buff, c = 1/3, 3/4, 2/3, 1/4.
void lsrk4_step(Model& m, cplx* rho, double dt, vector<cplx>& buf, Stats& st) {
  const long ms = static_cast<long>(m.n_hrchy) * m.N2;
  cplx* b = &buf[0];
  const double beta[4] = {0.0, -1.0, -1.0, -1.0};
  const double c[4] = {1.0 / 3.0, 3.0 / 4.0, 2.0 / 3.0, 1.0 / 4.0};
  for (int s = 0; s < 4; ++s) {
    calc_diff(m, b, rho, dt, beta[s]);
    vaxpy(ms, c[s], b, rho);
  }
  st.n_rhs += 4;
  st.n_steps += 1;
}

// fixed_step_size_solver::solve_1 (fixed_step_size_solver.h:29-52).
void fixed_solve_1(Model& m, cplx* rho, double t0, double t1, double dt, vector<cplx>& buf,
                   Stats& st, bool low_storage = false) {
  double t = t0;
  double dt_1 = dt;
  bool brk = false;
  while (true) {
    if (t + dt_1 - t1 > -std::numeric_limits<double>::epsilon()) {
      dt_1 = t1 - t;
      brk = true;
    }
    if (low_storage)
      lsrk4_step(m, rho, dt_1, buf, st);
    else
      rk4_step(m, rho, dt_1, buf, st);
    t += dt_1;
    if (brk) break;
  }
}

// Dormand-Prince 5(4) tableau (rkdp.h:21-57).
const double DP_A[6][6] = {
    {0, 0, 0, 0, 0, 0},
    {1.0 / 5, 0, 0, 0, 0, 0},
    {3.0 / 40, 9.0 / 40, 0, 0, 0, 0},
    {44.0 / 45, -56.0 / 15, 32.0 / 9, 0, 0, 0},
    {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729, 0, 0},
    {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656, 0}};
const double DP_B[6] = {35.0 / 384, 0.0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84};
const double DP_E[7] = {-71.0 / 57600, 0.0, 71.0 / 16695, -71.0 / 1920,
                        17253.0 / 339200, -22.0 / 525, 1.0 / 40};

// rkdp::solve_adaptive_step (rkdp.h:81-179).  Returns false when the
// 1000-attempt limit is hit (the reference sets t = t_bound silently).
bool rkdp_step(Model& m, cplx* rho, double& t, double t_bound, double& dt_1, double atol,
               double rtol, vector<cplx>& buf, Stats& st) {
  const long ms = static_cast<long>(m.n_hrchy) * m.N2;
  cplx* rho_n = &buf[0];
  cplx* rho_old = &buf[ms];
  cplx* kh[7];
  for (int i = 0; i < 7; ++i) kh[i] = &buf[(i + 2) * ms];
  const double safety = 0.9, dt_min = 1e-16, expo = -1.0 / 5.0;
  vcopy(ms, rho, rho_old);
  int count = 0;
  bool accepted = false;
  while (!accepted) {
    ++count;
    if (count > 1000) {
      t = t_bound;
      st.gave_up = 1;
      return false;
    }
    if (t + dt_1 > t_bound) dt_1 = t_bound - t;
    for (int i = 0; i < 6; ++i) {
      vcopy(ms, rho, rho_n);
      for (int j = 0; j < i; ++j) vaxpy(ms, DP_A[i][j], kh[j], rho_n);
      calc_diff(m, kh[i], rho_n, dt_1, 0.0);
    }
    for (int i = 0; i < 6; ++i) vaxpy(ms, DP_B[i], kh[i], rho);
    calc_diff(m, kh[6], rho, dt_1, 0.0);
    st.n_rhs += 7;
    vscal(ms, DP_E[0], kh[0]);
    for (int i = 1; i < 7; ++i) vaxpy(ms, DP_E[i], kh[i], kh[0]);
    double err = errnrm1(kh[0], rho, rho_old, atol, rtol, ms);
    if (err < 1.0) {
      t += dt_1;
      dt_1 *= safety * std::min(10.0, std::pow(err, expo));
      accepted = true;
      st.n_accepted += 1;
    } else {
      dt_1 *= safety * std::max(0.2, std::pow(err, expo));
      if (dt_1 < dt_min) dt_1 = dt_min;
      vcopy(ms, rho_old, rho);
      st.n_rejected += 1;
    }
  }
  st.n_steps += 1;
  return true;
}

}  // namespace

extern "C" {

struct orc_stats {
  long n_rhs, n_steps, n_accepted, n_rejected;
  int gave_up;
};

void* orc_create(int kind, int n_level, int max_depth, int n_threads) {
  auto* m = new Model();
  m->kind = kind;
  m->N = n_level;
  m->N2 = n_level * n_level;
  m->max_depth = max_depth;
  m->n_threads = std::max(1, n_threads);
  m->H.assign(m->N2, 0.0);
  m->Hnz.assign(m->N2, 0);
  return m;
}

void orc_destroy(void* p) { delete static_cast<Model*>(p); }

// Sensitivity experiments: reverse == 1 sums the heom_hilb coupling terms of
// the noises in reverse order (a reordering of the same arithmetic).
void orc_set_sum_order(void* p, int reverse) {
  static_cast<Model*>(p)->reverse_noise_order = reverse != 0;
}

void orc_set_filter(void* p, int (*filter)(const int*, int, int, void*), void* user) {
  auto* m = static_cast<Model*>(p);
  m->filter = filter;
  m->filter_user = user;
  m->ready = false;
}

void orc_set_options(void* p, int parity_q1, int sqrt_norm) {
  auto* m = static_cast<Model*>(p);
  m->parity_q1 = parity_q1 != 0;
  m->sqrt_norm = sqrt_norm != 0;
}

void orc_set_threads(void* p, int n_threads) {
  auto* m = static_cast<Model*>(p);
  m->n_threads = std::max(1, n_threads);
  m->temp.assign(static_cast<size_t>(3) * m->N2 * m->n_threads, 0.0);
}

// Matrices arrive as interleaved (re, im) doubles, row-major; a stored entry
// is any entry that is not exactly zero (LIL push semantics).
void orc_set_H(void* p, const double* H) {
  auto* m = static_cast<Model*>(p);
  for (int i = 0; i < m->N2; ++i) {
    m->H[i] = cplx(H[2 * i], H[2 * i + 1]);
    m->Hnz[i] = (m->H[i] != cplx(0.0));
  }
}

void orc_alloc_noises(void* p, int n_noise) {
  auto* m = static_cast<Model*>(p);
  m->noise.assign(n_noise, Noise());
}

void orc_set_noise(void* p, int u, const double* V, int len, const double* gamma,
                   const double* phi0, const double* sigma, const double* S, const double* A,
                   double s_delta_re, double s_delta_im) {
  auto* m = static_cast<Model*>(p);
  Noise& z = m->noise[u];
  z.V.assign(m->N2, 0.0);
  z.Vnz.assign(m->N2, 0);
  for (int i = 0; i < m->N2; ++i) {
    z.V[i] = cplx(V[2 * i], V[2 * i + 1]);
    z.Vnz[i] = (z.V[i] != cplx(0.0));
  }
  z.len = len;
  auto load = [](const double* src, int cnt) {
    vector<cplx> out(cnt);
    for (int i = 0; i < cnt; ++i) out[i] = cplx(src[2 * i], src[2 * i + 1]);
    return out;
  };
  z.gamma = load(gamma, len * len);
  z.phi0 = load(phi0, len);
  z.sigma = load(sigma, len);
  z.S = load(S, len * len);
  z.A = load(A, len * len);
  z.s_delta = cplx(s_delta_re, s_delta_im);
}

int orc_set_param(void* p) {
  try {
    set_param(*static_cast<Model*>(p));
  } catch (std::exception& e) {
    std::fprintf(stderr, "oracle set_param: %s\n", e.what());
    return -1;
  }
  return 0;
}

// calc_hrchy_element_count (hrchy_space.cc:68-71) = C(dim+level, level).
long long orc_hrchy_count(int level, int dim) { return multicombination(dim + 1, level); }

int orc_n_hrchy(void* p) { return static_cast<Model*>(p)->n_hrchy; }
int orc_n_dim(void* p) { return static_cast<Model*>(p)->n_dim; }
long orc_main_size(void* p) {
  auto* m = static_cast<Model*>(p);
  return static_cast<long>(m->n_hrchy) * m->N2;
}
long orc_ado_nnz(void* p) { return static_cast<long>(static_cast<Model*>(p)->R_col.size()); }

// Redfield products: Lambda_u (N x N interleaved) and the eigenvalues of H.
int orc_redfield_lambda(void* p, int u, double* lambda, double* eig) {
  auto* m = static_cast<Model*>(p);
  if (m->kind != REDFIELD_HILB && m->kind != REDFIELD_LIOU) return -1;
  if (u < 0 || u >= static_cast<int>(m->rf_Lambda.size())) return -1;
  if (lambda)
    for (int i = 0; i < m->N2; ++i) {
      lambda[2 * i] = m->rf_Lambda[u][i].real();
      lambda[2 * i + 1] = m->rf_Lambda[u][i].imag();
    }
  if (eig)
    for (int i = 0; i < m->N; ++i) eig[i] = m->rf_w[i];
  return 0;
}

// Number of stored entries of redfield_liou's R after optimize.
long orc_redfield_nnz(void* p) {
  auto* m = static_cast<Model*>(p);
  long c = 0;
  for (char x : m->rf_R.nz) c += x;
  return c;
}

// n: n_hrchy*n_dim; p1/m1: (n_hrchy+1)*n_dim
void orc_hrchy_tables(void* p, int* n_out, int* p1_out, int* m1_out) {
  auto* m = static_cast<Model*>(p);
  const int K = m->n_dim;
  for (int l = 0; l < m->n_hrchy; ++l)
    for (int k = 0; k < K; ++k) n_out[static_cast<size_t>(l) * K + k] = m->n[l][k];
  for (int l = 0; l <= m->n_hrchy; ++l)
    for (int k = 0; k < K; ++k) {
      p1_out[static_cast<size_t>(l) * K + k] = m->ptr_p1[l][k];
      m1_out[static_cast<size_t>(l) * K + k] = m->ptr_m1[l][k];
    }
}

void orc_ngamma_diag(void* p, double* out) {
  auto* m = static_cast<Model*>(p);
  for (int l = 0; l < m->n_hrchy; ++l) {
    out[2 * l] = m->ngamma_diag[l].real();
    out[2 * l + 1] = m->ngamma_diag[l].imag();
  }
}

void orc_calc_diff(void* p, double* drho, const double* rho, double alpha_re, double alpha_im,
                   double beta_re, double beta_im) {
  auto* m = static_cast<Model*>(p);
  calc_diff(*m, reinterpret_cast<cplx*>(drho), reinterpret_cast<const cplx*>(rho),
            cplx(alpha_re, alpha_im), cplx(beta_re, beta_im));
}

double orc_errnrm1(const double* e, const double* x1, const double* x2, double atol, double rtol,
                   long n) {
  return errnrm1(reinterpret_cast<const cplx*>(e), reinterpret_cast<const cplx*>(x1),
                 reinterpret_cast<const cplx*>(x2), atol, rtol, n);
}

// One rk4 step (rk4.h:40-65) in place.
void orc_rk4_step(void* p, double* rho, double dt) {
  auto* m = static_cast<Model*>(p);
  const long ms = static_cast<long>(m->n_hrchy) * m->N2;
  vector<cplx> buf(4 * ms);
  Stats st;
  rk4_step(*m, reinterpret_cast<cplx*>(rho), dt, buf, st);
}

// One lsrk4 step (lsrk4.h:33-48) in place.
void orc_lsrk4_step(void* p, double* rho, double dt) {
  auto* m = static_cast<Model*>(p);
  const long ms = static_cast<long>(m->n_hrchy) * m->N2;
  vector<cplx> buf(ms);
  Stats st;
  lsrk4_step(*m, reinterpret_cast<cplx*>(rho), dt, buf, st);
}

// One rkdp adaptive step (rkdp.h:81-179) in place; t and dt are in/out.
int orc_rkdp_step(void* p, double* rho, double* t, double t_bound, double* dt, double atol,
                  double rtol, orc_stats* out) {
  auto* m = static_cast<Model*>(p);
  const long ms = static_cast<long>(m->n_hrchy) * m->N2;
  vector<cplx> buf(9 * ms);
  Stats st;
  bool ok = rkdp_step(*m, reinterpret_cast<cplx*>(rho), *t, t_bound, *dt, atol, rtol, buf, st);
  if (out) {
    out->n_rhs = st.n_rhs;
    out->n_steps = st.n_steps;
    out->n_accepted = st.n_accepted;
    out->n_rejected = st.n_rejected;
    out->gave_up = st.gave_up;
  }
  return ok ? 0 : 1;
}

// solver_base::solve (solver_base.h:64-79) with the callback replaced by a
// record of rho_S = rho[0:N2] at every t_list point (what the callback sees).
// solver: 0 = rk4 (fixed_step_size_solver.h:29-52), 1 = rkdp
// (adaptive_step_size_solver.h:32-56), 2 = lsrk4.  dt_inout returns dt_save
// for rkdp.
int orc_solve(void* p, int solver, double* rho, const double* t_list, int n_t, double dt,
              double atol, double rtol, double* rho_s_out, double* dt_inout, orc_stats* out) {
  auto* m = static_cast<Model*>(p);
  const long ms = static_cast<long>(m->n_hrchy) * m->N2;
  cplx* r = reinterpret_cast<cplx*>(rho);
  vector<cplx> buf((solver == 1 ? 9 : 4) * ms);
  Stats st;
  double dt_save = (dt_inout && *dt_inout > 0) ? *dt_inout : -1.0;
  auto record = [&](int i) {
    if (rho_s_out) std::memcpy(rho_s_out + 2 * static_cast<size_t>(i) * m->N2, rho,
                               sizeof(cplx) * m->N2);
  };
  for (int i = 0; i < n_t - 1; ++i) {
    record(i);
    if (solver != 1) {
      fixed_solve_1(*m, r, t_list[i], t_list[i + 1], dt, buf, st, solver == 2);
    } else {
      // Zero-length interval: no attempt, dt_save kept.  The reference would
      // take one attempt of dt = 0 (state unchanged), save dt_save = 0 and then
      // never finish the next interval; the library skips such intervals and so
      // does the oracle.
      if (!(t_list[i + 1] > t_list[i])) continue;
      if (dt_save == -1.0) dt_save = dt;
      double t = t_list[i];
      double dt_1 = dt_save;
      while (true) {
        rkdp_step(*m, r, t, t_list[i + 1], dt_1, atol, rtol, buf, st);
        if (t >= t_list[i + 1]) break;
      }
      dt_save = dt_1;
    }
  }
  record(n_t - 1);
  if (dt_inout) *dt_inout = dt_save;
  if (out) {
    out->n_rhs = st.n_rhs;
    out->n_steps = st.n_steps;
    out->n_accepted = st.n_accepted;
    out->n_rejected = st.n_rejected;
    out->gave_up = st.gave_up;
  }
  return st.gave_up ? 1 : 0;
}

}  // extern "C"
