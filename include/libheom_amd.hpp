// libheom_amd.hpp -- C++17 drop-in for libheom's QME / solver templates,
// implemented over the C ABI of libheom_amd.h (HIP engine on one MI355X).
//
// Mirrors (names, template parameters, member names, argument meaning):
//   qme_base            qme_base.h:23-80      (n_level, n_level_2, H, n_noise, V,
//                                              len_gamma, gamma, phi_0, sigma, S, A,
//                                              s_delta, alloc_noises, set_param,
//                                              calc_diff_impl, main_size, temp_size)
//   heom                heom.h:21-140         (ctor(max_depth, n_inner, n_outer),
//                                              get_n_hrchy)
//   heom_hilb / heom_liou / heom_ado           heom_hilb.h:26, heom_liou.h:27, heom_ado.h:27
//   solver_base, rk4, rkdp                     solver/solver_base.h, rk4.h, rkdp.h
//   qme_solver          qme_solver.h:18-53
// The engine tag `hip` replaces eigen/mkl/cuda (linalg_engine.h:100-134).
// Parameter matrices use lil_matrix with the reference's `data` / `push` /
// `set_shape` members (lil_matrix.h.j2:22-183).  Errors throw heom_error
// (status code + message); a missing kwarg throws std::out_of_range and a
// wrongly typed one std::bad_any_cast, as kwargs.at / std::any_cast do in the
// reference (fixed_step_size_solver.h:37, adaptive_step_size_solver.h:40-43).
#ifndef LIBHEOM_AMD_HPP
#define LIBHEOM_AMD_HPP

#include <any>
#include <complex>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "libheom_amd.h"

namespace libheom {

using complex64 = std::complex<float>;
using complex128 = std::complex<double>;
template <typename T>
using real_t = typename T::value_type;
using kwargs_t = std::map<std::string, std::any>;
using order_t = bool;
constexpr order_t row_major = true;
constexpr order_t col_major = false;
constexpr int dynamic = 0;

struct heom_error : std::runtime_error {
  int code;
  heom_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != HEOM_OK) throw heom_error(rc, heom_last_error());
}

// engine tag: one HIP device
class hip {
 public:
  explicit hip(int device_ = 0) : device(device_) {}
  int device;
};

// storage-format tags of the reference's engine matrices (dense_matrix.h,
// sparse_matrix.h); the HIP engine picks its own device layout
template <int n_level_c, typename dtype, order_t order, typename engine>
class dense_matrix {};
template <int n_level_c, typename dtype, order_t order, typename engine>
class sparse_matrix {};

// parameter container with the reference's LIL interface
template <typename dtype>
class lil_matrix {
 public:
  std::tuple<int, int> shape{0, 0};
  std::map<int, std::map<int, dtype>> data;
  void set_shape(int rows, int cols) {
    shape = std::make_tuple(rows, cols);
    data.clear();
  }
  void push(int row, int col, const dtype& v) { data[row][col] += v; }
  std::vector<double> dense(int rows, int cols) const {
    std::vector<double> out(2 * static_cast<size_t>(rows) * cols, 0.0);
    for (auto& r : data)
      for (auto& c : r.second) {
        if (r.first < 0 || r.first >= rows || c.first < 0 || c.first >= cols)
          throw heom_error(HEOM_ERR_ARG, "lil_matrix entry outside its shape");
        out[2 * (static_cast<size_t>(r.first) * cols + c.first)] = std::real(c.second);
        out[2 * (static_cast<size_t>(r.first) * cols + c.first) + 1] = std::imag(c.second);
      }
    return out;
  }
};

template <typename dtype>
std::vector<double> interleave(const std::vector<dtype>& v) {
  std::vector<double> out(2 * v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    out[2 * i] = std::real(v[i]);
    out[2 * i + 1] = std::imag(v[i]);
  }
  return out;
}

template <typename dtype, order_t order, typename engine>
class qme_base {
 public:
  static_assert(std::is_same<dtype, complex128>::value || std::is_same<dtype, complex64>::value,
                "dtype must be complex128 or complex64");
  qme_base() : n_level(0), n_level_2(0), n_noise(0) {}
  virtual ~qme_base() {
    if (h_) heom_destroy(h_);
  }
  qme_base(const qme_base&) = delete;
  qme_base& operator=(const qme_base&) = delete;

  int n_level;
  int n_level_2;
  lil_matrix<dtype> H;
  int n_noise;
  std::unique_ptr<lil_matrix<dtype>[]> V;
  std::unique_ptr<int[]> len_gamma;
  std::unique_ptr<lil_matrix<dtype>[]> gamma;
  std::unique_ptr<std::vector<dtype>[]> phi_0;
  std::unique_ptr<std::vector<dtype>[]> sigma;
  std::unique_ptr<lil_matrix<dtype>[]> S;
  std::unique_ptr<lil_matrix<dtype>[]> A;
  std::unique_ptr<dtype[]> s_delta;

  void alloc_noises(int n) {
    n_noise = n;
    V.reset(new lil_matrix<dtype>[n]);
    len_gamma.reset(new int[n]());
    gamma.reset(new lil_matrix<dtype>[n]);
    phi_0.reset(new std::vector<dtype>[n]);
    sigma.reset(new std::vector<dtype>[n]);
    S.reset(new lil_matrix<dtype>[n]);
    A.reset(new lil_matrix<dtype>[n]);
    s_delta.reset(new dtype[n]());
  }

  virtual int main_size() { return n_level * n_level * n_hrchy_; }
  virtual int temp_size() { return 0; }

  virtual void set_param(engine* obj) {
    heom_options opt;
    heom_default_options(&opt);
    opt.order = order == row_major ? 0 : 1;
    opt.device = obj ? obj->device : 0;
    opt.parity_q1 = parity_q1;
    opt.sqrt_norm = sqrt_normalization;
    opt.matrix_cores = matrix_cores;
    opt.renorm_rho0 = renorm_rho0;
    const int dt = std::is_same<dtype, complex64>::value ? HEOM_C64 : HEOM_C128;
    // the handle is set up again in place while its create-time options are
    // unchanged, so a qme_solver built on this object sees the new parameters
    // (the reference's qme_solver keeps a pointer to the same qme)
    if (h_ && (n_level != h_n_level_ || std::memcmp(&opt, &h_opt_, sizeof(opt)) != 0)) {
      heom_destroy(h_);  // solvers bound to it keep it alive and rebind on their next solve
      h_ = nullptr;
    }
    if (!h_) {
      check(heom_create(kind_, dt, n_level, max_depth_, &opt, &h_));
      h_opt_ = opt;
      h_n_level_ = n_level;
      h_gen_ = next_generation();
    }
    check(heom_set_hrchy_filter(h_, nullptr, nullptr));
    check(heom_set_H(h_, H.dense(n_level, n_level).data()));
    check(heom_alloc_noises(h_, n_noise));
    for (int u = 0; u < n_noise; ++u) {
      const int len = len_gamma[u];
      if (static_cast<int>(phi_0[u].size()) != len || static_cast<int>(sigma[u].size()) != len)
        throw heom_error(HEOM_ERR_ARG, "phi_0/sigma length differs from len_gamma");
      const auto Vd = V[u].dense(n_level, n_level);
      const auto g = gamma[u].dense(len, len);
      const auto p0 = interleave(phi_0[u]);
      const auto sg = interleave(sigma[u]);
      const auto Sd = S[u].dense(len, len);
      const auto Ad = A[u].dense(len, len);
      check(heom_set_noise(h_, u, Vd.data(), len, g.data(), p0.data(), sg.data(), Sd.data(),
                           Ad.data(), std::real(s_delta[u]), std::imag(s_delta[u])));
    }
    if (hrchy_filter) check(heom_set_hrchy_filter(h_, &filter_trampoline, this));
    check(heom_set_param(h_));
    check(heom_get_n_hrchy(h_, &n_hrchy_));
  }

  // drho_dt = beta*drho_dt - alpha*D(rho) on device buffers of main_size
  // elements (qme_base.h:74-79); temp is not needed by the HIP engine
  virtual void calc_diff_impl(engine*, dtype* drho_dt, dtype* rho, dtype alpha, dtype beta,
                              dtype* temp = nullptr) {
    (void)temp;
    check(heom_calc_diff(h_, drho_dt, rho, std::real(alpha), std::imag(alpha), std::real(beta),
                         std::imag(beta), nullptr));
  }

  heom_handle handle() const { return h_; }
  // changes whenever set_param replaces the handle (addresses may be reused)
  unsigned long long handle_generation() const { return h_gen_; }

  // hrchy_filter of alloc_hrchy_space (hrchy_space.h:29-36); empty: none
  std::function<bool(std::vector<int>, int)> hrchy_filter;

  // reference quirk switches (parity defaults)
  int parity_q1 = 1;
  int sqrt_normalization = 0;
  int matrix_cores = 1;  // MFMA commutator where provided (16 < n_level <= 64, diagonal V)
  int renorm_rho0 = 0;   // FP64 trace / hermiticity projection of rho_0 after every step

 protected:
  static int filter_trampoline(const int* index, int n_dim, int depth, void* user) {
    const auto* self = static_cast<const qme_base*>(user);
    return self->hrchy_filter(std::vector<int>(index, index + n_dim), depth) ? 1 : 0;
  }
  int kind_ = HEOM_HILB;
  int max_depth_ = 0;
  int n_hrchy_ = 0;
  heom_handle h_ = nullptr;
  heom_options h_opt_{};  // create-time options of h_
  int h_n_level_ = 0;
  unsigned long long h_gen_ = 0;
  static unsigned long long next_generation() {
    static unsigned long long g = 0;
    return ++g;
  }
};

template <typename dtype, order_t order, typename engine>
class heom : public qme_base<dtype, order, engine> {
 public:
  heom(int max_depth, int n_inner_threads, int n_outer_threads)
      : n_inner_threads(n_inner_threads), n_outer_threads(n_outer_threads) {
    this->max_depth_ = max_depth;
  }
  int get_n_hrchy() { return this->n_hrchy_; }
  int n_inner_threads, n_outer_threads;  // accepted for API parity; the device replaces them
};

template <int n_level_c, typename dtype,
          template <int, typename, order_t, typename> class matrix_base, order_t order,
          typename engine>
class heom_hilb : public heom<dtype, order, engine> {
 public:
  heom_hilb(int max_depth, int n_inner_threads, int n_outer_threads)
      : heom<dtype, order, engine>(max_depth, n_inner_threads, n_outer_threads) {
    this->kind_ = HEOM_HILB;
  }
};

template <int n_level_c, typename dtype,
          template <int, typename, order_t, typename> class matrix_base, order_t order,
          order_t order_liou, typename engine>
class heom_liou : public heom<dtype, order, engine> {
 public:
  heom_liou(int max_depth, int n_inner_threads, int n_outer_threads)
      : heom<dtype, order, engine>(max_depth, n_inner_threads, n_outer_threads) {
    this->kind_ = HEOM_LIOU;
  }
};

template <int n_level_c, typename dtype,
          template <int, typename, order_t, typename> class matrix_base, order_t order,
          order_t order_liou, typename engine>
class heom_ado : public heom<dtype, order, engine> {
 public:
  heom_ado(int max_depth, int n_inner_threads, int n_outer_threads)
      : heom<dtype, order, engine>(max_depth, n_inner_threads, n_outer_threads) {
    this->kind_ = HEOM_ADO;
  }
};

// Redfield QMEs (redfield.h:19-174): the same parameter members; set_param
// computes Lambda_u from the eigendecomposition of H and the bath correlations
template <typename dtype, order_t order, typename engine>
class redfield : public qme_base<dtype, order, engine> {
 public:
  // Lambda_u (row-major N x N, logical) after set_param (redfield.h:30, 166)
  std::vector<dtype> Lambda(int u) const {
    std::vector<double> buf(2 * static_cast<size_t>(this->n_level) * this->n_level);
    check(heom_redfield_lambda(this->h_, u, buf.data(), nullptr));
    std::vector<dtype> out(buf.size() / 2);
    for (size_t i = 0; i < out.size(); ++i) out[i] = dtype(buf[2 * i], buf[2 * i + 1]);
    return out;
  }
};

template <int n_level_c, typename dtype,
          template <int, typename, order_t, typename> class matrix_base, order_t order,
          typename engine>
class redfield_hilb : public redfield<dtype, order, engine> {
 public:
  redfield_hilb() { this->kind_ = HEOM_REDFIELD_HILB; }  // redfield_hilb.h:36-37
};

template <int n_level_c, typename dtype,
          template <int, typename, order_t, typename> class matrix_base, order_t order,
          order_t order_liou, typename engine>
class redfield_liou : public redfield<dtype, order, engine> {
 public:
  redfield_liou() { this->kind_ = HEOM_REDFIELD_LIOU; }  // redfield_liou.h:41-42
};

template <typename dtype, order_t order, typename engine>
class solver_base {
 public:
  virtual ~solver_base() = default;
  virtual int kind() const = 0;
  virtual bool adaptive() const = 0;
};

template <typename dtype, order_t order, typename engine>
class rk4 : public solver_base<dtype, order, engine> {
 public:
  int kind() const override { return HEOM_RK4; }
  bool adaptive() const override { return false; }
};

template <typename dtype, order_t order, typename engine>
class rkdp : public solver_base<dtype, order, engine> {
 public:
  int kind() const override { return HEOM_RKDP; }
  bool adaptive() const override { return true; }
};

// solver/lsrk4.h
template <typename dtype, order_t order, typename engine>
class lsrk4 : public solver_base<dtype, order, engine> {
 public:
  int kind() const override { return HEOM_LSRK4; }
  bool adaptive() const override { return false; }
};

template <typename dtype, order_t order, typename engine>
class qme_solver {
 public:
  engine* eng;
  qme_base<dtype, order, engine>* qme;
  solver_base<dtype, order, engine>* solver;

  qme_solver(engine* engine_, qme_base<dtype, order, engine>* qme_,
             solver_base<dtype, order, engine>* solver_)
      : eng(engine_), qme(qme_), solver(solver_) {
    check(heom_solver_create(qme->handle(), solver->kind(), &s_));
    qgen_ = qme->handle_generation();
  }
  ~qme_solver() {
    if (s_) heom_solver_destroy(s_);
  }
  qme_solver(const qme_solver&) = delete;
  qme_solver& operator=(const qme_solver&) = delete;

  // one derivative evaluation drho_dt = -D(rho) on host buffers; the
  // reference forwards to an empty stub (solver_base.h:55-62)
  void calc_diff(dtype* drho_dt, dtype* rho) {
    const long long n = qme->main_size();
    std::vector<double> in(2 * n), out(2 * n, 0.0);
    for (long long i = 0; i < n; ++i) {
      in[2 * i] = std::real(rho[i]);
      in[2 * i + 1] = std::imag(rho[i]);
    }
    check(heom_calc_diff_host(qme->handle(), out.data(), in.data(), 1.0, 0.0, 0.0, 0.0));
    for (long long i = 0; i < n; ++i) drho_dt[i] = dtype(out[2 * i], out[2 * i + 1]);
  }

  // solver_base::solve (solver_base.h:64-79): rho (host, main_size) is
  // propagated in place; callback(t_i) sees rho(t_i)
  void solve(dtype* rho, const real_t<dtype>* t_list, int n_t,
             std::function<void(real_t<dtype>)> callback, const kwargs_t& kwargs) {
    const double dt = static_cast<double>(std::any_cast<real_t<dtype>>(kwargs.at("dt")));
    double atol = 0.0, rtol = 0.0;
    if (solver->adaptive()) {
      atol = static_cast<double>(std::any_cast<real_t<dtype>>(kwargs.at("atol")));
      rtol = static_cast<double>(std::any_cast<real_t<dtype>>(kwargs.at("rtol")));
    }
    const long long n = qme->main_size();
    std::vector<double> buf(2 * n);
    std::vector<double> ts(t_list, t_list + n_t);
    auto to_buf = [&] {
      for (long long i = 0; i < n; ++i) {
        buf[2 * i] = std::real(rho[i]);
        buf[2 * i + 1] = std::imag(rho[i]);
      }
    };
    auto from_buf = [&] {
      for (long long i = 0; i < n; ++i)
        rho[i] = dtype(static_cast<real_t<dtype>>(buf[2 * i]),
                       static_cast<real_t<dtype>>(buf[2 * i + 1]));
    };
    to_buf();
    struct Ctx {
      std::function<void(real_t<dtype>)>* cb;
      std::function<void()>* sync;
    };
    std::function<void()> sync = from_buf;
    Ctx ctx{&callback, &sync};
    auto trampoline = [](double t, const double*, const double*, void* user) {
      Ctx* c = static_cast<Ctx*>(user);
      (*c->sync)();
      if (*c->cb) (*c->cb)(static_cast<real_t<dtype>>(t));
    };
    if (qme->handle_generation() != qgen_) {  // set_param replaced the qme's handle: rebind
      heom_solver_handle s2 = nullptr;
      check(heom_solver_create(qme->handle(), solver->kind(), &s2));
      heom_solver_destroy(s_);
      s_ = s2;
      qgen_ = qme->handle_generation();
    }
    int rc = heom_solve(s_, buf.data(), ts.data(), n_t, trampoline, &ctx, dt, atol, rtol, 1);
    from_buf();
    check(rc);
  }

  heom_stats stats() const {
    heom_stats st;
    check(heom_solver_stats(s_, &st));
    return st;
  }

 private:
  heom_solver_handle s_ = nullptr;
  unsigned long long qgen_ = 0;  // generation of the qme handle s_ was created on
};

}  // namespace libheom

#endif  // LIBHEOM_AMD_HPP
