// extern "C" boundary of libheom_amd (include/libheom_amd.h).  Every entry
// point catches all exceptions and turns them into status codes.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/libheom_amd.h"
#include "kernels.h"
#include "qme.h"
#include "mlp.h"
#include "noalloc.h"
#include "solver.h"

// The QME is shared with the solvers created on it: a solver keeps it alive
// after heom_destroy, and checks its size before every solve.
struct heom_qme {
  std::shared_ptr<hamd::Qme> q;
};
struct heom_mlp {
  std::unique_ptr<hamd::Mlp> m;
};
struct heom_solver {
  std::unique_ptr<hamd::Solver> s;
};

namespace hamd {
namespace {
thread_local std::string g_err;
}
void set_error(int code, const std::string& msg) {
  (void)code;
  g_err = msg;
}
}  // namespace hamd

namespace {

template <typename F>
int guard(F&& f) {
  try {
    return f();
  } catch (hamd::HeomError& e) {
    hamd::set_error(e.code, e.msg);
    return e.code;
  } catch (std::exception& e) {
    hamd::set_error(HEOM_ERR_ARG, e.what());
    return HEOM_ERR_ARG;
  } catch (...) {
    hamd::set_error(HEOM_ERR_ARG, "unknown error");
    return HEOM_ERR_ARG;
  }
}

hamd::Qme& Q(heom_handle h) {
  if (!h || !h->q) throw hamd::HeomError{HEOM_ERR_ARG, "null handle"};
  return *h->q;
}

// handles whose parameters the caller may still change (not batches)
hamd::Qme& Qcfg(heom_handle h) {
  hamd::Qme& q = Q(h);
  if (q.kind == hamd::kBatch)
    throw hamd::HeomError{HEOM_ERR_STATE, "a batch handle has no parameters of its own"};
  return q;
}

hamd::Mlp& M(heom_mlp_handle m) {
  if (!m || !m->m) throw hamd::HeomError{HEOM_ERR_ARG, "null network handle"};
  return *m->m;
}

hamd::Qme& Qready(heom_handle h) {
  hamd::Qme& q = Q(h);
  if (!q.ready) throw hamd::HeomError{HEOM_ERR_STATE, "set_param has not been called"};
  return q;
}

std::vector<hamd::z128> load_c(const double* p, size_t n) {
  std::vector<hamd::z128> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = hamd::z128{p[2 * i], p[2 * i + 1]};
  return v;
}

}  // namespace

extern "C" {

const char* heom_last_error(void) { return hamd::g_err.c_str(); }

int heom_version(int* major, int* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return HEOM_OK;
}

int heom_default_options(heom_options* opt) {
  if (!opt) return HEOM_ERR_ARG;
  opt->parity_q1 = 1;
  opt->sqrt_norm = 0;
  opt->order = 0;
  opt->device = 0;
  opt->matrix_cores = 1;
  opt->renorm_rho0 = 0;
  return HEOM_OK;
}

int heom_create(int kind, int dtype, int n_level, int max_depth, const heom_options* opt,
                heom_handle* out) {
  return guard([&] {
    if (!out) throw hamd::HeomError{HEOM_ERR_ARG, "heom_create: out is NULL"};
    *out = nullptr;
    if (kind < HEOM_HILB || kind > HEOM_REDFIELD_LIOU)
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_create: unknown kind"};
    const bool redfield = kind == HEOM_REDFIELD_HILB || kind == HEOM_REDFIELD_LIOU;
    if (redfield) max_depth = 0;
    if (dtype != HEOM_C128 && dtype != HEOM_C64 && dtype != HEOM_C64_BF16X3)
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_create: unknown dtype"};
    if (n_level < 1) throw hamd::HeomError{HEOM_ERR_ARG, "heom_create: n_level < 1"};
    if (max_depth < 0) throw hamd::HeomError{HEOM_ERR_ARG, "heom_create: max_depth < 0"};
    heom_options o;
    heom_default_options(&o);
    if (opt) o = *opt;
    if (o.order != 0 && o.order != 1)
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_create: order must be 0 or 1"};
    auto h = std::make_unique<heom_qme>();
    h->q = std::make_shared<hamd::Qme>(kind, dtype, n_level, max_depth, o);
    *out = h.release();
    return HEOM_OK;
  });
}

int heom_destroy(heom_handle h) {
  return guard([&] {
    delete h;
    return HEOM_OK;
  });
}

int heom_set_H(heom_handle h, const double* H) {
  return guard([&] {
    hamd::Qme& q = Qcfg(h);
    if (!H) throw hamd::HeomError{HEOM_ERR_ARG, "heom_set_H: NULL"};
    q.H = load_c(H, q.N2);
    q.H_set = true;
    q.ready = false;
    return HEOM_OK;
  });
}

int heom_alloc_noises(heom_handle h, int n_noise) {
  return guard([&] {
    hamd::Qme& q = Qcfg(h);
    if (n_noise < 1) throw hamd::HeomError{HEOM_ERR_ARG, "heom_alloc_noises: n_noise < 1"};
    q.noise.assign(n_noise, hamd::NoiseParams());
    q.ready = false;
    return HEOM_OK;
  });
}

int heom_set_noise(heom_handle h, int u, const double* V, int len_gamma, const double* gamma,
                   const double* phi_0, const double* sigma, const double* S, const double* A,
                   double s_delta_re, double s_delta_im) {
  return guard([&] {
    hamd::Qme& q = Qcfg(h);
    if (u < 0 || u >= static_cast<int>(q.noise.size()))
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_set_noise: noise index out of range"};
    if (len_gamma < 1) throw hamd::HeomError{HEOM_ERR_ARG, "heom_set_noise: len_gamma < 1"};
    if (!V || !gamma || !phi_0 || !sigma || !S || !A)
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_set_noise: NULL parameter"};
    hamd::NoiseParams& z = q.noise[u];
    const size_t L = static_cast<size_t>(len_gamma);
    z.V = load_c(V, q.N2);
    z.len = len_gamma;
    z.gamma = load_c(gamma, L * L);
    z.phi0 = load_c(phi_0, L);
    z.sigma = load_c(sigma, L);
    z.S = load_c(S, L * L);
    z.A = load_c(A, L * L);
    z.s_delta = hamd::z128{s_delta_re, s_delta_im};
    z.set = true;
    q.ready = false;
    return HEOM_OK;
  });
}

int heom_set_hrchy_filter(heom_handle h, heom_hrchy_filter filter, void* user) {
  return guard([&] {
    hamd::Qme& q = Qcfg(h);
    q.hrchy_filter = filter;
    q.hrchy_filter_user = filter ? user : nullptr;
    q.ready = false;
    return HEOM_OK;
  });
}

int heom_set_param(heom_handle h) {
  return guard([&] {
    Qcfg(h).set_param();
    return HEOM_OK;
  });
}

int heom_get_n_hrchy(heom_handle h, int* n) {
  return guard([&] {
    if (!n) throw hamd::HeomError{HEOM_ERR_ARG, "NULL output"};
    *n = Qready(h).hs.n_hrchy;
    return HEOM_OK;
  });
}

int heom_get_n_dim(heom_handle h, int* n) {
  return guard([&] {
    if (!n) throw hamd::HeomError{HEOM_ERR_ARG, "NULL output"};
    *n = Qready(h).hs.n_dim;
    return HEOM_OK;
  });
}

int heom_main_size(heom_handle h, long long* m) {
  return guard([&] {
    if (!m) throw hamd::HeomError{HEOM_ERR_ARG, "NULL output"};
    *m = Qready(h).main_size();
    return HEOM_OK;
  });
}

int heom_get_hrchy(heom_handle h, int* n, int* p1, int* m1) {
  return guard([&] {
    hamd::Qme& q = Qready(h);
    if (n) std::memcpy(n, q.hs.n.data(), sizeof(int) * q.hs.n.size());
    if (p1) std::memcpy(p1, q.hs.ptr_p1.data(), sizeof(int) * q.hs.ptr_p1.size());
    if (m1) std::memcpy(m1, q.hs.ptr_m1.data(), sizeof(int) * q.hs.ptr_m1.size());
    return HEOM_OK;
  });
}

int heom_redfield_lambda(heom_handle h, int u, double* lambda, double* eig) {
  return guard([&] {
    hamd::Qme& q = Qready(h);
    if (!q.redfield()) throw hamd::HeomError{HEOM_ERR_ARG, "not a Redfield QME"};
    if (u < 0 || u >= static_cast<int>(q.rf_lambda.size()))
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_redfield_lambda: noise index out of range"};
    if (lambda) std::memcpy(lambda, q.rf_lambda[u].data(), sizeof(hamd::z128) * q.N2);
    if (eig) std::memcpy(eig, q.rf_eig.data(), sizeof(double) * q.N);
    return HEOM_OK;
  });
}

int heom_calc_diff(heom_handle h, void* drho_dev, const void* rho_dev, double alpha_re,
                   double alpha_im, double beta_re, double beta_im, void* stream) {
  return guard([&] {
    hamd::Qme& q = Qready(h);
    if (!drho_dev || !rho_dev) throw hamd::HeomError{HEOM_ERR_ARG, "heom_calc_diff: NULL"};
    if (drho_dev == rho_dev)
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_calc_diff: drho and rho must not alias"};
    HAMD_HIP(hipSetDevice(q.opt.device));
    const hamd::StageArgs A = q.diff_args(drho_dev, rho_dev, hamd::z128{alpha_re, alpha_im},
                                          hamd::z128{beta_re, beta_im});
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : q.stream;
    q.operate(q.lo() ? 4 : 8, A, st);
    if (!stream) HAMD_HIP(hipStreamSynchronize(st));
    return HEOM_OK;
  });
}

int heom_calc_diff_host(heom_handle h, double* drho, const double* rho, double alpha_re,
                        double alpha_im, double beta_re, double beta_im) {
  return guard([&] {
    hamd::Qme& q = Qready(h);
    if (!drho || !rho) throw hamd::HeomError{HEOM_ERR_ARG, "heom_calc_diff_host: NULL"};
    const long long ms = q.main_size();
    const bool lo = q.lo();
    const size_t eb = lo ? 8 : 16;
    hamd::DevBuf din, dout;
    din.alloc(eb * ms);
    dout.alloc(eb * ms);
    if (lo) {
      std::vector<float> a(2 * ms), b(2 * ms);
      for (long long i = 0; i < 2 * ms; ++i) {
        a[i] = static_cast<float>(rho[i]);
        b[i] = static_cast<float>(drho[i]);
      }
      HAMD_HIP(hipMemcpy(din.p, a.data(), eb * ms, hipMemcpyHostToDevice));
      HAMD_HIP(hipMemcpy(dout.p, b.data(), eb * ms, hipMemcpyHostToDevice));
    } else {
      HAMD_HIP(hipMemcpy(din.p, rho, eb * ms, hipMemcpyHostToDevice));
      HAMD_HIP(hipMemcpy(dout.p, drho, eb * ms, hipMemcpyHostToDevice));
    }
    int rc = heom_calc_diff(h, dout.p, din.p, alpha_re, alpha_im, beta_re, beta_im, nullptr);
    if (rc != HEOM_OK) return rc;
    if (lo) {
      std::vector<float> b(2 * ms);
      HAMD_HIP(hipMemcpy(b.data(), dout.p, eb * ms, hipMemcpyDeviceToHost));
      for (long long i = 0; i < 2 * ms; ++i) drho[i] = b[i];
    } else {
      HAMD_HIP(hipMemcpy(drho, dout.p, eb * ms, hipMemcpyDeviceToHost));
    }
    return HEOM_OK;
  });
}

int heom_batch_create(const heom_handle* members, int n_members, heom_handle* out) {
  return guard([&] {
    if (!out || !members || n_members < 1)
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_batch_create: bad arguments"};
    *out = nullptr;
    std::vector<hamd::Qme*> ms;
    for (int m = 0; m < n_members; ++m) ms.push_back(&Qready(members[m]));
    auto h = std::make_unique<heom_qme>();
    h->q = std::make_shared<hamd::Qme>(HEOM_HILB, ms[0]->dtype, 1, 0, ms[0]->opt);
    h->q->set_param_batch(ms);
    *out = h.release();
    return HEOM_OK;
  });
}

int heom_batch_layout(heom_handle batch, long long* elem_base, int* rho_s_offset, int* n_level) {
  return guard([&] {
    hamd::Qme& q = Qready(batch);
    if (q.kind != hamd::kBatch) throw hamd::HeomError{HEOM_ERR_ARG, "not a batch handle"};
    for (size_t m = 0; m < q.batch_inst.size(); ++m) {
      if (elem_base) elem_base[m] = q.batch_inst[m].elem_base;
      if (rho_s_offset) rho_s_offset[m] = q.batch_inst[m].rs_off;
      if (n_level) n_level[m] = q.batch_inst[m].N;
    }
    return HEOM_OK;
  });
}

int heom_batch_tuning(heom_handle batch, int* shape, int* tried, float* tried_us, int cap,
                      int* n_tried) {
  return guard([&] {
    hamd::Qme& q = Qready(batch);
    if (q.kind != hamd::kBatch) throw hamd::HeomError{HEOM_ERR_ARG, "not a batch handle"};
    if (shape) {
      shape[0] = q.dev.batch_ipw;
      shape[1] = q.dev.batch_wpb;
      shape[2] = q.dev.batch_grid;
    }
    const int n = static_cast<int>(q.tune_log.size());
    for (int k = 0; k < n && k < cap; ++k) {
      if (tried) {
        tried[3 * k] = q.tune_log[k].ipw;
        tried[3 * k + 1] = q.tune_log[k].wpb;
        tried[3 * k + 2] = q.tune_log[k].grid;
      }
      if (tried_us) tried_us[k] = q.tune_log[k].us;
    }
    if (n_tried) *n_tried = n;
    return HEOM_OK;
  });
}

int heom_batch_set_tuning(heom_handle batch, int items_per_step, int waves_per_group, int grid) {
  return guard([&] {
    hamd::Qme& q = Qready(batch);
    if (q.kind != hamd::kBatch) throw hamd::HeomError{HEOM_ERR_ARG, "not a batch handle"};
    if ((items_per_step != 1 && items_per_step != 2) ||
        (waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4) || grid < 1 ||
        grid > hamd::kMaxGrid / waves_per_group)
      throw hamd::HeomError{HEOM_ERR_ARG,
                            "heom_batch_set_tuning: items_per_step 1 or 2, waves_per_group 1, 2 "
                            "or 4, 1 <= grid <= 262144 / waves_per_group"};
    q.dev.batch_ipw = items_per_step;
    q.dev.batch_wpb = waves_per_group;
    q.dev.batch_grid = grid;
    return HEOM_OK;
  });
}

int heom_mlp_create(int n_in, int n_hidden, int n_out, unsigned long long seed, int device,
                    heom_mlp_handle* out) {
  return guard([&] {
    if (!out) throw hamd::HeomError{HEOM_ERR_ARG, "heom_mlp_create: out is NULL"};
    *out = nullptr;
    auto h = std::make_unique<heom_mlp>();
    h->m = std::make_unique<hamd::Mlp>(n_in, n_hidden, n_out, seed, device);
    *out = h.release();
    return HEOM_OK;
  });
}

int heom_mlp_destroy(heom_mlp_handle m) {
  return guard([&] {
    delete m;
    return HEOM_OK;
  });
}

int heom_mlp_n_params(heom_mlp_handle m, long long* n) {
  return guard([&] {
    if (!n) throw hamd::HeomError{HEOM_ERR_ARG, "NULL output"};
    *n = M(m).n_params();
    return HEOM_OK;
  });
}

int heom_mlp_get_params(heom_mlp_handle m, float* host) {
  return guard([&] {
    if (!host) throw hamd::HeomError{HEOM_ERR_ARG, "NULL output"};
    M(m).get_params(host);
    return HEOM_OK;
  });
}

int heom_mlp_set_params(heom_mlp_handle m, const float* host) {
  return guard([&] {
    if (!host) throw hamd::HeomError{HEOM_ERR_ARG, "NULL input"};
    M(m).set_params(host);
    return HEOM_OK;
  });
}

int heom_mlp_set_scaling(heom_mlp_handle m, const float* xs, const float* xc, const float* ys,
                         const float* yc) {
  return guard([&] {
    if (!xs || !xc || !ys || !yc) throw hamd::HeomError{HEOM_ERR_ARG, "NULL input"};
    M(m).set_scaling(xs, xc, ys, yc);
    return HEOM_OK;
  });
}

int heom_mlp_get_scaling(heom_mlp_handle m, float* xs, float* xc, float* ys, float* yc) {
  return guard([&] {
    M(m).get_scaling(xs, xc, ys, yc);
    return HEOM_OK;
  });
}

int heom_mlp_fit_scaling(heom_mlp_handle m, const float* X, const float* Y, long long n_rows) {
  return guard([&] {
    if (!X || !Y) throw hamd::HeomError{HEOM_ERR_ARG, "NULL input"};
    M(m).fit_scaling(X, Y, n_rows);
    return HEOM_OK;
  });
}

int heom_mlp_forward(heom_mlp_handle m, const float* X, float* Y, int n_rows) {
  return guard([&] {
    if (!X || !Y || n_rows < 0) throw hamd::HeomError{HEOM_ERR_ARG, "heom_mlp_forward: bad input"};
    M(m).forward(X, Y, n_rows);
    return HEOM_OK;
  });
}

int heom_mlp_loss_grad(heom_mlp_handle m, const float* X, const float* Y, int n_rows,
                       double* loss, float* grad_host) {
  return guard([&] {
    if (!X || !Y || n_rows < 1) throw hamd::HeomError{HEOM_ERR_ARG, "heom_mlp_loss_grad: bad input"};
    const double l = M(m).loss_grad(X, Y, n_rows, nullptr);
    if (loss) *loss = l;
    if (grad_host) M(m).get_grad(grad_host);
    return HEOM_OK;
  });
}

int heom_mlp_adam_step(heom_mlp_handle m, double lr, double beta1, double beta2, double eps) {
  return guard([&] {
    M(m).adam_step(lr, beta1, beta2, eps);
    HAMD_HIP(hipStreamSynchronize(M(m).stream));
    return HEOM_OK;
  });
}

int heom_mlp_train(heom_mlp_handle m, const float* X, const float* Y, long long n_rows, int batch,
                   int epochs, double lr, unsigned long long seed, double* epoch_loss) {
  return guard([&] {
    if (!X || !Y) throw hamd::HeomError{HEOM_ERR_ARG, "NULL input"};
    M(m).train(X, Y, n_rows, batch, epochs, lr, seed, epoch_loss);
    return HEOM_OK;
  });
}

int heom_set_closure(heom_handle h, heom_mlp_handle m, const float* features, int n_features,
                     int layout) {
  return guard([&] {
    hamd::Qme& q = Qready(h);
    if (q.kind == hamd::kBatch)
      throw hamd::HeomError{HEOM_ERR_UNSUPPORTED, "closure: not for batch handles"};
    q.clear_closure();
    if (m) {
      auto c = std::make_unique<hamd::Closure>();
      c->bind(q, M(m), features, n_features, layout);
      q.closure = std::move(c);
    }
    return HEOM_OK;
  });
}

int heom_closure_dims(heom_handle h, int layout, int n_features, int* rows, int* n_in,
                      int* n_out) {
  return guard([&] {
    const hamd::ClosureDims c = hamd::closure_dims(Qready(h), layout, n_features);
    if (rows) *rows = c.rows;
    if (n_in) *n_in = c.n_in;
    if (n_out) *n_out = c.n_out;
    return HEOM_OK;
  });
}

int heom_closure_rows(heom_handle deep, heom_handle shallow, const void* rho_deep_dev,
                      const float* features, int n_features, int layout, float* x_dev,
                      float* y_dev) {
  return guard([&] {
    if (!rho_deep_dev || !x_dev || !y_dev)
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_closure_rows: NULL"};
    hamd::closure_rows(Qready(deep), Qready(shallow), rho_deep_dev, features, n_features, layout,
                       x_dev, y_dev);
    return HEOM_OK;
  });
}

int heom_solver_create(heom_handle h, int kind, heom_solver_handle* out) {
  return guard([&] {
    if (!out) throw hamd::HeomError{HEOM_ERR_ARG, "heom_solver_create: out is NULL"};
    *out = nullptr;
    (void)Qready(h);
    auto s = std::make_unique<heom_solver>();
    s->s = std::make_unique<hamd::Solver>(h->q, kind);
    *out = s.release();
    return HEOM_OK;
  });
}

int heom_solver_destroy(heom_solver_handle s) {
  return guard([&] {
    delete s;
    return HEOM_OK;
  });
}

int heom_solve_dev(heom_solver_handle s, void* rho_dev, const double* t_list, int n_t, double dt,
                   double atol, double rtol, double* rho_s_out) {
  return guard([&] {
    if (!s || !s->s) throw hamd::HeomError{HEOM_ERR_ARG, "null solver"};
    if (!rho_dev || !t_list) throw hamd::HeomError{HEOM_ERR_ARG, "heom_solve_dev: NULL"};
    hamd::Solver& sv = *s->s;
    sv.check_qme();
    const bool batch = sv.q->kind == hamd::kBatch;
    const int N2 = batch ? sv.q->batch_rs_size : sv.q->N2;
    hamd::DevBuf rs;
    if (rho_s_out) rs.alloc(sizeof(hamd::z128) * N2 * static_cast<size_t>(n_t));
    auto out = [&](int i, const void* cur) {
      if (!rho_s_out) return;
      char* dst = static_cast<char*>(rs.p) + sizeof(hamd::z128) * N2 * i;
      if (batch)
        hamd::launch_gather_rho_s(sv.q->dev, dst, cur, sv.q->stream);
      else
        HAMD_HIP(hipMemcpyAsync(dst, cur, sizeof(hamd::z128) * N2, hipMemcpyDeviceToDevice,
                                sv.q->stream));
    };
    int rc = (rho_s_out && !batch)
                 ? sv.solve_gather(rho_dev, t_list, n_t, dt, atol, rtol, rs.p)
                 : sv.solve(rho_dev, t_list, n_t, dt, atol, rtol, out);
    if (rho_s_out)
      HAMD_HIP(hipMemcpy(rho_s_out, rs.p, sizeof(hamd::z128) * N2 * static_cast<size_t>(n_t),
                         hipMemcpyDeviceToHost));
    if (rc != HEOM_OK)
      hamd::set_error(rc, "rkdp: step rejected 1000 times; interval skipped as in the reference");
    return rc;
  });
}

int heom_solve_dev_noalloc(heom_solver_handle s, void* rho_dev, const double* t_list, int n_t,
                           double dt, double atol, double rtol, double* rho_s_out) {
  return guard([&] {
    if (!s || !s->s) throw hamd::HeomError{HEOM_ERR_ARG, "null solver"};
    if (!rho_dev || !t_list) throw hamd::HeomError{HEOM_ERR_ARG, "heom_solve_dev_noalloc: NULL"};
    if (n_t < 1) throw hamd::HeomError{HEOM_ERR_ARG, "heom_solve_dev_noalloc: n_t < 1"};
    hamd::Solver& sv = *s->s;
    sv.check_qme();
    const bool batch = sv.q->kind == hamd::kBatch;
    const size_t N2 = static_cast<size_t>(batch ? sv.q->batch_rs_size : sv.q->N2);
    const size_t rs_bytes = sizeof(hamd::z128) * N2 * static_cast<size_t>(n_t);
    if (rs_bytes > hamd::kNoallocRhoSBytes)
      throw hamd::HeomError{HEOM_ERR_ARG,
                            "heom_solve_dev_noalloc: n_t x n_level^2 x 16 bytes of rho_S exceed "
                            "the scratch (" + std::to_string(hamd::kNoallocRhoSBytes) +
                                " bytes)"};
    char* base = static_cast<char*>(hamd::noalloc_scratch(sv.q->opt.device));
    auto* ctrl = reinterpret_cast<hamd::Ctrl*>(base + hamd::kNoallocCtrlOff);
    auto* partials = reinterpret_cast<double*>(base + hamd::kNoallocPartialsOff);
    void* rs = base + hamd::kNoallocRhoSOff;
    int rc = sv.solve_noalloc(rho_dev, t_list, n_t, dt, atol, rtol, ctrl, partials, rs);
    if (rho_s_out) HAMD_HIP(hipMemcpy(rho_s_out, rs, rs_bytes, hipMemcpyDeviceToHost));
    if (rc != HEOM_OK)
      hamd::set_error(rc, "rkdp: step rejected 1000 times; interval skipped as in the reference");
    return rc;
  });
}

int heom_solve(heom_solver_handle s, double* rho_host, const double* t_list, int n_t,
               heom_callback cb, void* user, double dt, double atol, double rtol,
               int full_copy) {
  return guard([&] {
    if (!s || !s->s) throw hamd::HeomError{HEOM_ERR_ARG, "null solver"};
    if (!rho_host || !t_list) throw hamd::HeomError{HEOM_ERR_ARG, "heom_solve: NULL"};
    hamd::Solver& sv = *s->s;
    sv.check_qme();
    const long long ms = sv.q->main_size();
    const int N2 = sv.q->N2;
    const size_t bytes = sizeof(hamd::z128) * ms;
    hamd::DevBuf d;
    d.alloc(bytes);
    HAMD_HIP(hipMemcpy(d.p, rho_host, bytes, hipMemcpyHostToDevice));
    std::vector<double> rs(2 * N2);
    auto out = [&](int i, const void* cur) {
      if (full_copy) {
        HAMD_HIP(hipMemcpyAsync(rho_host, cur, bytes, hipMemcpyDeviceToHost, sv.q->stream));
        HAMD_HIP(hipStreamSynchronize(sv.q->stream));
        if (cb) cb(t_list[i], rho_host, rho_host, user);
      } else {
        HAMD_HIP(hipMemcpyAsync(rs.data(), cur, sizeof(hamd::z128) * N2, hipMemcpyDeviceToHost,
                                sv.q->stream));
        HAMD_HIP(hipStreamSynchronize(sv.q->stream));
        if (cb) cb(t_list[i], rs.data(), nullptr, user);
      }
    };
    int rc = sv.solve(d.p, t_list, n_t, dt, atol, rtol, out);
    HAMD_HIP(hipMemcpy(rho_host, d.p, bytes, hipMemcpyDeviceToHost));
    if (rc != HEOM_OK)
      hamd::set_error(rc, "rkdp: step rejected 1000 times; interval skipped as in the reference");
    return rc;
  });
}

int heom_solver_stats(heom_solver_handle s, heom_stats* out) {
  return guard([&] {
    if (!s || !s->s || !out) throw hamd::HeomError{HEOM_ERR_ARG, "heom_solver_stats: NULL"};
    *out = s->s->stats;
    return HEOM_OK;
  });
}

int heom_solver_reset(heom_solver_handle s) {
  return guard([&] {
    if (!s || !s->s) throw hamd::HeomError{HEOM_ERR_ARG, "null solver"};
    s->s->dt_save = -1.0;
    std::memset(&s->s->stats, 0, sizeof(heom_stats));
    return HEOM_OK;
  });
}

int heom_solver_set_graphs(heom_solver_handle s, int enable) {
  return guard([&] {
    if (!s || !s->s) throw hamd::HeomError{HEOM_ERR_ARG, "null solver handle"};
    s->s->use_graphs = enable != 0;
    return HEOM_OK;
  });
}

int heom_solver_time_stages(heom_solver_handle s, int reps, double* ms_per_stage, int* n_stage) {
  return guard([&] {
    if (!s || !s->s || !ms_per_stage) throw hamd::HeomError{HEOM_ERR_ARG, "NULL"};
    if (reps < 1) throw hamd::HeomError{HEOM_ERR_ARG, "reps < 1"};
    std::vector<double> v;
    s->s->time_stages(reps, v);
    for (size_t i = 0; i < v.size(); ++i) ms_per_stage[i] = v[i];
    if (n_stage) *n_stage = static_cast<int>(v.size());
    return HEOM_OK;
  });
}

int heom_solver_stage_bytes(heom_solver_handle s, double* bytes, int* n_stage) {
  return guard([&] {
    if (!s || !s->s || !bytes) throw hamd::HeomError{HEOM_ERR_ARG, "NULL"};
    std::vector<double> v;
    s->s->stage_bytes(v);
    for (size_t i = 0; i < v.size(); ++i) bytes[i] = v[i];
    if (n_stage) *n_stage = static_cast<int>(v.size());
    return HEOM_OK;
  });
}

int heom_dev_alloc(void** ptr, long long bytes) {
  return guard([&] {
    if (!ptr || bytes < 0) throw hamd::HeomError{HEOM_ERR_ARG, "heom_dev_alloc"};
    hipError_t e = hipMalloc(ptr, static_cast<size_t>(bytes));
    if (e != hipSuccess) throw hamd::HeomError{HEOM_ERR_NOMEM, hipGetErrorString(e)};
    return HEOM_OK;
  });
}

int heom_dev_free(void* ptr) {
  return guard([&] {
    HAMD_HIP(hipFree(ptr));
    return HEOM_OK;
  });
}

int heom_memcpy_h2d(void* dst, const void* src, long long bytes) {
  return guard([&] {
    HAMD_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice));
    return HEOM_OK;
  });
}

int heom_memcpy_d2h(void* dst, const void* src, long long bytes) {
  return guard([&] {
    HAMD_HIP(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost));
    return HEOM_OK;
  });
}

int heom_stream_copy(long long bytes, int grid, int reps, double* gbs) {
  return guard([&] {
    if (!gbs || bytes < 16 || grid < 1 || reps < 1)
      throw hamd::HeomError{HEOM_ERR_ARG, "heom_stream_copy: bad argument"};
    const long long n = bytes / 16;
    hamd::DevBuf src, dst;
    src.alloc(static_cast<size_t>(n) * 16);
    dst.alloc(static_cast<size_t>(n) * 16);
    HAMD_HIP(hipMemset(src.p, 0, static_cast<size_t>(n) * 16));
    hipEvent_t e0, e1;
    HAMD_HIP(hipEventCreate(&e0));
    HAMD_HIP(hipEventCreate(&e1));
    hamd::launch_copy_grid(dst.p, src.p, n, grid, nullptr);  // warm-up
    HAMD_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) hamd::launch_copy_grid(dst.p, src.p, n, grid, nullptr);
    HAMD_HIP(hipEventRecord(e1, nullptr));
    HAMD_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    HAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbs = 2.0 * 16.0 * static_cast<double>(n) * reps / (static_cast<double>(ms) * 1e-3) / 1e9;
    return HEOM_OK;
  });
}

int heom_device_sync(void) {
  return guard([&] {
    HAMD_HIP(hipDeviceSynchronize());
    return HEOM_OK;
  });
}

}  // extern "C"
