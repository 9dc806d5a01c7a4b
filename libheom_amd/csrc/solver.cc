// RK4 (rk4.h:40-65) and Dormand-Prince 5(4) (rkdp.h:21-179) drivers.  Every
// stage is one fused operator pass (kernels.hip); the Dormand-Prince step
// control runs on the device, so the host only reads the control block back
// once per batch of attempts instead of once per attempt.
#include "solver.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <set>

#include "kernels.h"

namespace hamd {

namespace {
// slot numbers
enum {
  SL_Y = 0, SL_YNEW = 1, SL_Y32 = 2, SL_YNEW32 = 3,
  SL_SA = 4, SL_SB = 5, SL_ACC = 6, SL_K1 = 7,  // K1..K5 = 7..11
  SL_E = 13  // Dormand-Prince: sum_{j<6} E_j k_j, always FP64
};
}  // namespace

int StageSpec::pre_index(int slot) {
  for (size_t s = 0; s < pre.size(); ++s)
    if (pre[s] == slot) return static_cast<int>(s);
  if (static_cast<int>(pre.size()) >= kMaxPre)
    throw HeomError{HEOM_ERR_STATE, "stage needs too many operands"};
  pre.push_back(slot);
  return static_cast<int>(pre.size()) - 1;
}

// dst (dst2) <- sum of terms (in the given order) + ckv*k
void StageSpec::combo(const std::vector<std::pair<int, double>>& terms, double ckv, int d,
                      int d2) {
  const int c = ncombo++;
  int last = -1;
  for (auto& t : terms) {
    const int s = pre_index(t.first);
    if (s <= last) throw HeomError{HEOM_ERR_STATE, "combination order differs from operands"};
    last = s;
    cc[c][s] = t.second;
  }
  ck[c] = ckv;
  dst[c] = d;
  dst2[c] = d2;
}

Solver::Solver(std::shared_ptr<Qme> q_, int kind_)
    : q_owner(std::move(q_)), q(q_owner.get()), kind(kind_) {
  if (!q) throw HeomError{HEOM_ERR_ARG, "solver: null QME"};
  if (!q->ready) throw HeomError{HEOM_ERR_STATE, "solver: set_param has not been called"};
  if (kind != HEOM_RK4 && kind != HEOM_RKDP && kind != HEOM_LSRK4)
    throw HeomError{HEOM_ERR_ARG, "solver: unknown solver kind"};
  HAMD_HIP(hipSetDevice(q->opt.device));
  in_bytes = q->lo() ? 4 : 8;
  ms = q->main_size();
  const size_t z128b = sizeof(z128) * ms;
  const size_t sinb = (in_bytes == 4 ? sizeof(z64) : sizeof(z128)) * ms;
  const int lo = in_bytes == 4 ? 1 : 0;
  auto mk = [&](int s, size_t b, int is_lo) {
    P[s].alloc(b);
    HAMD_HIP(hipMemset(P[s].p, 0, b));
    sp[s] = P[s].p;
    slo[s] = is_lo;
  };
  mk(SL_Y, z128b, 0);
  // Dormand-Prince: the y / y_new pair; lsrk4: the state alternates between them
  if (kind != HEOM_RK4) mk(SL_YNEW, z128b, 0);
  if (lo) {
    mk(SL_Y32, sinb, 1);
    if (kind != HEOM_RK4) mk(SL_YNEW32, sinb, 1);
  }
  if (kind != HEOM_LSRK4) {
    mk(SL_SA, sinb, lo);
    mk(SL_SB, sinb, lo);
  }
  if (kind == HEOM_RK4 || kind == HEOM_LSRK4) {
    mk(SL_ACC, z128b, 0);  // rk4: the accumulator; lsrk4: buff
  } else {
    for (int k = 0; k < 5; ++k) mk(SL_K1 + k, sinb, lo);
    mk(SL_E, z128b, 0);
  }
  if (q->opt.renorm_rho0) {
    if (q->kind == kBatch)
      throw HeomError{HEOM_ERR_UNSUPPORTED, "renorm_rho0 is not provided for batches"};
    if (q->N > 64) throw HeomError{HEOM_ERR_UNSUPPORTED, "renorm_rho0: n_level > 64"};
    d_tr0.alloc(sizeof(double));
  }
  d_ctrl.alloc(sizeof(Ctrl));
  HAMD_HIP(hipMemset(d_ctrl.p, 0, sizeof(Ctrl)));
  d_partials.alloc(sizeof(double) * kMaxGrid);
  HAMD_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_ctrl), sizeof(Ctrl), hipHostMallocDefault));
  std::memset(h_ctrl, 0, sizeof(Ctrl));
  build_stages(0.0, 0.0);
}

Solver::~Solver() {
  if (h_ctrl) (void)hipHostFree(h_ctrl);
}

void Solver::check_qme() const {
  if (!q->ready)
    throw HeomError{HEOM_ERR_STATE, "solver: the QME's parameters changed and set_param has not "
                                    "been called again"};
  if (q->main_size() != ms)
    throw HeomError{HEOM_ERR_STATE, "solver: the QME's main_size changed since the solver was "
                                    "created; create a new solver"};
}

// Physical buffer of a slot for ping-pong parity `cur` (slots 0..3 swap
// within their pair when cur == 1).
void* Solver::slot_ptr(int s, int cur) const {
  if (s < 0) return nullptr;
  if (s < 4) return sp[(s & 2) + ((s & 1) ^ cur)];
  return sp[s];
}

StageArgs Solver::resolve(const StageSpec& s) const {
  StageArgs a;
  std::memset(&a, 0, sizeof(a));
  a.mode = MODE_STAGE;
  const bool pp = kind == HEOM_RKDP;  // ping-pong pairs only with the control block
  for (int cur = 0; cur < 2; ++cur) {
    const int c = pp ? cur : 0;
    a.in[cur] = slot_ptr(s.in_slot, c);
    a.k[cur] = slot_ptr(s.k_slot, c);
    for (size_t j = 0; j < s.pre.size(); ++j) a.pre[cur][j] = slot_ptr(s.pre[j], c);
    for (int k = 0; k < 2; ++k) {
      a.dst[k][cur] = slot_ptr(s.dst[k], c);
      a.dst2[k][cur] = slot_ptr(s.dst2[k], c);
    }
  }
  a.in_lo = slo[s.in_slot];
  a.npre = static_cast<int>(s.pre.size());
  for (size_t j = 0; j < s.pre.size(); ++j)
    if (s.pre[j] >= 0 && slo[s.pre[j]]) a.pre_lo |= 1u << j;
  a.k_lo = s.k_slot >= 0 ? slo[s.k_slot] : 0;
  a.ncombo = s.ncombo;
  std::memcpy(a.cc, s.cc, sizeof(a.cc));
  std::memcpy(a.ck, s.ck, sizeof(a.ck));
  for (int k = 0; k < 2; ++k) {
    a.dst_lo[k] = s.dst[k] >= 0 ? slo[s.dst[k]] : 0;
    a.dst2_lo[k] = s.dst2[k] >= 0 ? slo[s.dst2[k]] : 0;
  }
  a.do_err = s.do_err;
  std::memcpy(a.ec, s.ec, sizeof(a.ec));
  a.ek = s.ek;
  a.ynew_pre = s.ynew_pre;
  a.yold_pre = s.yold_pre;
  a.control = s.control;
  a.inv_n = 1.0 / static_cast<double>(ms);
  if (kind == HEOM_RKDP) a.ctrl = ctrl_dev();
  a.partials = partials_dev();
  a.max_partials = kMaxGrid;
  return a;
}

void Solver::build_stages(double atol, double rtol) {
  specs.clear();
  stages.clear();
  const bool lo = in_bytes == 4;
  const int y_in = lo ? SL_Y32 : SL_Y;
  if (kind == HEOM_RK4) {
    // rk4.h:47-64; k2 and k3 enter the accumulator as soon as they exist
    StageSpec a;
    a.in_slot = y_in;
    a.combo({{SL_Y, 1.0}}, 0.5, SL_SA);
    a.combo({{SL_Y, 1.0}}, 1.0 / 6.0, SL_ACC);
    specs.push_back(a);

    a = StageSpec();
    a.in_slot = SL_SA;
    a.combo({{SL_Y, 1.0}}, 0.5, SL_SB);
    a.combo({{SL_ACC, 1.0}}, 2.0 / 6.0, SL_ACC);
    specs.push_back(a);

    a = StageSpec();
    a.in_slot = SL_SB;
    a.combo({{SL_Y, 1.0}}, 1.0, SL_SA);
    a.combo({{SL_ACC, 1.0}}, 2.0 / 6.0, SL_ACC);
    specs.push_back(a);

    a = StageSpec();
    a.in_slot = SL_SA;
    a.combo({{SL_ACC, 1.0}}, 1.0 / 6.0, SL_Y, lo ? SL_Y32 : -1);
    specs.push_back(a);
  } else if (kind == HEOM_LSRK4) {
    // lsrk4.h:33-48: buff = beta*buff + k (k = -dt*D) with beta = 0, -1, -1, -1,
    // then rho += c*buff, c = 1/3, 3/4, 2/3, 1/4.  The state alternates between
    // y and y_new (neighbour reads never see a half-written state); rho + c*buff
    // is formed as rho + c*beta*buff + c*k from the operands before the update.
    const double beta[4] = {0.0, -1.0, -1.0, -1.0};
    const double c[4] = {1.0 / 3.0, 3.0 / 4.0, 2.0 / 3.0, 1.0 / 4.0};
    const int src[4] = {SL_Y, SL_YNEW, SL_Y, SL_YNEW};
    const int dst[4] = {SL_YNEW, SL_Y, SL_YNEW, SL_Y};
    const int src32[4] = {SL_Y32, SL_YNEW32, SL_Y32, SL_YNEW32};
    for (int s = 0; s < 4; ++s) {
      StageSpec a;
      a.in_slot = lo ? src32[s] : src[s];
      if (beta[s] == 0.0) {
        a.combo({{src[s], 1.0}}, c[s], dst[s], lo ? src32[(s + 1) % 4] : -1);
        a.combo({}, 1.0, SL_ACC);
      } else {
        a.combo({{src[s], 1.0}, {SL_ACC, c[s] * beta[s]}}, c[s], dst[s],
                lo ? src32[(s + 1) % 4] : -1);
        a.combo({{SL_ACC, beta[s]}}, 1.0, SL_ACC);
      }
      specs.push_back(a);
    }
  } else {
    // Dormand-Prince: stage i (0-based) evaluates k_i at its input and writes
    // the input of stage i+1 (rkdp.h:111-127); stage 5 also forms the new
    // state (rkdp.h:130-132) and, from the same operands, the error sum
    // e6 = sum_{j<6} E_j k_j in the order of rkdp.h:142-145 (k_6 itself is
    // not stored); stage 6 evaluates k_7 at the new state, completes
    // e = e6 + E_7 k_7 and the error norm (rkdp.h:144-153).  The error vector
    // is the same fma chain as summing all seven terms in the last stage.
    const int sin_[6] = {y_in, SL_SA, SL_SB, SL_SA, SL_SB, SL_SA};
    const int sout[5] = {SL_SA, SL_SB, SL_SA, SL_SB, SL_SA};
    for (int i = 0; i < 6; ++i) {
      StageSpec a;
      a.in_slot = sin_[i];
      std::vector<std::pair<int, double>> terms{{SL_Y, 1.0}};
      if (i < 5) {
        a.k_slot = SL_K1 + i;
        for (int j = 0; j < i; ++j) terms.push_back({SL_K1 + j, kDpA[i + 1][j]});
        a.combo(terms, kDpA[i + 1][i], sout[i]);
      } else {
        for (int j = 0; j < 5; ++j)
          if (kDpB[j] != 0.0) terms.push_back({SL_K1 + j, kDpB[j]});
        a.combo(terms, kDpB[5], SL_YNEW, lo ? SL_YNEW32 : -1);
        std::vector<std::pair<int, double>> eterms;
        for (int j = 0; j < 5; ++j)
          if (kDpE[j] != 0.0) eterms.push_back({SL_K1 + j, kDpE[j]});
        a.combo(eterms, kDpE[5], SL_E);
      }
      specs.push_back(a);
    }
    StageSpec a;
    a.in_slot = lo ? SL_YNEW32 : SL_YNEW;
    a.do_err = 1;
    a.ec[a.pre_index(SL_E)] = 1.0;
    a.ek = kDpE[6];
    // the error weights read y_new / y_old from the last two operand positions
    a.pre.push_back(SL_YNEW);
    a.pre.push_back(SL_Y);
    a.ynew_pre = 1;
    a.yold_pre = 2;
    a.control = 1;
    specs.push_back(a);
  }
  for (auto& s : specs) {
    StageArgs r = resolve(s);
    r.atol = atol;
    r.rtol = rtol;
    stages.push_back(r);
  }
}

// The fixed-step stage program in LDS-buffer form for the resident integrator
// (slot SL_Y -> buffer 0, the other slots in order of first use).
void Solver::build_resident() {
  res = ResidentArgs{};
  if (!use_graphs || kind == HEOM_RKDP || in_bytes != 8 || q->opt.renorm_rho0 ||
      !resident_ok(q->dev, q->dev.max_links))
    return;
  std::map<int, int> bmap{{SL_Y, 0}};
  auto b = [&](int slot) {
    auto it = bmap.find(slot);
    if (it != bmap.end()) return it->second;
    const int id = static_cast<int>(bmap.size());
    bmap[slot] = id;
    return id;
  };
  ResidentArgs r{};
  r.n_stages = static_cast<int>(specs.size());
  if (r.n_stages > 4) return;
  for (int s = 0; s < r.n_stages; ++s) {
    const StageSpec& sp = specs[s];
    if (static_cast<int>(sp.pre.size()) > kResidentPre || sp.ncombo > 2 || sp.k_slot >= 0 ||
        sp.do_err)
      return;
    ResidentStage& t = r.st[s];
    t.in = b(sp.in_slot);
    t.npre = static_cast<int>(sp.pre.size());
    for (int p = 0; p < t.npre; ++p) t.pre[p] = b(sp.pre[p]);
    t.ncombo = sp.ncombo;
    for (int c = 0; c < sp.ncombo; ++c) {
      for (int p = 0; p < t.npre; ++p) t.cc[c][p] = sp.cc[c][p];
      t.ck[c] = sp.ck[c];
      t.dst[c] = b(sp.dst[c]);
      if (sp.dst2[c] >= 0) return;
    }
  }
  r.n_buf = static_cast<int>(bmap.size());
  if (resident_lds_bytes(q->dev, r.n_buf) > kResidentLds) return;
  std::vector<ResidentStage> prog(r.st, r.st + r.n_stages);
  d_res_prog.upload(prog);
  r.prog = d_res_prog.as<ResidentStage>();
  // the compiled-in program when the solver's equals it field by field
  auto same = [&](const ResidentProg& P) {
    if (P.n_stages != r.n_stages || P.n_buf != r.n_buf) return false;
    for (int s = 0; s < r.n_stages; ++s) {
      const ResidentStage &x = P.st[s], &y = r.st[s];
      if (x.in != y.in || x.npre != y.npre || x.ncombo != y.ncombo) return false;
      for (int p = 0; p < x.npre; ++p)
        if (x.pre[p] != y.pre[p]) return false;
      for (int c = 0; c < x.ncombo; ++c) {
        if (x.ck[c] != y.ck[c] || x.dst[c] != y.dst[c]) return false;
        for (int p = 0; p < x.npre; ++p)
          if (x.cc[c][p] != y.cc[c][p]) return false;
      }
    }
    return true;
  };
  r.scheme = same(kResidentRk4) ? 1 : same(kResidentLsrk4) ? 2 : 0;
  res = r;
}

// rho_0 projection after a step / attempt (heom_options.renorm_rho0)
void Solver::renorm(bool with_ctrl) {
  const bool lo = in_bytes == 4;
  launch_rho0_renorm(slot_ptr(SL_Y, 0), slot_ptr(SL_Y, with_ctrl ? 1 : 0),
                     lo ? slot_ptr(SL_Y32, 0) : nullptr,
                     lo ? slot_ptr(SL_Y32, with_ctrl ? 1 : 0) : nullptr,
                     with_ctrl ? ctrl_dev() : nullptr, q->N, d_tr0.as<double>(), q->stream);
}

void Solver::launch_stage(int s, double dt) {
  StageArgs a = stages[s];
  a.a_re = dt;
  int grid = q->operate(in_bytes, a, q->stream);
  if (a.do_err) launch_err_finalize(a, grid, q->stream);
}

// Record the kernels `body` launches on the library stream into g (stream
// capture; the stage launches enqueue kernels only, no memory operations).
template <typename F>
void Solver::capture(GraphExec& g, F&& body) {
  g.reset();
  hipStream_t st = q->stream;
  HAMD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipGraph_t graph = nullptr;
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(st, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HAMD_HIP(hipStreamEndCapture(st, &graph));
  const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HAMD_HIP(e);
}

void Solver::reset_graphs() {
  g_rk4.clear();
  g_dopri.clear();
  g_rk4_dt = 0.0;
}

// Graph of n repetitions of `one` (cached by n; the cache holds a few sizes).
GraphExec& Solver::graph_for(std::map<long long, std::unique_ptr<GraphExec>>& cache, long long n,
                             const std::function<void()>& one) {
  auto it = cache.find(n);
  if (it != cache.end()) return *it->second;
  if (cache.size() >= kMaxGraphs) cache.erase(cache.begin());
  auto g = std::make_unique<GraphExec>();
  capture(*g, [&] {
    for (long long k = 0; k < n; ++k) one();
  });
  return *cache.emplace(n, std::move(g)).first->second;
}

// `count` repetitions of `one` as graph launches of power-of-two sizes
void Solver::replay(std::map<long long, std::unique_ptr<GraphExec>>& cache, long long count,
                    const std::function<void()>& one) {
  while (count > 0) {
    long long n = kMaxGraphSteps;
    while (n > count) n >>= 1;
    HAMD_HIP(hipGraphLaunch(graph_for(cache, n, one).exec, q->stream));
    count -= n;
  }
}

// fixed_step_size_solver::solve_1 (fixed_step_size_solver.h:29-52): steps of
// dt, the last one shortened to end exactly at t1.  The step sequence is
// computed on the host with the reference's arithmetic; the steps of the full
// dt are replayed from graphs of kGraphSteps steps and of one step.
void Solver::fixed_steps(double t0, double t1, double dt, long long& n_full, double& dt_last) {
  double t = t0, dt_1 = dt;
  n_full = 0;
  while (true) {
    if (t + dt_1 - t1 > -std::numeric_limits<double>::epsilon()) {
      dt_1 = t1 - t;
      break;
    }
    ++n_full;
    t += dt_1;
  }
  dt_last = dt_1;
}

void Solver::rk4_interval(double t0, double t1, double dt) {
  long long n_full = 0;
  double dt_1 = dt;
  fixed_steps(t0, t1, dt, n_full, dt_1);
  if (res.n_stages > 0) {  // the whole interval in one launch
    ResidentArgs a = res;
    a.y = static_cast<z128*>(P[SL_Y].p);
    a.n_full = n_full;
    a.dt = dt;
    a.dt_last = dt_1;
    launch_resident(q->dev, q->dev.max_links, a, q->stream);
    stats.n_rhs += 4 * (n_full + 1);
    stats.n_attempts += n_full + 1;
    stats.n_accepted += n_full + 1;
    return;
  }
  long long done = 0;
  if (use_graphs && n_full > 0) {
    if (g_rk4_dt != dt) {
      g_rk4.clear();
      g_rk4_dt = dt;
    }
    replay(g_rk4, n_full, [&] {
      for (int s = 0; s < 4; ++s) launch_stage(s, dt);
      if (q->opt.renorm_rho0) renorm(false);
    });
    done = n_full;
  }
  for (; done < n_full; ++done) {
    for (int s = 0; s < 4; ++s) launch_stage(s, dt);
    if (q->opt.renorm_rho0) renorm(false);
  }
  for (int s = 0; s < 4; ++s) launch_stage(s, dt_1);  // the last (shortened) step
  if (q->opt.renorm_rho0) renorm(false);
  stats.n_rhs += 4 * (n_full + 1);
  stats.n_attempts += n_full + 1;
  stats.n_accepted += n_full + 1;
}

// adaptive_step_size_solver::solve_1 (adaptive_step_size_solver.h:32-56) with
// the attempts of rkdp::solve_adaptive_step launched in batches.
int Solver::rkdp_interval(double t0, double t1, double dt_first, int& batch) {
  launch_ctrl_init(ctrl_dev(), t0, t1, dt_first, q->stream);
  long long a0 = h_ctrl->attempts;
  // the projection kernel runs after every attempt and acts only when the
  // control block counts a new accepted step
  const auto attempt = [&] {
    for (int s = 0; s < 7; ++s) launch_stage(s, 0.0);
    if (q->opt.renorm_rho0) renorm(true);
  };
  while (true) {
    if (use_graphs) {
      replay(g_dopri, batch, attempt);
    } else {
      for (int b = 0; b < batch; ++b) attempt();
    }
    HAMD_HIP(hipMemcpyAsync(h_ctrl, ctrl_dev(), sizeof(Ctrl), hipMemcpyDeviceToHost, q->stream));
    HAMD_HIP(hipStreamSynchronize(q->stream));
    if (h_ctrl->done) break;
    // attempts still needed at the current step (attempts after the interval
    // is done are no-ops of a few microseconds each; a short batch costs one
    // more host round trip)
    const double rem = (h_ctrl->t_bound - h_ctrl->t) / h_ctrl->dt;
    const long long est =
        (rem > 0.0 && rem < static_cast<double>(kMaxAttemptBatch)) ? static_cast<long long>(
                                                                         std::ceil(rem))
                                                                   : 2LL * batch;
    batch = static_cast<int>(std::min<long long>(kMaxAttemptBatch, std::max<long long>(1, est)));
  }
  // the next interval starts with as many attempts as this one took
  const long long used = h_ctrl->attempts - a0;
  batch = static_cast<int>(std::max<long long>(1, std::min<long long>(used, 4096)));
  return h_ctrl->gave_up;
}

int Solver::solve(void* rho_dev, const double* t_list, int n_t, double dt, double atol,
                  double rtol, const std::function<void(int, const void*)>& on_output) {
  check_qme();
  if (n_t < 1) throw HeomError{HEOM_ERR_ARG, "solve: n_t < 1"};
  for (int i = 0; i + 1 < n_t; ++i)
    if (!(t_list[i + 1] >= t_list[i]))
      throw HeomError{HEOM_ERR_ARG, "solve: t_list must be non-decreasing"};
  if (kind != HEOM_RKDP && !(dt > 0.0))
    throw HeomError{HEOM_ERR_ARG, "rk4/lsrk4: dt must be > 0"};
  if (kind == HEOM_RKDP) {
    if (!(dt > 0.0) && dt_save <= 0.0) throw HeomError{HEOM_ERR_ARG, "rkdp: dt must be > 0"};
    if (!(atol >= 0.0) || !(rtol >= 0.0) || (atol == 0.0 && rtol == 0.0))
      throw HeomError{HEOM_ERR_ARG, "rkdp: atol/rtol must be >= 0 and not both 0"};
  }
  HAMD_HIP(hipSetDevice(q->opt.device));
  build_stages(atol, rtol);
  build_resident();  // nothing without graphs
  if (!noalloc) reset_graphs();  // stage arguments (tolerances, closure) may have changed
  const size_t bytes = sizeof(z128) * ms;
  hipStream_t st = q->stream;
  // the caller's state enters the FP64 state slot
  HAMD_HIP(hipMemcpyAsync(P[SL_Y].p, rho_dev, bytes, hipMemcpyDeviceToDevice, st));
  if (in_bytes == 4) launch_narrow(P[SL_Y32].p, P[SL_Y].p, ms, st);
  if (q->opt.renorm_rho0) launch_rho0_trace(P[SL_Y].p, q->N, d_tr0.as<double>(), st);
  int gave_up = 0;
  if (kind != HEOM_RKDP) {  // fixed step: rk4, lsrk4
    for (int i = 0; i < n_t - 1; ++i) {
      on_output(i, P[SL_Y].p);
      rk4_interval(t_list[i], t_list[i + 1], dt);
    }
    on_output(n_t - 1, P[SL_Y].p);
    HAMD_HIP(hipMemcpyAsync(rho_dev, P[SL_Y].p, bytes, hipMemcpyDeviceToDevice, st));
  } else {
    std::memset(h_ctrl, 0, sizeof(Ctrl));
    HAMD_HIP(hipMemcpyAsync(ctrl_dev(), h_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, st));
    HAMD_HIP(hipStreamSynchronize(st));
    double first = dt_save > 0.0 ? dt_save : dt;
    int batch = 4;
    int cur = 0;
    bool started = false;
    for (int i = 0; i < n_t - 1; ++i) {
      on_output(i, P[cur].p);
      // a zero-length interval takes no attempt and keeps the step size (the
      // reference would clamp the step to 0 and never finish the next interval)
      if (!(t_list[i + 1] > t_list[i])) continue;
      gave_up |= rkdp_interval(t_list[i], t_list[i + 1], started ? -1.0 : first, batch);
      started = true;
      cur = h_ctrl->cur;
    }
    on_output(n_t - 1, P[cur].p);
    if (started) dt_save = h_ctrl->dt;
    stats.n_rhs += 7 * h_ctrl->attempts;
    stats.n_attempts += h_ctrl->attempts;
    stats.n_accepted += h_ctrl->accepted;
    stats.n_rejected += h_ctrl->rejected;
    stats.last_err = h_ctrl->err;
    stats.last_dt = dt_save;
    HAMD_HIP(hipMemcpyAsync(rho_dev, P[cur].p, bytes, hipMemcpyDeviceToDevice, st));
  }
  HAMD_HIP(hipStreamSynchronize(st));
  stats.ado_stage_updates = stats.n_rhs * q->hs.n_hrchy;
  stats.gave_up |= gave_up;
  return gave_up ? HEOM_ERR_STEP_REJECTED : HEOM_OK;
}

int Solver::solve_gather(void* rho_dev, const double* t_list, int n_t, double dt, double atol,
                         double rtol, void* rho_s_dev) {
  check_qme();
  const size_t rs_bytes = sizeof(z128) * q->N2;
  hipStream_t st = q->stream;
  if (kind == HEOM_RKDP && n_t >= 2 && use_graphs && in_bytes == 8 && !q->opt.renorm_rho0 &&
      resident_dopri_ok(q->dev, q->dev.max_links))
    return solve_resident_dopri(rho_dev, t_list, n_t, dt, atol, rtol, rho_s_dev);
  if (kind == HEOM_RKDP || n_t < 2 || q->kind == kBatch) {
    return solve(rho_dev, t_list, n_t, dt, atol, rtol, [&](int i, const void* cur) {
      HAMD_HIP(hipMemcpyAsync(static_cast<char*>(rho_s_dev) + rs_bytes * i, cur, rs_bytes,
                              hipMemcpyDeviceToDevice, st));
    });
  }
  for (int i = 0; i + 1 < n_t; ++i)
    if (!(t_list[i + 1] >= t_list[i]))
      throw HeomError{HEOM_ERR_ARG, "solve: t_list must be non-decreasing"};
  if (!(dt > 0.0)) throw HeomError{HEOM_ERR_ARG, "rk4/lsrk4: dt must be > 0"};
  HAMD_HIP(hipSetDevice(q->opt.device));
  build_stages(atol, rtol);
  build_resident();
  if (res.n_stages == 0) {
    reset_graphs();
    return solve(rho_dev, t_list, n_t, dt, atol, rtol, [&](int i, const void* cur) {
      HAMD_HIP(hipMemcpyAsync(static_cast<char*>(rho_s_dev) + rs_bytes * i, cur, rs_bytes,
                              hipMemcpyDeviceToDevice, st));
    });
  }
  // the whole solve in one resident launch
  std::vector<long long> nf(n_t - 1);
  std::vector<double> dl(n_t - 1);
  long long steps = 0;
  for (int i = 0; i + 1 < n_t; ++i) {
    fixed_steps(t_list[i], t_list[i + 1], dt, nf[i], dl[i]);
    steps += nf[i] + 1;
  }
  d_res_nf.upload(nf);
  d_res_dl.upload(dl);
  ResidentArgs a = res;
  a.y = static_cast<z128*>(P[SL_Y].p);
  a.dt = dt;
  a.n_int = n_t - 1;
  a.n_full_i = d_res_nf.as<long long>();
  a.dt_last_i = d_res_dl.as<double>();
  a.rho_s = static_cast<z128*>(rho_s_dev);
  const size_t bytes = sizeof(z128) * ms;
  HAMD_HIP(hipMemcpyAsync(P[SL_Y].p, rho_dev, bytes, hipMemcpyDeviceToDevice, st));
  launch_resident(q->dev, q->dev.max_links, a, st);
  HAMD_HIP(hipMemcpyAsync(rho_dev, P[SL_Y].p, bytes, hipMemcpyDeviceToDevice, st));
  HAMD_HIP(hipStreamSynchronize(st));
  stats.n_rhs += 4 * steps;
  stats.n_attempts += steps;
  stats.n_accepted += steps;
  stats.ado_stage_updates = stats.n_rhs * q->hs.n_hrchy;
  return HEOM_OK;
}

// adaptive_step_size_solver::solve (adaptive_step_size_solver.h:32-56) for a
// small problem: every interval, attempt and output in one resident launch.
int Solver::solve_resident_dopri(void* rho_dev, const double* t_list, int n_t, double dt,
                                 double atol, double rtol, void* rho_s_dev) {
  check_qme();
  for (int i = 0; i + 1 < n_t; ++i)
    if (!(t_list[i + 1] >= t_list[i]))
      throw HeomError{HEOM_ERR_ARG, "solve: t_list must be non-decreasing"};
  if (!(dt > 0.0) && dt_save <= 0.0) throw HeomError{HEOM_ERR_ARG, "rkdp: dt must be > 0"};
  if (!(atol >= 0.0) || !(rtol >= 0.0) || (atol == 0.0 && rtol == 0.0))
    throw HeomError{HEOM_ERR_ARG, "rkdp: atol/rtol must be >= 0 and not both 0"};
  HAMD_HIP(hipSetDevice(q->opt.device));
  build_stages(atol, rtol);  // kept current for time_stages / stage_bytes
  reset_graphs();
  hipStream_t st = q->stream;
  std::vector<double> t(t_list, t_list + n_t);
  d_res_t.upload(t);
  ResidentDopriArgs a{};
  a.y = static_cast<z128*>(P[SL_Y].p);
  a.n_int = n_t - 1;
  a.t_list = d_res_t.as<double>();
  a.dt_first = dt_save > 0.0 ? dt_save : dt;
  a.atol = atol;
  a.rtol = rtol;
  a.inv_n = 1.0 / static_cast<double>(ms);
  a.rho_s = static_cast<z128*>(rho_s_dev);
  a.ctrl = d_ctrl.as<Ctrl>();
  const size_t bytes = sizeof(z128) * ms;
  HAMD_HIP(hipMemcpyAsync(P[SL_Y].p, rho_dev, bytes, hipMemcpyDeviceToDevice, st));
  launch_resident_dopri(q->dev, q->dev.max_links, a, st);
  HAMD_HIP(hipMemcpyAsync(rho_dev, P[SL_Y].p, bytes, hipMemcpyDeviceToDevice, st));
  HAMD_HIP(hipMemcpyAsync(h_ctrl, d_ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
  HAMD_HIP(hipStreamSynchronize(st));
  // only intervals of positive length touch the step size (as in solve)
  for (int i = 0; i + 1 < n_t; ++i)
    if (t_list[i + 1] > t_list[i]) {
      dt_save = h_ctrl->dt;
      break;
    }
  stats.n_rhs += 7 * h_ctrl->attempts;
  stats.n_attempts += h_ctrl->attempts;
  stats.n_accepted += h_ctrl->accepted;
  stats.n_rejected += h_ctrl->rejected;
  stats.last_err = h_ctrl->err;
  stats.last_dt = dt_save;
  stats.ado_stage_updates = stats.n_rhs * q->hs.n_hrchy;
  stats.gave_up |= h_ctrl->gave_up;
  return h_ctrl->gave_up ? HEOM_ERR_STEP_REJECTED : HEOM_OK;
}

int Solver::solve_noalloc(void* rho_dev, const double* t_list, int n_t, double dt, double atol,
                          double rtol, Ctrl* ctrl, double* partials, void* rho_s_dev) {
  check_qme();
  if (q->closure) throw HeomError{HEOM_ERR_UNSUPPORTED, "solve_noalloc: a closure is bound"};
  const bool batch = q->kind == kBatch;
  const size_t rs_bytes = sizeof(z128) * (batch ? q->batch_rs_size : q->N2);
  hipStream_t st = q->stream;
  const bool graphs = use_graphs;
  use_graphs = false;
  noalloc = true;
  ext_ctrl = ctrl;
  ext_partials = partials;
  int rc = HEOM_OK;
  try {
    rc = solve(rho_dev, t_list, n_t, dt, atol, rtol, [&](int i, const void* cur) {
      char* dst = static_cast<char*>(rho_s_dev) + rs_bytes * i;
      if (batch)
        launch_gather_rho_s(q->dev, dst, cur, st);
      else
        HAMD_HIP(hipMemcpyAsync(dst, cur, rs_bytes, hipMemcpyDeviceToDevice, st));
    });
  } catch (...) {
    use_graphs = graphs;
    noalloc = false;
    ext_ctrl = nullptr;
    ext_partials = nullptr;
    throw;
  }
  use_graphs = graphs;
  noalloc = false;
  ext_ctrl = nullptr;
  ext_partials = nullptr;
  return rc;
}

void Solver::time_stages(int reps, std::vector<double>& out) {
  check_qme();
  HAMD_HIP(hipSetDevice(q->opt.device));
  hipStream_t st = q->stream;
  const int ns = n_stages();
  out.assign(ns, 0.0);
  if (kind == HEOM_RKDP) {
    // a control block that never finishes; the last stage only reduces the norm
    Ctrl c{};
    c.t = 0.0;
    c.t_bound = 1e300;
    c.dt = dt_save > 0.0 ? dt_save : 1e-3;
    c.cur = h_ctrl ? h_ctrl->cur : 0;
    std::memcpy(h_ctrl, &c, sizeof(Ctrl));
    HAMD_HIP(hipMemcpyAsync(d_ctrl.p, h_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, st));
  }
  hipEvent_t e0, e1;
  HAMD_HIP(hipEventCreate(&e0));
  HAMD_HIP(hipEventCreate(&e1));
  for (int s = 0; s < ns; ++s) {
    StageArgs a = stages[s];
    a.control = 0;
    a.a_re = 1e-3;
    launch_operator(q->dev, in_bytes, a, st);  // warm-up
    HAMD_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch_operator(q->dev, in_bytes, a, st);
    HAMD_HIP(hipEventRecord(e1, st));
    HAMD_HIP(hipEventSynchronize(e1));
    float msf = 0.f;
    HAMD_HIP(hipEventElapsedTime(&msf, e0, e1));
    out[s] = static_cast<double>(msf) / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

// Algorithmic HBM bytes of each stage pass: every state-sized buffer the
// pass reads or writes counts once per element (neighbour reads of the stage
// input are another ADO's compulsory read), plus the operator tables.
void Solver::stage_bytes(std::vector<double>& out) {
  const int ns = n_stages();
  out.assign(ns, 0.0);
  auto size_of = [&](int slot) { return slo[slot] ? 8.0 : 16.0; };
  double tables = 0.0;
  if (q->kind == HEOM_ADO || q->kind == HEOM_REDFIELD_LIOU) {
    tables = 20.0 * q->ado_nnz + 8.0 * (ms + 1);
  } else if (q->kind == HEOM_REDFIELD_HILB) {
    // H, and V_u, Lambda_u, Lambda_u^+ per noise
    tables = 16.0 * q->N2 * (1.0 + 3.0 * static_cast<double>(q->noise.size()));
  } else {
    // the fast Hilbert kernel (diagonal V, 6 <= N <= 16) reads 8-byte links and
    // the coefficient table, the other diagonal-V Hilbert kernels LinkFlat
    const bool fast = q->kind == HEOM_HILB && q->diag_v && q->N >= 6 && q->N <= 16;
    const double lrec = q->kind == kBatch                  ? sizeof(BatchLink)
                        : fast                              ? sizeof(Link)
                        : (q->kind == HEOM_HILB && q->diag_v) ? sizeof(LinkFlat)
                                                               : sizeof(Link);
    tables = lrec * q->n_links + 4.0 * (q->hs.n_hrchy + 1) + 16.0 * q->hs.n_hrchy;
    if (q->kind == kBatch) tables += sizeof(BatchCoef) * static_cast<double>(q->batch_n_coef);
    if (fast) tables += sizeof(Coef) * static_cast<double>(q->coef.size());
  }
  for (int s = 0; s < ns; ++s) {
    const StageSpec& a = specs[s];
    std::set<int> rd, wr;
    rd.insert(a.in_slot);
    for (int j : a.pre)
      if (j >= 0) rd.insert(j);
    if (a.k_slot >= 0) wr.insert(a.k_slot);
    for (int c = 0; c < a.ncombo; ++c) {
      wr.insert(a.dst[c]);
      if (a.dst2[c] >= 0) wr.insert(a.dst2[c]);
    }
    double per = 0.0;
    for (int r : rd) per += size_of(r);
    for (int w : wr) per += size_of(w);
    out[s] = per * static_cast<double>(ms) + tables;
  }
}

}  // namespace hamd
