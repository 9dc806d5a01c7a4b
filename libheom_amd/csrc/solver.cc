// RK4 (rk4.h:40-65) and Dormand-Prince 5(4) (rkdp.h:21-179) drivers.  Every
// stage is one fused operator pass (kernels.hip); the Dormand-Prince step
// control runs on the device, so the host only reads the control block back
// once per batch of attempts instead of once per attempt.
#include "solver.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <set>

#include "kernels.h"

namespace hamd {

namespace {
// Dormand-Prince tableau (rkdp.h:21-57)
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

// slot numbers
enum {
  SL_Y = 0, SL_YNEW = 1, SL_Y32 = 2, SL_YNEW32 = 3,
  SL_SA = 4, SL_SB = 5, SL_ACC = 6, SL_K1 = 7  // K1..K6 = 7..12
};

StageArgs base_args() {
  StageArgs a;
  std::memset(&a, 0, sizeof(a));
  a.mode = MODE_STAGE;
  a.k_slot = -1;
  a.out_slot = -1;
  a.combo[0].dst2 = -1;
  a.combo[1].dst2 = -1;
  return a;
}

void add_src(Combo& c, int slot, double coef) {
  c.src[c.nsrc] = slot;
  c.c[c.nsrc] = coef;
  ++c.nsrc;
}
}  // namespace

Solver::Solver(Qme* q_, int kind_) : q(q_), kind(kind_) {
  if (!q->ready) throw HeomError{HEOM_ERR_STATE, "solver: set_param has not been called"};
  if (kind != HEOM_RK4 && kind != HEOM_RKDP)
    throw HeomError{HEOM_ERR_ARG, "solver: unknown solver kind"};
  HAMD_HIP(hipSetDevice(q->opt.device));
  in_bytes = q->dtype == HEOM_C64 ? 4 : 8;
  ms = q->main_size();
  const size_t z128b = sizeof(z128) * ms, sinb = (in_bytes == 4 ? sizeof(z64) : sizeof(z128)) * ms;
  const int tin = in_bytes == 4 ? ST_C64 : ST_C128;
  for (int s = 0; s < kMaxSlots; ++s) {
    slots.p[s] = nullptr;
    slots.type[s] = ST_C128;
  }
  auto mk = [&](int s, size_t b, int t) {
    P[s].alloc(b);
    HAMD_HIP(hipMemset(P[s].p, 0, b));
    slots.p[s] = P[s].p;
    slots.type[s] = t;
  };
  mk(SL_Y, z128b, ST_C128);
  if (kind == HEOM_RKDP) mk(SL_YNEW, z128b, ST_C128);
  if (in_bytes == 4) {
    mk(SL_Y32, sinb, tin);
    if (kind == HEOM_RKDP) mk(SL_YNEW32, sinb, tin);
  }
  mk(SL_SA, sinb, tin);
  mk(SL_SB, sinb, tin);
  if (kind == HEOM_RK4) {
    mk(SL_ACC, z128b, ST_C128);
  } else {
    for (int k = 0; k < 6; ++k) mk(SL_K1 + k, sinb, tin);
  }
  d_ctrl.alloc(sizeof(Ctrl));
  HAMD_HIP(hipMemset(d_ctrl.p, 0, sizeof(Ctrl)));
  d_partials.alloc(sizeof(double) * kMaxGrid);
  d_counter.alloc(sizeof(unsigned) * 4);
  HAMD_HIP(hipMemset(d_counter.p, 0, sizeof(unsigned) * 4));
  HAMD_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_ctrl), sizeof(Ctrl), hipHostMallocDefault));
  std::memset(h_ctrl, 0, sizeof(Ctrl));
  build_stages(0.0, 0.0);
}

Solver::~Solver() {
  if (h_ctrl) (void)hipHostFree(h_ctrl);
}

void Solver::build_stages(double atol, double rtol) {
  stages.clear();
  const bool lo = in_bytes == 4;
  const int y_in = lo ? SL_Y32 : SL_Y;
  if (kind == HEOM_RK4) {
    // rk4.h:47-64 with k2/k3 folded into the accumulator as they appear
    StageArgs a = base_args();
    a.in_slot = y_in;
    a.ncombo = 2;
    add_src(a.combo[0], SL_Y, 1.0);
    a.combo[0].ck = 0.5;
    a.combo[0].dst = SL_SA;
    add_src(a.combo[1], SL_Y, 1.0);
    a.combo[1].ck = 1.0 / 6.0;
    a.combo[1].dst = SL_ACC;
    stages.push_back(a);

    a = base_args();
    a.in_slot = SL_SA;
    a.ncombo = 2;
    add_src(a.combo[0], SL_Y, 1.0);
    a.combo[0].ck = 0.5;
    a.combo[0].dst = SL_SB;
    add_src(a.combo[1], SL_ACC, 1.0);
    a.combo[1].ck = 2.0 / 6.0;
    a.combo[1].dst = SL_ACC;
    stages.push_back(a);

    a = base_args();
    a.in_slot = SL_SB;
    a.ncombo = 2;
    add_src(a.combo[0], SL_Y, 1.0);
    a.combo[0].ck = 1.0;
    a.combo[0].dst = SL_SA;
    add_src(a.combo[1], SL_ACC, 1.0);
    a.combo[1].ck = 2.0 / 6.0;
    a.combo[1].dst = SL_ACC;
    stages.push_back(a);

    a = base_args();
    a.in_slot = SL_SA;
    a.ncombo = 1;
    add_src(a.combo[0], SL_ACC, 1.0);
    a.combo[0].ck = 1.0 / 6.0;
    a.combo[0].dst = SL_Y;
    a.combo[0].dst2 = lo ? SL_Y32 : -1;
    stages.push_back(a);
    return;
  }
  // Dormand-Prince: stage i (0-based) evaluates k_i at its input and writes
  // the input of stage i+1 (rkdp.h:111-127); stage 5 also forms the new state
  // (rkdp.h:130-132); stage 6 evaluates k_7 at the new state and the error
  // norm (rkdp.h:136-153).
  const int sin_[6] = {y_in, SL_SA, SL_SB, SL_SA, SL_SB, SL_SA};
  const int sout[5] = {SL_SA, SL_SB, SL_SA, SL_SB, SL_SA};
  for (int i = 0; i < 6; ++i) {
    StageArgs a = base_args();
    a.in_slot = sin_[i];
    a.k_slot = SL_K1 + i;
    a.ncombo = 1;
    Combo& c = a.combo[0];
    add_src(c, SL_Y, 1.0);
    if (i < 5) {
      for (int j = 0; j < i; ++j) add_src(c, SL_K1 + j, DP_A[i + 1][j]);
      c.ck = DP_A[i + 1][i];
      c.dst = sout[i];
    } else {
      for (int j = 0; j < 5; ++j)
        if (DP_B[j] != 0.0) add_src(c, SL_K1 + j, DP_B[j]);
      c.ck = DP_B[5];
      c.dst = SL_YNEW;
      c.dst2 = lo ? SL_YNEW32 : -1;
    }
    stages.push_back(a);
  }
  StageArgs a = base_args();
  a.in_slot = lo ? SL_YNEW32 : SL_YNEW;
  a.do_err = 1;
  for (int j = 0; j < 6; ++j)
    if (DP_E[j] != 0.0) {
      a.esrc[a.ne] = SL_K1 + j;
      a.ec[a.ne] = DP_E[j];
      ++a.ne;
    }
  a.ek = DP_E[6];
  a.ynew_slot = SL_YNEW;
  a.yold_slot = SL_Y;
  a.atol = atol;
  a.rtol = rtol;
  a.control = 1;
  a.inv_n = 1.0 / static_cast<double>(ms);
  stages.push_back(a);
  for (auto& s : stages) {
    s.ctrl = d_ctrl.as<Ctrl>();
    s.partials = d_partials.as<double>();
    s.counter = d_counter.as<unsigned>();
    s.max_partials = kMaxGrid;
  }
}

void Solver::launch_stage(int s, double dt) {
  StageArgs a = stages[s];
  a.a_re = dt;
  launch_operator(q->dev, in_bytes, slots, a, q->stream);
}

// fixed_step_size_solver::solve_1 (fixed_step_size_solver.h:29-52): steps of
// dt, the last one shortened to end exactly at t1.
void Solver::rk4_interval(double t0, double t1, double dt) {
  double t = t0, dt_1 = dt;
  bool brk = false;
  while (true) {
    if (t + dt_1 - t1 > -std::numeric_limits<double>::epsilon()) {
      dt_1 = t1 - t;
      brk = true;
    }
    for (int s = 0; s < 4; ++s) launch_stage(s, dt_1);
    stats.n_rhs += 4;
    stats.n_attempts += 1;
    stats.n_accepted += 1;
    t += dt_1;
    if (brk) break;
  }
}

// adaptive_step_size_solver::solve_1 (adaptive_step_size_solver.h:32-56) with
// the attempts of rkdp::solve_adaptive_step launched in batches.
int Solver::rkdp_interval(double t0, double t1, double dt_first, int& batch) {
  launch_ctrl_init(d_ctrl.as<Ctrl>(), t0, t1, dt_first, q->stream);
  long long a0 = h_ctrl->attempts;
  while (true) {
    for (int b = 0; b < batch; ++b)
      for (int s = 0; s < 7; ++s) launch_stage(s, 0.0);
    HAMD_HIP(hipMemcpyAsync(h_ctrl, d_ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, q->stream));
    HAMD_HIP(hipStreamSynchronize(q->stream));
    if (h_ctrl->done) break;
    batch = std::max(1, batch * 2);
  }
  long long used = h_ctrl->attempts - a0;
  batch = static_cast<int>(std::max<long long>(1, std::min<long long>(used + 1, 4096)));
  return h_ctrl->gave_up;
}

int Solver::solve(void* rho_dev, const double* t_list, int n_t, double dt, double atol,
                  double rtol, const std::function<void(int, const void*)>& on_output) {
  if (n_t < 1) throw HeomError{HEOM_ERR_ARG, "solve: n_t < 1"};
  for (int i = 0; i + 1 < n_t; ++i)
    if (!(t_list[i + 1] >= t_list[i]))
      throw HeomError{HEOM_ERR_ARG, "solve: t_list must be non-decreasing"};
  if (kind == HEOM_RK4 && !(dt > 0.0)) throw HeomError{HEOM_ERR_ARG, "rk4: dt must be > 0"};
  if (kind == HEOM_RKDP) {
    if (!(dt > 0.0) && dt_save <= 0.0) throw HeomError{HEOM_ERR_ARG, "rkdp: dt must be > 0"};
    if (!(atol >= 0.0) || !(rtol >= 0.0) || (atol == 0.0 && rtol == 0.0))
      throw HeomError{HEOM_ERR_ARG, "rkdp: atol/rtol must be >= 0 and not both 0"};
  }
  HAMD_HIP(hipSetDevice(q->opt.device));
  build_stages(atol, rtol);
  const size_t bytes = sizeof(z128) * ms;
  hipStream_t st = q->stream;
  // the caller's state enters the FP64 state slot
  HAMD_HIP(hipMemcpyAsync(P[SL_Y].p, rho_dev, bytes, hipMemcpyDeviceToDevice, st));
  if (in_bytes == 4) launch_narrow(P[SL_Y32].p, P[SL_Y].p, ms, st);
  int gave_up = 0;
  if (kind == HEOM_RK4) {
    for (int i = 0; i < n_t - 1; ++i) {
      on_output(i, P[SL_Y].p);
      rk4_interval(t_list[i], t_list[i + 1], dt);
    }
    on_output(n_t - 1, P[SL_Y].p);
    HAMD_HIP(hipMemcpyAsync(rho_dev, P[SL_Y].p, bytes, hipMemcpyDeviceToDevice, st));
  } else {
    std::memset(h_ctrl, 0, sizeof(Ctrl));
    HAMD_HIP(hipMemcpyAsync(d_ctrl.p, h_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, st));
    HAMD_HIP(hipStreamSynchronize(st));
    double first = dt_save > 0.0 ? dt_save : dt;
    int batch = 4;
    int cur = 0;
    for (int i = 0; i < n_t - 1; ++i) {
      on_output(i, P[cur].p);
      gave_up |= rkdp_interval(t_list[i], t_list[i + 1], i == 0 ? first : -1.0, batch);
      cur = h_ctrl->cur;
    }
    on_output(n_t - 1, P[cur].p);
    dt_save = h_ctrl->dt;
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

void Solver::time_stages(int reps, std::vector<double>& out) {
  HAMD_HIP(hipSetDevice(q->opt.device));
  hipStream_t st = q->stream;
  const int ns = n_stages();
  out.assign(ns, 0.0);
  if (kind == HEOM_RKDP) {
    // a control block that never finishes; stage 7 only reduces the norm
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
    launch_operator(q->dev, in_bytes, slots, a, st);  // warm-up
    HAMD_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch_operator(q->dev, in_bytes, slots, a, st);
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
// input are another ADO's compulsory read), plus the per-ADO link tables.
void Solver::stage_bytes(std::vector<double>& out) {
  const int ns = n_stages();
  out.assign(ns, 0.0);
  auto size_of = [&](int slot) { return slots.type[slot] == ST_C128 ? 16.0 : 8.0; };
  const double tables = 8.0 * q->links.size() + 4.0 * (q->hs.n_hrchy + 1) +
                        16.0 * q->hs.n_hrchy + sizeof(Coef) * q->coef.size();
  for (int s = 0; s < ns; ++s) {
    const StageArgs& a = stages[s];
    std::set<int> rd, wr;
    rd.insert(a.in_slot);
    if (a.k_slot >= 0) wr.insert(a.k_slot);
    for (int c = 0; c < a.ncombo; ++c) {
      for (int j = 0; j < a.combo[c].nsrc; ++j) rd.insert(a.combo[c].src[j]);
      wr.insert(a.combo[c].dst);
      if (a.combo[c].dst2 >= 0) wr.insert(a.combo[c].dst2);
    }
    if (a.do_err) {
      for (int j = 0; j < a.ne; ++j) rd.insert(a.esrc[j]);
      rd.insert(a.ynew_slot);
      rd.insert(a.yold_slot);
    }
    double per = 0.0;
    for (int r : rd) per += size_of(r);
    for (int w : wr) per += size_of(w);
    out[s] = per * static_cast<double>(ms) + (q->kind == HEOM_ADO ? 0.0 : tables);
  }
}

}  // namespace hamd
