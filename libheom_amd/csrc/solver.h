// Runge-Kutta drivers with device-resident state (solver/*.h of the reference).
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <vector>

#include "qme.h"

namespace hamd {

// Slot-level description of one stage pass; resolved into a StageArgs with
// device pointers for both ping-pong parities.
struct StageSpec {
  int in_slot = 0;
  int k_slot = -1;
  std::vector<int> pre;  // operand slots, in summation order
  int ncombo = 0;
  double cc[2][kMaxPre] = {};
  double ck[2] = {};
  int dst[2] = {-1, -1};
  int dst2[2] = {-1, -1};
  int do_err = 0;
  double ec[kMaxPre] = {};
  double ek = 0.0;
  int ynew_pre = -1, yold_pre = -1;
  int control = 0;
  int pre_index(int slot);
  void combo(const std::vector<std::pair<int, double>>& terms, double ckv, int d, int d2 = -1);
};

// An instantiated HIP graph owned by the solver.
struct GraphExec {
  hipGraphExec_t exec = nullptr;
  GraphExec() = default;
  GraphExec(const GraphExec&) = delete;
  GraphExec& operator=(const GraphExec&) = delete;
  ~GraphExec() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
};

class Solver {
 public:
  Solver(std::shared_ptr<Qme> q, int kind);
  ~Solver();

  std::shared_ptr<Qme> q_owner;  // keeps the QME alive while the solver exists
  Qme* q;
  int kind;        // HEOM_RK4 / HEOM_RKDP / HEOM_LSRK4
  int in_bytes;    // real size of the operator arithmetic and stage inputs (8 or 4)
  long long ms;    // main_size
  double dt_save = -1.0;
  heom_stats stats{};
  // replay fixed stage sequences from HIP graphs (RK4 steps of the full dt,
  // Dormand-Prince attempts); 0 launches every kernel individually
  bool use_graphs = true;

  // Propagate the FP64 device state rho_dev over t_list.  on_output(i, cur_ptr)
  // is called at every t_i with a device pointer to the FP64 state.
  int solve(void* rho_dev, const double* t_list, int n_t, double dt, double atol, double rtol,
            const std::function<void(int, const void*)>& on_output);
  // The same with rho_S of every output time gathered into rho_s_dev (n_t N^2
  // complex128 in device memory) instead of a callback; a small fixed-step
  // problem then runs the whole solve in one resident launch.
  int solve_gather(void* rho_dev, const double* t_list, int n_t, double dt, double atol,
                   double rtol, void* rho_s_dev);

  // solve_gather without any device allocation, for callers that cannot
  // allocate: every stage kernel is launched directly (no graphs, no resident
  // integrator), the control block and the error-norm partials are the
  // caller's (ctrl, partials: kMaxGrid doubles), and rho_S of every output time
  // is copied to rho_s_dev.  No closure may be bound.
  int solve_noalloc(void* rho_dev, const double* t_list, int n_t, double dt, double atol,
                    double rtol, Ctrl* ctrl, double* partials, void* rho_s_dev);

  // Average duration (ms) of each stage kernel over `reps` replays.
  void time_stages(int reps, std::vector<double>& ms_out);
  // throws HEOM_ERR_STATE unless the QME is set up with the size the solver's
  // buffers were allocated for (set_param may have been called again)
  void check_qme() const;
  void stage_bytes(std::vector<double>& bytes_out);
  int n_stages() const { return kind == HEOM_RKDP ? 7 : 4; }

 private:
  static constexpr int kSlots = 16;
  // slot table: slots 0..3 are ping-pong pairs (see common.h), resolved per
  // parity when a stage's pointers are built
  DevBuf P[kSlots];
  void* sp[kSlots] = {};
  int slo[kSlots] = {};  // 1: complex64 buffer
  DevBuf d_ctrl, d_partials;
  // control block and partials of the current solve (the solver's own, or a
  // caller's in solve_noalloc); noalloc: no graph, upload or resident launch
  Ctrl* ext_ctrl = nullptr;
  double* ext_partials = nullptr;
  bool noalloc = false;
  Ctrl* ctrl_dev() const { return ext_ctrl ? ext_ctrl : d_ctrl.as<Ctrl>(); }
  double* partials_dev() const { return ext_partials ? ext_partials : d_partials.as<double>(); }
  DevBuf d_tr0;  // Tr rho_0 at the start of the solve (renorm_rho0)
  void renorm(bool with_ctrl);
  Ctrl* h_ctrl = nullptr;  // pinned mirror
  std::vector<StageArgs> stages;

  void build_stages(double atol, double rtol);
  void* slot_ptr(int s, int cur) const;
  StageArgs resolve(const StageSpec& s) const;
  std::vector<StageSpec> specs;
  void launch_stage(int s, double dt);
  // graphs of 2^j full RK4 steps / Dormand-Prince attempts (j <= 8), by 2^j;
  // a count is replayed as its binary decomposition
  static constexpr int kMaxGraphSteps = 256;
  // most Dormand-Prince attempts enqueued before the host reads the control block
  static constexpr int kMaxAttemptBatch = 1 << 14;
  static constexpr size_t kMaxGraphs = 9;
  std::map<long long, std::unique_ptr<GraphExec>> g_rk4, g_dopri;
  double g_rk4_dt = 0.0;
  GraphExec& graph_for(std::map<long long, std::unique_ptr<GraphExec>>& cache, long long n,
                       const std::function<void()>& one);
  void replay(std::map<long long, std::unique_ptr<GraphExec>>& cache, long long count,
              const std::function<void()>& one);
  template <typename F>
  void capture(GraphExec& g, F&& body);
  void reset_graphs();
  void rk4_interval(double t0, double t1, double dt);
  // resident integrator (small heom_hilb problems): stage program built by
  // build_stages when the problem qualifies, else n_stages == 0
  ResidentArgs res{};
  DevBuf d_res_prog, d_res_nf, d_res_dl, d_res_t;
  int solve_resident_dopri(void* rho_dev, const double* t_list, int n_t, double dt, double atol,
                           double rtol, void* rho_s_dev);
  // fixed_step_size_solver::solve_1's step sequence over [t0, t1]: n_full steps
  // of dt, then one step of dt_last
  static void fixed_steps(double t0, double t1, double dt, long long& n_full, double& dt_last);
  void build_resident();
  int rkdp_interval(double t0, double t1, double dt_first, int& batch);
};

}  // namespace hamd
