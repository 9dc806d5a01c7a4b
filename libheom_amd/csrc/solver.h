// Runge-Kutta drivers with device-resident state (solver/*.h of the reference).
#pragma once

#include <functional>
#include <vector>

#include "qme.h"

namespace hamd {

class Solver {
 public:
  Solver(Qme* q, int kind);
  ~Solver();

  Qme* q;
  int kind;        // HEOM_RK4 / HEOM_RKDP
  int in_bytes;    // real size of the operator arithmetic and stage inputs (8 or 4)
  long long ms;    // main_size
  double dt_save = -1.0;
  heom_stats stats{};

  // Propagate the FP64 device state rho_dev over t_list.  on_output(i, cur_ptr)
  // is called at every t_i with a device pointer to the FP64 state.
  int solve(void* rho_dev, const double* t_list, int n_t, double dt, double atol, double rtol,
            const std::function<void(int, const void*)>& on_output);

  // Average duration (ms) of each stage kernel over `reps` replays.
  void time_stages(int reps, std::vector<double>& ms_out);
  void stage_bytes(std::vector<double>& bytes_out);
  int n_stages() const { return kind == HEOM_RK4 ? 4 : 7; }

 private:
  DevBuf P[kMaxSlots];
  Slots slots{};
  DevBuf d_ctrl, d_partials, d_counter;
  Ctrl* h_ctrl = nullptr;  // pinned mirror
  std::vector<StageArgs> stages;

  void build_stages(double atol, double rtol);
  void launch_stage(int s, double dt);
  void rk4_interval(double t0, double t1, double dt);
  int rkdp_interval(double t0, double t1, double dt_first, int& batch);
};

}  // namespace hamd
