// Host-side QME model: parameters, hierarchy, link/coefficient tables and the
// HBM-resident copies used by the kernels.
#pragma once

#include <map>
#include <memory>
#include <vector>

#include "../../include/libheom_amd.h"
#include "common.h"

namespace hamd {

struct NoiseParams {
  std::vector<z128> V;  // N x N
  int len = 0;
  std::vector<z128> gamma, phi0, sigma, S, A;  // len x len / len
  z128 s_delta{0.0, 0.0};
  bool set = false;
};

// Device allocation owned by a QME or a solver.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void alloc(size_t b);
  template <typename T>
  void upload(const std::vector<T>& v) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) HAMD_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Hierarchy {
  int n_dim = 0, n_hrchy = 0;
  std::vector<int> n;       // n_hrchy x n_dim
  std::vector<int> ptr_p1;  // (n_hrchy + 1) x n_dim, sentinel n_hrchy
  std::vector<int> ptr_m1;
};

// Breadth-first enumeration with the reference's visiting rule
// (hrchy_space.cc:137-231): pop an index, then push index + e_k for every
// k >= the mode last incremented, while the total depth stays <= max_depth.
// filter (hrchy_space.h:29-36): an index is enumerated only if it returns
// non-zero (hrchy_space.cc:183-187); null: every index up to max_depth.
Hierarchy build_hierarchy(int n_dim, int max_depth, heom_hrchy_filter filter = nullptr,
                          void* user = nullptr);

struct Closure;

class Qme {
 public:
  Qme(int kind, int dtype, int n_level, int max_depth, const heom_options& opt);
  ~Qme();

  int kind, dtype, N, N2, max_depth;
  // FP32 operator / stage-input storage (HEOM_C64 and HEOM_C64_BF16X3)
  bool lo() const { return dtype != HEOM_C128; }
  heom_options opt;
  std::vector<z128> H;
  bool H_set = false;
  std::vector<NoiseParams> noise;
  bool ready = false;
  heom_hrchy_filter hrchy_filter = nullptr;
  void* hrchy_filter_user = nullptr;

  // heom::set_param products
  Hierarchy hs;
  std::vector<std::vector<int>> lk;
  std::vector<z128> ngamma;               // Gamma_n
  std::vector<std::vector<z128>> s, a;    // per noise, per mode
  bool diag_v = false;
  bool real_v = false;  // diagonal V with real entries

  // neighbour links
  std::vector<int> link_ptr;
  std::vector<Link> links;
  std::vector<Coef> coef;

  // device copies
  FormDev dev{};
  hipStream_t stream = nullptr;
  DevBuf d_link_ptr, d_links, d_flat, d_cmac, d_coef, d_gam, d_H, d_V, d_Vd, d_dterm, d_sdelta;
  DevBuf d_bmap;
  DevBuf d_sop_ptr, d_sop_col, d_sop_val;
  DevBuf d_R_ptr, d_R_col, d_R_val;
  long long ado_nnz = 0;

  void set_param();
  bool redfield() const { return kind == HEOM_REDFIELD_HILB || kind == HEOM_REDFIELD_LIOU; }
  // Redfield set-up products: eigenvalues of H (ascending order not implied)
  // and Lambda_u (logical N x N row-major, before any col_major transform)
  std::vector<double> rf_eig;
  std::vector<std::vector<z128>> rf_lambda;
  DevBuf d_rf_L, d_rf_Ld, d_rf_T, d_rf_HL, d_rf_HR;
  // learned truncation closure (mlp.h), null when none is bound
  std::unique_ptr<Closure> closure;
  void clear_closure();
  // arguments of one derivative pass drho = beta*drho - alpha*D(rho)
  StageArgs diff_args(void* drho, const void* rho, z128 alpha, z128 beta) const;
  DevBuf scratch;  // closure training rows: one state of this QME
  // closure forward (when bound) followed by the operator pass
  int operate(int real_bytes, const StageArgs& A, hipStream_t st);
  // pack already set-up Hilbert-form members (diagonal V, n_level <= 8) into
  // one batch whose state is the concatenation of the members' states
  void set_param_batch(const std::vector<Qme*>& members);
  long long main_size() const {
    return kind == kBatch ? batch_main_size : static_cast<long long>(hs.n_hrchy) * N2;
  }

  // host copies kept for packing into a batch
  std::vector<LinkFlat> flat_host;
  std::vector<z128> cmac_host, dterm_host, H_host;
  long long n_links = 0;

  // batch
  std::vector<BatchInst> batch_inst;
  long long batch_main_size = 0;
  int batch_rs_size = 0;  // sum of the members' N^2
  DevBuf d_inst, d_items, d_blink, d_bcoef;
  long long batch_n_coef = 0;  // entries of the members' coefficient tables
  DevBuf d_Hre, d_Him, d_Hre32, d_Him32;
  DevBuf d_Vre, d_Vim, d_Vre32, d_Vim32;  // dense V planes of the matrix-core path
  struct TuneEntry {
    int ipw, wpb, grid;
    float us;  // mean microseconds per stage pass
  };
  std::vector<TuneEntry> tune_log;  // the batch autotuner's candidates, in the order tried
  void autotune_batch();

 private:
  // parameters after the col_major transform (identical for row_major)
  std::vector<NoiseParams> noise_work_;
  std::vector<z128> H_work_;
  std::vector<int> sop_ptr_, sop_col_;
  std::vector<z128> sop_val_;
  std::vector<long long> R_ptr_;
  std::vector<int> R_col_;
  std::vector<z128> R_val_;

  void build_links();
  // Work-unit order of the one-ADO-per-workgroup kernels (see BlockMap)
  std::vector<BlockMap> build_block_map() const;
  void set_param_redfield();
  void build_superops(std::vector<int>& ptr, std::vector<int>& col, std::vector<z128>& val,
                      std::vector<std::vector<std::map<int, z128>>>& ops);
  void build_ado(const std::vector<std::vector<std::map<int, z128>>>& ops);
  void upload();
};

}  // namespace hamd
