// Host side of the QME: hierarchy enumeration (hrchy_space.cc:137-231),
// heom::set_param (heom.h:57-139), the per-form set_param
// (heom_hilb.h:57-68, heom_liou.h:66-119, heom_ado.h:62-189) and the upload
// of every table to HBM.
#include "qme.h"
#include "kernels.h"
#include "mlp.h"
#include "supop.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <limits>
#include <numeric>
#include <string>
#include <tuple>
#include <unordered_map>

namespace hamd {

DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

void DevBuf::alloc(size_t b) {
  if (p) {
    HAMD_HIP(hipFree(p));
    p = nullptr;
  }
  bytes = b;
  if (b == 0) return;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    throw HeomError{HEOM_ERR_NOMEM, "hipMalloc of " + std::to_string(b) + " bytes failed: " +
                                        hipGetErrorString(e)};
  }
}

// ---------------------------------------------------------------------------
// hierarchy
// ---------------------------------------------------------------------------
namespace {
struct KeyHash {
  size_t operator()(const std::string& s) const { return std::hash<std::string>()(s); }
};
inline std::string key_of(const int* v, int K) {
  std::string s(static_cast<size_t>(K) * 2, '\0');
  for (int k = 0; k < K; ++k) {
    s[2 * k] = static_cast<char>(v[k] & 0xff);
    s[2 * k + 1] = static_cast<char>((v[k] >> 8) & 0xff);
  }
  return s;
}
}  // namespace

Hierarchy build_hierarchy(int K, int max_depth, heom_hrchy_filter filter, void* user) {
  Hierarchy hs;
  hs.n_dim = K;
  struct Item {
    std::vector<int> idx;
    int last;
    int depth;
  };
  std::deque<Item> queue;
  queue.push_back(Item{std::vector<int>(K, 0), 0, 0});
  std::unordered_map<std::string, int, KeyHash> book;
  while (!queue.empty()) {
    Item it = std::move(queue.front());
    queue.pop_front();
    int l = hs.n_hrchy++;
    book.emplace(key_of(it.idx.data(), K), l);
    hs.n.insert(hs.n.end(), it.idx.begin(), it.idx.end());
    if (it.depth + 1 <= max_depth) {
      for (int k = it.last; k < K; ++k) {
        Item nx{it.idx, k, it.depth + 1};
        ++nx.idx[k];
        if (filter && !filter(nx.idx.data(), K, nx.depth, user)) continue;
        queue.push_back(std::move(nx));
      }
    }
  }
  const int nh = hs.n_hrchy;
  hs.ptr_p1.assign(static_cast<size_t>(nh + 1) * K, nh);
  hs.ptr_m1.assign(static_cast<size_t>(nh + 1) * K, nh);
  std::vector<int> idx(K);
  for (int l = 0; l < nh; ++l) {
    std::memcpy(idx.data(), &hs.n[static_cast<size_t>(l) * K], sizeof(int) * K);
    for (int k = 0; k < K; ++k) {
      ++idx[k];
      auto f = book.find(key_of(idx.data(), K));
      if (f != book.end()) hs.ptr_p1[static_cast<size_t>(l) * K + k] = f->second;
      idx[k] -= 2;
      if (idx[k] >= 0) {
        f = book.find(key_of(idx.data(), K));
        if (f != book.end()) hs.ptr_m1[static_cast<size_t>(l) * K + k] = f->second;
      }
      ++idx[k];
    }
  }
  return hs;
}

// ---------------------------------------------------------------------------
// QME
// ---------------------------------------------------------------------------
Qme::Qme(int kind_, int dtype_, int n_level, int max_depth_, const heom_options& opt_)
    : kind(kind_), dtype(dtype_), N(n_level), N2(n_level * n_level), max_depth(max_depth_),
      opt(opt_) {
  H.assign(N2, z128{0.0, 0.0});
  HAMD_HIP(hipSetDevice(opt.device));
  HAMD_HIP(hipStreamCreate(&stream));
}

Qme::~Qme() {
  if (stream) (void)hipStreamDestroy(stream);
}

using namespace supop;

void Qme::clear_closure() {
  closure.reset();
  dev.clos = nullptr;
  dev.clos_lo = dev.clos_n = 0;
}

StageArgs Qme::diff_args(void* drho, const void* rho, z128 alpha, z128 beta) const {
  const int l = lo() ? 1 : 0;
  StageArgs A;
  std::memset(&A, 0, sizeof(A));
  A.mode = MODE_DIFF;
  A.in[0] = A.in[1] = rho;
  A.in_lo = l;
  A.dst[0][0] = A.dst[0][1] = drho;
  A.dst_lo[0] = l;
  A.ynew_pre = A.yold_pre = -1;
  if (beta.re != 0.0 || beta.im != 0.0) {
    A.npre = 1;
    A.pre[0][0] = A.pre[1][0] = drho;
    A.pre_lo = l ? 1u : 0u;
  }
  A.a_re = alpha.re;
  A.a_im = alpha.im;
  A.b_re = beta.re;
  A.b_im = beta.im;
  return A;
}

int Qme::operate(int real_bytes, const StageArgs& A, hipStream_t st) {
  if (closure) closure->forward(*this, A, st);
  return launch_operator(dev, real_bytes, A, st);
}

void Qme::set_param() {
  clear_closure();
  const int nn = static_cast<int>(noise.size());
  if (!H_set) throw HeomError{HEOM_ERR_ARG, "set_param: H not set"};
  if (nn == 0) throw HeomError{HEOM_ERR_ARG, "set_param: alloc_noises not called"};
  for (int u = 0; u < nn; ++u)
    if (!noise[u].set)
      throw HeomError{HEOM_ERR_ARG, "set_param: noise " + std::to_string(u) + " not set"};
  if (redfield()) {
    set_param_redfield();
    return;
  }

  // col_major storage is the transposed row-major problem with
  // H -> -H^T, V -> -V^T, A -> -A, gamma -> gamma^T (heom_hilb.h:148-162).
  std::vector<z128> Hw = H;
  std::vector<NoiseParams> nz = noise;
  if (opt.order == 1) {
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) Hw[i * N + j] = zscale(-1.0, H[j * N + i]);
    for (auto& z : nz) {
      std::vector<z128> V2(N2);
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) V2[i * N + j] = zscale(-1.0, z.V[j * N + i]);
      z.V = V2;
      for (auto& x : z.A) x = zscale(-1.0, x);
      std::vector<z128> g2(z.gamma.size());
      for (int i = 0; i < z.len; ++i)
        for (int j = 0; j < z.len; ++j) g2[i * z.len + j] = z.gamma[j * z.len + i];
      z.gamma = g2;
    }
  }

  // heom::set_param: lk linearisation (heom.h:66-74)
  lk.assign(nn, {});
  int K = 0;
  for (int u = 0; u < nn; ++u)
    for (int k = 0; k < nz[u].len; ++k) lk[u].push_back(K++);
  if (K == 0) throw HeomError{HEOM_ERR_ARG, "set_param: no bath modes"};
  if (max_depth < 0) throw HeomError{HEOM_ERR_ARG, "set_param: max_depth < 0"};

  hs = build_hierarchy(K, max_depth, hrchy_filter, hrchy_filter_user);
  const int nh = hs.n_hrchy;
  if (static_cast<long long>(nh) * N2 > (1LL << 31) - 1)
    throw HeomError{HEOM_ERR_UNSUPPORTED, "set_param: main_size exceeds 2^31"};

  // Gamma_n (heom.h:93-104)
  ngamma.assign(nh, z128{0.0, 0.0});
  for (int l = 0; l < nh; ++l) {
    z128 g{0.0, 0.0};
    for (int u = 0; u < nn; ++u)
      for (int k = 0; k < nz[u].len; ++k)
        g = g + zscale(static_cast<double>(hs.n[static_cast<size_t>(l) * K + lk[u][k]]),
                       nz[u].gamma[k * nz[u].len + k]);
    ngamma[l] = g;
  }

  // s = S phi_0, a = A phi_0 (heom.h:123-138)
  s.assign(nn, {});
  a.assign(nn, {});
  for (int u = 0; u < nn; ++u) {
    const int len = nz[u].len;
    s[u].assign(len, z128{0.0, 0.0});
    a[u].assign(len, z128{0.0, 0.0});
    for (int i = 0; i < len; ++i)
      for (int j = 0; j < len; ++j) {
        s[u][i] = s[u][i] + zmul(nz[u].S[i * len + j], nz[u].phi0[j]);
        a[u][i] = a[u][i] + zmul(nz[u].A[i * len + j], nz[u].phi0[j]);
      }
  }

  diag_v = true;
  for (int u = 0; u < nn && diag_v; ++u)
    for (int i = 0; i < N && diag_v; ++i)
      for (int j = 0; j < N; ++j)
        if (i != j && !zzero(nz[u].V[i * N + j])) {
          diag_v = false;
          break;
        }

  real_v = diag_v;
  for (int u = 0; u < nn && real_v; ++u)
    for (int i = 0; i < N; ++i)
      if (nz[u].V[i * N + i].im != 0.0) {
        real_v = false;
        break;
      }

  noise_work_ = nz;
  H_work_ = Hw;
  build_links();

  // superoperators for heom_liou / heom_ado (heom_liou.h:66-107)
  std::vector<std::vector<std::map<int, z128>>> ops;
  std::vector<int> sop_ptr, sop_col;
  std::vector<z128> sop_val;
  if (kind == HEOM_LIOU || kind == HEOM_ADO) {
    if (N > 32) throw HeomError{HEOM_ERR_UNSUPPORTED, "heom_liou/heom_ado: n_level > 32"};
    build_superops(sop_ptr, sop_col, sop_val, ops);
  }
  if (kind == HEOM_ADO) build_ado(ops);
  sop_ptr_ = std::move(sop_ptr);
  sop_col_ = std::move(sop_col);
  sop_val_ = std::move(sop_val);
  upload();
  ready = true;
}

void Qme::build_links() {
  const int nn = static_cast<int>(noise_work_.size());
  const int K = hs.n_dim, nh = hs.n_hrchy, pv = nh;
  const bool sq = opt.sqrt_norm != 0;
  // coefficient dedup: (type, mode, n_j, n_k, j, k) -> cid
  std::map<std::tuple<int, int, int, int, int, int>, int> cmap;
  coef.clear();
  auto get = [&](std::tuple<int, int, int, int, int, int> key, const Coef& c) {
    auto f = cmap.find(key);
    if (f != cmap.end()) return f->second;
    int id = static_cast<int>(coef.size());
    coef.push_back(c);
    cmap.emplace(key, id);
    return id;
  };
  link_ptr.assign(nh + 1, 0);
  links.clear();
  for (int l = 0; l < nh; ++l) {
    const int* nl = &hs.n[static_cast<size_t>(l) * K];
    for (int u = 0; u < nn; ++u) {
      const NoiseParams& z = noise_work_[u];
      // +1 terms (heom_hilb.h:170-182)
      for (int k = 0; k < z.len; ++k) {
        int mode = lk[u][k];
        int p = hs.ptr_p1[static_cast<size_t>(l) * K + mode];
        if (p == pv) continue;
        int nk = nl[mode];
        Coef c{};
        c.phi = sq ? zscale(std::sqrt(nk + 1.0), z.sigma[k]) : z.sigma[k];
        c.u = u;
        links.push_back(Link{p, get(std::make_tuple(0, mode, 0, sq ? nk : 0, 0, 0), c)});
      }
      // -1 terms (heom_hilb.h:184-198)
      for (int k = 0; k < z.len; ++k) {
        int mode = lk[u][k];
        int m = hs.ptr_m1[static_cast<size_t>(l) * K + mode];
        if (m == pv) continue;
        int nk = nl[mode];
        double f = sq ? std::sqrt(static_cast<double>(nk)) : static_cast<double>(nk);
        Coef c{};
        c.phi = zscale(f, s[u][k]);
        c.psi = zscale(-f, a[u][k]);
        c.u = u;
        links.push_back(Link{m, get(std::make_tuple(1, mode, 0, nk, 0, 0), c)});
      }
    }
    // off-diagonal gamma (heom_hilb.h:128-147); dropped in parity mode
    if (!opt.parity_q1) {
      for (int u = 0; u < nn; ++u) {
        const NoiseParams& z = noise_work_[u];
        for (int j = 0; j < z.len; ++j)
          for (int k = 0; k < z.len; ++k) {
            z128 v = z.gamma[j * z.len + k];
            if (j == k || zzero(v)) continue;
            int mj = lk[u][j], mk = lk[u][k];
            int q1 = hs.ptr_m1[static_cast<size_t>(l) * K + mj];
            if (q1 == pv) continue;
            int q = hs.ptr_p1[static_cast<size_t>(q1) * K + mk];
            if (q == pv) continue;
            int nj = nl[mj], nk = nl[mk];
            double f = sq ? std::sqrt(nj * (nk + 1.0)) : static_cast<double>(nj);
            Coef c{};
            c.dir = zscale(f, v);
            c.u = nn;
            links.push_back(Link{q, get(std::make_tuple(2, u, nj, nk, j, k), c)});
          }
      }
    }
    link_ptr[l + 1] = static_cast<int>(links.size());
  }
}

// Workgroups are dealt to the eight XCDs round-robin, and the neighbours an
// ADO reads are cheap only when they sit in the same XCD's L2.  ADOs are
// therefore listed in depth-first preorder of the tree in which every ADO hangs
// from the one that generated it in the breadth-first enumeration (n - e_k, k
// its highest occupied mode: hrchy_space.cc:160-198 pushes n + e_k only for k
// at or above the mode last raised), cut into eight consecutive parts of equal
// size (the ADOs and their links then mix the tiers evenly), and each part is
// given back in breadth-first order to workgroups x, x + 8, x + 16, ...
// Subtrees stay together, so most +-1 links join ADOs of one part.
std::vector<BlockMap> Qme::build_block_map() const {
  const int nh = hs.n_hrchy, K = hs.n_dim;
  constexpr int P = 8;
  std::vector<int> order;
  order.reserve(nh);
  std::vector<char> seen(nh, 0);
  std::vector<int> stack{0};
  while (!stack.empty()) {
    const int g = stack.back();
    stack.pop_back();
    if (g < 0 || g >= nh || seen[g]) continue;
    seen[g] = 1;
    order.push_back(g);
    const int* ng = &hs.n[static_cast<size_t>(g) * K];
    int top = 0;
    for (int k = 0; k < K; ++k)
      if (ng[k] > 0) top = k;
    for (int k = K - 1; k >= top; --k) {  // lower modes are visited first
      const int c = hs.ptr_p1[static_cast<size_t>(g) * K + k];
      if (c != nh) stack.push_back(c);
    }
  }
  for (int g = 0; g < nh; ++g)  // anything the tree walk missed keeps its order
    if (!seen[g]) order.push_back(g);
  std::vector<std::vector<int>> part(P);
  size_t pos = 0;
  for (int x = 0; x < P; ++x) {
    const int cnt = (nh - x + P - 1) / P;  // workgroups b < nh with b % 8 == x
    part[x].assign(order.begin() + pos, order.begin() + pos + cnt);
    std::sort(part[x].begin(), part[x].end());
    pos += cnt;
  }
  std::vector<BlockMap> map(nh);
  for (int x = 0; x < P; ++x)
    for (size_t i = 0; i < part[x].size(); ++i) {
      const int g = part[x][i];
      map[i * P + x] = BlockMap{g, link_ptr[g], link_ptr[g + 1], 0};
    }
  return map;
}

void Qme::build_superops(std::vector<int>& ptr, std::vector<int>& col, std::vector<z128>& val,
                         std::vector<std::vector<std::map<int, z128>>>& ops) {
  const int nn = static_cast<int>(noise_work_.size());
  SupOp L(N2);
  kron_x_1(L, kI, H_work_, N);
  kron_1_x_T(L, zscale(-1.0, kI), H_work_, N);
  std::vector<SupOp> Phi(nn, SupOp(N2)), Psi(nn, SupOp(N2)), Xi(nn);
  for (int u = 0; u < nn; ++u) {
    kron_x_1(Phi[u], kI, noise_work_[u].V, N);
    kron_1_x_T(Phi[u], zscale(-1.0, kI), noise_work_[u].V, N);
    optimize(Phi[u]);
    kron_x_1(Psi[u], kOne, noise_work_[u].V, N);
    kron_1_x_T(Psi[u], kOne, noise_work_[u].V, N);
    optimize(Psi[u]);
    Xi[u] = supop_gemm(zscale(-1.0, noise_work_[u].s_delta), Phi[u], Phi[u]);
  }
  SupOp R0(N2);
  supop_axpy(R0, kOne, L);
  for (int u = 0; u < nn; ++u) supop_axpy(R0, kOne, Xi[u]);
  ops.clear();
  ops.push_back(R0);
  for (int u = 0; u < nn; ++u) {
    ops.push_back(Phi[u]);
    ops.push_back(Psi[u]);
  }
  ptr.clear();
  col.clear();
  val.clear();
  for (auto& op : ops) {
    for (int r = 0; r < N2; ++r) {
      ptr.push_back(static_cast<int>(col.size()));
      for (auto& kv : op[r]) {
        col.push_back(kv.first);
        val.push_back(kv.second);
      }
    }
    ptr.push_back(static_cast<int>(col.size()));
  }
}

void Qme::build_ado(const std::vector<std::vector<std::map<int, z128>>>& ops) {
  // heom_ado::set_param (heom_ado.h:62-189)
  const int nn = static_cast<int>(noise_work_.size());
  const int K = hs.n_dim, nh = hs.n_hrchy, pv = nh;
  const bool sq = opt.sqrt_norm != 0;
  const SupOp& R0 = ops[0];
  std::vector<std::vector<SupOp>> Theta(nn);
  for (int u = 0; u < nn; ++u) {
    const SupOp& Phi = ops[1 + 2 * u];
    const SupOp& Psi = ops[2 + 2 * u];
    for (int k = 0; k < noise_work_[u].len; ++k) {
      SupOp T(N2);
      supop_axpy(T, s[u][k], Phi);
      supop_axpy(T, zscale(-1.0, a[u][k]), Psi);
      optimize(T);
      Theta[u].push_back(std::move(T));
    }
  }
  std::vector<long long> rptr;
  std::vector<int> rcol;
  std::vector<z128> rval;
  rptr.reserve(static_cast<size_t>(nh) * N2 + 1);
  rptr.push_back(0);
  std::vector<std::pair<int, z128>> row;
  for (int l = 0; l < nh; ++l) {
    const int* nl = &hs.n[static_cast<size_t>(l) * K];
    for (int a_ = 0; a_ < N2; ++a_) {
      row.clear();
      for (int u = 0; u < nn; ++u)
        for (int k = 0; k < noise_work_[u].len; ++k) {
          int m = hs.ptr_m1[static_cast<size_t>(l) * K + lk[u][k]];
          if (m == pv) continue;
          double nk = nl[lk[u][k]];
          for (auto& kv : Theta[u][k][a_]) {
            z128 v = zscale(sq ? std::sqrt(nk) : nk, kv.second);
            if (!zzero(v)) row.emplace_back(m * N2 + kv.first, v);
          }
        }
      row.emplace_back(l * N2 + a_, ngamma[l]);
      if (!opt.parity_q1) {
        for (int u = 0; u < nn; ++u) {
          const NoiseParams& z = noise_work_[u];
          for (int j = 0; j < z.len; ++j)
            for (int k = 0; k < z.len; ++k) {
              z128 v = z.gamma[j * z.len + k];
              if (j == k || zzero(v)) continue;
              int q1 = hs.ptr_m1[static_cast<size_t>(l) * K + lk[u][j]];
              if (q1 == pv) continue;
              int q = hs.ptr_p1[static_cast<size_t>(q1) * K + lk[u][k]];
              if (q == pv) continue;
              double nj = nl[lk[u][j]], nk = nl[lk[u][k]];
              row.emplace_back(q * N2 + a_, zscale(sq ? std::sqrt(nj * (nk + 1)) : nj, v));
            }
        }
      }
      for (auto& kv : R0[a_]) row.emplace_back(l * N2 + kv.first, kv.second);
      for (int u = 0; u < nn; ++u)
        for (int k = 0; k < noise_work_[u].len; ++k) {
          int p = hs.ptr_p1[static_cast<size_t>(l) * K + lk[u][k]];
          if (p == pv) continue;
          double nk = nl[lk[u][k]];
          for (auto& kv : ops[1 + 2 * u][a_]) {
            z128 v = kv.second;
            if (sq) v = zscale(std::sqrt(nk + 1), v);
            v = zmul(v, noise_work_[u].sigma[k]);
            if (!zzero(v)) row.emplace_back(p * N2 + kv.first, v);
          }
        }
      std::stable_sort(row.begin(), row.end(),
                       [](const std::pair<int, z128>& x, const std::pair<int, z128>& y) {
                         return x.first < y.first;
                       });
      for (size_t i = 0; i < row.size();) {
        size_t j = i;
        z128 acc{0.0, 0.0};
        while (j < row.size() && row[j].first == row[i].first) {
          acc = acc + row[j].second;
          ++j;
        }
        rcol.push_back(row[i].first);
        rval.push_back(acc);
        i = j;
      }
      rptr.push_back(static_cast<long long>(rcol.size()));
    }
  }
  ado_nnz = static_cast<long long>(rcol.size());
  R_ptr_ = std::move(rptr);
  R_col_ = std::move(rcol);
  R_val_ = std::move(rval);
}

void Qme::upload() {
  const int nn = static_cast<int>(noise_work_.size());
  HAMD_HIP(hipSetDevice(opt.device));
  d_link_ptr.upload(link_ptr);
  d_links.upload(links);
  d_bmap.upload(build_block_map());
  std::vector<LinkFlat> flat(links.size());
  for (size_t l = 0; l < links.size(); ++l) {
    const Coef& cf = coef[links[l].cid];
    LinkFlat f{};
    f.nbr = links[l].nbr;
    f.u = cf.u;
    f.iphi = z128{-cf.phi.im, cf.phi.re};  // i*phi
    f.psi = cf.psi;
    f.dir = cf.dir;
    flat[l] = f;
  }
  d_flat.upload(flat);
  flat_host = flat;
  n_links = static_cast<long long>(links.size());
  {
    // (v_i - v_j, v_i + v_j) per noise and element, one zero row for direct links
    std::vector<z128> tab;
    if (diag_v) {
      const size_t rows = static_cast<size_t>(nn + 1) * N2;
      tab.assign(real_v ? rows : 2 * rows, z128{0.0, 0.0});
      for (int u = 0; u < nn; ++u)
        for (int i = 0; i < N; ++i)
          for (int j = 0; j < N; ++j) {
            const z128 vi = noise_work_[u].V[i * N + i], vj = noise_work_[u].V[j * N + j];
            const size_t b = static_cast<size_t>(u) * N2 + i * N + j;
            if (real_v) {
              tab[b] = z128{vi.re - vj.re, vi.re + vj.re};
            } else {
              tab[2 * b] = vi - vj;
              tab[2 * b + 1] = vi + vj;
            }
          }
    }
    d_cmac.upload(tab);
    cmac_host = tab;
  }
  d_coef.upload(coef);
  d_gam.upload(ngamma);
  d_H.upload(H_work_);
  std::vector<z128> V(static_cast<size_t>(nn) * N2), Vd(static_cast<size_t>(nn) * N),
      sd(nn), dterm(N2, z128{0.0, 0.0});
  for (int u = 0; u < nn; ++u) {
    std::copy(noise_work_[u].V.begin(), noise_work_[u].V.end(), V.begin() + u * N2);
    for (int i = 0; i < N; ++i) Vd[u * N + i] = noise_work_[u].V[i * N + i];
    sd[u] = noise_work_[u].s_delta;
  }
  if (diag_v) {
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        z128 acc{0.0, 0.0};
        for (int u = 0; u < nn; ++u) {
          z128 c = Vd[u * N + i] - Vd[u * N + j];
          acc = acc + zmul(noise_work_[u].s_delta, zmul(c, c));
        }
        dterm[i * N + j] = acc;
      }
  }
  d_V.upload(V);
  d_Vd.upload(Vd);
  d_dterm.upload(dterm);
  dterm_host = dterm;
  H_host = H_work_;
  d_sdelta.upload(sd);
  if (!sop_ptr_.empty()) {
    d_sop_ptr.upload(sop_ptr_);
    d_sop_col.upload(sop_col_);
    d_sop_val.upload(sop_val_);
  }
  if (!R_ptr_.empty()) {
    d_R_ptr.upload(R_ptr_);
    d_R_col.upload(R_col_);
    d_R_val.upload(R_val_);
    R_ptr_.clear();
    R_col_.clear();
    R_val_.clear();
  }
  dev.kind = kind;
  dev.N = N;
  dev.N2 = N2;
  dev.n_hrchy = hs.n_hrchy;
  dev.n_noise = nn;
  dev.main_size = main_size();
  dev.link_ptr = d_link_ptr.as<int>();
  dev.links = d_links.as<Link>();
  dev.coef = d_coef.as<Coef>();
  dev.gam = d_gam.as<z128>();
  dev.H = d_H.as<z128>();
  dev.V = d_V.as<z128>();
  dev.Vd = d_Vd.as<z128>();
  dev.dterm = d_dterm.as<z128>();
  dev.sdelta = d_sdelta.as<z128>();
  dev.diag_v = diag_v ? 1 : 0;
  dev.real_v = real_v ? 1 : 0;
  dev.flat = d_flat.as<LinkFlat>();
  dev.bmap = d_bmap.as<BlockMap>();
  dev.cmac_tab = d_cmac.as<z128>();
  dev.max_links = 0;
  for (int l = 0; l < hs.n_hrchy; ++l)
    dev.max_links = std::max(dev.max_links, link_ptr[l + 1] - link_ptr[l]);
  dev.sop_ptr = d_sop_ptr.as<int>();
  dev.sop_col = d_sop_col.as<int>();
  dev.sop_val = d_sop_val.as<z128>();
  dev.R_ptr = d_R_ptr.as<long long>();
  dev.R_col = d_R_col.as<int>();
  dev.R_val = d_R_val.as<z128>();
  // matrix-core commutator: heom_hilb, 16 < N <= 64, diagonal V (tiles of 32 or
  // 64 rows; H is zero-padded to the tile in its planes)
  const bool big = kind == HEOM_HILB && N > 16 && N <= 64;
  const bool mfma_ok = big && diag_v;
  const bool mfma_dense = big && !diag_v;
  dev.mfma = 0;
  if (opt.matrix_cores && mfma_ok) dev.mfma = dtype == HEOM_C64_BF16X3 ? 2 : 1;
  if (opt.matrix_cores && mfma_dense && dtype != HEOM_C64_BF16X3) dev.mfma = 3;
  if (dtype == HEOM_C64_BF16X3 && dev.mfma != 2)
    throw HeomError{HEOM_ERR_UNSUPPORTED,
                    "HEOM_C64_BF16X3 is provided for heom_hilb with 16 < n_level <= 64, diagonal "
                    "coupling operators and matrix_cores = 1"};
  dev.real_h = 1;
  {
    // real / imaginary planes of H with leading dimension P (the MFMA tile,
    // zero-padded), P = N outside the matrix-core path
    const int P = (mfma_ok || mfma_dense) ? (N <= 32 ? 32 : 64) : N;
    std::vector<double> hr(static_cast<size_t>(P) * P, 0.0), hi(hr.size(), 0.0);
    std::vector<float> hr32(hr.size(), 0.f), hi32(hr.size(), 0.f);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        const z128 h = H_work_[i * N + j];
        const size_t e = static_cast<size_t>(i) * P + j;
        hr[e] = h.re;
        hi[e] = h.im;
        hr32[e] = static_cast<float>(h.re);
        hi32[e] = static_cast<float>(h.im);
        if (h.im != 0.0) dev.real_h = 0;
      }
    d_Hre.upload(hr);
    d_Him.upload(hi);
    d_Hre32.upload(hr32);
    d_Him32.upload(hi32);
  }
  dev.real_vd = 1;
  if (dev.mfma == 3) {
    // every V_u zero-padded to the tile, as real / imaginary planes
    const int P = N <= 32 ? 32 : 64;
    const size_t pl = static_cast<size_t>(P) * P;
    std::vector<double> vr(pl * nn, 0.0), vi(vr.size(), 0.0);
    std::vector<float> vr32(vr.size(), 0.f), vi32(vr.size(), 0.f);
    for (int u = 0; u < nn; ++u)
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          const z128 v = noise_work_[u].V[i * N + j];
          const size_t e = u * pl + static_cast<size_t>(i) * P + j;
          vr[e] = v.re;
          vi[e] = v.im;
          vr32[e] = static_cast<float>(v.re);
          vi32[e] = static_cast<float>(v.im);
          if (v.im != 0.0) dev.real_vd = 0;
        }
    d_Vre.upload(vr);
    d_Vim.upload(vi);
    d_Vre32.upload(vr32);
    d_Vim32.upload(vi32);
  }
  dev.Vre = d_Vre.as<double>();
  dev.Vim = d_Vim.as<double>();
  dev.Vre32 = d_Vre32.as<float>();
  dev.Vim32 = d_Vim32.as<float>();
  dev.Hre = d_Hre.as<double>();
  dev.Him = d_Him.as<double>();
  dev.Hre32 = d_Hre32.as<float>();
  dev.Him32 = d_Him32.as<float>();
}

}  // namespace hamd
