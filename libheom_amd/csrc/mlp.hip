// Kernels and host code of the closure network (mlp.h).  Every reduction is
// done in a fixed order (no atomics), so training is reproducible run to run.
#include "mlp.h"

#include <cmath>
#include <cstring>
#include <numeric>
#include <string>

#include "kernels.h"

namespace hamd {

namespace {

// ---------------------------------------------------------------------------
// SGEMM: 64 x 64 output tile per 256-thread workgroup, k in slices of 16
// staged in LDS; thread (ty, tx) owns rows ty + 16 i and columns tx + 16 j
// ---------------------------------------------------------------------------
constexpr int kT = 64, kK = 16;

template <bool TA, bool TB>
__global__ void __launch_bounds__(256)
    sgemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                 const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                 const float* __restrict__ bias, int act) {
  __shared__ float As[kK][kT + 4];
  __shared__ float Bs[kK][kT + 4];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += kK) {
    for (int idx = tid; idx < kT * kK; idx += 256) {
      int mm, kk;
      if (TA) {
        mm = idx % kT;
        kk = idx / kT;
      } else {
        mm = idx / kK;
        kk = idx % kK;
      }
      const int m = m0 + mm, k = k0 + kk;
      float v = 0.f;
      if (m < M && k < K) v = TA ? A[static_cast<long long>(k) * lda + m]
                                 : A[static_cast<long long>(m) * lda + k];
      As[kk][mm] = v;
    }
    for (int idx = tid; idx < kT * kK; idx += 256) {
      int nn, kk;
      if (TB) {
        nn = idx / kK;
        kk = idx % kK;
      } else {
        nn = idx % kT;
        kk = idx / kT;
      }
      const int n = n0 + nn, k = k0 + kk;
      float v = 0.f;
      if (n < N && k < K) v = TB ? B[static_cast<long long>(n) * ldb + k]
                                 : B[static_cast<long long>(k) * ldb + n];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kK; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + ty + 16 * i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tx + 16 * j;
      if (n >= N) continue;
      float v = acc[i][j];
      if (bias) v += bias[n];
      if (act) v = tanhf(v);
      C[static_cast<long long>(m) * ldc + n] = v;
    }
  }
}

// out[j] = sum over rows of X[r*ld + j] (bias gradients), row order fixed
__global__ void colsum_kernel(const float* X, int rows, int cols, int ld, float* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; ++r) s += X[static_cast<long long>(r) * ld + j];
  out[j] = s;
}

// dH *= 1 - H^2 (derivative of tanh at the stored activation)
__global__ void tanh_back_kernel(float* dH, const float* H, long long n) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    dH[i] *= 1.f - H[i] * H[i];
}

// dY = 2 (Yhat - T) / n and per-workgroup partial sums of (Yhat - T)^2
__global__ void __launch_bounds__(256) mse_kernel(const float* Yhat, const float* T, float* dY,
                                                  long long n, double* partials) {
  __shared__ double s[256];
  double acc = 0.0;
  const float scale = 2.f / static_cast<float>(n);
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float d = Yhat[i] - T[i];
    dY[i] = scale * d;
    acc += static_cast<double>(d) * d;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

// Adam (Kingma & Ba, 2015) with bias-corrected moments
__global__ void adam_kernel(float* p, const float* g, float* m, float* v, long long n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// per-column mean and standard deviation over rows (two passes, fixed order);
// a constant column gets scale 1
__global__ void colstats_kernel(const float* X, long long rows, int cols, float* mean,
                                float* scale) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (long long r = 0; r < rows; ++r) s += X[r * cols + j];
  const double mu = s / static_cast<double>(rows);
  double q = 0.0;
  for (long long r = 0; r < rows; ++r) {
    const double d = X[r * cols + j] - mu;
    q += d * d;
  }
  const double sd = sqrt(q / static_cast<double>(rows));
  mean[j] = static_cast<float>(mu);
  scale[j] = sd > 1e-30 * (fabs(mu) + 1e-30) && sd > 0.0 ? static_cast<float>(sd) : 1.f;
}

// normalised rows: out[r][j] = (X[perm[r] or r][j] - mean[j]) / scale[j]
__global__ void gather_norm_kernel(const float* X, const int* perm, int rows, int cols,
                                   const float* mean, const float* scale, float* out) {
  const long long n = static_cast<long long>(rows) * cols;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int r = static_cast<int>(i / cols), j = static_cast<int>(i - static_cast<long long>(r) * cols);
    const long long src = perm ? perm[r] : r;
    out[i] = (X[src * cols + j] - mean[j]) / scale[j];
  }
}

// Y = Yn * scale + mean (column-wise)
__global__ void denorm_kernel(const float* Yn, int rows, int cols, const float* mean,
                              const float* scale, float* Y) {
  const long long n = static_cast<long long>(rows) * cols;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int j = static_cast<int>(i % cols);
    Y[i] = Yn[i] * scale[j] + mean[j];
  }
}

// x: Re / Im blocks of the last-tier ADOs (ADO r at r * x_stride) and the
// constant columns; y: the dropped terms, y = -buf with
// buf = -D_full + D_trunc, in blocks of 2 N^2
__global__ void closure_pack_kernel(const z128* rho, const z128* buf, long long base, int n_ado,
                                    int N2, int x_stride, float* x, float* y, const float* cst,
                                    int cst_rows, int n_cst, int cst_off, int row_len) {
  const long long n = static_cast<long long>(n_ado) * N2;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long r = i / N2;
    const int e = static_cast<int>(i - r * N2);
    const z128 v = rho[base + i], b = buf[base + i];
    x[r * x_stride + e] = static_cast<float>(v.re);
    x[r * x_stride + N2 + e] = static_cast<float>(v.im);
    y[r * 2 * N2 + e] = static_cast<float>(-b.re);
    y[r * 2 * N2 + N2 + e] = static_cast<float>(-b.im);
  }
  const long long nc = static_cast<long long>(cst_rows) * n_cst;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < nc;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long r = i / n_cst;
    x[r * row_len + cst_off + (i - r * n_cst)] = cst[i];
  }
}

int grid_for(long long n) {
  const long long g = (n + 255) / 256;
  return static_cast<int>(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

void launch_sgemm(int M, int N, int K, const float* A, int lda, bool ta, const float* B, int ldb,
                  bool tb, float* C, int ldc, const float* bias, bool act, hipStream_t st) {
  if (M <= 0 || N <= 0) return;
  const dim3 grid((N + kT - 1) / kT, (M + kT - 1) / kT);
  const int a = act ? 1 : 0;
  if (ta && tb)
    hipLaunchKernelGGL((sgemm_kernel<true, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, C, ldc, bias, a);
  else if (ta)
    hipLaunchKernelGGL((sgemm_kernel<true, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, C, ldc, bias, a);
  else if (tb)
    hipLaunchKernelGGL((sgemm_kernel<false, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, C, ldc, bias, a);
  else
    hipLaunchKernelGGL((sgemm_kernel<false, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, C, ldc, bias, a);
  HAMD_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Mlp
// ---------------------------------------------------------------------------
Mlp::Mlp(int n_in_, int n_hidden, int n_out_, uint64_t seed, int device_)
    : n_in(n_in_), n_hid(n_hidden), n_out(n_out_), device(device_) {
  if (n_in < 1 || n_hid < 1 || n_out < 1)
    throw HeomError{HEOM_ERR_ARG, "mlp: layer sizes must be >= 1"};
  HAMD_HIP(hipSetDevice(device));
  HAMD_HIP(hipStreamCreate(&stream));
  n_par = static_cast<long long>(n_hid) * n_in + n_hid + static_cast<long long>(n_hid) * n_hid +
          n_hid + static_cast<long long>(n_out) * n_hid + n_out;
  // Glorot-uniform weights, zero biases
  std::vector<float> h(n_par, 0.f);
  uint64_t s = seed;
  auto fill = [&](long long off, int fan_out, int fan_in) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    for (long long i = 0; i < static_cast<long long>(fan_out) * fan_in; ++i)
      h[off + i] = static_cast<float>((2.0 * uniform01(s) - 1.0) * a);
  };
  long long o = 0;
  fill(o, n_hid, n_in);
  o += static_cast<long long>(n_hid) * n_in + n_hid;
  fill(o, n_hid, n_hid);
  o += static_cast<long long>(n_hid) * n_hid + n_hid;
  fill(o, n_out, n_hid);
  params.upload(h);
  std::vector<float> z(n_par, 0.f);
  grads.upload(z);
  adam_m.upload(z);
  adam_v.upload(z);
  std::vector<float> zi(n_in, 0.f), oi(n_in, 1.f), zo(n_out, 0.f), oo(n_out, 1.f);
  xm_.upload(zi);
  xs_.upload(oi);
  ym_.upload(zo);
  ys_.upload(oo);
}

Mlp::~Mlp() {
  if (stream) (void)hipStreamDestroy(stream);
}

float* Mlp::W1() const { return params.as<float>(); }
float* Mlp::b1() const { return W1() + static_cast<long long>(n_hid) * n_in; }
float* Mlp::W2() const { return b1() + n_hid; }
float* Mlp::b2() const { return W2() + static_cast<long long>(n_hid) * n_hid; }
float* Mlp::W3() const { return b2() + n_hid; }
float* Mlp::b3() const { return W3() + static_cast<long long>(n_out) * n_hid; }

void Mlp::get_params(float* host) const {
  HAMD_HIP(hipMemcpy(host, params.p, sizeof(float) * n_par, hipMemcpyDeviceToHost));
}

void Mlp::set_params(const float* host) {
  HAMD_HIP(hipMemcpy(params.p, host, sizeof(float) * n_par, hipMemcpyHostToDevice));
}

void Mlp::get_grad(float* host) const {
  HAMD_HIP(hipMemcpy(host, grads.p, sizeof(float) * n_par, hipMemcpyDeviceToHost));
}

void Mlp::set_scaling(const float* xm, const float* xs, const float* ym, const float* ys) {
  for (int j = 0; j < n_in; ++j)
    if (!(xs[j] > 0.f)) throw HeomError{HEOM_ERR_ARG, "mlp: input scale must be > 0"};
  for (int j = 0; j < n_out; ++j)
    if (!(ys[j] > 0.f)) throw HeomError{HEOM_ERR_ARG, "mlp: output scale must be > 0"};
  HAMD_HIP(hipMemcpy(xm_.p, xm, sizeof(float) * n_in, hipMemcpyHostToDevice));
  HAMD_HIP(hipMemcpy(xs_.p, xs, sizeof(float) * n_in, hipMemcpyHostToDevice));
  HAMD_HIP(hipMemcpy(ym_.p, ym, sizeof(float) * n_out, hipMemcpyHostToDevice));
  HAMD_HIP(hipMemcpy(ys_.p, ys, sizeof(float) * n_out, hipMemcpyHostToDevice));
}

void Mlp::get_scaling(float* xm, float* xs, float* ym, float* ys) const {
  if (xm) HAMD_HIP(hipMemcpy(xm, xm_.p, sizeof(float) * n_in, hipMemcpyDeviceToHost));
  if (xs) HAMD_HIP(hipMemcpy(xs, xs_.p, sizeof(float) * n_in, hipMemcpyDeviceToHost));
  if (ym) HAMD_HIP(hipMemcpy(ym, ym_.p, sizeof(float) * n_out, hipMemcpyDeviceToHost));
  if (ys) HAMD_HIP(hipMemcpy(ys, ys_.p, sizeof(float) * n_out, hipMemcpyDeviceToHost));
}

void Mlp::fit_scaling(const float* X, const float* Y, long long n_rows) {
  if (n_rows < 1) throw HeomError{HEOM_ERR_ARG, "mlp: no rows"};
  HAMD_HIP(hipSetDevice(device));
  hipLaunchKernelGGL(colstats_kernel, dim3((n_in + 255) / 256), dim3(256), 0, stream, X, n_rows,
                     n_in, xm_.as<float>(), xs_.as<float>());
  hipLaunchKernelGGL(colstats_kernel, dim3((n_out + 255) / 256), dim3(256), 0, stream, Y, n_rows,
                     n_out, ym_.as<float>(), ys_.as<float>());
  HAMD_HIP(hipGetLastError());
  HAMD_HIP(hipStreamSynchronize(stream));
}

void Mlp::ensure_rows(int rows) {
  if (rows <= cap_rows) return;
  xb.alloc(sizeof(float) * rows * static_cast<size_t>(n_in));
  tb.alloc(sizeof(float) * rows * static_cast<size_t>(n_out));
  h1.alloc(sizeof(float) * rows * static_cast<size_t>(n_hid));
  h2.alloc(sizeof(float) * rows * static_cast<size_t>(n_hid));
  dh1.alloc(sizeof(float) * rows * static_cast<size_t>(n_hid));
  dh2.alloc(sizeof(float) * rows * static_cast<size_t>(n_hid));
  yb.alloc(sizeof(float) * rows * static_cast<size_t>(n_out));
  dyb.alloc(sizeof(float) * rows * static_cast<size_t>(n_out));
  partials.alloc(sizeof(double) * 4096);
  cap_rows = rows;
}

void Mlp::forward_normalised(int rows) {
  launch_sgemm(rows, n_hid, n_in, xb.as<float>(), n_in, false, W1(), n_in, true, h1.as<float>(),
               n_hid, b1(), true, stream);
  launch_sgemm(rows, n_hid, n_hid, h1.as<float>(), n_hid, false, W2(), n_hid, true,
               h2.as<float>(), n_hid, b2(), true, stream);
  launch_sgemm(rows, n_out, n_hid, h2.as<float>(), n_hid, false, W3(), n_hid, true,
               yb.as<float>(), n_out, b3(), false, stream);
}

void Mlp::forward(const float* X, float* Y, int n_rows) {
  HAMD_HIP(hipSetDevice(device));
  const int chunk = 8192;
  for (int r0 = 0; r0 < n_rows; r0 += chunk) {
    const int rows = std::min(chunk, n_rows - r0);
    ensure_rows(rows);
    hipLaunchKernelGGL(gather_norm_kernel, dim3(grid_for(static_cast<long long>(rows) * n_in)),
                       dim3(256), 0, stream, X + static_cast<long long>(r0) * n_in,
                       static_cast<const int*>(nullptr), rows, n_in, xm_.as<float>(),
                       xs_.as<float>(), xb.as<float>());
    forward_normalised(rows);
    hipLaunchKernelGGL(denorm_kernel, dim3(grid_for(static_cast<long long>(rows) * n_out)),
                       dim3(256), 0, stream, yb.as<float>(), rows, n_out, ym_.as<float>(),
                       ys_.as<float>(), Y + static_cast<long long>(r0) * n_out);
    HAMD_HIP(hipGetLastError());
  }
  HAMD_HIP(hipStreamSynchronize(stream));
}

double Mlp::loss_grad(const float* X, const float* Y, int rows, const int* perm_dev) {
  HAMD_HIP(hipSetDevice(device));
  ensure_rows(rows);
  const long long ni = static_cast<long long>(rows) * n_in, no = static_cast<long long>(rows) * n_out;
  hipLaunchKernelGGL(gather_norm_kernel, dim3(grid_for(ni)), dim3(256), 0, stream, X, perm_dev,
                     rows, n_in, xm_.as<float>(), xs_.as<float>(), xb.as<float>());
  hipLaunchKernelGGL(gather_norm_kernel, dim3(grid_for(no)), dim3(256), 0, stream, Y, perm_dev,
                     rows, n_out, ym_.as<float>(), ys_.as<float>(), tb.as<float>());
  forward_normalised(rows);
  const int gm = grid_for(no);
  hipLaunchKernelGGL(mse_kernel, dim3(gm), dim3(256), 0, stream, yb.as<float>(), tb.as<float>(),
                     dyb.as<float>(), no, partials.as<double>());
  float* g = grads.as<float>();
  float* gW1 = g;
  float* gb1 = gW1 + static_cast<long long>(n_hid) * n_in;
  float* gW2 = gb1 + n_hid;
  float* gb2 = gW2 + static_cast<long long>(n_hid) * n_hid;
  float* gW3 = gb2 + n_hid;
  float* gb3 = gW3 + static_cast<long long>(n_out) * n_hid;
  const long long nh = static_cast<long long>(rows) * n_hid;
  // layer 3
  launch_sgemm(n_out, n_hid, rows, dyb.as<float>(), n_out, true, h2.as<float>(), n_hid, false,
               gW3, n_hid, nullptr, false, stream);
  hipLaunchKernelGGL(colsum_kernel, dim3((n_out + 255) / 256), dim3(256), 0, stream,
                     dyb.as<float>(), rows, n_out, n_out, gb3);
  launch_sgemm(rows, n_hid, n_out, dyb.as<float>(), n_out, false, W3(), n_hid, false,
               dh2.as<float>(), n_hid, nullptr, false, stream);
  hipLaunchKernelGGL(tanh_back_kernel, dim3(grid_for(nh)), dim3(256), 0, stream, dh2.as<float>(),
                     h2.as<float>(), nh);
  // layer 2
  launch_sgemm(n_hid, n_hid, rows, dh2.as<float>(), n_hid, true, h1.as<float>(), n_hid, false,
               gW2, n_hid, nullptr, false, stream);
  hipLaunchKernelGGL(colsum_kernel, dim3((n_hid + 255) / 256), dim3(256), 0, stream,
                     dh2.as<float>(), rows, n_hid, n_hid, gb2);
  launch_sgemm(rows, n_hid, n_hid, dh2.as<float>(), n_hid, false, W2(), n_hid, false,
               dh1.as<float>(), n_hid, nullptr, false, stream);
  hipLaunchKernelGGL(tanh_back_kernel, dim3(grid_for(nh)), dim3(256), 0, stream, dh1.as<float>(),
                     h1.as<float>(), nh);
  // layer 1
  launch_sgemm(n_hid, n_in, rows, dh1.as<float>(), n_hid, true, xb.as<float>(), n_in, false, gW1,
               n_in, nullptr, false, stream);
  hipLaunchKernelGGL(colsum_kernel, dim3((n_hid + 255) / 256), dim3(256), 0, stream,
                     dh1.as<float>(), rows, n_hid, n_hid, gb1);
  HAMD_HIP(hipGetLastError());
  std::vector<double> part(gm);
  HAMD_HIP(hipMemcpyAsync(part.data(), partials.p, sizeof(double) * gm, hipMemcpyDeviceToHost,
                          stream));
  HAMD_HIP(hipStreamSynchronize(stream));
  double s = 0.0;
  for (double v : part) s += v;
  return s / static_cast<double>(no);
}

void Mlp::adam_step(double lr, double beta1, double beta2, double eps) {
  HAMD_HIP(hipSetDevice(device));
  t_adam += 1;
  const float c1 = static_cast<float>(1.0 - std::pow(beta1, static_cast<double>(t_adam)));
  const float c2 = static_cast<float>(1.0 - std::pow(beta2, static_cast<double>(t_adam)));
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n_par)), dim3(256), 0, stream, params.as<float>(),
                     grads.as<float>(), adam_m.as<float>(), adam_v.as<float>(), n_par,
                     static_cast<float>(lr), static_cast<float>(beta1), static_cast<float>(beta2),
                     static_cast<float>(eps), c1, c2);
  HAMD_HIP(hipGetLastError());
}

void Mlp::train(const float* X, const float* Y, long long n_rows, int batch, int epochs,
                double lr, uint64_t seed, double* epoch_loss) {
  if (n_rows < 1 || batch < 1 || epochs < 0)
    throw HeomError{HEOM_ERR_ARG, "mlp train: bad sizes"};
  if (n_rows > (1LL << 31) - 1) throw HeomError{HEOM_ERR_UNSUPPORTED, "mlp train: too many rows"};
  HAMD_HIP(hipSetDevice(device));
  std::vector<int> order(static_cast<size_t>(n_rows));
  perm.alloc(sizeof(int) * order.size());
  uint64_t s = seed;
  for (int ep = 0; ep < epochs; ++ep) {
    std::iota(order.begin(), order.end(), 0);
    for (long long i = n_rows - 1; i > 0; --i) {  // Fisher-Yates
      const long long j = static_cast<long long>(splitmix(s) % static_cast<uint64_t>(i + 1));
      std::swap(order[i], order[j]);
    }
    HAMD_HIP(hipMemcpyAsync(perm.p, order.data(), sizeof(int) * order.size(),
                            hipMemcpyHostToDevice, stream));
    double tot = 0.0;
    for (long long r0 = 0; r0 < n_rows; r0 += batch) {
      const int rows = static_cast<int>(std::min<long long>(batch, n_rows - r0));
      tot += loss_grad(X, Y, rows, perm.as<int>() + r0) * rows;
      adam_step(lr, 0.9, 0.999, 1e-8);
    }
    if (epoch_loss) epoch_loss[ep] = tot / static_cast<double>(n_rows);
  }
  HAMD_HIP(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// Closure
// ---------------------------------------------------------------------------
void last_tier(const Qme& q, int& first, int& count) {
  const int K = q.hs.n_dim;
  first = q.hs.n_hrchy;
  for (int a = 0; a < q.hs.n_hrchy; ++a) {
    int tier = 0;
    for (int k = 0; k < K; ++k) tier += q.hs.n[static_cast<size_t>(a) * K + k];
    if (tier == q.max_depth) {
      first = a;
      break;
    }
  }
  count = q.hs.n_hrchy - first;  // breadth-first order: the last tier is a suffix
}

ClosureDims closure_dims(const Qme& q, int layout, int n_features) {
  if (layout != HEOM_CLOSURE_PER_ADO && layout != HEOM_CLOSURE_GLOBAL)
    throw HeomError{HEOM_ERR_ARG, "closure: unknown layout"};
  if (n_features < 0) throw HeomError{HEOM_ERR_ARG, "closure: n_features < 0"};
  ClosureDims c{};
  last_tier(q, c.lo, c.n_ado);
  const int K = q.hs.n_dim, N2 = q.N2;
  if (layout == HEOM_CLOSURE_PER_ADO) {
    c.rows = c.n_ado;
    c.n_in = 2 * N2 + K + n_features;
    c.n_out = 2 * N2;
    c.ado_stride = c.n_in;
    c.cst_rows = c.n_ado;
    c.n_cst = K + n_features;
    c.cst_off = 2 * N2;
  } else {
    c.rows = 1;
    c.n_in = 2 * N2 * c.n_ado + n_features;
    c.n_out = 2 * N2 * c.n_ado;
    c.ado_stride = 2 * N2;
    c.cst_rows = 1;
    c.n_cst = n_features;
    c.cst_off = 2 * N2 * c.n_ado;
  }
  return c;
}

namespace {
// constant input columns of a closure row buffer: index vectors and features
std::vector<float> closure_constants(const Qme& q, const ClosureDims& c, int layout,
                                     const float* features, int n_features) {
  const int K = q.hs.n_dim;
  std::vector<float> cst(static_cast<size_t>(c.cst_rows) * c.n_cst);
  for (int r = 0; r < c.cst_rows; ++r) {
    float* row = cst.data() + static_cast<size_t>(r) * c.n_cst;
    int o = 0;
    if (layout == HEOM_CLOSURE_PER_ADO)
      for (int k = 0; k < K; ++k)
        row[o++] = static_cast<float>(q.hs.n[static_cast<size_t>(c.lo + r) * K + k]);
    for (int f = 0; f < n_features; ++f) row[o++] = features[f];
  }
  return cst;
}
}  // namespace

void closure_rows(Qme& deep, Qme& shallow, const void* rho_deep, const float* features,
                  int n_features, int layout, float* x, float* y) {
  if (deep.kind != HEOM_HILB || shallow.kind != HEOM_HILB)
    throw HeomError{HEOM_ERR_UNSUPPORTED, "closure rows: heom_hilb only"};
  if (deep.lo() || shallow.lo())
    throw HeomError{HEOM_ERR_ARG, "closure rows: both QMEs must be complex128"};
  if (deep.N != shallow.N || deep.hs.n_dim != shallow.hs.n_dim ||
      shallow.max_depth >= deep.max_depth || shallow.max_depth < 1)
    throw HeomError{HEOM_ERR_ARG, "closure rows: shallow must be the same model at a smaller "
                                  "depth >= 1"};
  // the shallow hierarchy must be a prefix of the deep one (breadth-first order)
  for (size_t i = 0; i < shallow.hs.n.size(); ++i)
    if (shallow.hs.n[i] != deep.hs.n[i])
      throw HeomError{HEOM_ERR_ARG, "closure rows: hierarchies differ"};
  if (n_features > 0 && !features) throw HeomError{HEOM_ERR_ARG, "closure rows: bad features"};
  const ClosureDims c = closure_dims(shallow, layout, n_features);
  const int N2 = shallow.N2;
  HAMD_HIP(hipSetDevice(deep.opt.device));
  hipStream_t st = deep.stream;
  HAMD_HIP(hipStreamSynchronize(shallow.stream));
  const size_t bytes = sizeof(z128) * static_cast<size_t>(deep.main_size());
  if (deep.scratch.bytes != bytes) deep.scratch.alloc(bytes);
  // buf = -D_full(rho); buf[shallow] += D_trunc(rho[shallow])
  const StageArgs a1 = deep.diff_args(deep.scratch.p, rho_deep, z128{1.0, 0.0}, z128{0.0, 0.0});
  launch_operator(deep.dev, 8, a1, st);
  const StageArgs a2 =
      shallow.diff_args(deep.scratch.p, rho_deep, z128{-1.0, 0.0}, z128{1.0, 0.0});
  launch_operator(shallow.dev, 8, a2, st);
  DevBuf dc;
  dc.upload(closure_constants(shallow, c, layout, features, n_features));
  const long long n = std::max(static_cast<long long>(c.n_ado) * N2,
                               static_cast<long long>(c.cst_rows) * c.n_cst);
  hipLaunchKernelGGL(closure_pack_kernel, dim3(grid_for(n)), dim3(256), 0, st,
                     static_cast<const z128*>(rho_deep), deep.scratch.as<z128>(),
                     static_cast<long long>(c.lo) * N2, c.n_ado, N2, c.ado_stride, x, y,
                     dc.as<float>(), c.cst_rows, c.n_cst, c.cst_off, c.n_in);
  HAMD_HIP(hipGetLastError());
  HAMD_HIP(hipStreamSynchronize(st));
}

void Closure::bind(Qme& q, const Mlp& m, const float* features, int n_features, int layout) {
  if (q.kind != HEOM_HILB)
    throw HeomError{HEOM_ERR_UNSUPPORTED, "closure: heom_hilb only"};
  if (q.max_depth < 1) throw HeomError{HEOM_ERR_ARG, "closure: max_depth must be >= 1"};
  if (n_features > 0 && !features) throw HeomError{HEOM_ERR_ARG, "closure: bad features"};
  dm = closure_dims(q, layout, n_features);
  if (dm.n_ado < 1) throw HeomError{HEOM_ERR_ARG, "closure: no ADO at the last tier"};
  const int n_in = dm.n_in, n_out = dm.n_out;
  n_hid = m.n_hid;
  if (m.n_in != n_in || m.n_out != n_out)
    throw HeomError{HEOM_ERR_ARG, "closure: this layout needs a network with n_in = " +
                                      std::to_string(n_in) + " and n_out = " +
                                      std::to_string(n_out)};
  // fold the normalisation into the first and last layers
  std::vector<float> p(m.n_params()), xm(n_in), xs(n_in), ym(n_out), ys(n_out);
  m.get_params(p.data());
  m.get_scaling(xm.data(), xs.data(), ym.data(), ys.data());
  const long long oW1 = 0, ob1 = oW1 + static_cast<long long>(n_hid) * n_in, oW2 = ob1 + n_hid,
                  ob2 = oW2 + static_cast<long long>(n_hid) * n_hid, oW3 = ob2 + n_hid,
                  ob3 = oW3 + static_cast<long long>(n_out) * n_hid;
  std::vector<float> w1(static_cast<size_t>(n_hid) * n_in), bb1(n_hid),
      w3(static_cast<size_t>(n_out) * n_hid), bb3(n_out);
  for (int j = 0; j < n_hid; ++j) {
    double b = p[ob1 + j];
    for (int i = 0; i < n_in; ++i) {
      const double w = p[oW1 + static_cast<long long>(j) * n_in + i];
      w1[static_cast<size_t>(j) * n_in + i] = static_cast<float>(w / xs[i]);
      b -= w * xm[i] / xs[i];
    }
    bb1[j] = static_cast<float>(b);
  }
  for (int o = 0; o < n_out; ++o) {
    for (int j = 0; j < n_hid; ++j)
      w3[static_cast<size_t>(o) * n_hid + j] =
          p[oW3 + static_cast<long long>(o) * n_hid + j] * ys[o];
    bb3[o] = p[ob3 + o] * ys[o] + ym[o];
  }
  HAMD_HIP(hipSetDevice(q.opt.device));
  W1.upload(w1);
  b1.upload(bb1);
  W2.upload(std::vector<float>(p.begin() + oW2, p.begin() + ob2));
  b2.upload(std::vector<float>(p.begin() + ob2, p.begin() + oW3));
  W3.upload(w3);
  b3.upload(bb3);
  // the constant input columns are written once
  std::vector<float> xh(static_cast<size_t>(dm.rows) * n_in, 0.f);
  const std::vector<float> cst = closure_constants(q, dm, layout, features, n_features);
  for (int r = 0; r < dm.cst_rows; ++r)
    for (int c = 0; c < dm.n_cst; ++c)
      xh[static_cast<size_t>(r) * n_in + dm.cst_off + c] = cst[static_cast<size_t>(r) * dm.n_cst + c];
  x.upload(xh);
  h1.alloc(sizeof(float) * dm.rows * static_cast<size_t>(n_hid));
  h2.alloc(sizeof(float) * dm.rows * static_cast<size_t>(n_hid));
  y.upload(std::vector<float>(static_cast<size_t>(dm.rows) * n_out, 0.f));
  q.dev.clos = y.as<float>();
  q.dev.clos_lo = dm.lo;
  q.dev.clos_n = dm.n_ado;
  // per-ADO rows of a network that fits: one fused launch per operator pass
  fused = layout == HEOM_CLOSURE_PER_ADO && closure_mlp_fits(n_in, n_hid, n_out);
}

void Closure::forward(const Qme& q, const StageArgs& A, hipStream_t st) {
  const int n_in = dm.n_in, n_out = dm.n_out, rows = dm.rows;
  if (fused) {
    launch_closure_mlp(q.dev, A, x.as<float>(), n_in, n_hid, n_out, W1.as<float>(),
                       b1.as<float>(), W2.as<float>(), b2.as<float>(), W3.as<float>(),
                       b3.as<float>(), y.as<float>(), st);
    return;
  }
  launch_closure_gather(q.dev, A, x.as<float>(), dm.ado_stride, st);
  launch_sgemm(rows, n_hid, n_in, x.as<float>(), n_in, false, W1.as<float>(), n_in, true,
               h1.as<float>(), n_hid, b1.as<float>(), true, st);
  launch_sgemm(rows, n_hid, n_hid, h1.as<float>(), n_hid, false, W2.as<float>(), n_hid, true,
               h2.as<float>(), n_hid, b2.as<float>(), true, st);
  launch_sgemm(rows, n_out, n_hid, h2.as<float>(), n_hid, false, W3.as<float>(), n_hid, true,
               y.as<float>(), n_out, b3.as<float>(), false, st);
}

}  // namespace hamd
