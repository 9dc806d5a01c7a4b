// Kernel launchers of the libheom_amd engine.
#pragma once

#include "common.h"

namespace hamd {

constexpr int kMaxGrid = 1 << 18;  // grid cap = partial-sum slots of the error norm

// Redfield Hilbert form: LDS elements (complex, operator precision) the
// one-workgroup kernel stages -- rho, H or HL, HR (a block kept for both), per
// noise Lambda_u and Lambda_u^+ (and V_u for dense V), then the diagonals of V
// (diagonal V) or one T block (dense V).  Problems with n_level <= 16 whose tables fit in
// kRedfieldLds bytes run in one workgroup; the others take two passes through
// a scratch block of n_noise N^2 elements.
constexpr size_t kRedfieldLds = 60 * 1024;  // leaves room for the static LDS of the error norm
__host__ __device__ inline size_t redfield_lds_elems(int N2, int N, int n_noise, bool diag) {
  return diag ? static_cast<size_t>(3 + 2 * n_noise) * N2 + static_cast<size_t>(n_noise) * N
              : static_cast<size_t>(4 + 3 * n_noise) * N2;
}

// Launch one fused operator pass (calc_diff + epilogue) for the form in `d`.
// real_bytes: 8 = FP64 operator arithmetic, 4 = FP32.
// Returns the grid size used (the number of error-norm partials written).
int launch_operator(const FormDev& d, int real_bytes, const StageArgs& A, hipStream_t stream);

// Combine the error-norm partials of the last stage pass (nparts = its grid)
// and run the Dormand-Prince step control.
void launch_err_finalize(const StageArgs& A, int nparts, hipStream_t stream);

// Set up the Dormand-Prince control block for one output interval.
void launch_ctrl_init(Ctrl* c, double t0, double t_bound, double dt, hipStream_t stream);

// rho_S of every member of a batch (d.kind == kBatch), gathered in member
// order from an FP64 state
// closure network input: Re / Im (N^2 each) of every last-tier ADO of the
// stage input, ADO r's block starting at x[r * ado_stride]
void launch_closure_gather(const FormDev& d, const StageArgs& A, float* x, int ado_stride,
                           hipStream_t stream);
void launch_gather_rho_s(const FormDev& d, void* dst, const void* src, hipStream_t stream);
// the whole per-ADO closure network in one launch (gather + three layers on
// the matrix cores): widths n_in <= 64, n_hid a multiple of 16 up to 256,
// n_out a multiple of 16 up to 64 (closure_mlp_fits)
bool closure_mlp_fits(int n_in, int n_hid, int n_out);
void launch_closure_mlp(const FormDev& d, const StageArgs& A, const float* xc, int n_in,
                        int n_hid, int n_out, const float* W1, const float* b1, const float* W2,
                        const float* b2, const float* W3, const float* b3, float* y,
                        hipStream_t stream);

// Resident fixed-step integrator (heom_hilb with diagonal V and at most 8
// links per ADO, or redfield_hilb with diagonal V; FP64, at most kResidentMax
// elements, no closure): one launch runs a whole output interval.
// resident_ok tells whether a problem qualifies.
bool resident_ok(const FormDev& d, int max_links);
// LDS bytes of a resident launch with n_buf state-sized buffers
size_t resident_lds_bytes(const FormDev& d, int n_buf);
void launch_resident(const FormDev& d, int max_links, const ResidentArgs& a, hipStream_t stream);
// Resident Dormand-Prince (the fixed-step conditions and at most
// kResidentDopriMax elements): the whole adaptive solve in one launch.
bool resident_dopri_ok(const FormDev& d, int max_links);
void launch_resident_dopri(const FormDev& d, int max_links, const ResidentDopriArgs& a,
                           hipStream_t stream);
// Tr rho_0 (real part) of the FP64 state y into *tr0
void launch_rho0_trace(const void* y, int N, double* tr0, hipStream_t stream);
// rho_0 <- (rho_0 + rho_0^+)/2 * tr0 / Tr rho_0 in the FP64 state of parity
// ctrl->cur (y0 / y1; parity 0 without ctrl), and its FP32 copy f0 / f1 if given
void launch_rho0_renorm(void* y0, void* y1, void* f0, void* f1, Ctrl* ctrl, int N,
                        const double* tr0, hipStream_t stream);
// dst (FP32 complex) <- src (FP64 complex), n elements
void launch_narrow(void* dst, const void* src, long long n, hipStream_t stream);
// dst <- src for n complex elements of the given slot type
void launch_copy(void* dst, const void* src, long long n, int type, hipStream_t stream);
// dst <- src for n complex128 elements on a grid-stride grid of `grid` x 256 threads
void launch_copy_grid(void* dst, const void* src, long long n, int grid, hipStream_t stream);

}  // namespace hamd
