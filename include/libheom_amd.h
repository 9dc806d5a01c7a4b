/* libheom_amd -- MI355X-native HEOM propagation engine, C ABI.
 *
 * This is the drop-in boundary for libheom's time-propagation path.  The
 * reference exposes it as C++ templates (no C ABI); every entry point below
 * names the reference interface it replaces.  Conventions:
 *   - complex numbers are interleaved (re, im) doubles; matrices are N x N
 *     row-major (order = row_major in the reference);
 *   - every function returns a status (HEOM_OK == 0); heom_last_error()
 *     gives the message of the last failure on the calling thread;
 *   - no C++ exception crosses this boundary;
 *   - "dev" pointers are HIP device pointers, "host" pointers are host memory.
 *
 * The derivative computed everywhere is the reference's
 *   drho = beta*drho - alpha*D(rho)                 (heom_hilb.h:210-211)
 * with D the HEOM operator of heom_hilb.h:108-209.
 */
#ifndef LIBHEOM_AMD_H
#define LIBHEOM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HEOM_OK 0
#define HEOM_ERR_ARG 1            /* bad argument / missing parameter */
#define HEOM_ERR_STATE 2          /* call out of order (e.g. before set_param) */
#define HEOM_ERR_HIP 3            /* HIP runtime failure */
#define HEOM_ERR_STEP_REJECTED 4  /* rkdp gave up after 1000 rejected attempts */
#define HEOM_ERR_UNSUPPORTED 5    /* size / form combination not provided */
#define HEOM_ERR_NOMEM 6          /* device allocation failed */

/* QME representation (heom_hilb.h:26, heom_liou.h:27, heom_ado.h:27) */
#define HEOM_HILB 0
#define HEOM_LIOU 1
#define HEOM_ADO 2
/* Redfield QMEs (redfield_hilb.h:24, redfield_liou.h:25): the same parameter
 * members; max_depth is ignored and main_size = n_level^2 (redfield_hilb.h:39-49,
 * redfield_liou.h:44-54) */
#define HEOM_REDFIELD_HILB 3
#define HEOM_REDFIELD_LIOU 4

/* arithmetic (dtype template parameter: complex128 / complex64) */
#define HEOM_C128 0  /* FP64 complex everywhere */
#define HEOM_C64 1   /* FP32 complex operator + stage inputs; FP64 stage
                        combination, state and error norm */
#define HEOM_C64_BF16X3 2 /* as HEOM_C64, with the commutator's matrix products
                             on bf16 matrix cores: each FP32 operand split into
                             hi + lo bf16 parts, hi*hi + hi*lo + lo*hi
                             accumulated in FP32 (matrix-core path only) */

/* solvers (solver/rk4.h, solver/rkdp.h, solver/lsrk4.h) */
#define HEOM_RK4 0
#define HEOM_RKDP 1
#define HEOM_LSRK4 2

typedef struct heom_qme* heom_handle;
typedef struct heom_solver* heom_solver_handle;

typedef struct heom_options {
  int parity_q1;  /* 1 (default): drop off-diagonal gamma like the reference
                     (heom.h:106-121 clears gamma at set_shape, line 110, before copying it) */
  int sqrt_norm;  /* 1: LIBHEOM_SQRT_NORMALIZATION factors (heom_hilb.h:140-196) */
  int order;      /* 0: row_major (default), 1: col_major (matrix_base.h) */
  int device;     /* HIP device ordinal */
  int matrix_cores; /* 1 (default): heom_hilb with 16 < n_level <= 64 and diagonal
                       V computes the commutator on MFMA (blocks zero-padded to 32
                       or 64 rows); 0: vector ALU kernels */
  int renorm_rho0;  /* 1: after every RK step / accepted Dormand-Prince step, replace rho_0
                       (rho_S) by (rho_0 + rho_0^+)/2 scaled to the trace it had when
                       the solve started, in FP64 on the device (n_level <= 64, not
                       for batches).  0 (default, parity mode): off.  No counterpart
                       in the reference; for the complex64 variants. */
} heom_options;

typedef struct heom_stats {
  long long n_rhs;          /* calc_diff evaluations (RK stages) */
  long long n_attempts;     /* rkdp attempts / rk4 steps */
  long long n_accepted;
  long long n_rejected;
  long long ado_stage_updates; /* n_hrchy * n_rhs */
  int gave_up;              /* rkdp hit the 1000-attempt limit (rkdp.h:95-102) */
  double last_dt;           /* dt_save after the last solve (adaptive_step_size_solver.h:55) */
  double last_err;          /* last error norm (rkdp.h:153) */
} heom_stats;

/* called once per output time t_i, before propagating past it
 * (solver_base.h:73,78); rho_s is rho[0:N*N] (host, complex interleaved) and
 * rho_full is the whole host state when the solve copies it, else NULL. */
typedef void (*heom_callback)(double t, const double* rho_s, const double* rho_full,
                              void* user);

const char* heom_last_error(void);
int heom_version(int* major, int* minor);
int heom_default_options(heom_options* opt);

/* heom_hilb / heom_liou / heom_ado constructor (heom.h:44-50:
 * heom(int max_depth, int n_inner_threads, int n_outer_threads)); threads are
 * replaced by the device.  kind: HEOM_HILB|LIOU|ADO, dtype: HEOM_C128|C64.
 * kind HEOM_REDFIELD_HILB|LIOU: the redfield_hilb / redfield_liou default
 * constructors (redfield_hilb.h:36-37, redfield_liou.h:41-42); max_depth ignored. */
int heom_create(int kind, int dtype, int n_level, int max_depth, const heom_options* opt,
                heom_handle* out);
int heom_destroy(heom_handle h);

/* qme_base public members (qme_base.h:32-48) */
int heom_set_H(heom_handle h, const double* H);
int heom_alloc_noises(heom_handle h, int n_noise);           /* qme_base.h:54-67 */
int heom_set_noise(heom_handle h, int u, const double* V, int len_gamma,
                   const double* gamma, const double* phi_0, const double* sigma,
                   const double* S, const double* A, double s_delta_re,
                   double s_delta_im);

/* hrchy_filter of alloc_hrchy_space (hrchy_space.h:29-36): when set, an ADO
 * index is enumerated only if filter(index, n_dim, depth, user) returns
 * non-zero, tested where the breadth-first enumeration pushes it
 * (hrchy_space.cc:183-187), so indices reached only through a rejected one are
 * pruned too and their links become void.  The reference's heom classes pass
 * none (heom.h:77-91); here any QME can take one.  NULL removes it; takes
 * effect at the next heom_set_param. */
typedef int (*heom_hrchy_filter)(const int* index, int n_dim, int depth, void* user);
int heom_set_hrchy_filter(heom_handle h, heom_hrchy_filter filter, void* user);

/* set_param (heom.h:57-139 + heom_hilb.h:57-68 / heom_liou.h:66-119 /
 * heom_ado.h:62-189): builds the hierarchy, neighbour links and coefficient
 * tables and uploads them to HBM.  heom_liou / heom_ado take n_level <= 32
 * (their N^2 x N^2 superoperators grow as N^4) and heom_hilb any n_level with
 * diagonal V_u, dense V_u up to 64; beyond, HEOM_ERR_UNSUPPORTED. */
int heom_set_param(heom_handle h);
int heom_get_n_hrchy(heom_handle h, int* n_hrchy);           /* heom.h:52-55 */
int heom_get_n_dim(heom_handle h, int* n_dim);
int heom_main_size(heom_handle h, long long* main_size);     /* heom_hilb.h:45-49 */
/* hierarchy tables as built (hrchy_space.h:18-27): n (n_hrchy x n_dim),
 * ptr_p1 / ptr_m1 ((n_hrchy+1) x n_dim, sentinel = n_hrchy).  Any may be NULL. */
int heom_get_hrchy(heom_handle h, int* n, int* ptr_p1, int* ptr_m1);

/* Redfield set-up products after heom_set_param (redfield.h:43-174): the
 * relaxation operator Lambda_u = U (U^+ V_u U o Ct_u) U^+ of noise u as an
 * N x N row-major complex matrix (logical, whatever the storage order), and
 * the eigenvalues of H (N doubles, in the eigensolver's order).  Either output
 * may be NULL.  The reference keeps them as the public members Lambda and the
 * local w of redfield::set_param. */
int heom_redfield_lambda(heom_handle h, int u, double* lambda, double* eig);

/* calc_diff_impl (qme_base.h:74-79): drho_dev = beta*drho_dev - alpha*D(rho_dev)
 * on device buffers of main_size complex elements of the handle's dtype.
 * stream: hipStream_t or NULL for the library stream. */
int heom_calc_diff(heom_handle h, void* drho_dev, const void* rho_dev, double alpha_re,
                   double alpha_im, double beta_re, double beta_im, void* stream);
/* same on host buffers (complex128), staging through HBM */
int heom_calc_diff_host(heom_handle h, double* drho, const double* rho, double alpha_re,
                        double alpha_im, double beta_re, double beta_im);

/* Batch of independent heom_hilb members (diagonal V, n_level <= 8; no
 * counterpart in the reference, which runs one hierarchy per object): the
 * members' tables are packed so that one operator launch per Runge-Kutta stage
 * advances all of them.  The batch behaves as one QME whose state is the
 * concatenation of the members' states (member order); members must have had
 * set_param; their tables are copied, so they may be destroyed afterwards.
 * For a batch, heom_solve_dev's rho_s_out receives n_t blocks of all members'
 * rho_S (member order) and the heom_solve callback sees member 0's rho_S. */
int heom_batch_create(const heom_handle* members, int n_members, heom_handle* out);
/* per member: first state element, offset of its rho_S in a gathered block,
 * n_level (any array may be NULL) */
int heom_batch_layout(heom_handle batch, long long* elem_base, int* rho_s_offset,
                      int* n_level);
/* The batch kernel's autotuner (no counterpart in the reference): at
 * heom_batch_create it times one Runge-Kutta stage pass of the whole batch for
 * every tile shape -- work items per wavefront step (1 or 2; a work item is
 * 64/n_level^2 ADOs of one member) and wavefronts per workgroup (1, 2 or 4) --
 * and grid size, and keeps the fastest.  shape receives {items per step,
 * wavefronts per workgroup, grid}; tried receives up to `cap` such triples of
 * the candidates in the order timed, tried_us their microseconds per pass, and
 * *n_tried their count. */
int heom_batch_tuning(heom_handle batch, int* shape, int* tried, float* tried_us, int cap,
                      int* n_tried);
/* Overrides the autotuner's choice (items_per_step 1 or 2, waves_per_group 1,
 * 2 or 4, 1 <= grid <= 262144 / waves_per_group); the results do not depend on
 * the shape beyond rounding, only the speed does. */
int heom_batch_set_tuning(heom_handle batch, int items_per_step, int waves_per_group, int grid);

/* Learned truncation closure (no counterpart in the reference, which drops
 * every ADO beyond max_depth; heom.h:77-91).  A 3-layer network (n_in ->
 * n_hidden -> n_hidden -> n_out, tanh, FP32) maps the ADOs of the last tier L
 * to the tier-(L+1) coupling terms that truncation drops, added to their D_n.
 * Row layouts:
 *   HEOM_CLOSURE_PER_ADO: one row per last-tier ADO n = (Re rho_n, Im rho_n
 *     (N^2 each), the index vector n, features) -> (Re, Im of n's term);
 *   HEOM_CLOSURE_GLOBAL: one row per state = (Re, Im blocks of every
 *     last-tier ADO in hierarchy order, features) -> all their terms.
 * Training and inference run in the library's own kernels; X / Y are device
 * arrays of rows (row-major FP32). */
#define HEOM_CLOSURE_PER_ADO 0
#define HEOM_CLOSURE_GLOBAL 1
typedef struct heom_mlp* heom_mlp_handle;
int heom_mlp_create(int n_in, int n_hidden, int n_out, unsigned long long seed, int device,
                    heom_mlp_handle* out);
int heom_mlp_destroy(heom_mlp_handle m);
int heom_mlp_n_params(heom_mlp_handle m, long long* n_params);
/* parameter order: W1 (n_hidden x n_in, row-major), b1, W2, b2, W3 (n_out x n_hidden), b3 */
int heom_mlp_get_params(heom_mlp_handle m, float* host);
int heom_mlp_set_params(heom_mlp_handle m, const float* host);
/* x' = (x - x_shift) / x_scale, y = y' * y_scale + y_shift (host arrays) */
int heom_mlp_set_scaling(heom_mlp_handle m, const float* x_shift, const float* x_scale,
                         const float* y_shift, const float* y_scale);
int heom_mlp_get_scaling(heom_mlp_handle m, float* x_shift, float* x_scale, float* y_shift,
                         float* y_scale);
/* per-column mean / standard deviation of the training rows */
int heom_mlp_fit_scaling(heom_mlp_handle m, const float* X_dev, const float* Y_dev,
                         long long n_rows);
/* raw rows in, raw rows out (device, row-major) */
int heom_mlp_forward(heom_mlp_handle m, const float* X_dev, float* Y_dev, int n_rows);
/* mean squared error of the normalised outputs and its gradient (host array
 * of n_params, may be NULL); the gradient also stays on the device for
 * heom_mlp_adam_step */
int heom_mlp_loss_grad(heom_mlp_handle m, const float* X_dev, const float* Y_dev, int n_rows,
                       double* loss, float* grad_host);
int heom_mlp_adam_step(heom_mlp_handle m, double lr, double beta1, double beta2, double eps);
/* shuffled minibatch Adam (beta 0.9 / 0.999, eps 1e-8); epoch_loss[epochs] may be NULL */
int heom_mlp_train(heom_mlp_handle m, const float* X_dev, const float* Y_dev, long long n_rows,
                   int batch, int epochs, double lr, unsigned long long seed,
                   double* epoch_loss);
/* bind a snapshot of the network to heom_hilb h (m == NULL removes it); every
 * later calc_diff / solve adds the predicted term to the last tier */
int heom_set_closure(heom_handle h, heom_mlp_handle m, const float* features, int n_features,
                     int layout);
/* network rows per state and widths of a layout for heom_hilb h */
int heom_closure_dims(heom_handle h, int layout, int n_features, int* rows_per_state, int* n_in,
                      int* n_out);
/* training rows from a state of a deeper hierarchy: shallow is the same model
 * at depth L < depth(deep), both heom_hilb complex128; writes the rows of the
 * last tier of shallow: x (n_in wide) and y = the tier-(L+1) term of D_n that
 * truncation at L drops */
int heom_closure_rows(heom_handle deep, heom_handle shallow, const void* rho_deep_dev,
                      const float* features, int n_features, int layout, float* x_dev,
                      float* y_dev);

/* solver_base::init via qme_solver (qme_solver.h:26-35).  The solver shares
 * ownership of the QME: heom_destroy(h) leaves it usable until
 * heom_solver_destroy.  If set_param is called again on h, the solver follows
 * the new parameters; a solve returns HEOM_ERR_STATE when h is not set up or
 * its main_size differs from the one the solver was created for. */
int heom_solver_create(heom_handle h, int solver_kind, heom_solver_handle* out);
int heom_solver_destroy(heom_solver_handle s);

/* qme_solver::solve / solver_base::solve (solver_base.h:64-79): rho_host
 * (main_size complex128) is propagated in place over t_list; kwargs dt, atol,
 * rtol as in fixed_step_size_solver.h:37 / adaptive_step_size_solver.h:40-43.
 * t_list must be non-decreasing; under Dormand-Prince a zero-length interval
 * takes no attempt and keeps the step size (the reference would clamp the step
 * to 0 there and never finish the next interval).
 * full_copy = 1 copies the whole state to rho_host at every output time
 * (reference behaviour, solver_base.h:74-76); 0 copies rho_S per output and the
 * whole state only at the end. */
int heom_solve(heom_solver_handle s, double* rho_host, const double* t_list, int n_t,
               heom_callback cb, void* user, double dt, double atol, double rtol,
               int full_copy);
/* device-resident variant: rho_dev (main_size, complex128) is propagated in
 * place; rho_s_out (host, n_t x N*N complex128) receives rho_S at every t_i. */
int heom_solve_dev(heom_solver_handle s, void* rho_dev, const double* t_list, int n_t,
                   double dt, double atol, double rtol, double* rho_s_out);
/* No counterpart in the reference: heom_solve_dev for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * heom_solve_dev with its stage kernels launched directly (no graph replay, no
 * resident integrator), and no hipMalloc / hipFree / graph instantiation on the
 * call: the Dormand-Prince control block, the error-norm partials and rho_S of
 * every output time live in one fixed 8 MiB device array per GPU, held until
 * the process ends.  Limit: n_t * n_level^2 * 16 bytes of rho_S <= 6 MiB (for a
 * batch, n_level^2 is the sum over its members); larger requests return
 * HEOM_ERR_ARG.  The solver's own state buffers come from heom_solver_create.
 * Every call on one GPU shares that array: calls that overlap on one GPU (from
 * several host threads or streams) race; calls in sequence are independent.
 * HEOM_ERR_UNSUPPORTED when a closure network is bound. */
int heom_solve_dev_noalloc(heom_solver_handle s, void* rho_dev, const double* t_list, int n_t,
                           double dt, double atol, double rtol, double* rho_s_out);
int heom_solver_stats(heom_solver_handle s, heom_stats* out);
int heom_solver_reset(heom_solver_handle s); /* forget dt_save (adaptive_step_size_solver.h:28) */
/* 1 (default): replay the fixed stage sequences (RK4 steps of the full dt,
 * Dormand-Prince attempts) from HIP graphs, and run a small heom_hilb problem
 * (diagonal V, FP64, at most 1024 elements -- 960 for Dormand-Prince -- and at
 * most 8 links per ADO) or a small redfield_hilb problem (diagonal V, FP64,
 * its tables within LDS) in resident launches with the state in LDS: one per
 * RK4 / lsrk4 output interval, or one per heom_solve_dev call (RK4, lsrk4 and
 * Dormand-Prince with the step control on the device); 0: launch every stage
 * kernel directly.  Fixed-step results are identical either way; Dormand-Prince
 * sums its error norm in another order, so a borderline attempt may be decided
 * differently (the states agree within the tolerance).  No counterpart in the
 * reference. */
int heom_solver_set_graphs(heom_solver_handle s, int enable);

/* measurement helpers: average device duration (ms, HIP events on the
 * library stream) of each RK stage kernel, replayed `reps` times on the
 * current state; n_stage receives 4 (rk4) or 7 (rkdp). */
int heom_solver_time_stages(heom_solver_handle s, int reps, double* ms_per_stage,
                            int* n_stage);
/* algorithmic HBM bytes moved by each stage kernel (same order) */
int heom_solver_stage_bytes(heom_solver_handle s, double* bytes_per_stage, int* n_stage);

/* device memory helpers for callers without their own allocator */
int heom_dev_alloc(void** ptr, long long bytes);
int heom_dev_free(void* ptr);
int heom_memcpy_h2d(void* dst_dev, const void* src_host, long long bytes);
int heom_memcpy_d2h(void* dst_host, const void* src_dev, long long bytes);
int heom_device_sync(void);
/* Measurement helper, no counterpart in the reference: a grid-stride copy of
 * `bytes` between two fresh device buffers (grid x 256 threads, 16-byte
 * elements), timed with HIP events over `reps` launches after one warm-up;
 * *gbs = (bytes read + bytes written) per second / 1e9. */
int heom_stream_copy(long long bytes, int grid, int reps, double* gbs);

#ifdef __cplusplus
}
#endif

#endif /* LIBHEOM_AMD_H */
