/*
 * qlles.h -- C ABI of the MI355X-native spectral Navier-Stokes step and
 * particle interpolation (gfx950, HIP).
 *
 * The reference (tumaer/ql-les-data, package l3es) exposes this path as
 * closure factories over static parameters that return pure array functions;
 * each entry point below replaces one of them.  Citations are file:line in
 * that repository.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers owned by the caller, in the
 *    reference's C-order, component-major layouts:
 *        3-D: u (3,N,N,N) real, u_hat (3,N,N,N/2+1) complex (interleaved re,im)
 *        2-D: u (3,N,N,1) real, u_hat (3,N,N/2+1)   complex, component 2 == 0
 *    Element type is float or double as chosen by desc.dtype (complex =
 *    2 consecutive elements).
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *    stream-ordered and asynchronous; they return once work is enqueued.
 *  - Return value: QLLES_OK or an error code; qlles_last_error() gives a
 *    thread-local message.  No exceptions cross the ABI.
 *  - A plan owns its device workspace.  A plan is not thread-safe; separate
 *    plans on separate streams are.
 */
#ifndef QLLES_H
#define QLLES_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QLLES_OK 0
#define QLLES_ERR_ARG 1         /* bad argument / unsupported combination */
#define QLLES_ERR_SHAPE 2       /* unsupported N / dim */
#define QLLES_ERR_HIP 3         /* HIP runtime error */
#define QLLES_ERR_NONFINITE 4   /* non-finite value detected by a check */

#define QLLES_F32 0
#define QLLES_F64 1

#define QLLES_FORCING_NONE 0
#define QLLES_FORCING_EKIN_TOT 1   /* spectral_solver.py:82-112 */
#define QLLES_FORCING_POWER 2      /* fixed injection power on 0<|k|<=kf (no reference counterpart) */
#define QLLES_FORCING_KOLMOGOROV 3 /* f_x = amp sin(k y), linear drag (no reference counterpart) */

#define QLLES_VISC_EXPLICIT 0      /* -nu k^2 u_hat inside RK4, spectral_solver.py:59 */
#define QLLES_VISC_IF 1            /* integrating factor exp(-nu k^2 t) (no reference counterpart) */

typedef struct qlles_plan qlles_plan;

/* Static parameters of rhs_wrapper(N, nu, fft_axes) and
 * rk4_wrapper(dt, rhs_fn, forcing_mask, e_kin_init, forcing_type)
 * (spectral_solver.py:27,65). */
typedef struct qlles_desc {
  int dim;            /* 2 or 3 (fft_axes (1,2) or (1,2,3)) */
  int N;              /* grid points per axis: 8,16,32,64,96,128,256 (3-D); also 512 (2-D) */
  int dtype;          /* QLLES_F32 or QLLES_F64 */
  double nu;          /* kinematic viscosity */
  double dt;          /* RK4 time step */
  int forcing;        /* QLLES_FORCING_* */
  int kf;             /* forcing shell radius (ekin_tot / power), 0 < kf < 16 */
  double e_kin_init;  /* ekin_tot target energy (spectral_solver.py:209) */
  double power;       /* fixed injection power */
  int kolm_k;         /* Kolmogorov forcing wavenumber */
  double kolm_amp;    /* Kolmogorov forcing amplitude */
  double drag;        /* linear drag coefficient */
  int viscous;        /* QLLES_VISC_* */
} qlles_desc;

/* Plan = rhs_wrapper + rk4_wrapper + custom_jit_integrate
 * (spectral_solver.py:27-48, :65-71, :130-153): validates the descriptor
 * and builds twiddle tables.  Device workspace is allocated on first use:
 * the reduction partials at creation, the transform scratch by the first
 * qlles_irfftn / qlles_filter / qlles_energy_spectrum / qlles_dissipation,
 * and the step workspace (1.85 GiB at 256^3 FP32) by the first qlles_step,
 * qlles_step_advect, qlles_rhs, qlles_plan_tune, tangent or adjoint call.
 * A caller that captures a plan's calls into a HIP graph makes one such call
 * eagerly first (qlles_batch_add does this for its members), so that no
 * allocation happens inside the capture.  A plan used only for qlles_rfftn,
 * qlles_diagnostics or qlles_energy_hat holds a few KiB. */
int qlles_plan_create(const qlles_desc* desc, qlles_plan** plan);
int qlles_plan_destroy(qlles_plan* plan);
/* bytes of device workspace the plan holds now */
size_t qlles_plan_workspace_bytes(const qlles_plan* plan);

/* Kernel variants (no reference counterpart; the reference leaves FFT
 * planning to XLA): xv picks the x / y-pass tile width (16 or 8 columns) and
 * zv the radix plan of the z-pass (2-D: y-pass) half-length FFTs.  Grids up
 * to 128 have 2 choices each, larger grids 1 (*nvar).  Variants change the
 * rounding of the FFTs, not the algorithm.  qlles_plan_tune times every
 * (xv, zv) over trial_steps steps on a scratch copy of u_hat (u_hat does not
 * move) and keeps the fastest; *best_us (optional) = its time per step. */
int qlles_plan_set_variant(qlles_plan* plan, int xv, int zv);
int qlles_plan_get_variant(const qlles_plan* plan, int* xv, int* zv, int* nvar);
int qlles_plan_tune(qlles_plan* plan, const void* u_hat, int trial_steps, double* best_us, void* stream);

/* `nsteps` applications of step_fn(u, u_hat) -> (u, u_hat, e_inj)
 * (spectral_solver.py:73-118).  u_hat is updated in place.
 *   u_in   : optional physical field used by the first RK stage (the
 *            reference's stage 0 uses the passed-in u, :51); NULL = irfftn(u_hat)
 *   u_out  : optional; receives irfftn(u_hat) after the last step (:114-116)
 *   e_inj  : optional device array of nsteps doubles (injection rate, :109) */
int qlles_step(qlles_plan* plan, void* u_out, void* u_hat, const void* u_in, int nsteps,
               double* e_inj, void* stream);

/* The combined-mode hot loop (combined.py:97,127,130,199): `nsteps` times
 *   u_hat <- step(u_hat); u_r <- interpolate(spectral_filtering(irfftn(u_hat), n), r);
 *   r <- (r + dt u_r) mod 2 pi
 * with the interpolator `it` (n = it's grid, n <= N, same dim and dtype as the
 * plan).  r: (np, dim) FP64, updated in place; u_r: (np, dim) in the plan dtype
 * (the last step's velocities); u_out / e_inj as in qlles_step. */
struct qlles_interp;
int qlles_step_advect(qlles_plan* plan, struct qlles_interp* it, void* u_out, void* u_hat, int nsteps, double* r,
                      int np, void* u_r, double* e_inj, void* stream);

/* rhs_fn(u, u_hat, rk) -> du_hat (spectral_solver.py:50-60).  For rk == 0
 * the physical u is used when non-NULL; otherwise u = irfftn(u_hat). */
int qlles_rhs(qlles_plan* plan, const void* u, const void* u_hat, int rk, void* du_hat, void* stream);

/* rfftn(u, axes=fft_axes) (spectral_solver.py:206) and irfftn(u_hat)
 * (:114); unnormalized forward, 1/N^dim inverse as numpy. */
int qlles_rfftn(qlles_plan* plan, const void* u, void* u_hat, void* stream);
int qlles_irfftn(qlles_plan* plan, const void* u_hat, void* u, void* stream);

/* FP64 diagnostics of a physical field u (spectral_solver.py:318-320,
 * :123-127): out[0] = 0.5 <|u|^2>, out[1] = max|u|, out[2..4] = max|u_i| per
 * component.  `out` is a device array of 5 doubles. */
int qlles_diagnostics(qlles_plan* plan, const void* u, double* out, void* stream);

/* FP64 dissipation nu <|curl u|^2> through irfftn of i k x u_hat
 * (utils.py:300-315); out is a device double. */
int qlles_dissipation(qlles_plan* plan, const void* u_hat, double* out, void* stream);

/* FP64 Parseval energy of u_hat: out[0] = 0.5 <|u|^2>, out[1] = same over
 * the forcing shell 0 < |k|^2 <= kf^2 (spectral_solver.py:94,:98). */
int qlles_energy_hat(qlles_plan* plan, const void* u_hat, double* out, void* stream);

/* energy_spectrum(vel, mul_fac, dim) (utils.py:35-97): shell-binned E(k) of
 * the physical field u (3-D (3,N,N,N); 2-D (3,N,N,1), components 0-1 used)
 * into ek[0..N), a device array of N doubles; FP64 binning with the
 * reference's half-space weights (2 inside the halved axis, 1 on its 0 and
 * N/2 planes), shell = int(|k| + 0.5), ek *= 4 pi k^2 / (samples + eps). */
int qlles_energy_spectrum(qlles_plan* plan, const void* u, double mul_fac, double* ek, void* stream);

/* spectral_filtering(u, ckp_N) (utils.py:100-138) evaluated directly from
 * u_hat: crop to [-n/2, n/2)^dim, scale (n/N)^dim, take the real part.
 * u_lres: (3,n,n,n) / (2,n,n) real, n = ckp_N <= N, n even. */
int qlles_filter(qlles_plan* plan, const void* u_hat, int ckp_N, void* u_lres, void* stream);

/* ---- tangent-linear and adjoint maps -----------------------------------------
 * No reference counterpart (the reference computes no gradients); they make
 * the step differentiable.  Defined for 3-D plans with the explicit viscous
 * term and forcing none or fixed power (the power term a(u) M u with
 * a = P / (2 E_f(u) + eps) is differentiated through E_f).  Inner products: sum_x a.b on physical fields and
 * Re sum_k conj(a_k) b_k over the stored half-spectrum (torch's convention
 * for complex gradients), so that for the map y -> f(y):
 *   <jvp(v), w> = <v, vjp(w)>.
 * All spectra are (3,N,N,N/2+1); `out` may not alias an input.
 *   qlles_rhs_jvp / _vjp : R'(u_hat) v and R'(u_hat)^T lam for
 *                          rhs_fn(u, u_hat, rk>0) (spectral_solver.py:50-60)
 *   qlles_step_jvp / _vjp: the same for one classical RK4 step
 *                          (spectral_solver.py:73-80) taken from u_hat
 *   qlles_irfftn_adjoint : adjoint of irfftn: out (spectrum) from g (physical)
 *   qlles_rfftn_adjoint  : adjoint of rfftn: out (physical) from g_hat (spectrum) */
int qlles_rhs_jvp(qlles_plan* plan, const void* u_hat, const void* v, void* out, void* stream);
/* dpower: optional device double; the VJPs ADD <lam, d(map)/dP> to it (the
 * cotangent of the injection power of a fixed-power plan) */
int qlles_rhs_vjp(qlles_plan* plan, const void* u_hat, const void* lam, void* out, double* dpower, void* stream);
int qlles_step_jvp(qlles_plan* plan, const void* u_hat, const void* v, void* out, void* stream);
int qlles_step_vjp(qlles_plan* plan, const void* u_hat, const void* lam, void* out, double* dpower, void* stream);
/* Recorded steps for the adjoint of a multi-step integration (no reference
 * counterpart).  qlles_step_record takes one step u_hat_out = step(u_hat)
 * with the kernels of qlles_step (bit-identical to it; classical RK4,
 * spectral_solver.py:70-80) and keeps its three stage states and the
 * physical u and omega of all four stage states in `record` (device memory
 * of qlles_step_record_bytes; u_hat_out may be u_hat).
 * qlles_step_vjp_record is qlles_step_vjp at that u_hat, reading the record
 * instead of recomputing the stages.  Same scope as qlles_step_vjp. */
size_t qlles_step_record_bytes(const qlles_plan* plan);
int qlles_step_record(qlles_plan* plan, const void* u_hat, void* u_hat_out, void* record, void* stream);
int qlles_step_vjp_record(qlles_plan* plan, const void* u_hat, const void* record, const void* lam, void* out,
                          double* dpower, void* stream);
int qlles_irfftn_adjoint(qlles_plan* plan, const void* g, void* out, void* stream);
int qlles_rfftn_adjoint(qlles_plan* plan, const void* g_hat, void* out, void* stream);

/* ---- measurement ----------------------------------------------------------
 * No reference counterpart (the reference times whole steps with wall-clock
 * around block_until_ready, spectral_solver.py:310-313).  When enabled, every
 * kernel launch of the step is bracketed by HIP events on its stream and
 * charged with the algorithmic bytes it must move.  qlles_profile_read
 * synchronizes the recorded events, returns per-kind totals and resets them.
 * Kinds: 0 x-pass (fused RK stage), 1 inverse y-pass, 2 z-pass (c2r / cross
 * product / r2c), 3 forward y-pass, 4 other. */
#define QLLES_NKIND 5
int qlles_profile(qlles_plan* plan, int enable);
int qlles_profile_read(qlles_plan* plan, double* ms, int* count, double* bytes);

/* ---- particle interpolation (integrator.py:26-93) ------------------------
 * An interpolator = spectral_interpolator_wrapper(N, fft_axes, splits,
 * backend) with N = the interpolation grid (ckp_N).  backend:
 *   QLLES_INTERP_NUFFT : type-2 NUFFT, exponential-of-semicircle kernel
 *                        (the reference calls finufft nufft{2,3}d2, isign=+1,
 *                        eps 1e-6, integrator.py:53,55)
 *   QLLES_INTERP_DFT   : exact direct sum (interpolate_dft, integrator.py:62-91)
 * Positions r are (np, dim) FP64; velocities u_r are (np, dim) in the plan
 * dtype. */
#define QLLES_INTERP_NUFFT 0
#define QLLES_INTERP_DFT 1

typedef struct qlles_interp qlles_interp;

int qlles_interp_create(int dim, int n, int dtype, int backend, double tol, qlles_interp** out);
int qlles_interp_destroy(qlles_interp* it);

/* interpolate(u, r) -> u_r from the physical (dim, n,...) field u on the
 * n-grid (integrator.py:36-60). */
int qlles_interp_field(qlles_interp* it, const void* u, const double* r, int np, void* u_r, void* stream);

/* Fused filter + interpolate + shift (combined.py:127,130,199):
 * u_r = interpolate(spectral_filtering(irfftn(u_hat), n), r) straight from the
 * N-grid spectrum, then, when dt != 0, r <- (r + dt u_r) mod 2 pi in FP64
 * (shift_fn, integrator.py:21-23).  N is the solver grid. */
int qlles_interp_hat(qlles_interp* it, int N, const void* u_hat, double* r, int np, void* u_r, double dt,
                     void* stream);

/* ---- batched trajectories ------------------------------------------------
 * Replaces the reference's one-trajectory-per-process SLURM arrays
 * (slurm_*.sh:5; each array task runs main.py -> simulate / combined on its
 * own GPU) for grids too small to fill an MI355X.  A batch holds members
 * (plan, optional interpolator + particles, state arrays); qlles_batch_run
 * advances member i by nsteps[i] steps (uneven lengths allowed), exactly as
 *   qlles_step_advect(plan, it, NULL, u_hat, nsteps[i], r, np, u_r, NULL, s)
 * (or qlles_step without particles) would alone: the same kernels in the
 * same order with fixed-order reductions, so every trajectory is
 * bit-identical to its solo run.  Per-member modes replay HIP graphs of
 * each member's steps over `lanes` streams; grouped modes step all members
 * of one kind (dim, N, precision, viscous mode, kernel variants and
 * interpolator settings; ekin_tot forcing runs alone) with one launch per
 * kernel of the step for the whole group.  Members need their own plan and
 * interpolator.  Not thread-safe. */
typedef struct qlles_batch qlles_batch;
int qlles_batch_create(qlles_batch** out);
int qlles_batch_destroy(qlles_batch* b);
/* r, u_r, np as in qlles_step_advect; it == NULL: no particles.  *index
 * (optional) receives the member's position. */
int qlles_batch_add(qlles_batch* b, qlles_plan* plan, qlles_interp* it, void* u_hat, double* r, int np, void* u_r,
                    int* index);
/* manual configuration: 1 <= lanes <= 16 streams, 1 <= chunk <= 1000 steps
 * per graph.  single_graph = 1 captures every member's chunk into ONE graph
 * (parallel branches over the lanes), launched once per chunk while every
 * member still needs a whole chunk; single_graph = 2 captures every member's
 * whole run (its own step count, uneven lengths included) into ONE graph, so
 * each qlles_batch_run is a single graph launch (captured again when the
 * step counts change; chunk is unused).  single_graph = 3: grouped launches
 * (each kernel of the step runs once per group of members of one kind; the
 * groups' streams are interleaved `chunk` steps at a time), eager;
 * single_graph = 4: the grouped run of every group captured into ONE graph,
 * one launch per qlles_batch_run (captured again when the step counts
 * change).  Grouped tables are remade when the order of the step counts
 * within a group or a member's kernel variants change. */
int qlles_batch_configure(qlles_batch* b, int lanes, int chunk, int single_graph);
/* autotuner: for each member, its plan's kernel variants (qlles_plan_tune);
 * then each kind's group alone, timed per step (grouped modes place groups on
 * streams longest first by these times); then single_graph in {0,1} x lanes
 * in {1,2,4,8,16} x chunk in {1,4,16,64}, single_graph = 3 x lanes in
 * {1,2,3,4,6,8,16} x chunk in {1,4,16} and single_graph = 4 x lanes (lanes up
 * to the number of groups),
 * timed over about trial_steps steps per member on scratch copies of the
 * member states (the trajectories do not move); keeps the fastest */
int qlles_batch_tune(qlles_batch* b, int trial_steps, void* stream);
/* current configuration and the tuned time per trajectory-step (0 if untuned) */
int qlles_batch_get_config(const qlles_batch* b, int* lanes, int* chunk, int* single_graph, double* us_per_traj_step);
/* the kinds the grouped modes step together: *nkinds, kind_of[i] (optional,
 * one per member) the kind of member i, and us_per_step[k] (optional, up to
 * cap entries) the time per step of kind k's group alone as the tuner
 * measured it (0 before tuning).  No reference counterpart (a scheduling
 * report). */
int qlles_batch_kinds(const qlles_batch* b, int* nkinds, int* kind_of, double* us_per_step, int cap);
/* nsteps: per-member step counts (NULL: every member takes nsteps_all).
 * Stream-ordered with respect to `stream` (work forks onto the batch's
 * streams and joins back). */
int qlles_batch_run(qlles_batch* b, const int* nsteps, int nsteps_all, void* stream);

/* Adjoint of qlles_interp_hat's interpolation u_r(u_hat, r) (dt = 0; 3-D,
 * NUFFT backend; no reference counterpart).  For the cotangent u_bar (np, 3)
 * in the plan dtype it ADDS
 *   lam_hat += (d u_r / d u_hat)^T u_bar   ((3,N,N,N/2+1) spectrum, the
 *              adjoint of the filter + type-2 NUFFT: type-1 spreading with
 *              the same kernel, forward fine-grid FFT, window adjoint)
 *   r_bar   += (d u_r / d r)^T u_bar       ((np, 3) FP64, optional)
 * with the inner products of qlles_step_vjp. */
int qlles_interp_hat_vjp(qlles_interp* it, int N, const void* u_hat, const double* r, int np, const void* u_bar,
                         void* lam_hat, double* r_bar, void* stream);

/* ---- SPH density and particle relaxation ---------------------------------
 * An SPH context = relax_wrapper(n, dim, box, is_physical, u_ref)
 * (relax.py:16-74) and rho_computer(n, dim, box) (utils.py:169-206): the
 * Morris quintic kernel with h = box/n (support 3h), particle mass
 * dx^dim rho_ref, Tait EoS p = p_eos (rho/rho_ref - 1) with
 * p_eos = (10 u_ref / l_ref)^2 rho_ref, l_ref = 1 (is_physical) or dx.
 * Neighbours are every particle, itself included, closer than 3h under the
 * periodic minimum image (the reference's jax_md neighbour list with
 * mask_self=False); a cell grid replaces the list.  Positions r are (np, dim)
 * FP64 in physical units; all outputs are FP64.  7 <= n <= 256 (3-D),
 * 4096 (2-D); np may differ from n^dim (the workspace grows on demand).
 * Results do not depend on run-to-run scheduling (fixed summation order). */
typedef struct qlles_sph qlles_sph;
int qlles_sph_create(int dim, int n, double box, int is_physical, double u_ref, qlles_sph** out);
int qlles_sph_destroy(qlles_sph* sph);
/* comp_rho(r) (utils.py:208-216): rho (np,) = m sum_j W(|r_ij|) */
int qlles_sph_density(qlles_sph* sph, const double* r, int np, double* rho, void* stream);
/* loop_body(r) (relax.py:76-120): acc (np, dim) = pressure + 0.5 background
 * (transport-velocity) acceleration in physical units; rho (optional) gets
 * the density it used */
int qlles_sph_relax_acc(qlles_sph* sph, const double* r, int np, double* acc, double* rho, void* stream);
/* sph_body(r, u) (relax.py:123-176; is_physical only): acc_p, acc_visc,
 * acc_tvf (np, dim) for velocities u (np, dim) FP64 */
int qlles_sph_forces(qlles_sph* sph, const double* r, const double* u, int np, double* acc_p, double* acc_visc,
                     double* acc_tvf, void* stream);
/* The relaxation branch of the combined loop plus its particle shift
 * (combined.py:132-142, :199; is_physical only), in place:
 *   r_t = shift(r, dt u_r); a = loop_body(r_t); r_t = shift(r_t, (factor dt)^2 a)
 *   u_r = displacement(r_t, r) / dt;  r = shift(r, dt u_r)
 * u_r (np, dim) is in u_dtype (QLLES_F32 / QLLES_F64; the shift uses the FP64
 * value); acc (optional) receives a. */
int qlles_sph_relax_advect(qlles_sph* sph, double* r, int np, void* u_r, int u_dtype, double dt, double factor,
                           double* acc, void* stream);

/* Particle -> grid transfers of the point-cloud spectrum (turbulence.py),
 * on the context's n-grid of side box; u_r (np, dim) FP64.
 * ur_to_u_mls_wrapper(n, box, dim)(r, u_r) (turbulence.py:145-295):
 * second-order moving least squares with the quintic kernel, h = 2/3 box/n,
 * at the nodes (i + 1/2) box/n; u_grid is (dim, n, n[, n]) FP64, 'ij' order.
 * A node with too few particles in reach has a singular system (NaN). */
int qlles_sph_mls_grid(qlles_sph* sph, const double* r, const double* u_r, int np, double* u_grid, void* stream);
/* ur_to_u_dft_wrapper (turbulence.py:298-336) up to its inverse transform:
 * u_hat receives the half spectrum of ifftn(U).real, U_c(k) = sum_j u_jc
 * exp(-i k.r_j) over the fftfreq modes, in the solver layout
 * (3, n, n, n/2+1) / (3, n, n/2+1) complex FP64 (component 2 of a 2-D field
 * zero); qlles_irfftn of an FP64 plan on the n-grid then gives the field on
 * the nodes m box/n. */
int qlles_sph_dft_hat(qlles_sph* sph, const double* r, const double* u_r, int np, void* u_hat, void* stream);

/* Dataset statistics (gen_dataset.py:143-239, compute_statistics_h5): the
 * moments of one trajectory.  r and u are (T, np, dim) FP32, C order, on the
 * device.  Per element, in FP32 as the reference computes them:
 *   vel_t = mod(d + box/2, box) - box/2, d = r_{t+1} - r_t   (T-1 frames;
 *           jax_md's periodic displacement, mod with the divisor's sign)
 *   acc_t = vel_{t+1} - vel_t                                  (T-2 frames)
 *   u_t                                                         (T frames)
 *   au_t  = u_{t+1} - u_t                                      (T-1 frames)
 * out (device, 8*dim FP64) receives, for the quantities q = vel, acc, u, au in
 * that order, the sums of the values (out[q*dim + c]) and then the sums of
 * their squared deviations from center (out[(4 + q)*dim + c]); center
 * (device, 4*dim FP64, same order) may be NULL for zeros.  Sums are FP64 in a
 * fixed order, so results are bitwise reproducible.  A context holds the
 * per-workgroup partial sums and is used from one stream at a time. */
typedef struct qlles_stats qlles_stats;
int qlles_stats_create(int dim, double box, qlles_stats** out);
int qlles_stats_destroy(qlles_stats* st);
int qlles_stats_moments(qlles_stats* st, const float* r, const float* u, long long T, long long np,
                        const double* center, double* out, void* stream);

/* qlles_stats_moments without a context and without allocating: no
 * counterpart in the reference.  It is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  dim and box are those of qlles_stats_create; the other
 * arguments, the launches and the summation order are those of
 * qlles_stats_moments, so the results are bitwise equal.  The per-workgroup
 * partial sums live in a static device array of QLLES_STATS_SCRATCH_DOUBLES
 * doubles (8 MiB) that each GPU the library runs on holds until the process
 * ends; a trajectory whose partials need more (8 dim doubles per workgroup
 * of 256 particles x up to 256 frames: 3840 workgroups for 7501 frames of
 * 32^3 particles) is refused with QLLES_ERR_ARG.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race and give wrong sums. */
#define QLLES_STATS_SCRATCH_DOUBLES (1 << 20)
int qlles_stats_moments_noalloc(int dim, double box, const float* r, const float* u, long long T, long long np,
                                const double* center, double* out, void* stream);

/* message of the last failed call on this thread */
const char* qlles_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* QLLES_H */
