// C ABI (include/qlles.h): argument validation, error codes, thread-local
// error message.  No C++ exception crosses this boundary.
#include <cstring>

#include "capi_common.hpp"

namespace {
using qlapi::fail;
using qlapi::guarded;

bool grid_ok(int dim, int N) {
  if (dim == 3) return N == 8 || N == 16 || N == 32 || N == 64 || N == 96 || N == 128 || N == 256;
  if (dim == 2) return N == 8 || N == 16 || N == 32 || N == 64 || N == 96 || N == 128 || N == 256 || N == 512;
  return false;
}
}  // namespace

extern "C" {

const char* qlles_last_error(void) { return qlapi::g_err.c_str(); }

int qlles_plan_create(const qlles_desc* d, qlles_plan** out) {
  if (!d || !out) return fail(QLLES_ERR_ARG, "null argument");
  *out = nullptr;
  if (!grid_ok(d->dim, d->N))
    return fail(QLLES_ERR_SHAPE, "unsupported grid: dim=" + std::to_string(d->dim) + " N=" + std::to_string(d->N));
  if (d->dtype != QLLES_F32 && d->dtype != QLLES_F64) return fail(QLLES_ERR_ARG, "dtype must be QLLES_F32 or QLLES_F64");
  if (!(d->nu >= 0.0) || !(d->dt > 0.0)) return fail(QLLES_ERR_ARG, "need nu >= 0 and dt > 0");
  if (d->forcing < QLLES_FORCING_NONE || d->forcing > QLLES_FORCING_KOLMOGOROV)
    return fail(QLLES_ERR_ARG, "unknown forcing type");
  if ((d->forcing == QLLES_FORCING_EKIN_TOT || d->forcing == QLLES_FORCING_POWER) && (d->kf <= 0 || d->kf >= 16))
    return fail(QLLES_ERR_ARG, "forcing shell radius kf must be in [1, 15]");
  if ((d->forcing == QLLES_FORCING_EKIN_TOT || d->forcing == QLLES_FORCING_POWER) && 2 * d->kf >= d->N)
    return fail(QLLES_ERR_ARG, "forcing shell radius kf must be below N/2");
  if (d->forcing == QLLES_FORCING_EKIN_TOT && d->dim != 3)
    return fail(QLLES_ERR_ARG, "ekin_tot forcing is defined for 3-D HIT only");
  if (d->forcing == QLLES_FORCING_KOLMOGOROV && (d->dim != 2 || d->kolm_k <= 0 || 2 * d->kolm_k >= d->N))
    return fail(QLLES_ERR_ARG, "Kolmogorov forcing needs dim=2 and 0 < kolm_k < N/2");
  if (d->viscous != QLLES_VISC_EXPLICIT && d->viscous != QLLES_VISC_IF) return fail(QLLES_ERR_ARG, "unknown viscous mode");
  return guarded([&] {
    ql::Desc q{d->dim, d->N, d->dtype, d->nu, d->dt, d->forcing, d->kf, d->e_kin_init, d->power,
               d->kolm_k, d->kolm_amp, d->drag, d->viscous};
    ql::Engine* e = ql::make_engine(q);
    if (!e) throw std::invalid_argument("no engine for this configuration");
    *out = new qlles_plan{q, e};
  });
}

int qlles_plan_destroy(qlles_plan* p) {
  if (!p) return QLLES_OK;
  return guarded([&] {
    delete p->eng;
    delete p;
  });
}

size_t qlles_plan_workspace_bytes(const qlles_plan* p) { return p ? p->eng->workspace_bytes() : 0; }

int qlles_step(qlles_plan* p, void* u_out, void* u_hat, const void* u_in, int nsteps, double* e_inj, void* stream) {
  if (!p || !u_hat) return fail(QLLES_ERR_ARG, "null plan or u_hat");
  if (nsteps < 0) return fail(QLLES_ERR_ARG, "nsteps must be >= 0");
  return guarded([&] {
    if (nsteps == 0) {
      if (u_out) p->eng->irfftn(u_hat, u_out, (hipStream_t)stream);
      return;
    }
    p->eng->step(u_out, u_hat, u_in, nsteps, e_inj, (hipStream_t)stream);
  });
}

int qlles_step_advect(qlles_plan* p, qlles_interp* it, void* u_out, void* u_hat, int nsteps, double* r, int np,
                      void* u_r, double* e_inj, void* stream) {
  if (!p || !it || !u_hat || !r || !u_r) return fail(QLLES_ERR_ARG, "null argument");
  if (nsteps < 0 || np < 0) return fail(QLLES_ERR_ARG, "nsteps and np must be >= 0");
  if (ql::interp_dim(it) != p->desc.dim || ql::interp_dtype(it) != p->desc.dtype || ql::interp_n(it) > p->desc.N)
    return fail(QLLES_ERR_ARG, "interpolator does not match the plan (dim, dtype, n <= N)");
  if (nsteps == 0) return QLLES_OK;
  return guarded([&] { qlapi::run_steps(p, it, u_out, u_hat, nsteps, r, np, u_r, e_inj, (hipStream_t)stream); });
}

int qlles_rhs(qlles_plan* p, const void* u, const void* u_hat, int rk, void* du_hat, void* stream) {
  if (!p || !u_hat || !du_hat) return fail(QLLES_ERR_ARG, "null argument");
  if (rk < 0 || rk > 3) return fail(QLLES_ERR_ARG, "rk must be in 0..3");
  return guarded([&] { p->eng->rhs(u, u_hat, rk, du_hat, (hipStream_t)stream); });
}

int qlles_rfftn(qlles_plan* p, const void* u, void* u_hat, void* stream) {
  if (!p || !u || !u_hat) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->rfftn(u, u_hat, (hipStream_t)stream); });
}

int qlles_irfftn(qlles_plan* p, const void* u_hat, void* u, void* stream) {
  if (!p || !u || !u_hat) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->irfftn(u_hat, u, (hipStream_t)stream); });
}

int qlles_diagnostics(qlles_plan* p, const void* u, double* out, void* stream) {
  if (!p || !u || !out) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->diagnostics(u, out, (hipStream_t)stream); });
}

int qlles_dissipation(qlles_plan* p, const void* u_hat, double* out, void* stream) {
  if (!p || !u_hat || !out) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->dissipation(u_hat, out, (hipStream_t)stream); });
}

int qlles_energy_hat(qlles_plan* p, const void* u_hat, double* out, void* stream) {
  if (!p || !u_hat || !out) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->energy_hat(u_hat, out, (hipStream_t)stream); });
}

int qlles_filter(qlles_plan* p, const void* u_hat, int ckp_N, void* u_lres, void* stream) {
  if (!p || !u_hat || !u_lres) return fail(QLLES_ERR_ARG, "null argument");
  if (ckp_N <= 0 || ckp_N > p->desc.N || (ckp_N % 2) != 0)
    return fail(QLLES_ERR_ARG, "ckp_N must be even and <= N (combined.py:123-126)");
  return guarded([&] { p->eng->filter(u_hat, ckp_N, u_lres, (hipStream_t)stream); });
}

int qlles_energy_spectrum(qlles_plan* p, const void* u, double mul_fac, double* ek, void* stream) {
  if (!p || !u || !ek) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->spectrum(u, mul_fac, ek, (hipStream_t)stream); });
}

int qlles_plan_set_variant(qlles_plan* p, int xv, int zv) {
  if (!p) return fail(QLLES_ERR_ARG, "null plan");
  return guarded([&] { p->eng->set_variant(xv, zv); });
}

int qlles_plan_get_variant(const qlles_plan* p, int* xv, int* zv, int* nvar) {
  if (!p) return fail(QLLES_ERR_ARG, "null plan");
  p->eng->get_variant(xv, zv);
  if (nvar) *nvar = p->eng->variants();
  return QLLES_OK;
}

int qlles_plan_tune(qlles_plan* p, const void* u_hat, int trial_steps, double* best_us, void* stream) {
  if (!p || !u_hat) return fail(QLLES_ERR_ARG, "null argument");
  if (trial_steps < 1) return fail(QLLES_ERR_ARG, "trial_steps must be >= 1");
  return guarded([&] { qlapi::tune_plan(p, u_hat, trial_steps, best_us, (hipStream_t)stream); });
}

int qlles_rhs_jvp(qlles_plan* p, const void* u_hat, const void* v, void* out, void* stream) {
  if (!p || !u_hat || !v || !out) return fail(QLLES_ERR_ARG, "null argument");
  if (out == u_hat || out == v) return fail(QLLES_ERR_ARG, "out may not alias an input");
  return guarded([&] { p->eng->rhs_jvp(u_hat, v, out, (hipStream_t)stream); });
}

int qlles_rhs_vjp(qlles_plan* p, const void* u_hat, const void* lam, void* out, double* dpower, void* stream) {
  if (!p || !u_hat || !lam || !out) return fail(QLLES_ERR_ARG, "null argument");
  if (out == u_hat || out == lam) return fail(QLLES_ERR_ARG, "out may not alias an input");
  return guarded([&] { p->eng->rhs_vjp(u_hat, lam, out, dpower, (hipStream_t)stream); });
}

int qlles_step_jvp(qlles_plan* p, const void* u_hat, const void* v, void* out, void* stream) {
  if (!p || !u_hat || !v || !out) return fail(QLLES_ERR_ARG, "null argument");
  if (out == u_hat || out == v) return fail(QLLES_ERR_ARG, "out may not alias an input");
  return guarded([&] { p->eng->step_jvp(u_hat, v, out, (hipStream_t)stream); });
}

int qlles_step_vjp(qlles_plan* p, const void* u_hat, const void* lam, void* out, double* dpower, void* stream) {
  if (!p || !u_hat || !lam || !out) return fail(QLLES_ERR_ARG, "null argument");
  if (out == u_hat || out == lam) return fail(QLLES_ERR_ARG, "out may not alias an input");
  return guarded([&] { p->eng->step_vjp(u_hat, lam, out, dpower, (hipStream_t)stream); });
}

size_t qlles_step_record_bytes(const qlles_plan* p) { return p ? p->eng->record_bytes() : 0; }

int qlles_step_record(qlles_plan* p, const void* u_hat, void* u_hat_out, void* record, void* stream) {
  if (!p || !u_hat || !u_hat_out || !record) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->step_record(u_hat, u_hat_out, record, (hipStream_t)stream); });
}

int qlles_step_vjp_record(qlles_plan* p, const void* u_hat, const void* record, const void* lam, void* out,
                          double* dpower, void* stream) {
  if (!p || !u_hat || !record || !lam || !out) return fail(QLLES_ERR_ARG, "null argument");
  if (out == u_hat || out == lam) return fail(QLLES_ERR_ARG, "out may not alias an input");
  return guarded([&] { p->eng->step_vjp_record(u_hat, record, lam, out, dpower, (hipStream_t)stream); });
}

int qlles_irfftn_adjoint(qlles_plan* p, const void* g, void* out, void* stream) {
  if (!p || !g || !out) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->irfftn_adjoint(g, out, (hipStream_t)stream); });
}

int qlles_rfftn_adjoint(qlles_plan* p, const void* g_hat, void* out, void* stream) {
  if (!p || !g_hat || !out) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] { p->eng->rfftn_adjoint(g_hat, out, (hipStream_t)stream); });
}

int qlles_profile(qlles_plan* p, int enable) {
  if (!p) return fail(QLLES_ERR_ARG, "null plan");
  return guarded([&] { p->eng->profile_enable(enable != 0); });
}

int qlles_profile_read(qlles_plan* p, double* ms, int* count, double* bytes) {
  if (!p) return fail(QLLES_ERR_ARG, "null plan");
  return guarded([&] { p->eng->profile_read(ms, count, bytes); });
}

int qlles_interp_create(int dim, int n, int dtype, int backend, double tol, qlles_interp** out) {
  if (!out) return fail(QLLES_ERR_ARG, "null argument");
  *out = nullptr;
  if (dim != 2 && dim != 3) return fail(QLLES_ERR_SHAPE, "dim must be 2 or 3");
  if (dtype != QLLES_F32 && dtype != QLLES_F64) return fail(QLLES_ERR_ARG, "bad dtype");
  if (backend != QLLES_INTERP_NUFFT && backend != QLLES_INTERP_DFT) return fail(QLLES_ERR_ARG, "bad backend");
  if (n < 4 || (n % 2) != 0 || n > 256) return fail(QLLES_ERR_SHAPE, "interpolation grid n must be even, 4..256");
  if (backend == QLLES_INTERP_NUFFT && !(tol > 0.0 && tol < 0.1)) return fail(QLLES_ERR_ARG, "tol must be in (0, 0.1)");
  return guarded([&] { *out = ql::make_interp(dim, n, dtype, backend, tol); });
}

int qlles_interp_destroy(qlles_interp* it) {
  if (!it) return QLLES_OK;
  return guarded([&] { ql::destroy_interp(it); });
}

int qlles_interp_field(qlles_interp* it, const void* u, const double* r, int np, void* u_r, void* stream) {
  if (!it || !u || !r || !u_r) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 0) return fail(QLLES_ERR_ARG, "np must be >= 0");
  return guarded([&] { ql::interp_field(it, u, r, np, u_r, (hipStream_t)stream); });
}

int qlles_interp_hat(qlles_interp* it, int N, const void* u_hat, double* r, int np, void* u_r, double dt,
                     void* stream) {
  if (!it || !u_hat || !r || !u_r) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 0) return fail(QLLES_ERR_ARG, "np must be >= 0");
  return guarded([&] { ql::interp_hat(it, N, u_hat, N / 2 + 1, 0, r, np, u_r, dt, (hipStream_t)stream); });
}

int qlles_interp_hat_vjp(qlles_interp* it, int N, const void* u_hat, const double* r, int np, const void* u_bar,
                         void* lam_hat, double* r_bar, void* stream) {
  if (!it || !u_hat || !r || !u_bar || !lam_hat) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 0) return fail(QLLES_ERR_ARG, "np must be >= 0");
  if (lam_hat == u_hat) return fail(QLLES_ERR_ARG, "lam_hat may not alias u_hat");
  return guarded([&] { ql::interp_hat_vjp(it, N, u_hat, r, np, u_bar, lam_hat, r_bar, (hipStream_t)stream); });
}

}  // extern "C"
