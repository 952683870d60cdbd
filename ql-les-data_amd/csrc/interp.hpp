// Particle interpolation of the spectral velocity (integrator.py:26-93).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/qlles.h"
#include "engine.hpp"

namespace ql {
qlles_interp* make_interp(int dim, int n, int dtype, int backend, double tol);
void destroy_interp(qlles_interp* it);
void interp_field(qlles_interp* it, const void* u, const double* r, int np, void* u_r, hipStream_t s);
int interp_dim(const qlles_interp* it);
int interp_dtype(const qlles_interp* it);
int interp_n(const qlles_interp* it);
// u_hat: rfft-layout spectrum of the N-grid field, rows of ld complex; the
// kz = N/2 entries sit in column N/2 of the rows (q == 0, ld >= N/2+1) or,
// for a 3-D field, in a separate [ncomp][N][N] plane at element q (ld >= N/2)
void interp_hat(qlles_interp* it, int N, const void* u_hat, int ld, size_t q, double* r, int np, void* u_r,
                double dt, hipStream_t s);
// adjoint of u_r = interp_hat(u_hat, r) (3-D NUFFT): lam_hat += (du_r/du_hat)^T u_bar,
// r_bar (optional) += (du_r/dr)^T u_bar
void interp_hat_vjp(qlles_interp* it, int N, const void* u_hat, const double* r, int np, const void* u_bar,
                    void* lam_hat, double* r_bar, hipStream_t s);

// Particle advection of a group of batch members (see EngineGroup): one
// filter + type-2 NUFFT + shift for all members stepping, with the members'
// coefficient and fine grids side by side so the fine-grid transforms are one
// batched call.  Members share dim, n, precision and kernel (NUFFT backend);
// member i's particles r[i] (np[i], FP64, updated in place when dt[i] != 0)
// and velocities ur[i] are its own, as is its interpolator's correction
// table.  Results equal interp_hat's for each member bit for bit.
struct InterpGroup;
bool interp_groupable(const qlles_interp* a, const qlles_interp* b);
InterpGroup* make_interp_group(const std::vector<qlles_interp*>& its, int N, const std::vector<SpecRef>& cont,
                               const std::vector<SpecRef>& last, const std::vector<double*>& r,
                               const std::vector<int>& np, const std::vector<void*>& ur,
                               const std::vector<double>& dt);
// members [0, f) read their spec_cont spectrum, [f, m) their spec_last one
void interp_group_run(InterpGroup* g, int m, int f, hipStream_t s);
void destroy_interp_group(InterpGroup* g);
}  // namespace ql
