// Shared pieces of the C ABI implementation: the plan object, the
// thread-local error message and the exception guard.
#pragma once
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/qlles.h"
#include "engine.hpp"
#include "fftlib.hpp"
#include "interp.hpp"

struct qlles_plan {
  ql::Desc desc;
  ql::Engine* eng;
};

namespace qlapi {
inline thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return QLLES_OK;
  } catch (const ql::HipError& e) {
    return fail(QLLES_ERR_HIP, e.what());
  } catch (const std::invalid_argument& e) {
    return fail(QLLES_ERR_ARG, e.what());
  } catch (const std::bad_alloc&) {
    return fail(QLLES_ERR_HIP, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(QLLES_ERR_ARG, e.what());
  }
}

// nsteps of the combined-mode loop body on one plan (combined.py:97,127,130,199);
// without an interpolator, nsteps plain RK4 steps
inline void run_steps(qlles_plan* p, qlles_interp* it, void* u_out, void* u_hat, int nsteps, double* r, int np,
                      void* u_r, double* e_inj, hipStream_t s) {
  if (it == nullptr) {
    p->eng->step(u_out, u_hat, nullptr, nsteps, e_inj, s);
    return;
  }
  const int N = p->desc.N;
  const double dt = p->desc.dt;
  struct Reset {
    ql::Engine* e;
    ~Reset() { e->after_step = nullptr; }
  } reset{p->eng};
  p->eng->after_step = [=](hipStream_t st, const void* uh, int ld, size_t q) {
    ql::interp_hat(it, N, uh, ld, q, r, np, u_r, dt, st);
  };
  p->eng->step(u_out, u_hat, nullptr, nsteps, e_inj, s);
}
// time every kernel variant of the plan over `trials` steps on a scratch copy
// of u_hat (the caller's state does not move) and keep the fastest
inline void tune_plan(qlles_plan* p, const void* u_hat, int trials, double* best_us, hipStream_t s) {
  const int N = p->desc.N, dim = p->desc.dim;
  const size_t rows = dim == 3 ? (size_t)N * N : (size_t)N;
  const size_t bytes = 3 * rows * (N / 2 + 1) * 2 * (p->desc.dtype == QLLES_F32 ? 4 : 8);
  void* tmp = nullptr;
  QL_HIP(hipMalloc(&tmp, bytes));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int bx = 0, bz = 0, ox = 0, oz = 0;
  p->eng->get_variant(&ox, &oz);
  double best = -1.0;
  try {
    QL_HIP(hipEventCreate(&e0));
    QL_HIP(hipEventCreate(&e1));
    const int nv = p->eng->variants();
    for (int xv = 0; xv < nv; ++xv)
      for (int zv = 0; zv < nv; ++zv) {
        p->eng->set_variant(xv, zv);
        QL_HIP(hipMemcpyAsync(tmp, u_hat, bytes, hipMemcpyDeviceToDevice, s));
        p->eng->step(nullptr, tmp, nullptr, 1, nullptr, s);  // warm-up
        QL_HIP(hipEventRecord(e0, s));
        p->eng->step(nullptr, tmp, nullptr, trials, nullptr, s);
        QL_HIP(hipEventRecord(e1, s));
        QL_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        QL_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1e3 * ms / trials;
        if (best < 0 || us < best) {
          best = us;
          bx = xv;
          bz = zv;
        }
      }
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipFree(tmp);
    p->eng->set_variant(ox, oz);  // leave the plan as it was
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s);
  (void)hipFree(tmp);
  p->eng->set_variant(bx, bz);
  if (best_us) *best_us = best;
}
}  // namespace qlapi
