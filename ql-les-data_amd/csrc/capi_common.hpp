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
}  // namespace qlapi
