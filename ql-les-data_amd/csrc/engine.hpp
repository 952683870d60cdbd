// Engine: one solver configuration (dim, N, precision, nu, dt, forcing) with
// its device workspace; the C ABI plan wraps one of these.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <functional>
#include <vector>

namespace ql {

struct Desc {
  int dim, N, dtype;
  double nu, dt;
  int forcing, kf;
  double e_kin_init, power;
  int kolm_k;
  double kolm_amp, drag;
  int viscous;
};

// Where a member's spectrum of u_{n+1} sits after a step of a group (see
// EngineGroup): rows of ld complex entries, the kz = N/2 entries in column
// N/2 (q == 0) or in a separate plane at element q.
struct SpecRef {
  const void* p;
  int ld;
  size_t q;
};

// Engines of one kind -- same dim, N, precision, viscous mode and kernel
// variants, forcing other than ekin_tot -- stepped together: each kernel of
// the RK4 step runs once for all members still stepping, z-slice m of its
// grid running member m with the arguments of entry m of a device table.
// Member order is fixed when the group is made; the step counts given to
// steps() must not increase along it, so the members still stepping at any
// step are a prefix.  Kernels, arguments and reduction orders are those of
// Engine::step, so every member's trajectory is bit-identical to its own.
class EngineGroup {
 public:
  virtual ~EngineGroup() {}
  virtual int size() const = 0;
  // steps [from, to) of a run in which member i takes n[i] steps; from == 0
  // first loads the states.  after(s, m, f) follows each step: members
  // [0, f) go on stepping, members [f, m) have just taken their last step.
  virtual void steps(const int* n, int from, int to, hipStream_t s,
                     const std::function<void(hipStream_t, int, int)>& after) = 0;
  // member i's u_{n+1} after a step that is not (cont) / is (last) its last
  virtual SpecRef spec_cont(int i) const = 0;
  virtual SpecRef spec_last(int i) const = 0;
};

class Engine {
 public:
  virtual ~Engine() {}
  virtual size_t workspace_bytes() const = 0;
  // allocate the step workspace now (it is otherwise allocated by the first
  // step, RHS, tangent or adjoint call): done before a graph capture
  virtual void prepare() = 0;
  virtual void step(void* u_out, void* uhat, const void* u_in, int nsteps, double* e_inj, hipStream_t s) = 0;
  virtual void rhs(const void* u, const void* uhat, int rk, void* du, hipStream_t s) = 0;
  virtual void rfftn(const void* u, void* uhat, hipStream_t s) = 0;
  virtual void irfftn(const void* uhat, void* u, hipStream_t s) = 0;
  virtual void diagnostics(const void* u, double* out, hipStream_t s) = 0;
  virtual void dissipation(const void* uhat, double* out, hipStream_t s) = 0;
  virtual void energy_hat(const void* uhat, double* out, hipStream_t s) = 0;
  virtual void filter(const void* uhat, int n, void* ulres, hipStream_t s) = 0;
  // shell-binned energy spectrum of the physical field u into ek[0..N) (FP64)
  virtual void spectrum(const void* u, double mul_fac, double* ek, hipStream_t s) = 0;
  // kernel variants: xv = x / y-pass tile width, zv = z-pass (2-D y-pass)
  // radix plan; 2 choices each for grids up to 128, 1 above
  virtual int variants() const = 0;
  virtual void set_variant(int xv, int zv) = 0;
  virtual void get_variant(int* xv, int* zv) const = 0;
  // whether o can join a group with this engine, and the group of this engine
  // (first member) and others of its kind stepping the states uhat[i]
  virtual bool groupable_with(const Engine* o) const = 0;
  virtual EngineGroup* make_group(const std::vector<Engine*>& members, const std::vector<void*>& uhat) = 0;
  // tangent-linear and adjoint maps of the RHS and of one RK4 step at u_hat
  // (3-D, explicit viscous term, forcing none or fixed power); inputs and
  // outputs are [3][N][N][N/2+1] spectra and may not alias each other.
  // dpow (device double, optional) accumulates the cotangent of the power.
  virtual void rhs_jvp(const void* uhat, const void* v, void* out, hipStream_t s) = 0;
  virtual void rhs_vjp(const void* uhat, const void* lam, void* out, double* dpow, hipStream_t s) = 0;
  virtual void step_jvp(const void* uhat, const void* v, void* out, hipStream_t s) = 0;
  virtual void step_vjp(const void* uhat, const void* lam, void* out, double* dpow, hipStream_t s) = 0;
  // one RK4 step taken stage by stage, keeping what step_vjp_record needs
  virtual size_t record_bytes() const = 0;
  virtual void step_record(const void* uhat, void* uhat_out, void* rec, hipStream_t s) = 0;
  virtual void step_vjp_record(const void* uhat, const void* rec, const void* lam, void* out, double* dpow,
                               hipStream_t s) = 0;
  // adjoints of irfftn (c/N^d rfftn) and rfftn (N^d irfftn(. / c))
  virtual void irfftn_adjoint(const void* g, void* out, hipStream_t s) = 0;
  virtual void rfftn_adjoint(const void* gh, void* out, hipStream_t s) = 0;

  // per-kernel-kind timing with HIP events on the launch stream, plus the
  // algorithmic bytes each launch must move (reads + writes of its arrays)
  enum Kind { K_XPASS = 0, K_YINV = 1, K_ZPASS = 2, K_YFWD = 3, K_OTHER = 4, NKIND = 5 };
  void profile_enable(bool on);
  bool profiling() const { return prof_; }
  // called on the step's stream after every completed step (particle advection)
  // (stream, spectrum of u_{n+1}, its row stride in complex elements, offset
  // of its separate Nyquist plane or 0 when kz = N/2 sits in the rows)
  std::function<void(hipStream_t, const void*, int, size_t)> after_step;
  void profile_read(double* ms, int* count, double* bytes);

 protected:
  struct Rec {
    int kind;
    double bytes;
    hipEvent_t a, b;
  };
  bool prof_ = false;
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> pool_;
  double acc_ms_[NKIND] = {0, 0, 0, 0, 0};
  double acc_bytes_[NKIND] = {0, 0, 0, 0, 0};
  int acc_count_[NKIND] = {0, 0, 0, 0, 0};
  hipEvent_t take_event();
  void prof_begin(hipStream_t s, hipEvent_t* a);
  void prof_end(hipStream_t s, int kind, double bytes, hipEvent_t a);
  void prof_drain();
};

Engine* make_engine(const Desc& d);
template <typename T, int DIM, int N> Engine* make_engine_t(const Desc& d);

// dealias bookkeeping shared with the host: kz kept iff kz < 2/3 (N/2+1)
// evaluated in double exactly as spectral_solver.py:42 does
inline int dealias_count(int N) {
  const double kmax = 2.0 / 3.0 * (double)(N / 2 + 1);
  int c = 0;
  for (int k = 0; k <= N / 2; ++k)
    if ((double)k < kmax) ++c;
  return c;
}
inline int round_up(int a, int m) { return (a + m - 1) / m * m; }

}  // namespace ql
