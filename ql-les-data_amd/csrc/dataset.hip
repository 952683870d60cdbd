// Moments of particle trajectories for the dataset statistics
// (gen_dataset.py:143-239, compute_statistics_h5).
//
// One call reads a trajectory r, u of T frames x np particles x dim (FP32,
// C order) once and returns, per quantity and component, the FP64 sum of the
// values and of their squared deviations from a given centre.  The quantities
// are the ones the reference derives per frame pair, in FP32 as it does:
//   vel_t = periodic displacement(r_{t+1}, r_t)   gen_dataset.py:177
//   acc_t = vel_{t+1} - vel_t                     :180
//   u_t                                           :182
//   au_t  = u_{t+1} - u_t                         :183
// The periodic displacement is jax_md's space.periodic (:153):
// mod(d + box/2, box) - box/2 with d = r_{t+1} - r_t, mod taking the sign of
// the divisor.
//
// Layout of the work: a workgroup takes 256 consecutive particles over TC
// consecutive frames (plus the two frames after them), and each thread walks
// one particle through them in groups of ST_U frames whose loads are all in
// flight at once.  Consecutive lanes read consecutive particles, so a
// wavefront's frame load is 64*dim*4 contiguous bytes.  HBM sees each element
// about once; the frames shared by neighbouring groups come from the cache.
// Per-workgroup sums (wavefront xor butterflies, then the four wavefronts in
// order) go to a partials array, one row per summed quantity, and one
// workgroup per row adds it up in block order: the result does not depend on
// scheduling.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "capi_common.hpp"

using qlapi::fail;
using qlapi::guarded;

// Partial sums of qlles_stats_moments_noalloc: one array per GPU for the
// life of the process (8 MiB).  In host code `scratch` names a host-side
// stand-in, not a device address; the device address comes from
// hipGetSymbolAddress, once per device, and only that address is passed on.
static __device__ __align__(256) double scratch[QLLES_STATS_SCRATCH_DOUBLES];

namespace {

constexpr int ST_THREADS = 256;
constexpr int ST_WAVES = ST_THREADS / 64;
constexpr int ST_TC = 256;  // frames per workgroup (at most, unless T needs more)
#ifndef QL_ST_U
#define QL_ST_U 8
#endif
constexpr int ST_U = QL_ST_U;  // frames whose loads are in flight together

// jnp.mod(d + half, box) - half in FP32, box > 0 (fmod is exact; the sign
// fix-up adds the divisor, as numpy and jax do).  For positions inside the
// box, x = d + half lies in (-box, 2 box), where that mod is x, x + box or
// x - box with the same rounding (x - box is exact there), so only other x
// take the fmodf path.
__device__ __forceinline__ float periodic_disp(float d, float box, float half) {
  const float x = d + half;
  float m;
  if (x >= 0.0f && x < box) {
    m = x;
  } else if (x < 0.0f && x > -box) {
    m = x + box;
  } else if (x >= box && x < 2.0f * box) {
    m = x - box;
  } else {
    m = fmodf(x, box);
    if (m != 0.0f && m < 0.0f) m += box;
  }
  return m - half;
}

template <int NV>
__device__ __forceinline__ void wave_sum(double (&v)[NV]) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] += __shfl_xor(v[i], off, 64);
}

template <int DIM>
__global__ void __launch_bounds__(ST_THREADS) k_traj_moments(const float* __restrict__ r, const float* __restrict__ u,
                                                             long long T, long long np, long long tc, float box,
                                                             const double* __restrict__ center,
                                                             double* __restrict__ partials) {
  constexpr int NV = 8 * DIM;
  __shared__ double s_red[ST_WAVES][NV];
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  double cen[4 * DIM];
#pragma unroll
  for (int i = 0; i < 4 * DIM; ++i) cen[i] = center ? center[i] : 0.0;
  const float half = box * 0.5f;
  const long long p = (long long)blockIdx.x * ST_THREADS + threadIdx.x;
  const long long t0 = (long long)blockIdx.y * tc;
  const long long t1 = t0 + tc < T ? t0 + tc : T;
  auto add = [&](int q, int c, float x) {
    const double xd = (double)x, dd = xd - cen[q * DIM + c];
    acc[q * DIM + c] += xd;
    acc[(4 + q) * DIM + c] += dd * dd;
  };
  if (p < np && t0 < T) {
    const size_t fs = (size_t)np * DIM;  // floats per frame
    const float* rp = r + (size_t)p * DIM;
    const float* up = u + (size_t)p * DIM;
    // ST_U frames at a time: every load of the group is issued before the
    // first use (frames t .. t+ST_U+1 of r, t .. t+ST_U of u; the last two
    // of r and the last of u are re-read from the cache by the next group)
    for (long long tb = t0; tb < t1; tb += ST_U) {
      float R[ST_U + 2][DIM], Uf[ST_U + 1][DIM];
#pragma unroll
      for (int i = 0; i < ST_U + 2; ++i)
#pragma unroll
        for (int c = 0; c < DIM; ++c) R[i][c] = (tb + i < T) ? rp[(tb + i) * fs + c] : 0.0f;
#pragma unroll
      for (int i = 0; i < ST_U + 1; ++i)
#pragma unroll
        for (int c = 0; c < DIM; ++c) Uf[i][c] = (tb + i < T) ? up[(tb + i) * fs + c] : 0.0f;
      // the group's displacements, each formed once (V[i] = vel at frame tb + i)
      float V[ST_U + 1][DIM];
#pragma unroll
      for (int i = 0; i < ST_U + 1; ++i)
#pragma unroll
        for (int c = 0; c < DIM; ++c) V[i][c] = periodic_disp(R[i + 1][c] - R[i][c], box, half);
#pragma unroll
      for (int i = 0; i < ST_U; ++i) {
        const long long t = tb + i;
        if (t < t1) {
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            add(2, c, Uf[i][c]);
            if (t + 1 < T) {
              add(0, c, V[i][c]);
              add(3, c, Uf[i + 1][c] - Uf[i][c]);
              if (t + 2 < T) add(1, c, V[i + 1][c] - V[i][c]);
            }
          }
        }
      }
    }
  }
  wave_sum<NV>(acc);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) s_red[w][i] = acc[i];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < ST_WAVES; ++k) s += s_red[k][threadIdx.x];
    const size_t nb = (size_t)gridDim.x * gridDim.y;
    partials[threadIdx.x * nb + (size_t)blockIdx.y * gridDim.x + blockIdx.x] = s;  // [NV][nb]
  }
}

// out[v] = sum of row v of the [nv][nb] partials: workgroup v, each thread
// striding over the row in order, then a fixed butterfly
__global__ void __launch_bounds__(ST_THREADS) k_traj_moments_final(const double* __restrict__ partials, long long nb,
                                                                   double* __restrict__ out) {
  __shared__ double s_red[ST_WAVES];
  const double* row = partials + (size_t)blockIdx.x * nb;
  double s[1] = {0.0};
  for (long long b = threadIdx.x; b < nb; b += ST_THREADS) s[0] += row[b];
  wave_sum<1>(s);
  if (threadIdx.x % 64 == 0) s_red[threadIdx.x / 64] = s[0];
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < ST_WAVES; ++k) t += s_red[k];
    out[blockIdx.x] = t;
  }
}

// Launch shape of one trajectory: gx workgroups of 256 particles times gy
// groups of tc frames.  Frames per workgroup: up to ST_TC, fewer while that
// leaves under ~4096 workgroups (16 per CU), more if T would overflow grid.y.
// Both entry points use it, so they sum in the same order.
struct StatsGrid {
  long long gx, gy, tc, nb;
};
StatsGrid stats_grid(long long T, long long np) {
  StatsGrid g{};
  g.gx = (np + ST_THREADS - 1) / ST_THREADS;
  g.tc = std::min<long long>(ST_TC, std::max<long long>(ST_TC / 4, (T * g.gx + 4095) / 4096));
  g.tc = std::max<long long>(g.tc, (T + 65534) / 65535);
  g.gy = (T + g.tc - 1) / g.tc;
  if (g.gx > 0x7fffffffLL) throw std::invalid_argument("np too large");
  g.nb = g.gx * g.gy;
  return g;
}

void launch_moments(int dim, double box_d, const float* r, const float* u, long long T, long long np,
                    const double* center, double* partials, double* out, bool empty, hipStream_t s) {
  long long nb = 0;
  if (!empty) {
    const StatsGrid g = stats_grid(T, np);
    nb = g.nb;
    const dim3 grid((unsigned)g.gx, (unsigned)g.gy);
    const float box = (float)box_d;
    if (dim == 3)
      hipLaunchKernelGGL(k_traj_moments<3>, grid, dim3(ST_THREADS), 0, s, r, u, T, np, g.tc, box, center, partials);
    else
      hipLaunchKernelGGL(k_traj_moments<2>, grid, dim3(ST_THREADS), 0, s, r, u, T, np, g.tc, box, center, partials);
    QL_LAUNCHED();
  }
  hipLaunchKernelGGL(k_traj_moments_final, dim3(8 * dim), dim3(ST_THREADS), 0, s, partials, nb, out);
  QL_LAUNCHED();
}

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

double* scratch_on_current_device() {
  int dev = -1;
  QL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw std::invalid_argument("device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    QL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw ql::HipError("hipGetSymbolAddress returned a null address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

struct qlles_stats {
  int dim = 0;
  double box = 0.0;
  double* partials = nullptr;
  size_t cap = 0;  // doubles
};

extern "C" {

int qlles_stats_create(int dim, double box, qlles_stats** out) {
  if (!out) return fail(QLLES_ERR_ARG, "null argument");
  *out = nullptr;
  if (dim != 2 && dim != 3) return fail(QLLES_ERR_SHAPE, "dim must be 2 or 3");
  if (!(box > 0.0) || !std::isfinite(box)) return fail(QLLES_ERR_ARG, "box side must be positive");
  qlles_stats* S = new (std::nothrow) qlles_stats();
  if (!S) return fail(QLLES_ERR_HIP, "host allocation failed");
  S->dim = dim;
  S->box = box;
  *out = S;
  return QLLES_OK;
}

int qlles_stats_destroy(qlles_stats* S) {
  if (!S) return QLLES_OK;
  if (S->partials) {
    (void)hipDeviceSynchronize();
    (void)hipFree(S->partials);
  }
  delete S;
  return QLLES_OK;
}

int qlles_stats_moments(qlles_stats* S, const float* r, const float* u, long long T, long long np,
                        const double* center, double* out, void* stream) {
  if (!S || !out) return fail(QLLES_ERR_ARG, "null argument");
  if (T < 0 || np < 0) return fail(QLLES_ERR_ARG, "T and np must be >= 0");
  const bool empty = T == 0 || np == 0;
  if (!empty && (!r || !u)) return fail(QLLES_ERR_ARG, "null argument");
  const int dim = S->dim, nv = 8 * dim;
  hipStream_t s = (hipStream_t)stream;
  return guarded([&] {
    if (!empty) {
      const size_t need = (size_t)stats_grid(T, np).nb * nv;
      if (need > S->cap) {
        if (S->partials) {
          QL_HIP(hipStreamSynchronize(s));
          QL_HIP(hipFree(S->partials));
          S->partials = nullptr;
          S->cap = 0;
        }
        QL_HIP(hipMalloc(&S->partials, need * sizeof(double)));
        S->cap = need;
      }
    }
    launch_moments(dim, S->box, r, u, T, np, center, S->partials, out, empty, s);
  });
}

int qlles_stats_moments_noalloc(int dim, double box, const float* r, const float* u, long long T, long long np,
                                const double* center, double* out, void* stream) {
  if (!out) return fail(QLLES_ERR_ARG, "null argument");
  if (dim != 2 && dim != 3) return fail(QLLES_ERR_SHAPE, "dim must be 2 or 3");
  if (!(box > 0.0) || !std::isfinite(box)) return fail(QLLES_ERR_ARG, "box side must be positive");
  if (T < 0 || np < 0) return fail(QLLES_ERR_ARG, "T and np must be >= 0");
  const bool empty = T == 0 || np == 0;
  if (!empty && (!r || !u)) return fail(QLLES_ERR_ARG, "null argument");
  hipStream_t s = (hipStream_t)stream;
  return guarded([&] {
    if (!empty) {
      const long long need = stats_grid(T, np).nb * 8 * dim;
      if (need > (long long)QLLES_STATS_SCRATCH_DOUBLES)
        throw std::invalid_argument("trajectory too large for qlles_stats_moments_noalloc: its partial sums need " +
                                    std::to_string(need) + " doubles, the scratch array holds " +
                                    std::to_string((long long)QLLES_STATS_SCRATCH_DOUBLES) +
                                    " (use qlles_stats_moments)");
    }
    // every partial the final pass reads is written by this call's first
    // kernel, so nothing needs clearing
    launch_moments(dim, box, r, u, T, np, center, scratch_on_current_device(), out, empty, s);
  });
}

}  // extern "C"
