// SPH density and particle relaxation (relax.py:16-178, utils.py:169-218,
// combined.py:132-142,199).
//
// The reference sums the Morris quintic kernel (support 3h) over a jax_md
// neighbour list.  Here each call bins the particles into a periodic cell
// grid of nc >= 2 cells per axis, with cells no smaller than the cutoff:
//   1. k_sph_bin      cell of every particle (+ the shifted positions of the
//                     fused relax/advect entry), per-cell counts (atomics)
//   2. k_sph_scan     exclusive scan of the counts (one block)
//   3. k_sph_scatter  particle ids into their cell's range (atomics)
//   4. k_sph_rank     each id's rank among its cell's ids gives its slot, so
//                     cells are in ascending id order; positions (and
//                     velocities) gathered into cell order, SoA FP64
//   5. k_sph_pass     density rho_i = m sum_j W(r_ij) over the 3^dim cells
//                     around i, one workgroup per cell, candidates in LDS
//   6. k_sph_pass     pressure / background / viscous accelerations, and for
//                     the relax entry the velocity correction and the shift
// Step 4 makes the order of every sum independent of the atomics, so results
// are bit-reproducible from run to run.  In steps 5-6 LPP consecutive lanes
// share one particle: lane l takes candidates l, l+LPP, ... of the cell's
// neighbourhood, and the LPP partial sums are combined by an xor butterfly
// (equal bits on every lane).  Everything is FP64, as in the reference.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "capi_common.hpp"

using qlapi::fail;
using qlapi::guarded;

namespace {

struct SphGeom {
  double box;      // periodic side in working units (L / l_ref)
  double half;     // box / 2
  double inv_cell; // cells per working length unit
  double cutoff;   // 3 h
  double cutoff2;  // (3 h)^2
  double inv_h;    // 1 / h
  double sigma;    // kernel normalisation
  double mass;     // particle mass dx^dim rho_ref
  double p_eos;    // reference pressure (10 u_ref / l_ref)^2 rho_ref
  double lref;     // length unit
  int nc;          // cells per axis
  int cap;         // stride of the SoA arrays
};

}  // namespace

struct qlles_sph {
  int dim = 0, n = 0;
  double L = 0.0;
  SphGeom g{};
  int ncells = 0;
  int cap = 0;
  int* cnt = nullptr;      // [ncells]
  int* start = nullptr;    // [ncells + 1]
  int* fill = nullptr;     // [ncells]
  int* perm = nullptr;     // [cap] particle id per cell-ordered slot
  int* cell_of = nullptr;  // [cap] cell per particle id
  int* tmp = nullptr;      // [cap] ids scattered into their cell's range, any order
  double* xw = nullptr;    // [cap][dim] wrapped positions in working units, by particle id
  double* rt = nullptr;    // [cap][dim] shifted physical positions (relax entry)
  double* pos = nullptr;   // [dim][cap] cell-ordered positions
  double* vel = nullptr;   // [dim][cap] cell-ordered velocities (forces entry)
  double* rho = nullptr;   // [cap] cell-ordered densities
  std::vector<void*> cell_bufs, part_bufs;
  // particle -> grid DFT (allocated on first use): phase tables [dim][np][n]
  // and the full spectrum [dim][n^dim], complex FP64
  void* dft_E = nullptr;
  void* dft_U = nullptr;
  int dft_ps = 1;  // particle subsets of the point-cloud DFT sum
  size_t dft_E_elems = 0;
};

namespace {

constexpr double kEps = 2.220446049250313e-16;  // finfo(float64).eps, relax.py:13

__device__ __forceinline__ double ipow4(double x) {
  const double x2 = x * x;
  return x2 * x2;
}
__device__ __forceinline__ double ipow5(double x) { return x * ipow4(x); }

// kernel_fn.w, kernel_fn.grad_w of the quintic spline
__device__ __forceinline__ double quintic_w(double r, const SphGeom& g) {
  const double q = r * g.inv_h;
  const double q1 = fmax(0.0, 1.0 - q), q2 = fmax(0.0, 2.0 - q), q3 = fmax(0.0, 3.0 - q);
  return g.sigma * (ipow5(q3) - 6.0 * ipow5(q2) + 15.0 * ipow5(q1));
}
__device__ __forceinline__ double quintic_dw(double r, const SphGeom& g) {
  const double q = r * g.inv_h;
  const double q1 = fmax(0.0, 1.0 - q), q2 = fmax(0.0, 2.0 - q), q3 = fmax(0.0, 3.0 - q);
  return g.sigma * g.inv_h * (-5.0 * ipow4(q3) + 30.0 * ipow4(q2) - 75.0 * ipow4(q1));
}

// jnp.mod(x, side) for side > 0 (space.periodic shift)
__device__ __forceinline__ double mod_side(double x, double side) {
  const double m = fmod(x, side);
  return m < 0.0 ? m + side : m;
}
// displacement mod(d + side/2, side) - side/2 for |d| <= side (both points
// wrapped into [0, side]); the two branches equal fmod's exact result
__device__ __forceinline__ double min_image(double d, const SphGeom& g) {
  double t = d + g.half;
  if (t >= g.box) t -= g.box;
  else if (t < 0.0) t += g.box;
  return t - g.half;
}

__device__ __forceinline__ double wrap_box(double x, double box) {
  const double w = x - box * floor(x / box);
  return w;
}

__device__ __forceinline__ int cell_coord(double x, const SphGeom& g) {
  int c = (int)floor(x * g.inv_cell);
  return c < 0 ? 0 : (c >= g.nc ? g.nc - 1 : c);
}

// 1. cells (+ optional shift r + dt u, physical units)
template <int DIM>
__global__ void k_sph_bin(const double* __restrict__ r, const void* __restrict__ u, int u_f32, double dt, double L,
                          int np, SphGeom g, double* __restrict__ rt, double* __restrict__ xw,
                          int* __restrict__ cell_of, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int c = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const size_t e = (size_t)i * DIM + d;
    double x = r[e];
    if (u) {
      const double v = u_f32 ? (double)static_cast<const float*>(u)[e] : static_cast<const double*>(u)[e];
      x = mod_side(x + dt * v, L);  // combined.py:136
      rt[e] = x;
    }
    const double w = wrap_box(x / g.lref, g.box);  // relax.py:78
    xw[e] = w;
    c = c * g.nc + cell_coord(w, g);
  }
  cell_of[i] = c;
  atomicAdd(&cnt[c], 1);
}

// 2. exclusive scan (one block of 1024); clears the scatter cursors and the
// counts for the next call
__global__ __launch_bounds__(1024) void k_sph_scan(int* __restrict__ cnt, int* __restrict__ start,
                                                   int* __restrict__ fill, int ncells) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (ncells + 1023) / 1024;
  const int b = min(t * per, ncells), e = min(b + per, ncells);
  int s = 0;
  for (int k = b; k < e; ++k) s += cnt[k];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t ? part[t - 1] : 0;
  for (int k = b; k < e; ++k) {
    start[k] = run;
    run += cnt[k];
    cnt[k] = 0;
    fill[k] = 0;
  }
  if (t == 1023) start[ncells] = part[1023];
}

// 3. ids into their cell's range (the order inside a cell is fixed in step 4)
__global__ void k_sph_scatter(const int* __restrict__ cell_of, const int* __restrict__ start, int* __restrict__ fill,
                              int* __restrict__ tmp, int np) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int c = cell_of[i];
  tmp[start[c] + atomicAdd(&fill[c], 1)] = i;
}

// 4. one thread per particle: its rank among the ids of its cell gives its
// slot, so every cell ends up in ascending id order; cell-ordered SoA gathers
template <int DIM>
__global__ void k_sph_rank(const int* __restrict__ cell_of, const int* __restrict__ start,
                           const int* __restrict__ tmp, const double* __restrict__ xw, const double* __restrict__ u,
                           int np, int cap, int* __restrict__ perm, double* __restrict__ pos,
                           double* __restrict__ vel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int c = cell_of[i];
  const int b = start[c], e = start[c + 1];
  int rank = 0;
  for (int k = b; k < e; ++k) rank += tmp[k] < i;
  const int slot = b + rank;
  perm[slot] = i;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    pos[(size_t)d * cap + slot] = xw[(size_t)i * DIM + d];
    if (u) vel[(size_t)d * cap + slot] = u[(size_t)i * DIM + d];
  }
}

// neighbour cells of `cell`: 3 per axis (offsets -1, 0, 1), or 2 when the
// grid has only 2 cells per axis; each distinct cell appears once
template <int DIM>
__device__ int stencil(int cell, int nc, int* out) {
  const int m = nc >= 3 ? 3 : 2;
  int base[DIM];
  for (int d = DIM - 1; d >= 0; --d) {
    const int cd = cell % nc;
    cell /= nc;
    base[d] = nc >= 3 ? cd - 1 + nc : cd;  // + nc keeps the modulo non-negative
  }
  const int cnt = DIM == 3 ? m * m * m : m * m;
  for (int t = 0; t < cnt; ++t) {
    int c = 0, rest = t;
    int o[DIM];
    for (int d = DIM - 1; d >= 0; --d) {
      o[d] = rest % m;
      rest /= m;
    }
    for (int d = 0; d < DIM; ++d) c = c * nc + (base[d] + o[d]) % nc;
    out[t] = c;
  }
  return cnt;
}

// blocks are dealt round-robin over the 8 XCDs; give each XCD a contiguous
// run of cells so that neighbouring cells share an L2
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int per = n / 8, rem = n % 8;
  const int x = b % 8, k = b / 8;
  return x * per + min(x, rem) + k;
}

constexpr int kBlock = 256;
constexpr int kTile = 1024;  // candidates staged in LDS at a time
static_assert(kTile % kBlock == 0, "whole tile rows per thread");

enum { PASS_DENSITY = 0, ACC_RELAX = 1, ACC_FORCES = 2, ACC_ADVECT = 3 };

struct PassOut {
  double* rho;       // PASS_DENSITY: densities by particle id (optional)
  double* acc;       // relax acceleration (np, dim), physical units (optional for ACC_ADVECT); acc_p for forces
  double* acc_visc;  // ACC_FORCES: viscous and transport-velocity terms
  double* acc_tvf;
  // ACC_ADVECT
  const double* rt;  // shifted positions r_t (np, dim)
  double* r;         // in: r_0, out: shift(r_0, dt u)
  void* u_r;         // out: displacement(r_t2, r_0) / dt in the interpolator dtype
  int u_f32;
  double dt, fdt2, L;
};

template <int DIM, int MODE>
struct PassLds {
  static constexpr int NS = DIM == 3 ? 27 : 9;
  static constexpr bool ACC = MODE != PASS_DENSITY, VEL = MODE == ACC_FORCES;
  static constexpr int NC = MODE == PASS_DENSITY ? 1 : (MODE == ACC_FORCES ? 3 * DIM : DIM);  // sums per home
  int beg[NS], off[NS + 1], cnt;
  double sh[NS][DIM];  // periodic image shift of each stencil cell
  double x[DIM][kTile];
  double rho[ACC ? kTile : 1], p[ACC ? kTile : 1], ir[ACC ? kTile : 1];
  double v[VEL ? DIM : 1][VEL ? kTile : 1];
  double red[NC][kBlock];  // per-lane partial sums
};

// 5./6. pair passes, one workgroup per home cell.  The particles of the
// neighbour cells (the candidates, in stencil order) are staged in LDS in
// tiles.  With nc >= 3 cells per axis each stencil cell is staged at its
// periodic image next to the home cell (SHIFT): a pair closer than the cutoff
// (< one cell) is then exactly its minimum image, and farther pairs stay
// farther, so the displacement is a plain difference.  (nc == 2 lists each
// cell once for two images and keeps the minimum-image formula.)
// lpp lanes share a home particle, lane l taking candidates l, l+lpp, ...
// Per tile a lane first marks its candidates inside the cutoff in a register
// bitmask, then does the kernel arithmetic for the marked ones only.  The
// lanes' partial sums are added in lane order, so every sum has a fixed
// order.
//   PASS_DENSITY  rho_i = m sum_j W(r_ij)      (utils.py:208-216, relax.py:80-87)
//   ACC_*         accelerations                 (relax.py:90-119, :132-176)
//   ACC_ADVECT    + velocity correction and shift (combined.py:140-142, :199)
template <int DIM, int MODE, bool SHIFT>
__device__ void run_pass(PassLds<DIM, MODE>& L, const double* __restrict__ pos, const double* __restrict__ vel,
                         double* __restrict__ rho_s, const int* __restrict__ perm, const SphGeom& g,
                         const PassOut& out, int hb, int he) {
  constexpr int NW = 4;  // mask words: a lane holds at most kTile / 4 candidates of a tile
  constexpr bool ACC = MODE != PASS_DENSITY;
  constexpr bool VEL = MODE == ACC_FORCES;
  constexpr int NC = PassLds<DIM, MODE>::NC;
  static_assert(kTile <= 4 * 64 * NW, "a lane's candidates of one tile must fit its mask");
  const int nh = he - hb;
  const int lpp = max(4, min(64, kBlock / nh));  // lanes per home particle
  const int hpc = kBlock / lpp;                   // home particles per chunk
  const int T = L.off[L.cnt];
  const int lane = threadIdx.x % lpp;
  for (int h0 = hb; h0 < he; h0 += hpc) {
    const int hl = threadIdx.x / lpp;
    const int h = h0 + hl;
    const bool act = hl < hpc && h < he;
    double p[DIM], ui[DIM];
    double rho_i = 1.0, p_i = 0.0, ir_i = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      p[d] = act ? pos[(size_t)d * g.cap + h] : 0.0;
      ui[d] = VEL && act ? vel[(size_t)d * g.cap + h] : 0.0;
    }
    if (ACC && act) {
      rho_i = rho_s[h];
      p_i = g.p_eos * (rho_i / 1.0 - 1.0) + 0.0;  // TaitEoS, gamma 1, p_background 0 (relax.py:66)
      ir_i = 1.0 / rho_i;
    }
    double a0[DIM] = {}, a1[DIM] = {}, a2[DIM] = {};
    double w = 0.0;
    for (int t0 = 0; t0 < T; t0 += kTile) {
      const int tn = min(kTile, T - t0);
      __syncthreads();  // the previous tile is consumed
      {
        // all loads of this thread's (at most kTile / kBlock) candidates are
        // issued before the first LDS store
        constexpr int NL = kTile / kBlock;
        constexpr int NV = DIM + (ACC ? 1 : 0) + (VEL ? DIM : 0);
        double val[NL][NV];
        int cell_t[NL];
#pragma unroll
        for (int u = 0; u < NL; ++u) {
          const int k = (int)threadIdx.x + u * kBlock;
          cell_t[u] = -1;
          if (k >= tn) continue;
          const int gi = t0 + k;
          int lo = 0, hi = L.cnt;  // stencil cell t with off[t] <= gi < off[t+1]
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (L.off[mid] <= gi) lo = mid;
            else hi = mid;
          }
          cell_t[u] = lo;
          const int j = L.beg[lo] + (gi - L.off[lo]);
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            val[u][d] = pos[(size_t)d * g.cap + j];
            if (VEL) val[u][DIM + (ACC ? 1 : 0) + d] = vel[(size_t)d * g.cap + j];
          }
          if (ACC) val[u][DIM] = rho_s[j];
        }
#pragma unroll
        for (int u = 0; u < NL; ++u) {
          if (cell_t[u] < 0) continue;
          const int k = (int)threadIdx.x + u * kBlock;
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            L.x[d][k] = SHIFT ? val[u][d] + L.sh[cell_t[u]][d] : val[u][d];
            if (VEL) L.v[d][k] = val[u][DIM + (ACC ? 1 : 0) + d];
          }
          if (ACC) {
            const double rj = val[u][DIM];
            L.rho[k] = rj;
            L.p[k] = g.p_eos * (rj / 1.0 - 1.0) + 0.0;
            L.ir[k] = 1.0 / rj;
          }
        }
      }
      __syncthreads();
      if (!act) continue;
      // this lane's candidates of the tile: k = lane + i * lpp, i < nk
      const int nk = lane < tn ? (tn - lane + lpp - 1) / lpp : 0;
      uint64_t mask[NW];
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        uint64_t bits = 0;
        const int nb = min(64, nk - 64 * q);
#pragma unroll 4
        for (int b = 0; b < nb; ++b) {
          const int k = lane + (q * 64 + b) * lpp;
          double r2 = 0.0;
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            const double x = SHIFT ? p[d] - L.x[d][k] : min_image(p[d] - L.x[d][k], g);
            r2 += x * x;
          }
          bits |= (uint64_t)(r2 < g.cutoff2) << b;  // W and dW/dr vanish from 3h on
        }
        mask[q] = bits;
      }
      // one loop over the marked candidates of every word, in candidate order
      int q = 0;
      uint64_t bits = mask[0];
      while (true) {
        while (bits == 0 && q + 1 < NW) bits = mask[++q];
        if (bits == 0) break;
        {
          const int b = __builtin_ctzll(bits);
          bits &= bits - 1;
          const int k = lane + (q * 64 + b) * lpp;
          double x[DIM], r2 = 0.0;
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            x[d] = SHIFT ? p[d] - L.x[d][k] : min_image(p[d] - L.x[d][k], g);
            r2 += x[d] * x[d];
          }
          const double r = r2 > 0.0 ? sqrt(r2) : 0.0;
          if (MODE == PASS_DENSITY) {
            w += quintic_w(r, g);
            continue;
          }
          const double der = quintic_dw(r, g);
          const double rho_j = L.rho[k], p_j = L.p[k], ir_j = L.ir[k];
          const double p_ij = (rho_j * p_i + rho_i * p_j) / (rho_i + rho_j);  // relax.py:99
          const double pref = g.mass * (ir_i * ir_i + ir_j * ir_j);          // relax.py:102
          const double ire = 1.0 / (r + kEps);                                // e_ij = r_ij / (d + EPS), :92
          const double dk = der * ire;
          if (MODE == ACC_FORCES) {
            const double cp = -pref * p_ij * dk, ct = pref * (-g.p_eos) * dk, cv = pref * 1.0 * ire * der;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
              a0[d] += cp * x[d];                 // relax.py:149
              a1[d] += cv * (ui[d] - L.v[d][k]);  // relax.py:151-154
              a2[d] += ct * x[d];                 // relax.py:158
            }
          } else {
            const double ca = (-pref * p_ij + 0.5 * pref * (-g.p_eos)) * dk;  // relax.py:103,107
#pragma unroll
            for (int d = 0; d < DIM; ++d) a0[d] += ca * x[d];
          }
        }
      }
    }
    // partial sums through LDS, added in lane order by each home's first lane
    if (MODE == PASS_DENSITY) {
      L.red[0][threadIdx.x] = w;
    } else {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        L.red[d][threadIdx.x] = a0[d];
        if (MODE == ACC_FORCES) {
          L.red[DIM + d][threadIdx.x] = a1[d];
          L.red[2 * DIM + d][threadIdx.x] = a2[d];
        }
      }
    }
    __syncthreads();
    double sum[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      sum[c] = 0.0;
      if (act && lane == 0)
        for (int l = 0; l < lpp; ++l) sum[c] += L.red[c][threadIdx.x + l];
    }
    __syncthreads();  // L.red is rewritten by the next chunk
    if (!act || lane != 0) continue;
    const int i = perm[h];
    if (MODE == PASS_DENSITY) {
      const double rho = g.mass * sum[0];
      rho_s[h] = rho;
      if (out.rho) out.rho[i] = rho;
      continue;
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const size_t e = (size_t)i * DIM + d;
      if (MODE == ACC_FORCES) {
        out.acc[e] = sum[d];
        out.acc_visc[e] = sum[DIM + d];
        out.acc_tvf[e] = sum[2 * DIM + d];
        continue;
      }
      const double a = sum[d] * g.lref;  // relax.py:119
      if (out.acc) out.acc[e] = a;
      if (MODE == ACC_ADVECT) {
        const double r0 = out.r[e];
        const double r_t = mod_side(out.rt[e] + out.fdt2 * a, out.L);  // combined.py:140
        double dd = r_t - r0;                                            // combined.py:142
        dd = mod_side(dd + out.L * 0.5, out.L) - 0.5 * out.L;
        const double v = dd / out.dt;
        if (out.u_f32) static_cast<float*>(out.u_r)[e] = (float)v;
        else static_cast<double*>(out.u_r)[e] = v;
        out.r[e] = mod_side(r0 + out.dt * v, out.L);  // combined.py:199
      }
    }
  }
}

template <int DIM, int MODE>
__global__ __launch_bounds__(kBlock) void k_sph_pass(const double* __restrict__ pos, const double* __restrict__ vel,
                                                     double* __restrict__ rho_s, const int* __restrict__ start,
                                                     const int* __restrict__ perm, int ncells, SphGeom g,
                                                     PassOut out) {
  __shared__ PassLds<DIM, MODE> L;
  const int c = xcd_remap(blockIdx.x, ncells);
  const int hb = start[c], he = start[c + 1];
  if (hb == he) return;  // block-uniform
  if (threadIdx.x == 0) {
    int cells[PassLds<DIM, MODE>::NS];
    const int n = stencil<DIM>(c, g.nc, cells);
    // home cell coordinates, to tell which stencil cells wrap around the box
    int hc[DIM];
    for (int d = DIM - 1, cc = c; d >= 0; --d) {
      hc[d] = cc % g.nc;
      cc /= g.nc;
    }
    L.off[0] = 0;
    for (int t = 0; t < n; ++t) {
      const int b = start[cells[t]];
      L.beg[t] = b;
      L.off[t + 1] = L.off[t] + (start[cells[t] + 1] - b);
      for (int d = DIM - 1, cc = cells[t]; d >= 0; --d) {
        const int nd = cc % g.nc;
        cc /= g.nc;
        const int o = nd - hc[d];  // -1, 0, 1 or, across the box, +-(nc-1)
        L.sh[t][d] = o > 1 ? -g.box : (o < -1 ? g.box : 0.0);
      }
    }
    L.cnt = n;
  }
  __syncthreads();
  if (g.nc >= 3) run_pass<DIM, MODE, true>(L, pos, vel, rho_s, perm, g, out, hb, he);
  else run_pass<DIM, MODE, false>(L, pos, vel, rho_s, perm, g, out, hb, he);
}

// ---- particle -> grid transfers (turbulence.py:145-336) ---------------------

// Second-order moving least squares at the grid nodes (i + 1/2) box/n, 'ij'
// order (mls_2nd_order with the quintic kernel, h = 2/3 dx, support 2 dx;
// turbulence.py:145-250, :253-295).  One thread per node over the particles
// of its neighbour cells.  The reference solves D A x = D b with
// A = sum_j w r r^T, b = sum_j w f r and a diagonal D; this solves A x = b by
// Cholesky, with the monomials r = [1, s_a, s_a^2/2, s_a s_b] in units of h
// (s = r_ji / h), which leaves x[0], the value at the node, unchanged and
// keeps A well scaled.
template <int DIM>
__global__ __launch_bounds__(64) void k_sph_mls(const double* __restrict__ pos, const double* __restrict__ vel,
                                                const int* __restrict__ start, SphGeom g, int n, double h,
                                                double sigma, double* __restrict__ out) {
  constexpr int M = DIM == 3 ? 10 : 6;
  constexpr int NP = M * (M + 1) / 2;
  const int nodes = DIM == 3 ? n * n * n : n * n;
  const int node = blockIdx.x * 64 + threadIdx.x;
  if (node >= nodes) return;
  const double dxw = g.box / n;
  double x[DIM];
  int cell = 0;
  for (int d = DIM - 1, rest = node; d >= 0; --d) {
    x[d] = ((rest % n) + 0.5) * dxw;
    rest /= n;
  }
  for (int d = 0; d < DIM; ++d) cell = cell * g.nc + cell_coord(x[d], g);
  int cells[DIM == 3 ? 27 : 9];
  const int ns = stencil<DIM>(cell, g.nc, cells);
  const double cut = 3.0 * h, cut2 = cut * cut, ih = 1.0 / h;
  double A[NP], B[DIM][M];
#pragma unroll
  for (int e = 0; e < NP; ++e) A[e] = 0.0;
#pragma unroll
  for (int c = 0; c < DIM; ++c)
#pragma unroll
    for (int e = 0; e < M; ++e) B[c][e] = 0.0;
  for (int t = 0; t < ns; ++t) {
    const int ke = start[cells[t] + 1];
    for (int k = start[cells[t]]; k < ke; ++k) {
      double d[DIM], r2 = 0.0;
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        d[a] = min_image(pos[(size_t)a * g.cap + k] - x[a], g);  // r_ji = r_j - r_i (turbulence.py:196)
        r2 += d[a] * d[a];
      }
      if (!(r2 < cut2)) continue;
      const double q = sqrt(r2) * ih;
      const double q1 = fmax(0.0, 1.0 - q), q2 = fmax(0.0, 2.0 - q), q3 = fmax(0.0, 3.0 - q);
      const double w = sigma * (ipow5(q3) - 6.0 * ipow5(q2) + 15.0 * ipow5(q1));
      double row[M];
      row[0] = 1.0;
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        const double sa = d[a] * ih;
        row[1 + a] = sa;
        row[1 + DIM + a] = 0.5 * sa * sa;
      }
      {
        int e = 1 + 2 * DIM;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
          for (int b = a + 1; b < DIM; ++b) row[e++] = d[a] * ih * (d[b] * ih);
      }
      {
        int e = 0;
#pragma unroll
        for (int i = 0; i < M; ++i) {
          const double wi = w * row[i];
#pragma unroll
          for (int j = i; j < M; ++j) A[e++] += wi * row[j];
        }
      }
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        const double wf = w * vel[(size_t)c * g.cap + k];
#pragma unroll
        for (int i = 0; i < M; ++i) B[c][i] += wf * row[i];
      }
    }
  }
  // A = U^T U (packed upper rows), then U^T y = b, U x = y per component
  auto at = [](int i, int j) { return i * M - i * (i - 1) / 2 + (j - i); };
#pragma unroll
  for (int i = 0; i < M; ++i) {
    double sd = A[at(i, i)];
#pragma unroll
    for (int p = 0; p < i; ++p) sd -= A[at(p, i)] * A[at(p, i)];
    const double uii = sqrt(sd), inv = 1.0 / uii;
    A[at(i, i)] = uii;
#pragma unroll
    for (int j = i + 1; j < M; ++j) {
      double v = A[at(i, j)];
#pragma unroll
      for (int p = 0; p < i; ++p) v -= A[at(p, i)] * A[at(p, j)];
      A[at(i, j)] = v * inv;
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    double y[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double v = B[c][i];
#pragma unroll
      for (int p = 0; p < i; ++p) v -= A[at(p, i)] * y[p];
      y[i] = v / A[at(i, i)];
    }
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
      double v = y[i];
#pragma unroll
      for (int j = i + 1; j < M; ++j) v -= A[at(i, j)] * y[j];
      y[i] = v / A[at(i, i)];
    }
    out[(size_t)c * nodes + node] = y[0];
  }
}

// Exact type-1 sum of ur_to_u_dft_wrapper (turbulence.py:298-336),
//   U_c(k) = sum_j u_jc exp(-i k.r_j),  k over fftfreq(n) * n per axis,
// from per-axis phase tables E[a][j][k] = exp(-i f(k) r_ja) (FP64).
__global__ void k_dft_phase(const double* __restrict__ r, int np, int dim, int n, double2* __restrict__ E) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long long)dim * np * n) return;
  const int k = (int)(id % n);
  const long long aj = id / n;
  const int j = (int)(aj % np), a = (int)(aj / np);
  const int f = k < (n + 1) / 2 ? k : k - n;
  double sn, cs;
  sincos((double)f * r[(size_t)j * dim + a], &sn, &cs);
  E[id] = make_double2(cs, -sn);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Tiled sum.  A workgroup owns DFT_RB rows of the leading axes (kx, or kx and
// ky) times DFT_KW consecutive last-axis modes (one mode per thread) and one
// of DFT_PS interleaved particle subsets.  Particles are staged DFT_TJ at a
// time in LDS: per row the weights w_c = u_jc e_x(j) [e_y(j)], and the
// last-axis phases e_z(j, k) of the workgroup's modes, so each phase-table
// row is read from memory once per workgroup instead of once per mode row.
// Each thread then adds w_c e_z over the tile in particle order; the ps
// partial spectra are added in subset order by k_dft_reduce (reproducible).
// ps is chosen per context so that the grid holds ~2048 workgroups.
constexpr int DFT_THREADS = 256;
constexpr int DFT_TJ = 32;      // particles per LDS tile
constexpr int DFT_PS_MAX = 16;  // particle subsets (workgroups per mode block), at most

template <int DIM, int KW>
__global__ __launch_bounds__(DFT_THREADS) void k_dft_sum(const double2* __restrict__ E, const double* __restrict__ u,
                                                        int np, int n, int rows, int ps, double2* __restrict__ Up) {
  constexpr int RB = DFT_THREADS / KW;  // rows per workgroup
  __shared__ double2 s_w[DFT_TJ][RB][DIM];
  __shared__ double2 s_e[DFT_TJ][KW];
  const int tid = threadIdx.x, rl = tid / KW, kl = tid % KW;
  const int row = blockIdx.x * RB + rl, kz = blockIdx.y * KW + kl, sub = blockIdx.z;
  const size_t plane = (size_t)np * n;
  // particle subset `sub`: j = sub, sub + ps, ... (tiles of TJ consecutive members)
  const int nsub = np > sub ? (np - sub + ps - 1) / ps : 0;
  double2 acc[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) acc[c] = make_double2(0.0, 0.0);
  for (int t0 = 0; t0 < nsub; t0 += DFT_TJ) {
    const int tn = nsub - t0 < DFT_TJ ? nsub - t0 : DFT_TJ;
    __syncthreads();  // the previous tile is consumed
    // weights: TJ x RB entries, one (particle, row) pair per thread and step
    for (int q = tid; q < DFT_TJ * RB; q += DFT_THREADS) {
      const int jt = q / RB, rr = q % RB, rw = blockIdx.x * RB + rr;
      if (jt < tn && rw < rows) {
        const int j = sub + (t0 + jt) * ps;
        double2 e = make_double2(1.0, 0.0);
        for (int d = DIM - 2, rest = rw; d >= 0; --d) {
          e = cmul(E[(size_t)d * plane + (size_t)j * n + rest % n], e);
          rest /= n;
        }
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const double uc = u[(size_t)j * DIM + c];
          s_w[jt][rr][c] = make_double2(uc * e.x, uc * e.y);
        }
      }
    }
    // last-axis phases of the workgroup's modes
    for (int q = tid; q < DFT_TJ * KW; q += DFT_THREADS) {
      const int jt = q / KW, kk = q % KW, k = blockIdx.y * KW + kk;
      if (jt < tn && k < n) {
        const int j = sub + (t0 + jt) * ps;
        s_e[jt][kk] = E[(size_t)(DIM - 1) * plane + (size_t)j * n + k];
      }
    }
    __syncthreads();
    if (row < rows && kz < n) {
      for (int jt = 0; jt < tn; ++jt) {
        const double2 e = s_e[jt][kl];
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const double2 w = s_w[jt][rl][c];
          acc[c].x += w.x * e.x - w.y * e.y;
          acc[c].y += w.x * e.y + w.y * e.x;
        }
      }
    }
  }
  if (row < rows && kz < n)
#pragma unroll
    for (int c = 0; c < DIM; ++c) Up[(((size_t)sub * DIM + c) * rows + row) * n + kz] = acc[c];
}

// U = the ps partial spectra added in subset order
__global__ void k_dft_reduce(const double2* __restrict__ Up, size_t count, int ps, double2* __restrict__ U) {
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= count) return;
  double2 s = Up[id];
  for (int p = 1; p < ps; ++p) {
    const double2 v = Up[(size_t)p * count + id];
    s.x += v.x;
    s.y += v.y;
  }
  U[id] = s;
}

// half spectrum of ifftn(U).real: H(k) = (U(k) + conj(U(-k mod n))) / 2 in
// the solver layout (3, n, n, n/2+1) / (3, n, n/2+1), component 2 of a 2-D
// field zero
template <int DIM>
__global__ void k_dft_half(const double2* __restrict__ U, int n, double2* __restrict__ H) {
  const int nh = n / 2 + 1;
  const long long rows = DIM == 3 ? (long long)n * n : n;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 3 * rows * nh) return;
  const int kz = (int)(id % nh);
  const long long cr = id / nh;
  const int row = (int)(cr % rows), c = (int)(cr / rows);
  if (c >= DIM) {
    H[id] = make_double2(0.0, 0.0);
    return;
  }
  int neg = 0;
  for (int d = DIM - 2, rest = row, mul = 1; d >= 0; --d) {
    const int kd = rest % n;
    rest /= n;
    neg += ((n - kd) % n) * mul;
    mul *= n;
  }
  const double2 a = U[((size_t)c * rows + row) * n + kz];
  const double2 b = U[((size_t)c * rows + neg) * n + (n - kz) % n];
  H[id] = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
}

double quintic_sigma(double h, int dim) {
  // jax_sph QuinticKernel normalisation (Morris, Fox & Zhu 1997); the 3-D
  // value 3/(359 pi) is the library's and integrates to 360/359
  return dim == 2 ? 7.0 / 478.0 / M_PI / (h * h) : 3.0 / 359.0 / M_PI / (h * h * h);
}

template <class P>
P* sph_alloc(std::vector<void*>& list, size_t n) {
  void* p = nullptr;
  QL_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(P)));
  list.push_back(p);
  return static_cast<P*>(p);
}

void free_all(std::vector<void*>& list) {
  for (void* p : list) (void)hipFree(p);
  list.clear();
}

// particle arrays for np particles (grown on demand; a growth waits for the
// device before releasing the old arrays)
void ensure_cap(qlles_sph* S, int np) {
  if (np <= S->cap && S->perm) return;
  if (S->perm) QL_HIP(hipDeviceSynchronize());
  free_all(S->part_bufs);
  S->perm = nullptr;
  const int cap = std::max(np, S->cap);
  const size_t d = S->dim;
  auto& L = S->part_bufs;
  S->perm = sph_alloc<int>(L, cap);
  S->cell_of = sph_alloc<int>(L, cap);
  S->tmp = sph_alloc<int>(L, cap);
  S->xw = sph_alloc<double>(L, d * cap);
  S->rt = sph_alloc<double>(L, d * cap);
  S->pos = sph_alloc<double>(L, d * cap);
  S->vel = sph_alloc<double>(L, d * cap);
  S->rho = sph_alloc<double>(L, cap);
  S->cap = cap;
  S->g.cap = cap;
}

int blocks(long long n, int b) { return (int)((n + b - 1) / b); }

// steps 1-4 (cell list); u (shift velocity) optional, uvel (FP64 velocities
// to gather into cell order) optional
template <int DIM>
void build_cells(qlles_sph* S, const double* r, const void* u, int u_f32, double dt, int np, const double* uvel,
                 hipStream_t s) {
  ensure_cap(S, np);
  const SphGeom& g = S->g;
  // the counts start at zero: cleared at creation and by every scan
  hipLaunchKernelGGL((k_sph_bin<DIM>), dim3(blocks(np, 256)), dim3(256), 0, s, r, u, u_f32, dt, S->L, np, g, S->rt,
                     S->xw, S->cell_of, S->cnt);
  hipLaunchKernelGGL(k_sph_scan, dim3(1), dim3(1024), 0, s, S->cnt, S->start, S->fill, S->ncells);
  hipLaunchKernelGGL(k_sph_scatter, dim3(blocks(np, 256)), dim3(256), 0, s, S->cell_of, S->start, S->fill, S->tmp,
                     np);
  hipLaunchKernelGGL((k_sph_rank<DIM>), dim3(blocks(np, 256)), dim3(256), 0, s, S->cell_of, S->start, S->tmp, S->xw,
                     uvel, np, S->cap, S->perm, S->pos, S->vel);
  QL_LAUNCHED();
}

// steps 1-5; u (shift velocity) optional, uvel (FP64 velocities to gather) optional
template <int DIM>
void build_and_density(qlles_sph* S, const double* r, const void* u, int u_f32, double dt, int np,
                       const double* uvel, double* rho_out, hipStream_t s) {
  build_cells<DIM>(S, r, u, u_f32, dt, np, uvel, s);
  const SphGeom& g = S->g;
  PassOut o{};
  o.rho = rho_out;
  hipLaunchKernelGGL((k_sph_pass<DIM, PASS_DENSITY>), dim3(S->ncells), dim3(kBlock), 0, s, S->pos,
                     S->vel, S->rho, S->start, S->perm, S->ncells, g, o);
  QL_LAUNCHED();
}

template <int DIM, int MODE>
void accel(qlles_sph* S, const PassOut& o, hipStream_t s) {
  hipLaunchKernelGGL((k_sph_pass<DIM, MODE>), dim3(S->ncells), dim3(kBlock), 0, s, S->pos, S->vel,
                     S->rho, S->start, S->perm, S->ncells, S->g, o);
  QL_LAUNCHED();
}

template <class F>
void by_dim(const qlles_sph* S, F&& f) {
  if (S->dim == 3) f(std::integral_constant<int, 3>{});
  else f(std::integral_constant<int, 2>{});
}

}  // namespace

extern "C" {

int qlles_sph_create(int dim, int n, double box, int is_physical, double u_ref, qlles_sph** out) {
  if (!out) return fail(QLLES_ERR_ARG, "null argument");
  *out = nullptr;
  if (dim != 2 && dim != 3) return fail(QLLES_ERR_SHAPE, "dim must be 2 or 3");
  // cutoff 3 L/n below L/2; workspace ~170 B per particle
  if (n < 7 || n > (dim == 3 ? 256 : 4096)) return fail(QLLES_ERR_SHAPE, "n must be in [7, 256] (3-D) / [7, 4096] (2-D)");
  if (!(box > 0.0) || !std::isfinite(box)) return fail(QLLES_ERR_ARG, "box side must be positive");
  if (!std::isfinite(u_ref)) return fail(QLLES_ERR_ARG, "u_ref must be finite");
  qlles_sph* S = new (std::nothrow) qlles_sph();
  if (!S) return fail(QLLES_ERR_HIP, "host allocation failed");
  const int rc = guarded([&] {
    S->dim = dim;
    S->n = n;
    S->L = box;
    // relax.py:30-41, 61-64 (utils.py:181-186 with l_ref = 1)
    const double dx_phys = box / n;
    const double lref = is_physical ? 1.0 : dx_phys;
    const double dx = dx_phys / lref;
    SphGeom& g = S->g;
    g.lref = lref;
    g.box = box / lref;
    g.half = g.box * 0.5;
    g.inv_h = 1.0 / dx;
    g.cutoff = 3.0 * dx;
    g.cutoff2 = g.cutoff * g.cutoff;
    g.sigma = quintic_sigma(dx, dim);
    g.mass = std::pow(dx, dim) * 1.0;
    const double u_eos = u_ref / lref, c_eos = 10.0 * u_eos;
    g.p_eos = c_eos * c_eos * 1.0;
    // cells at least one cutoff wide: box / cutoff = n / 3, one fewer when exact
    g.nc = (n - 1) / 3;
    g.inv_cell = g.nc / g.box;
    S->ncells = dim == 3 ? g.nc * g.nc * g.nc : g.nc * g.nc;
    S->cnt = sph_alloc<int>(S->cell_bufs, S->ncells);
    S->start = sph_alloc<int>(S->cell_bufs, S->ncells + 1);
    S->fill = sph_alloc<int>(S->cell_bufs, S->ncells);
    QL_HIP(hipMemset(S->cnt, 0, sizeof(int) * S->ncells));  // later calls leave them cleared (k_sph_scan)
    QL_HIP(hipDeviceSynchronize());
    ensure_cap(S, dim == 3 ? n * n * n : n * n);
  });
  if (rc != QLLES_OK) {
    free_all(S->cell_bufs);
    free_all(S->part_bufs);
    delete S;
    return rc;
  }
  *out = S;
  return QLLES_OK;
}

int qlles_sph_destroy(qlles_sph* S) {
  if (!S) return QLLES_OK;
  (void)hipDeviceSynchronize();
  free_all(S->cell_bufs);
  free_all(S->part_bufs);
  if (S->dft_E) (void)hipFree(S->dft_E);
  if (S->dft_U) (void)hipFree(S->dft_U);
  delete S;
  return QLLES_OK;
}

int qlles_sph_density(qlles_sph* S, const double* r, int np, double* rho, void* stream) {
  if (!S) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 0) return fail(QLLES_ERR_ARG, "np must be >= 0");
  if (np == 0) return QLLES_OK;  // empty arrays may have NULL data
  if (!r || !rho) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] {
    by_dim(S, [&](auto D) {
      build_and_density<decltype(D)::value>(S, r, nullptr, 0, 0.0, np, nullptr, rho, (hipStream_t)stream);
    });
  });
}

int qlles_sph_relax_acc(qlles_sph* S, const double* r, int np, double* acc, double* rho, void* stream) {
  if (!S) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 0) return fail(QLLES_ERR_ARG, "np must be >= 0");
  if (np == 0) return QLLES_OK;
  if (!r || !acc) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] {
    by_dim(S, [&](auto D) {
      constexpr int DIM = decltype(D)::value;
      build_and_density<DIM>(S, r, nullptr, 0, 0.0, np, nullptr, rho, (hipStream_t)stream);
      PassOut o{};
      o.acc = acc;
      accel<DIM, ACC_RELAX>(S, o, (hipStream_t)stream);
    });
  });
}

int qlles_sph_forces(qlles_sph* S, const double* r, const double* u, int np, double* acc_p, double* acc_visc,
                     double* acc_tvf, void* stream) {
  if (!S) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 0) return fail(QLLES_ERR_ARG, "np must be >= 0");
  if (S->g.lref != 1.0) return fail(QLLES_ERR_ARG, "sph forces need physical units (is_physical)");
  if (np == 0) return QLLES_OK;
  if (!r || !u || !acc_p || !acc_visc || !acc_tvf) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] {
    by_dim(S, [&](auto D) {
      constexpr int DIM = decltype(D)::value;
      build_and_density<DIM>(S, r, nullptr, 0, 0.0, np, u, nullptr, (hipStream_t)stream);
      PassOut o{};
      o.acc = acc_p;
      o.acc_visc = acc_visc;
      o.acc_tvf = acc_tvf;
      accel<DIM, ACC_FORCES>(S, o, (hipStream_t)stream);
    });
  });
}

int qlles_sph_relax_advect(qlles_sph* S, double* r, int np, void* u_r, int u_dtype, double dt, double factor,
                           double* acc, void* stream) {
  if (!S) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 0) return fail(QLLES_ERR_ARG, "np must be >= 0");
  if (np > 0 && (!r || !u_r)) return fail(QLLES_ERR_ARG, "null argument");
  if (u_dtype != QLLES_F32 && u_dtype != QLLES_F64) return fail(QLLES_ERR_ARG, "u_dtype must be QLLES_F32/F64");
  if (!(dt != 0.0) || !std::isfinite(dt) || !std::isfinite(factor))
    return fail(QLLES_ERR_ARG, "dt must be finite and non-zero, factor finite");
  if (S->g.lref != 1.0) return fail(QLLES_ERR_ARG, "the relaxation step needs physical units (is_physical)");
  if (np == 0) return QLLES_OK;
  return guarded([&] {
    by_dim(S, [&](auto D) {
      constexpr int DIM = decltype(D)::value;
      const int f32 = u_dtype == QLLES_F32;
      build_and_density<DIM>(S, r, u_r, f32, dt, np, nullptr, nullptr, (hipStream_t)stream);
      PassOut o{};
      o.acc = acc;
      o.rt = S->rt;
      o.r = r;
      o.u_r = u_r;
      o.u_f32 = f32;
      o.dt = dt;
      const double fd = factor * dt;
      o.fdt2 = fd * fd;
      o.L = S->L;
      accel<DIM, ACC_ADVECT>(S, o, (hipStream_t)stream);
    });
  });
}

int qlles_sph_mls_grid(qlles_sph* S, const double* r, const double* u_r, int np, double* u_grid, void* stream) {
  if (!S || !r || !u_r || !u_grid) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 1) return fail(QLLES_ERR_ARG, "np must be >= 1");
  return guarded([&] {
    by_dim(S, [&](auto D) {
      constexpr int DIM = decltype(D)::value;
      hipStream_t s = (hipStream_t)stream;
      build_cells<DIM>(S, r, nullptr, 0, 0.0, np, u_r, s);
      const int n = S->n;
      const long long nodes = DIM == 3 ? (long long)n * n * n : (long long)n * n;
      const double h = 2.0 / 3.0 * (S->g.box / n);  // turbulence.py:167
      hipLaunchKernelGGL((k_sph_mls<DIM>), dim3(blocks(nodes, 64)), dim3(64), 0, s, S->pos, S->vel, S->start, S->g,
                         n, h, quintic_sigma(h, DIM), u_grid);
      QL_LAUNCHED();
    });
  });
}

int qlles_sph_dft_hat(qlles_sph* S, const double* r, const double* u_r, int np, void* u_hat, void* stream) {
  if (!S || !u_hat) return fail(QLLES_ERR_ARG, "null argument");
  if (np < 0) return fail(QLLES_ERR_ARG, "np must be >= 0");
  if (np > 0 && (!r || !u_r)) return fail(QLLES_ERR_ARG, "null argument");
  return guarded([&] {
    by_dim(S, [&](auto D) {
      constexpr int DIM = decltype(D)::value;
      hipStream_t s = (hipStream_t)stream;
      const int n = S->n;
      const size_t rows = DIM == 3 ? (size_t)n * n : (size_t)n;
      const size_t e_elems = (size_t)DIM * std::max(np, 1) * n;
      if (e_elems > S->dft_E_elems) {
        if (S->dft_E) {
          QL_HIP(hipDeviceSynchronize());
          QL_HIP(hipFree(S->dft_E));
          S->dft_E = nullptr;
        }
        QL_HIP(hipMalloc(&S->dft_E, e_elems * sizeof(double2)));
        S->dft_E_elems = e_elems;
      }
      // U (DIM x rows x n) followed by the dft_ps partial spectra; the subset
      // count fills the grid to ~2048 workgroups (n and dim are fixed per context)
      const size_t u_elems = (size_t)DIM * rows * n;
      const int kw = n <= 32 ? 32 : (n <= 64 ? 64 : 128);
      const long long mode_blocks = (long long)((rows + DFT_THREADS / kw - 1) / (DFT_THREADS / kw)) * ((n + kw - 1) / kw);
      if (!S->dft_U) {
        S->dft_ps = (int)std::max<long long>(1, std::min<long long>(DFT_PS_MAX, 2048 / mode_blocks));
        QL_HIP(hipMalloc(&S->dft_U, (S->dft_ps + 1) * u_elems * sizeof(double2)));
      }
      const int ps = S->dft_ps;
      double2* E = static_cast<double2*>(S->dft_E);
      double2* U = static_cast<double2*>(S->dft_U);
      double2* Up = U + u_elems;
      if (np > 0)
        hipLaunchKernelGGL(k_dft_phase, dim3(blocks((long long)DIM * np * n, 256)), dim3(256), 0, s, r, np, DIM, n,
                           E);
      auto sum = [&](auto kw) {
        constexpr int KW = decltype(kw)::value;
        const dim3 g((unsigned)((rows + DFT_THREADS / KW - 1) / (DFT_THREADS / KW)), (unsigned)((n + KW - 1) / KW),
                     (unsigned)ps);
        hipLaunchKernelGGL((k_dft_sum<DIM, KW>), g, dim3(DFT_THREADS), 0, s, E, u_r, np, n, (int)rows, ps, Up);
      };
      if (n <= 32) sum(std::integral_constant<int, 32>{});
      else if (n <= 64) sum(std::integral_constant<int, 64>{});
      else sum(std::integral_constant<int, 128>{});
      hipLaunchKernelGGL(k_dft_reduce, dim3(blocks((long long)u_elems, 256)), dim3(256), 0, s, Up, u_elems, ps, U);
      hipLaunchKernelGGL((k_dft_half<DIM>), dim3(blocks(3LL * rows * (n / 2 + 1), 256)), dim3(256), 0, s, U, n,
                         static_cast<double2*>(u_hat));
      QL_LAUNCHED();
    });
  });
}

}  // extern "C"
