// Grid -> particle interpolation u_r(x) = Re sum_{k in [-n/2,n/2)^d} G_k e^{i k.x}
// (integrator.py:36-91; finufft nufft{2,3}d2 with isign=+1 at integrator.py:53,55).
//
// Two backends:
//  * NUFFT (type 2): the coefficients G_k are divided by the kernel's Fourier
//    transform, zero-padded onto a fine grid of nf = 2n points per axis,
//    inverse FFT'd with the library's own pencil FFT, and gathered at each
//    particle with the exponential-of-semicircle kernel
//        phi(z) = exp(beta (sqrt(1 - z^2) - 1)),  |z| <= 1,
//    of width w = ceil(log10(1/tol)) + 2 grid points and beta = 2.30 w
//    (Barnett, Magland & af Klinteberg, SIAM J. Sci. Comput. 41 (2019)).
//    One 64-lane wavefront gathers one particle: lanes stride over the w^2
//    (dy, dz) stencil columns and loop over dx, then reduce across the wave.
//  * DFT: the exact direct sum (interpolate_dft, integrator.py:62-91), one
//    particle per lane with the coefficient planes staged through LDS and
//    FP64 phases.
// G_k comes either from a physical n-grid field (interpolate(u, r)) or, fused
// with spectral_filtering (utils.py:100-138), straight from the solver's N-grid
// rfft spectrum: G_k = (F_k + F_k*) / (2 N^d), F the Hermitian extension of
// u_hat and k* = k with every -n/2 component replaced by +n/2 (the .real
// projection of the asymmetric crop).  After interpolation an optional FP64
// Euler shift r <- (r + dt u_r) mod 2 pi (shift_fn, integrator.py:21-23).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "aux_kernels.hpp"
#include "fftlib.hpp"
#include "interp.hpp"

struct qlles_interp {
  int dim, n, dtype, backend;
  double tol;
  int w;        // kernel width (grid points)
  double beta;  // kernel shape
  int nf;       // fine grid points per axis
  void* corr;   // [n] deconvolution factors (T)
  void* coef;   // NUFFT: [ncomp][nf^d] zero-padded coefficients; DFT: [ncomp][n^d] shifted coefficients
  void* fine;   // NUFFT: [ncomp][nf^d] fine-grid values
  void* tmp;    // [ncomp][n^(d-1)][n/2+1] rfft of a physical n-grid field
  void* packed; // 3-D NUFFT: [nf^3][4] real fine-grid values (u, v, w, 0) for the gather
  void* adj_a;  // adjoint (3-D NUFFT, allocated on first use): [3][nf^3] complex
  void* adj_b;  //   second [3][nf^3] complex buffer
};

namespace ql {
using namespace qk;

namespace {

// Gauss-Legendre nodes/weights on [-1, 1] (Newton on P_m)
void gauss_legendre(int m, std::vector<double>& x, std::vector<double>& wts) {
  x.resize(m);
  wts.resize(m);
  for (int i = 0; i < m; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (m + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= m; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double dp = m * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= m; ++k) {
      const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    const double dp = m * (z * p1 - p0) / (z * z - 1.0);
    x[i] = z;
    wts[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

template <typename T>
__device__ __forceinline__ T es_kernel(T z, T beta) {
  const T a = (T)1 - z * z;
  return a > (T)0 ? exp(beta * (sqrt(a) - (T)1)) : (T)0;
}

__device__ __forceinline__ double wrap2pi(double v) {
  const double L = 6.283185307179586476925286766559;
  v = v - L * floor(v / L);
  return v >= L ? v - L : v;
}

// coefficients G_k from a source rfft spectrum (grid Ns, row stride ld)
template <typename T>
__device__ __forceinline__ void build_coef_run(const cx<T>* __restrict__ uh, int Ns, int ld, size_t qplane, int dim,
                                               int n, double scale, const T* __restrict__ corr,
                                               cx<T>* __restrict__ out, int mode, int nf) {
  const int ncomp = dim;
  const size_t per = dim == 3 ? (size_t)n * n * n : (size_t)n * n;
  const size_t cs_in = (dim == 3 ? (size_t)Ns * Ns : (size_t)Ns) * ld;
  const int h = n / 2;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < per * ncomp; m += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(m / per);
    const size_t q = m % per;
    const cx<T>* u = uh + c * cs_in;
    // 3-D source with a separate Nyquist plane at element qplane (0: in the rows)
    const cx<T>* uq = (dim == 3 && qplane != 0) ? uh + qplane + (size_t)c * Ns * Ns : nullptr;
    cx<T> g;
    int ix, iy, iz = 0;
    if (dim == 3) {
      ix = (int)(q / ((size_t)n * n));
      iy = (int)((q / n) % n);
      iz = (int)(q % n);
      const int kx = ix - h, ky = iy - h, kz = iz - h;
      const int sx = kx == -h ? h : kx, sy = ky == -h ? h : ky, sz = kz == -h ? h : kz;
      g = cscale(cadd(full_spectrum3(u, Ns, ld, kx, ky, kz, uq), full_spectrum3(u, Ns, ld, sx, sy, sz, uq)),
                 (T)(0.5 * scale));
      if (mode == 0) {
        g = cscale(g, corr[ix] * corr[iy] * corr[iz]);
        const size_t o = (((size_t)c * nf + (kx + nf) % nf) * nf + (ky + nf) % nf) * nf + (kz + nf) % nf;
        out[o] = g;
      } else {
        out[(size_t)c * per + q] = g;
      }
    } else {
      ix = (int)(q / n);
      iy = (int)(q % n);
      const int kx = ix - h, ky = iy - h;
      const int sx = kx == -h ? h : kx, sy = ky == -h ? h : ky;
      g = cscale(cadd(full_spectrum2(u, Ns, ld, kx, ky), full_spectrum2(u, Ns, ld, sx, sy)), (T)(0.5 * scale));
      if (mode == 0) {
        g = cscale(g, corr[ix] * corr[iy]);
        const size_t o = ((size_t)c * nf + (kx + nf) % nf) * nf + (ky + nf) % nf;
        out[o] = g;
      } else {
        out[(size_t)c * per + q] = g;
      }
    }
  }
}

// per-launch arguments of the coefficient builder and the gathers, passed in
// a Pack (ns3d.hpp): z-slice m of the grid runs entry m (one entry for one
// interpolator, one per member for a batch group)
template <typename T>
struct CoefArgs {
  const cx<T>* uh;
  int Ns, ld;
  size_t qplane;
  double scale;
  const T* corr;
  cx<T>* out;
};
template <typename T>
struct GatherArgs {
  const T* f;  // fine grid: [dim][nf^dim] complex (2-D) or the packed [nf^3][4] reals (3-D)
  const double* r;
  int np;
  T* ur;
  double dt;
  double* r_out;
};

template <typename T>
__global__ void k_build_coef(Pack<CoefArgs<T>> pk, int dim, int n, int mode, int nf) {
  const CoefArgs<T> a = pk.m[blockIdx.z];
  build_coef_run<T>(a.uh, a.Ns, a.ld, a.qplane, dim, n, a.scale, a.corr, a.out, mode, nf);
}

// one wavefront per particle; lanes own (dy, dz) columns of the w^3 stencil
template <typename T, int DIM>
__device__ __forceinline__ void gather_run(const cx<T>* __restrict__ f, int nf, int w, T beta,
                                           const double* __restrict__ r, int np, T* __restrict__ ur, double dt,
                                           double* __restrict__ r_out) {
  const int lane = threadIdx.x & 63;
  const int pid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pid >= np) return;  // whole wavefront exits together
  const double h = 6.283185307179586476925286766559 / nf;
  const double halfw = 0.5 * w;
  const T two_over_w = (T)(2.0 / w);
  double t[DIM];
  int i0[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    t[d] = r[(size_t)pid * DIM + d] / h;
    i0[d] = (int)ceil(t[d] - halfw);
  }
  T acc[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) acc[d] = 0;
  const size_t cs = DIM == 3 ? (size_t)nf * nf * nf : (size_t)nf * nf;
  if (DIM == 3) {
    for (int q = lane; q < w * w; q += 64) {
      const int dy = q / w, dz = q % w;
      const int iy = i0[1] + dy, iz = i0[2] + dz;
      const T wy = es_kernel((T)(t[1] - iy) * two_over_w, beta);
      const T wz = es_kernel((T)(t[2] - iz) * two_over_w, beta);
      const T wyz = wy * wz;
      const int jy = ((iy % nf) + nf) % nf, jz = ((iz % nf) + nf) % nf;
      for (int dx = 0; dx < w; ++dx) {
        const int ix = i0[0] + dx;
        const T wx = es_kernel((T)(t[0] - ix) * two_over_w, beta) * wyz;
        const int jx = ((ix % nf) + nf) % nf;
        const size_t o = ((size_t)jx * nf + jy) * nf + jz;
#pragma unroll
        for (int c = 0; c < DIM; ++c) acc[c] += wx * f[c * cs + o].re;
      }
    }
  } else {
    for (int dy = lane; dy < w; dy += 64) {
      const int iy = i0[1] + dy;
      const T wy = es_kernel((T)(t[1] - iy) * two_over_w, beta);
      const int jy = ((iy % nf) + nf) % nf;
      for (int dx = 0; dx < w; ++dx) {
        const int ix = i0[0] + dx;
        const T wx = es_kernel((T)(t[0] - ix) * two_over_w, beta) * wy;
        const int jx = ((ix % nf) + nf) % nf;
        const size_t o = (size_t)jx * nf + jy;
#pragma unroll
        for (int c = 0; c < DIM; ++c) acc[c] += wx * f[c * cs + o].re;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      ur[(size_t)pid * DIM + c] = acc[c];
      if (r_out) r_out[(size_t)pid * DIM + c] = wrap2pi(r[(size_t)pid * DIM + c] + dt * (double)acc[c]);
    }
  }
}

template <typename T, int DIM>
__global__ void __launch_bounds__(256)
k_gather(Pack<GatherArgs<T>> pk, int nf, int w, T beta) {
  const GatherArgs<T> a = pk.m[blockIdx.z];
  gather_run<T, DIM>(reinterpret_cast<const cx<T>*>(a.f), nf, w, beta, a.r, a.np, a.ur, a.dt, a.r_out);
}

// real parts of the three fine-grid components, interleaved per point; grid
// row blockIdx.y packs the blockIdx.y-th of consecutive [3][cs] grids
template <typename T>
__global__ void k_pack3(const cx<T>* __restrict__ f, T* __restrict__ out, size_t cs) {
  f += 3 * cs * blockIdx.y;
  out += 4 * cs * blockIdx.y;
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < cs; o += (size_t)gridDim.x * blockDim.x) {
    out[4 * o] = f[o].re;
    out[4 * o + 1] = f[cs + o].re;
    out[4 * o + 2] = f[2 * cs + o].re;
    out[4 * o + 3] = (T)0;
  }
}

// the 3-D gather on the packed grid: one 4-component load per stencil point;
// the kernel width W is a compile-time constant so each lane's W loads along
// x are issued back to back
template <typename T, int W>
__device__ __forceinline__ void gather3p_run(const T* __restrict__ f4, int nf, T beta, const double* __restrict__ r,
                                             int np, T* __restrict__ ur, double dt, double* __restrict__ r_out) {
  constexpr int w = W;
  const int lane = threadIdx.x & 63;
  // workgroups go round-robin over the 8 XCDs; give each XCD a contiguous run
  // of particles (neighbours in the caller's order), so its L2 holds their
  // part of the fine grid
  const int nb = gridDim.x, q8 = nb / 8, r8 = nb % 8;
  const int xcd = blockIdx.x % 8, kb = blockIdx.x / 8;
  const int blk = xcd * q8 + (xcd < r8 ? xcd : r8) + kb;
  const int pid = blk * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pid >= np) return;  // whole wavefront exits together
  const double h = 6.283185307179586476925286766559 / nf;
  const double halfw = 0.5 * w;
  const T two_over_w = (T)(2.0 / w);
  double t[3];
  int i0[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    t[d] = r[(size_t)pid * 3 + d] / h;
    i0[d] = (int)ceil(t[d] - halfw);
  }
  // One 1-D kernel weight and wrapped index per lane -- lanes [0, W) for x,
  // [W, 2W) for y, [2W, 3W) for z -- then broadcast by cross-lane moves, so
  // a wavefront evaluates 3W kernels instead of 64 * (W + 2).
  static_assert(3 * W <= 64, "kernel width");
  const bool pow2 = (nf & (nf - 1)) == 0;
  T wl = 0;
  int jl = 0;
  if (lane < 3 * W) {
    const int ax = lane / W, dl = lane % W;
    const double td = ax == 0 ? t[0] : (ax == 1 ? t[1] : t[2]);
    const int i = (ax == 0 ? i0[0] : (ax == 1 ? i0[1] : i0[2])) + dl;
    wl = es_kernel((T)(td - i) * two_over_w, beta);
    jl = pow2 ? (i & (nf - 1)) : ((i % nf) + nf) % nf;
  }
  T acc[3] = {0, 0, 0};
  T wxs[W];
  int jxs[W];
#pragma unroll
  for (int dx = 0; dx < W; ++dx) {
    wxs[dx] = __shfl(wl, dx, 64);
    jxs[dx] = __shfl(jl, dx, 64);
  }
  constexpr int NIT = (W * W + 63) / 64;
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int q = lane + 64 * it;
    const int qc = q < W * W ? q : 0;  // every lane takes part in the moves
    const int dy = qc / W, dz = qc % W;
    const T wyz = __shfl(wl, W + dy, 64) * __shfl(wl, 2 * W + dz, 64);
    const int jy = __shfl(jl, W + dy, 64), jz = __shfl(jl, 2 * W + dz, 64);
    if (q >= W * W) continue;
#pragma unroll
    for (int dx = 0; dx < W; ++dx) {
      const T wx = wxs[dx] * wyz;
      const T* p = f4 + 4 * ((jxs[dx] * nf + jy) * nf + jz);  // nf <= 128: 32-bit offsets
      if constexpr (sizeof(T) == 4) {
        const float4 v = *reinterpret_cast<const float4*>(p);  // 16-byte aligned point
        acc[0] += wx * v.x;
        acc[1] += wx * v.y;
        acc[2] += wx * v.z;
      } else {
        const double2 v = *reinterpret_cast<const double2*>(p);
        acc[0] += wx * v.x;
        acc[1] += wx * v.y;
        acc[2] += wx * p[2];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ur[(size_t)pid * 3 + c] = acc[c];
      if (r_out) r_out[(size_t)pid * 3 + c] = wrap2pi(r[(size_t)pid * 3 + c] + dt * (double)acc[c]);
    }
  }
}

template <typename T, int W>
__global__ void __launch_bounds__(256)
k_gather3p(Pack<GatherArgs<T>> pk, int nf, T beta) {
  const GatherArgs<T> a = pk.m[blockIdx.z];
  gather3p_run<T, W>(a.f, nf, beta, a.r, a.np, a.ur, a.dt, a.r_out);
}

// exact direct sum; one particle per lane, coefficient planes through LDS
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
k_dft(const cx<T>* __restrict__ G, int n, const double* __restrict__ r, int np, T* __restrict__ ur, double dt,
      double* __restrict__ r_out) {
  extern __shared__ unsigned char s_raw[];
  cx<T>* s_plane = reinterpret_cast<cx<T>*>(s_raw);  // [DIM][n^(DIM-1)]
  const int pid = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = pid < np;
  double x[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) x[d] = ok ? r[(size_t)pid * DIM + d] : 0.0;
  const int h = n / 2;
  const size_t plane = DIM == 3 ? (size_t)n * n : (size_t)n;
  const size_t per = plane * n;
  double acc[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) acc[d] = 0.0;
  for (int ixk = 0; ixk < n; ++ixk) {
    __syncthreads();
    for (size_t e = threadIdx.x; e < plane * DIM; e += blockDim.x) {
      const size_t c = e / plane, q = e % plane;
      s_plane[e] = G[c * per + (size_t)ixk * plane + q];
    }
    __syncthreads();
    double sx, cxv;
    sincos((double)(ixk - h) * x[0], &sx, &cxv);
    if (DIM == 3) {
      double ez_s, ez_c, s0, c0;
      sincos(x[2], &ez_s, &ez_c);
      sincos(-(double)h * x[2], &s0, &c0);
      for (int iyk = 0; iyk < n; ++iyk) {
        double sy, cy;
        sincos((double)(iyk - h) * x[1], &sy, &cy);
        const double pr = cxv * cy - sx * sy, pi = cxv * sy + sx * cy;  // e^{i(kx x + ky y)}
        double zr = pr * c0 - pi * s0, zi = pr * s0 + pi * c0;       // times e^{-i h z}
        const cx<T>* row = s_plane + (size_t)iyk * n;
        for (int izk = 0; izk < n; ++izk) {
#pragma unroll
          for (int c = 0; c < DIM; ++c) {
            const cx<T> g = row[c * plane + izk];
            acc[c] += (double)g.re * zr - (double)g.im * zi;
          }
          const double t = zr * ez_c - zi * ez_s;
          zi = zr * ez_s + zi * ez_c;
          zr = t;
        }
      }
    } else {
      double ey_s, ey_c, s0, c0;
      sincos(x[1], &ey_s, &ey_c);
      sincos(-(double)h * x[1], &s0, &c0);
      double zr = cxv * c0 - sx * s0, zi = cxv * s0 + sx * c0;
      for (int iyk = 0; iyk < n; ++iyk) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const cx<T> g = s_plane[c * plane + iyk];
          acc[c] += (double)g.re * zr - (double)g.im * zi;
        }
        const double t = zr * ey_c - zi * ey_s;
        zi = zr * ey_s + zi * ey_c;
        zr = t;
      }
    }
  }
  if (ok) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      ur[(size_t)pid * DIM + c] = (T)acc[c];
      if (r_out) r_out[(size_t)pid * DIM + c] = wrap2pi(x[c] + dt * (double)(T)acc[c]);
    }
  }
}

// ---- adjoint of the fused filter + type-2 NUFFT interpolation (3-D) ----
//
// Forward: u_r(p) = Re sum_{k in W} G_k e^{i k.r_p}, W = [-h, h)^3, h = n/2,
//   G = H_n(crop(FS(u_hat))) / N^3  (FS: Hermitian full spectrum of the
//   N-grid rfft state; H_n: the n-grid Hermitian projection that the
//   reference's .real implies, pairing -h with -h),
// evaluated as gather(ES weights) of ifft(pad(corr * G)) on the fine grid.
// Adjoint for the cotangent ub (np x 3):
//   fbar = spread(ub) (same ES weights), cbar = fft(fbar),
//   Gbar_k = corr(k) cbar_k (k in W),
//   lam_k += c_kz / N^3 * 0.5 (1[k in W] H_n(Gbar)_k + 1[-k in W] conj(H_n(Gbar)_{-k}))
// for every stored half-spectrum entry k (c = 1 on the kz = 0 / N/2 planes,
// 2 elsewhere: the adjoint of irfftn).  Position cotangent:
//   rbar_{p,j} += sum_c ub_{p,c} Re sum_W (i k_j) G_k e^{i k.r_p}.

// fine-grid spreading: one wavefront per particle, atomics into the real parts
template <typename T>
__global__ void __launch_bounds__(256)
k_spread3(cx<T>* __restrict__ f, int nf, int w, T beta, const double* __restrict__ r, const T* __restrict__ ub,
          int np) {
  const int lane = threadIdx.x & 63;
  const int pid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pid >= np) return;
  const double h = 6.283185307179586476925286766559 / nf;
  const double halfw = 0.5 * w;
  const T two_over_w = (T)(2.0 / w);
  double t[3];
  int i0[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    t[d] = r[(size_t)pid * 3 + d] / h;
    i0[d] = (int)ceil(t[d] - halfw);
  }
  const T u0 = ub[(size_t)pid * 3], u1 = ub[(size_t)pid * 3 + 1], u2 = ub[(size_t)pid * 3 + 2];
  const size_t cs = (size_t)nf * nf * nf;
  for (int q = lane; q < w * w; q += 64) {
    const int dy = q / w, dz = q % w;
    const int iy = i0[1] + dy, iz = i0[2] + dz;
    const T wyz = es_kernel((T)(t[1] - iy) * two_over_w, beta) * es_kernel((T)(t[2] - iz) * two_over_w, beta);
    const int jy = ((iy % nf) + nf) % nf, jz = ((iz % nf) + nf) % nf;
    for (int dx = 0; dx < w; ++dx) {
      const int ix = i0[0] + dx;
      const T wx = es_kernel((T)(t[0] - ix) * two_over_w, beta) * wyz;
      const int jx = ((ix % nf) + nf) % nf;
      const size_t o = ((size_t)jx * nf + jy) * nf + jz;
      atomicAdd(&f[o].re, wx * u0);
      atomicAdd(&f[cs + o].re, wx * u1);
      atomicAdd(&f[2 * cs + o].re, wx * u2);
    }
  }
}

template <typename T>
__device__ __forceinline__ cx<T> gbar_at(const cx<T>* cb, const T* corr, int n, int nf, size_t cs, int c, int kx,
                                          int ky, int kz) {
  const int h = n / 2;
  const size_t o = ((size_t)((kx + nf) % nf) * nf + (ky + nf) % nf) * nf + (kz + nf) % nf;
  return cscale(cb[c * cs + o], corr[kx + h] * corr[ky + h] * corr[kz + h]);
}
// H_n(Gbar)_k: pairs k with -k, where -(-h) wraps to -h
template <typename T>
__device__ __forceinline__ cx<T> hn_gbar(const cx<T>* cb, const T* corr, int n, int nf, size_t cs, int c, int kx,
                                          int ky, int kz) {
  const int h = n / 2;
  auto neg = [h](int v) { return v == -h ? -h : -v; };
  const cx<T> a = gbar_at(cb, corr, n, nf, cs, c, kx, ky, kz);
  const cx<T> b = cconj(gbar_at(cb, corr, n, nf, cs, c, neg(kx), neg(ky), neg(kz)));
  return cscale(cadd(a, b), (T)0.5);
}

// one thread per stored entry (kx, ky in (-h-1, h], kz in [0, h]) x component
template <typename T>
__global__ void k_window_adjoint3(const cx<T>* __restrict__ cb, const T* __restrict__ corr, int n, int nf, int N,
                                  cx<T>* __restrict__ lam) {
  const int h = n / 2, Nh = N / 2 + 1;
  const int ax = n + 1;  // kx, ky in [-h, h]
  const size_t total = (size_t)3 * ax * ax * (h + 1);
  const size_t cs = (size_t)nf * nf * nf, csl = (size_t)N * N * Nh;
  const double inv_n3 = 1.0 / ((double)N * N * N);
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(m / ((size_t)ax * ax * (h + 1)));
    const size_t q = m % ((size_t)ax * ax * (h + 1));
    const int kx = (int)(q / ((size_t)ax * (h + 1))) - h, ky = (int)((q / (h + 1)) % ax) - h, kz = (int)(q % (h + 1));
    // each stored entry once: its signed row indices are the fftfreq values
    // (when n == N, +N/2 is the entry of -N/2)
    const int ix = (kx + N) % N, iy = (ky + N) % N;
    if (kfreq(ix, N) != kx || kfreq(iy, N) != ky) continue;
    // -k on the N-periodic grid, as signed fftfreq values (the rfft axis: kz = N/2 -> -N/2)
    const int nx = kfreq((N - ix) % N, N), ny = kfreq((N - iy) % N, N), nz = -kz;
    // the rfft-axis entry kz = N/2 is the periodic index -N/2 (in W when n == N)
    const int kzc = 2 * kz == N ? -kz : kz;
    const bool in_w = kx < h && ky < h && kzc >= -h && kzc < h;  // k in W (kx, ky >= -h by construction)
    const bool neg_w = nx >= -h && nx < h && ny >= -h && ny < h && nz >= -h && nz < h;
    if (!in_w && !neg_w) continue;
    cx<T> v = {0, 0};
    if (in_w) v = cadd(v, hn_gbar(cb, corr, n, nf, cs, c, kx, ky, kzc));
    if (neg_w) v = cadd(v, cconj(hn_gbar(cb, corr, n, nf, cs, c, nx, ny, nz)));
    const double cw = (kz == 0 || 2 * kz == N) ? 1.0 : 2.0;
    const size_t o = (size_t)c * csl + ((size_t)ix * N + iy) * Nh + kz;
    lam[o] = cadd(lam[o], cscale(v, (T)(0.5 * cw * inv_n3)));
  }
}

// coefficients of d/dr_j: out = i k_j coef on the padded fine-grid window
template <typename T>
__global__ void k_coef_deriv(const cx<T>* __restrict__ co, cx<T>* __restrict__ out, int nf, int j) {
  const size_t cs = (size_t)nf * nf * nf, total = 3 * cs;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const size_t o = m % cs;
    const int i = j == 0 ? (int)(o / ((size_t)nf * nf)) : (j == 1 ? (int)((o / nf) % nf) : (int)(o % nf));
    const int k = i < nf / 2 ? i : i - nf;
    out[m] = cmuli(cscale(co[m], (T)k));
  }
}

// rbar_{p,j} += sum_c ub_{p,c} (gathered derivative field c)
template <typename T>
__global__ void __launch_bounds__(256)
k_gather_dot3(const cx<T>* __restrict__ f, int nf, int w, T beta, const double* __restrict__ r, int np,
              const T* __restrict__ ub, double* __restrict__ rbar, int j) {
  const int lane = threadIdx.x & 63;
  const int pid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pid >= np) return;
  const double h = 6.283185307179586476925286766559 / nf;
  const double halfw = 0.5 * w;
  const T two_over_w = (T)(2.0 / w);
  double t[3];
  int i0[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    t[d] = r[(size_t)pid * 3 + d] / h;
    i0[d] = (int)ceil(t[d] - halfw);
  }
  const size_t cs = (size_t)nf * nf * nf;
  T acc[3] = {0, 0, 0};
  for (int q = lane; q < w * w; q += 64) {
    const int dy = q / w, dz = q % w;
    const int iy = i0[1] + dy, iz = i0[2] + dz;
    const T wyz = es_kernel((T)(t[1] - iy) * two_over_w, beta) * es_kernel((T)(t[2] - iz) * two_over_w, beta);
    const int jy = ((iy % nf) + nf) % nf, jz = ((iz % nf) + nf) % nf;
    for (int dx = 0; dx < w; ++dx) {
      const int ix = i0[0] + dx;
      const T wx = es_kernel((T)(t[0] - ix) * two_over_w, beta) * wyz;
      const int jx = ((ix % nf) + nf) % nf;
      const size_t o = ((size_t)jx * nf + jy) * nf + jz;
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] += wx * f[c * cs + o].re;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
  }
  if (lane == 0) {
    double dot = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) dot += (double)ub[(size_t)pid * 3 + c] * (double)acc[c];
    rbar[(size_t)pid * 3 + j] += dot;
  }
}

template <typename T>
__global__ void k_zero_cx(cx<T>* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = cx<T>{0, 0};
}

template <typename T, int W>
void launch_gather3p_w(dim3 g, hipStream_t s, const Pack<GatherArgs<T>>& pk, int nf, T beta) {
  hipLaunchKernelGGL((k_gather3p<T, W>), g, dim3(256), 0, s, pk, nf, beta);
}
template <typename T>
void launch_gather3p(int w, dim3 g, hipStream_t s, const Pack<GatherArgs<T>>& pk, int nf, T beta) {
  switch (w) {
    case 2: launch_gather3p_w<T, 2>(g, s, pk, nf, beta); break;
    case 3: launch_gather3p_w<T, 3>(g, s, pk, nf, beta); break;
    case 4: launch_gather3p_w<T, 4>(g, s, pk, nf, beta); break;
    case 5: launch_gather3p_w<T, 5>(g, s, pk, nf, beta); break;
    case 6: launch_gather3p_w<T, 6>(g, s, pk, nf, beta); break;
    case 7: launch_gather3p_w<T, 7>(g, s, pk, nf, beta); break;
    case 8: launch_gather3p_w<T, 8>(g, s, pk, nf, beta); break;
    case 9: launch_gather3p_w<T, 9>(g, s, pk, nf, beta); break;
    case 10: launch_gather3p_w<T, 10>(g, s, pk, nf, beta); break;
    case 11: launch_gather3p_w<T, 11>(g, s, pk, nf, beta); break;
    case 12: launch_gather3p_w<T, 12>(g, s, pk, nf, beta); break;
    case 13: launch_gather3p_w<T, 13>(g, s, pk, nf, beta); break;
    case 14: launch_gather3p_w<T, 14>(g, s, pk, nf, beta); break;
    case 15: launch_gather3p_w<T, 15>(g, s, pk, nf, beta); break;
    case 16: launch_gather3p_w<T, 16>(g, s, pk, nf, beta); break;
    default: throw std::invalid_argument("NUFFT kernel width out of range");
  }
  QL_LAUNCHED();
}
template <class A>
Pack<A> pack1(const A& a) {
  Pack<A> pk{};
  pk.m[0] = a;
  return pk;
}

size_t tsize(int dtype) { return dtype == QLLES_F32 ? sizeof(float) : sizeof(double); }

template <typename T>
void setup_t(qlles_interp* it) {
  const int n = it->n, d = it->dim;
  const size_t per_f = d == 3 ? (size_t)it->nf * it->nf * it->nf : (size_t)it->nf * it->nf;
  const size_t per_n = d == 3 ? (size_t)n * n * n : (size_t)n * n;
  // deconvolution factors corr(k) = h / psi^(k), psi^(k) = (w h/2) int_{-1}^{1} phi(z) cos(k z w h/2) dz
  std::vector<T> corr(n, (T)1);
  if (it->backend == QLLES_INTERP_NUFFT) {
    std::vector<double> gx, gw;
    gauss_legendre(256, gx, gw);
    const double h = 2.0 * M_PI / it->nf;
    const double a = 0.5 * it->w * h;
    for (int i = 0; i < n; ++i) {
      const double k = i - n / 2;
      double s = 0.0;
      for (size_t q = 0; q < gx.size(); ++q) {
        const double z = gx[q];
        const double ph = std::exp(it->beta * (std::sqrt(std::max(0.0, 1.0 - z * z)) - 1.0));
        s += gw[q] * ph * std::cos(k * z * a);
      }
      corr[i] = (T)(h / (a * s));
    }
  }
  QL_HIP(hipMalloc(&it->corr, sizeof(T) * n));
  QL_HIP(hipMemcpy(it->corr, corr.data(), sizeof(T) * n, hipMemcpyHostToDevice));
  if (it->backend == QLLES_INTERP_NUFFT) {
    QL_HIP(hipMalloc(&it->coef, sizeof(cx<T>) * per_f * d));
    QL_HIP(hipMemset(it->coef, 0, sizeof(cx<T>) * per_f * d));
    // the fill runs on the null stream, which does not order itself with the
    // non-blocking streams a batch launches on: finish it here
    QL_HIP(hipDeviceSynchronize());
    QL_HIP(hipMalloc(&it->fine, sizeof(cx<T>) * per_f * d));
    if (d == 3) QL_HIP(hipMalloc(&it->packed, sizeof(T) * 4 * per_f));
  } else {
    QL_HIP(hipMalloc(&it->coef, sizeof(cx<T>) * per_n * d));
  }
  const size_t nrows = d == 3 ? (size_t)n * n : (size_t)n;
  QL_HIP(hipMalloc(&it->tmp, sizeof(cx<T>) * nrows * (n / 2 + 1) * d));
  // twiddle tables are built now, so no later call allocates (graph capture)
  (void)twiddles<T>(n);
  if (it->backend == QLLES_INTERP_NUFFT) (void)twiddles<T>(it->nf);
}

template <typename T>
void run_t(qlles_interp* it, const cx<T>* uh, int Ns, int ld, size_t q, double scale, const double* r, int np,
           T* ur, double dt, double* r_out, hipStream_t s) {
  const int n = it->n, d = it->dim;
  const size_t per_n = d == 3 ? (size_t)n * n * n : (size_t)n * n;
  const int nb = (int)std::min<size_t>((per_n * d + 255) / 256, 8192);
  const int mode = it->backend == QLLES_INTERP_NUFFT ? 0 : 1;
  hipLaunchKernelGGL((k_build_coef<T>), dim3(nb), dim3(256), 0, s,
                     pack1(CoefArgs<T>{uh, Ns, ld, q, scale, static_cast<const T*>(it->corr),
                                       static_cast<cx<T>*>(it->coef)}),
                     d, n, mode, it->nf);
  QL_LAUNCHED();
  if (np == 0) return;
  if (mode == 0) {
    const int nf = it->nf;
    const cx<T>* twf = twiddles<T>(nf);
    cx<T>* co = static_cast<cx<T>*>(it->coef);
    cx<T>* fi = static_cast<cx<T>*>(it->fine);
    if (d == 3) {
      c2c_strided<T>(nf, co, fi, 3, nf * nf, false, (T)1, twf, s);
      c2c_strided<T>(nf, fi, fi, 3 * nf, nf, false, (T)1, twf, s);
      c2c_contig<T>(nf, fi, fi, 3 * nf * nf, false, (T)1, twf, s);
      const size_t cs = (size_t)nf * nf * nf;
      T* pk = static_cast<T*>(it->packed);
      hipLaunchKernelGGL((k_pack3<T>), dim3((int)std::min<size_t>((cs + 255) / 256, 4096)), dim3(256), 0, s, fi, pk,
                         cs);
      QL_LAUNCHED();
      launch_gather3p<T>(it->w, dim3((np + 3) / 4), s, pack1(GatherArgs<T>{pk, r, np, ur, dt, r_out}), nf,
                         (T)it->beta);
    } else {
      c2c_strided<T>(nf, co, fi, 2, nf, false, (T)1, twf, s);
      c2c_contig<T>(nf, fi, fi, 2 * nf, false, (T)1, twf, s);
      hipLaunchKernelGGL((k_gather<T, 2>), dim3((np + 3) / 4), dim3(256), 0, s,
                         pack1(GatherArgs<T>{reinterpret_cast<const T*>(fi), r, np, ur, dt, r_out}), nf, it->w,
                         (T)it->beta);
    }
    QL_LAUNCHED();
  } else {
    const size_t shm = sizeof(cx<T>) * (d == 3 ? (size_t)n * n : (size_t)n) * d;
    if (d == 3)
      hipLaunchKernelGGL((k_dft<T, 3>), dim3((np + 255) / 256), dim3(256), shm, s,
                         static_cast<const cx<T>*>(it->coef), n, r, np, ur, dt, r_out);
    else
      hipLaunchKernelGGL((k_dft<T, 2>), dim3((np + 255) / 256), dim3(256), shm, s,
                         static_cast<const cx<T>*>(it->coef), n, r, np, ur, dt, r_out);
    QL_LAUNCHED();
  }
}

template <typename T>
void field_t(qlles_interp* it, const T* u, const double* r, int np, T* ur, hipStream_t s) {
  // n-grid rfft of the physical field, then the same coefficient builder with Ns = n
  const int n = it->n, d = it->dim;
  const int nh = n / 2 + 1;
  cx<T>* uh = static_cast<cx<T>*>(it->tmp);
  const cx<T>* twn = twiddles<T>(n);
  if (d == 3) {
    r2c_contig<T>(n, u, uh, 3 * n * n, nh, (T)1, twn, s);
    c2c_strided<T>(n, uh, uh, 3 * n, nh, true, (T)1, twn, s);
    c2c_strided<T>(n, uh, uh, 3, n * nh, true, (T)1, twn, s);
    run_t<T>(it, uh, n, nh, 0, 1.0 / ((double)n * n * n), r, np, ur, 0.0, nullptr, s);
  } else {
    r2c_contig<T>(n, u, uh, 2 * n, nh, (T)1, twn, s);
    c2c_strided<T>(n, uh, uh, 2, nh, true, (T)1, twn, s);
    run_t<T>(it, uh, n, nh, 0, 1.0 / ((double)n * n), r, np, ur, 0.0, nullptr, s);
  }
}

}  // namespace

qlles_interp* make_interp(int dim, int n, int dtype, int backend, double tol) {
  if (!supported_len(n)) throw std::invalid_argument("interpolation grid n must be one of 8,16,32,64,96,128,256");
  auto* it = new qlles_interp{};
  it->dim = dim;
  it->n = n;
  it->dtype = dtype;
  it->backend = backend;
  it->tol = tol;
  if (backend == QLLES_INTERP_NUFFT) {
    it->nf = 2 * n;
    if (!supported_len(it->nf)) {
      delete it;
      throw std::invalid_argument("NUFFT fine grid 2n unsupported for this n");
    }
    if (dim == 3 && n > 64) {
      delete it;
      throw std::invalid_argument("3-D NUFFT supports n <= 64");
    }
    // one point wider than finufft's ceil(log10(10/tol)) so the 3-D relative l2
    // error stays below tol (the 1-D errors of the three axes add up)
    it->w = (int)std::ceil(std::log10(1.0 / tol)) + 2;
    if (it->w < 2) it->w = 2;
    if (it->w > 16) it->w = 16;
    const double bo = it->w == 2 ? 2.20 : it->w == 3 ? 2.26 : it->w == 4 ? 2.38 : 2.30;
    it->beta = bo * it->w;
  } else {
    if (dim == 3 && n > (dtype == QLLES_F32 ? 64 : 32)) {
      delete it;
      throw std::invalid_argument("3-D DFT interpolation supports n <= 64 (FP32) / 32 (FP64)");
    }
    it->nf = n;
    it->w = 0;
    it->beta = 0;
  }
  try {
    if (dtype == QLLES_F32) setup_t<float>(it);
    else setup_t<double>(it);
  } catch (...) {
    destroy_interp(it);
    throw;
  }
  return it;
}

int interp_dim(const qlles_interp* it) { return it->dim; }
int interp_dtype(const qlles_interp* it) { return it->dtype; }
int interp_n(const qlles_interp* it) { return it->n; }

void destroy_interp(qlles_interp* it) {
  if (!it) return;
  if (it->adj_a) (void)hipFree(it->adj_a);
  if (it->adj_b) (void)hipFree(it->adj_b);
  if (it->corr) (void)hipFree(it->corr);
  if (it->coef) (void)hipFree(it->coef);
  if (it->fine) (void)hipFree(it->fine);
  if (it->packed) (void)hipFree(it->packed);
  if (it->tmp) (void)hipFree(it->tmp);
  delete it;
}

void interp_field(qlles_interp* it, const void* u, const double* r, int np, void* u_r, hipStream_t s) {
  if (it->dtype == QLLES_F32)
    field_t<float>(it, static_cast<const float*>(u), r, np, static_cast<float*>(u_r), s);
  else
    field_t<double>(it, static_cast<const double*>(u), r, np, static_cast<double*>(u_r), s);
}

template <typename T>
void hat_vjp_t(qlles_interp* it, int N, const cx<T>* uh, const double* r, int np, const T* ub, cx<T>* lam,
               double* rbar, hipStream_t s) {
  const int n = it->n, nf = it->nf;
  const size_t cs = (size_t)nf * nf * nf;
  if (!it->adj_a) {
    QL_HIP(hipMalloc(&it->adj_a, sizeof(cx<T>) * 3 * cs));
    QL_HIP(hipMalloc(&it->adj_b, sizeof(cx<T>) * 3 * cs));
  }
  cx<T>* fa = static_cast<cx<T>*>(it->adj_a);
  cx<T>* fb = static_cast<cx<T>*>(it->adj_b);
  const cx<T>* twf = twiddles<T>(nf);
  const int nbz = (int)std::min<size_t>((3 * cs + 255) / 256, 8192);
  // spectrum cotangent: spread, forward fine-grid FFT, window adjoint
  hipLaunchKernelGGL((k_zero_cx<T>), dim3(nbz), dim3(256), 0, s, fa, 3 * cs);
  QL_LAUNCHED();
  if (np > 0) {
    hipLaunchKernelGGL((k_spread3<T>), dim3((np + 3) / 4), dim3(256), 0, s, fa, nf, it->w, (T)it->beta, r, ub, np);
    QL_LAUNCHED();
  }
  c2c_strided<T>(nf, fa, fa, 3, nf * nf, true, (T)1, twf, s);
  c2c_strided<T>(nf, fa, fa, 3 * nf, nf, true, (T)1, twf, s);
  c2c_contig<T>(nf, fa, fa, 3 * nf * nf, true, (T)1, twf, s);
  const size_t nw = (size_t)3 * (n + 1) * (n + 1) * (n / 2 + 1);
  hipLaunchKernelGGL((k_window_adjoint3<T>), dim3((int)std::min<size_t>((nw + 255) / 256, 8192)), dim3(256), 0, s,
                     fa, static_cast<const T*>(it->corr), n, nf, N, lam);
  QL_LAUNCHED();
  if (rbar == nullptr || np == 0) return;
  // position cotangent: derivative coefficients of the forward interpolant
  const size_t per_n = (size_t)n * n * n;
  hipLaunchKernelGGL((k_build_coef<T>), dim3((int)std::min<size_t>((per_n * 3 + 255) / 256, 8192)), dim3(256), 0,
                     s,
                     pack1(CoefArgs<T>{uh, N, N / 2 + 1, (size_t)0, 1.0 / ((double)N * N * N),
                                       static_cast<const T*>(it->corr), static_cast<cx<T>*>(it->coef)}),
                     3, n, 0, nf);
  QL_LAUNCHED();
  for (int j = 0; j < 3; ++j) {
    hipLaunchKernelGGL((k_coef_deriv<T>), dim3(nbz), dim3(256), 0, s, static_cast<const cx<T>*>(it->coef), fb, nf, j);
    QL_LAUNCHED();
    c2c_strided<T>(nf, fb, fb, 3, nf * nf, false, (T)1, twf, s);
    c2c_strided<T>(nf, fb, fb, 3 * nf, nf, false, (T)1, twf, s);
    c2c_contig<T>(nf, fb, fb, 3 * nf * nf, false, (T)1, twf, s);
    hipLaunchKernelGGL((k_gather_dot3<T>), dim3((np + 3) / 4), dim3(256), 0, s, fb, nf, it->w, (T)it->beta, r, np,
                       ub, rbar, j);
    QL_LAUNCHED();
  }
}

void interp_hat_vjp(qlles_interp* it, int N, const void* u_hat, const double* r, int np, const void* u_bar,
                    void* lam_hat, double* r_bar, hipStream_t s) {
  if (it->dim != 3 || it->backend != QLLES_INTERP_NUFFT)
    throw std::invalid_argument("the interpolation adjoint is implemented for the 3-D NUFFT backend");
  if (N < it->n || (N % 2) != 0) throw std::invalid_argument("solver grid N must be even and >= n");
  if (it->dtype == QLLES_F32)
    hat_vjp_t<float>(it, N, static_cast<const cx<float>*>(u_hat), r, np, static_cast<const float*>(u_bar),
                     static_cast<cx<float>*>(lam_hat), r_bar, s);
  else
    hat_vjp_t<double>(it, N, static_cast<const cx<double>*>(u_hat), r, np, static_cast<const double*>(u_bar),
                      static_cast<cx<double>*>(lam_hat), r_bar, s);
}

void interp_hat(qlles_interp* it, int N, const void* u_hat, int ld, size_t q, double* r, int np, void* u_r,
                double dt, hipStream_t s) {
  if (N < it->n || (N % 2) != 0) throw std::invalid_argument("solver grid N must be even and >= n");
  if (q == 0 ? ld < N / 2 + 1 : (ld < N / 2 || it->dim != 3))
    throw std::invalid_argument("spectrum row stride too small for its layout");
  const double scale = 1.0 / (it->dim == 3 ? (double)N * N * N : (double)N * N);
  double* r_out = dt != 0.0 ? r : nullptr;
  if (it->dtype == QLLES_F32)
    run_t<float>(it, static_cast<const cx<float>*>(u_hat), N, ld, q, scale, r, np, static_cast<float*>(u_r), dt,
                 r_out, s);
  else
    run_t<double>(it, static_cast<const cx<double>*>(u_hat), N, ld, q, scale, r, np, static_cast<double*>(u_r), dt,
                  r_out, s);
}


// ---------------- groups of batch members ----------------

struct InterpGroup {
  int dim = 0, n = 0, nf = 0, w = 0, dtype = 0, M = 0, max_np = 0;
  double beta = 0;
  void* coef = nullptr;    // [M][dim][nf^dim] zero-padded, deconvolved coefficients
  void* fine = nullptr;    // [M][dim][nf^dim] fine-grid values
  void* packed = nullptr;  // 3-D: [M][nf^3][4]
  // per member: coefficient-builder arguments reading the spectrum after a
  // step that is not / is the member's last, and gather arguments
  std::vector<unsigned char> cont, last, gath;  // CoefArgs<T>[M], CoefArgs<T>[M], GatherArgs<T>[M]
};

bool interp_groupable(const qlles_interp* a, const qlles_interp* b) {
  return a->backend == QLLES_INTERP_NUFFT && b->backend == QLLES_INTERP_NUFFT && a->dim == b->dim && a->n == b->n &&
         a->dtype == b->dtype && a->w == b->w && a->beta == b->beta && a->nf == b->nf;
}

void destroy_interp_group(InterpGroup* g) {
  if (!g) return;
  if (g->coef) (void)hipFree(g->coef);
  if (g->fine) (void)hipFree(g->fine);
  if (g->packed) (void)hipFree(g->packed);
  delete g;
}

namespace {

template <class A>
const A* entries(const std::vector<unsigned char>& v) {
  return reinterpret_cast<const A*>(v.data());
}
template <class A>
void put_entries(std::vector<unsigned char>& v, const std::vector<A>& a) {
  v.resize(sizeof(A) * a.size());
  std::memcpy(v.data(), a.data(), v.size());
}

// launches for members [0, m), Pack<A>::K per launch; entry(i) gives member i's arguments
template <class A, class E, class L>
void packs(int m, E&& entry, L&& launch) {
  for (int lo = 0; lo < m; lo += Pack<A>::K) {
    Pack<A> pk{};
    const int c = std::min(Pack<A>::K, m - lo);
    for (int i = 0; i < c; ++i) pk.m[i] = entry(lo + i);
    launch(pk, c, lo);
  }
}

template <typename T>
void group_setup_t(InterpGroup* g, const std::vector<qlles_interp*>& its, int N, const std::vector<SpecRef>& cont,
                   const std::vector<SpecRef>& last, const std::vector<double*>& r, const std::vector<int>& np,
                   const std::vector<void*>& ur, const std::vector<double>& dt) {
  const int M = g->M, d = g->dim, nf = g->nf;
  const size_t per_f = d == 3 ? (size_t)nf * nf * nf : (size_t)nf * nf;
  QL_HIP(hipMalloc(&g->coef, sizeof(cx<T>) * per_f * d * M));
  QL_HIP(hipMemset(g->coef, 0, sizeof(cx<T>) * per_f * d * M));  // the padding stays zero
  // the fill runs on the null stream, unordered with the batch's non-blocking
  // streams that first use the grids: finish it before returning
  QL_HIP(hipDeviceSynchronize());
  QL_HIP(hipMalloc(&g->fine, sizeof(cx<T>) * per_f * d * M));
  if (d == 3) QL_HIP(hipMalloc(&g->packed, sizeof(T) * 4 * per_f * M));
  const double scale = 1.0 / (d == 3 ? (double)N * N * N : (double)N * N);
  std::vector<CoefArgs<T>> tc(M), tl(M);
  std::vector<GatherArgs<T>> tg(M);
  for (int i = 0; i < M; ++i) {
    cx<T>* out = static_cast<cx<T>*>(g->coef) + (size_t)i * d * per_f;
    const T* corr = static_cast<const T*>(its[i]->corr);
    tc[i] = CoefArgs<T>{static_cast<const cx<T>*>(cont[i].p), N, cont[i].ld, cont[i].q, scale, corr, out};
    tl[i] = CoefArgs<T>{static_cast<const cx<T>*>(last[i].p), N, last[i].ld, last[i].q, scale, corr, out};
    const T* f = d == 3 ? static_cast<const T*>(g->packed) + (size_t)i * 4 * per_f
                        : reinterpret_cast<const T*>(static_cast<const cx<T>*>(g->fine) + (size_t)i * d * per_f);
    tg[i] = GatherArgs<T>{f, r[i], np[i], static_cast<T*>(ur[i]), dt[i], dt[i] != 0.0 ? r[i] : nullptr};
  }
  put_entries(g->cont, tc);
  put_entries(g->last, tl);
  put_entries(g->gath, tg);
}

template <typename T>
void group_run_t(InterpGroup* g, int m, int f, hipStream_t s) {
  const int n = g->n, d = g->dim, nf = g->nf;
  const size_t per_n = d == 3 ? (size_t)n * n * n : (size_t)n * n;
  const int nb = (int)std::min<size_t>((per_n * d + 255) / 256, 8192);
  const auto* tc = entries<CoefArgs<T>>(g->cont);
  const auto* tl = entries<CoefArgs<T>>(g->last);
  const auto* tg = entries<GatherArgs<T>>(g->gath);
  packs<CoefArgs<T>>(m, [&](int i) -> const CoefArgs<T>& { return i < f ? tc[i] : tl[i]; },
                     [&](const Pack<CoefArgs<T>>& pk, int c, int) {
                       hipLaunchKernelGGL((k_build_coef<T>), dim3(nb, 1, c), dim3(256), 0, s, pk, d, n, 0, nf);
                       QL_LAUNCHED();
                     });
  if (g->max_np == 0) return;
  const cx<T>* twf = twiddles<T>(nf);
  cx<T>* co = static_cast<cx<T>*>(g->coef);
  cx<T>* fi = static_cast<cx<T>*>(g->fine);
  if (d == 3) {
    c2c_strided<T>(nf, co, fi, 3 * m, nf * nf, false, (T)1, twf, s);
    c2c_strided<T>(nf, fi, fi, 3 * m * nf, nf, false, (T)1, twf, s);
    c2c_contig<T>(nf, fi, fi, 3 * m * nf * nf, false, (T)1, twf, s);
    const size_t cs = (size_t)nf * nf * nf;
    hipLaunchKernelGGL((k_pack3<T>), dim3((int)std::min<size_t>((cs + 255) / 256, 4096), m), dim3(256), 0, s, fi,
                       static_cast<T*>(g->packed), cs);
    QL_LAUNCHED();
    packs<GatherArgs<T>>(m, [&](int i) -> const GatherArgs<T>& { return tg[i]; },
                         [&](const Pack<GatherArgs<T>>& pk, int c, int) {
                           launch_gather3p<T>(g->w, dim3((g->max_np + 3) / 4, 1, c), s, pk, nf, (T)g->beta);
                         });
  } else {
    c2c_strided<T>(nf, co, fi, 2 * m, nf, false, (T)1, twf, s);
    c2c_contig<T>(nf, fi, fi, 2 * m * nf, false, (T)1, twf, s);
    packs<GatherArgs<T>>(m, [&](int i) -> const GatherArgs<T>& { return tg[i]; },
                         [&](const Pack<GatherArgs<T>>& pk, int c, int) {
                           hipLaunchKernelGGL((k_gather<T, 2>), dim3((g->max_np + 3) / 4, 1, c), dim3(256), 0, s, pk,
                                              nf, g->w, (T)g->beta);
                           QL_LAUNCHED();
                         });
  }
}

}  // namespace

InterpGroup* make_interp_group(const std::vector<qlles_interp*>& its, int N, const std::vector<SpecRef>& cont,
                               const std::vector<SpecRef>& last, const std::vector<double*>& r,
                               const std::vector<int>& np, const std::vector<void*>& ur,
                               const std::vector<double>& dt) {
  const size_t M = its.size();
  if (M == 0 || cont.size() != M || last.size() != M || r.size() != M || np.size() != M || ur.size() != M ||
      dt.size() != M)
    throw std::invalid_argument("one spectrum, position and velocity array per group member");
  for (size_t i = 0; i < M; ++i) {
    if (!interp_groupable(its[0], its[i])) throw std::invalid_argument("interpolators of a group must match");
    if (np[i] < 0) throw std::invalid_argument("particle counts must be >= 0");
    for (const SpecRef* sp : {&cont[i], &last[i]})
      if (sp->q == 0 ? sp->ld < N / 2 + 1 : (sp->ld < N / 2 || its[0]->dim != 3))
        throw std::invalid_argument("spectrum row stride too small for its layout");
  }
  if (N < its[0]->n || (N % 2) != 0) throw std::invalid_argument("solver grid N must be even and >= n");
  auto* g = new InterpGroup();
  g->dim = its[0]->dim;
  g->n = its[0]->n;
  g->nf = its[0]->nf;
  g->w = its[0]->w;
  g->beta = its[0]->beta;
  g->dtype = its[0]->dtype;
  g->M = (int)M;
  for (int v : np) g->max_np = std::max(g->max_np, v);
  try {
    if (g->dtype == QLLES_F32) group_setup_t<float>(g, its, N, cont, last, r, np, ur, dt);
    else group_setup_t<double>(g, its, N, cont, last, r, np, ur, dt);
  } catch (...) {
    destroy_interp_group(g);
    throw;
  }
  return g;
}

void interp_group_run(InterpGroup* g, int m, int f, hipStream_t s) {
  if (m < 0 || m > g->M || f < 0 || f > m) throw std::invalid_argument("group member range out of bounds");
  if (m == 0) return;
  if (g->dtype == QLLES_F32) group_run_t<float>(g, m, f, s);
  else group_run_t<double>(g, m, f, s);
}

}  // namespace ql
