// Small kernels around the step: FP64 reductions (diagnostics, forcing
// normalisation), spectral windowing for the LES filter, curl formation.
// All reductions use a fixed grid and fixed tree order, so results are
// bitwise reproducible run to run.
#pragma once
#include "ns3d.hpp"

namespace qk {

constexpr int RED_BLOCKS = 512;
constexpr int RED_THREADS = 256;

static __global__ void k_fill_double(double* p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// partials[b*4 + 0] = sum u^2 over comps; [b*4 + 1 + c] = max |u_c|
template <typename T>
static __global__ void __launch_bounds__(RED_THREADS)
k_reduce_phys(const T* __restrict__ u, size_t npts, int ncomp, double* partials) {
  __shared__ double s_red[RED_THREADS];
  const size_t chunk = (npts + gridDim.x - 1) / gridDim.x;
  const size_t lo = (size_t)blockIdx.x * chunk;
  const size_t hi = lo + chunk < npts ? lo + chunk : npts;
  double ss = 0.0;
  double mx[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < ncomp; ++c) {
    const T* uc = u + (size_t)c * npts;
    double m = 0.0;
    for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      const double x = (double)uc[i];
      ss += x * x;
      m = fmax(m, fabs(x));
    }
    mx[c] = m;
  }
  double v[1] = {ss};
  block_sum<1>(v, s_red);
  for (int c = 0; c < 3; ++c) {
    s_red[threadIdx.x] = mx[c];
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) s_red[threadIdx.x] = fmax(s_red[threadIdx.x], s_red[threadIdx.x + h]);
      __syncthreads();
    }
    mx[c] = s_red[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 4 + 0] = v[0];
    for (int c = 0; c < 3; ++c) partials[blockIdx.x * 4 + 1 + c] = mx[c];
  }
}

// out[0] = 0.5 sum / npts; out[1] = max over comps; out[2..4] = per-comp max
static __global__ void __launch_bounds__(RED_THREADS)
k_reduce_phys_final(const double* partials, int nblk, double npts, double* out) {
  __shared__ double s_red[RED_THREADS];
  double v[1] = {0.0};
  double mx[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    v[0] += partials[b * 4];
    for (int c = 0; c < 3; ++c) mx[c] = fmax(mx[c], partials[b * 4 + 1 + c]);
  }
  block_sum<1>(v, s_red);
  for (int c = 0; c < 3; ++c) {
    s_red[threadIdx.x] = mx[c];
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) s_red[threadIdx.x] = fmax(s_red[threadIdx.x], s_red[threadIdx.x + h]);
      __syncthreads();
    }
    mx[c] = s_red[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = 0.5 * v[0] / npts;
    out[1] = fmax(mx[0], fmax(mx[1], mx[2]));
    for (int c = 0; c < 3; ++c) out[2 + c] = mx[c];
  }
}

// dst[0] = f0 * src[0]; dst[i] = src[i] for 0 < i < n
static __global__ void k_scale_copy(const double* src, double* dst, int n, double f0) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = i == 0 ? f0 * src[0] : src[i];
}

// sum of NV partial columns (partials[b*NV + q]) in fixed order
template <int NV>
__device__ void reduce_partials(const double* partials, int nblk, double (&v)[NV], double* s_red) {
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += partials[b * NV + q];
  }
  block_sum<NV>(v, s_red);
}

// ekin_tot forcing scalars (spectral_solver.py:93-109) from Parseval partials
static __global__ void __launch_bounds__(RED_THREADS)
k_ekin_alpha(const double* partials, int nblk, double inv_n2d, double e_init, double dt, double* alpha,
             double* e_inj) {
  __shared__ double s_red[2 * RED_THREADS];
  double v[2];
  reduce_partials<2>(partials, nblk, v, s_red);
  if (threadIdx.x == 0) {
    const double eps = 2.220446049250313e-16;
    const double e_new = 0.5 * v[0] * inv_n2d;
    const double e_low = 0.5 * v[1] * inv_n2d;
    const double e_high = e_new - e_low;
    *alpha = sqrt(fmax(0.0, (e_init - e_high) / (e_low + eps)));
    if (e_inj) *e_inj = fmax(0.0, (e_init - e_new) / dt);
  }
}

// Parseval partial sums of w|u^|^2 over all modes and over the shell
template <typename T>
static __global__ void __launch_bounds__(RED_THREADS)
k_energy_hat(const cx<T>* __restrict__ uh, int N, int dim, int ld, int kf2, double* partials) {
  __shared__ double s_red[2 * RED_THREADS];
  const int Nh = N / 2 + 1;
  const size_t nmodes = (dim == 3) ? (size_t)N * N * Nh : (size_t)N * Nh;
  const int ncomp = dim == 3 ? 3 : 2;
  const size_t chunk = (nmodes + gridDim.x - 1) / gridDim.x;
  const size_t lo = (size_t)blockIdx.x * chunk;
  const size_t hi = lo + chunk < nmodes ? lo + chunk : nmodes;
  double v[2] = {0.0, 0.0};
  for (size_t m = lo + threadIdx.x; m < hi; m += blockDim.x) {
    const int kz = (int)(m % Nh);
    const size_t r = m / Nh;  // row (x,y) or x
    int k2;
    if (dim == 3) {
      const int kx = kfreq((int)(r / N), N), ky = kfreq((int)(r % N), N);
      k2 = kx * kx + ky * ky + kz * kz;
    } else {
      const int kx = kfreq((int)r, N);
      k2 = kx * kx + kz * kz;
    }
    const double w = (kz == 0 || kz == N / 2) ? 1.0 : 2.0;
    double e = 0.0;
    const size_t cs = (dim == 3 ? (size_t)N * N : (size_t)N) * ld;
    for (int c = 0; c < ncomp; ++c) {
      const cx<T> a = uh[c * cs + r * ld + kz];
      e += (double)a.re * a.re + (double)a.im * a.im;
    }
    v[0] += w * e;
    if (k2 > 0 && k2 <= kf2) v[1] += w * e;
  }
  block_sum<2>(v, s_red);
  if (threadIdx.x == 0) {
    partials[blockIdx.x * 2] = v[0];
    partials[blockIdx.x * 2 + 1] = v[1];
  }
}

static __global__ void __launch_bounds__(RED_THREADS)
k_energy_hat_final(const double* partials, int nblk, double inv_n2d, double* out) {
  __shared__ double s_red[2 * RED_THREADS];
  double v[2];
  reduce_partials<2>(partials, nblk, v, s_red);
  if (threadIdx.x == 0) {
    out[0] = 0.5 * v[0] * inv_n2d;
    out[1] = 0.5 * v[1] * inv_n2d;
  }
}

// out = i k x u^ (utils.py:305-312 wavenumbers: kz Nyquist +N/2); 2-D: only
// omega_z is non-zero.  Layouts: in [3][rows][ld_in], out [3][rows][ld_out].
template <typename T>
__global__ void k_curl_hat(const cx<T>* __restrict__ uh, cx<T>* __restrict__ out, int N, int dim, int ld_in,
                           int ld_out) {
  const int Nh = N / 2 + 1;
  const size_t rows = dim == 3 ? (size_t)N * N : (size_t)N;
  const size_t total = rows * Nh;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const int kz = (int)(m % Nh);
    const size_t r = m / Nh;
    int kx, ky, kzz;
    if (dim == 3) {
      kx = kfreq((int)(r / N), N);
      ky = kfreq((int)(r % N), N);
      kzz = kz;
    } else {
      kx = kfreq((int)r, N);
      ky = kz;
      kzz = 0;
    }
    const cx<T> a = uh[r * ld_in + kz];
    const cx<T> b = uh[rows * ld_in + r * ld_in + kz];
    const cx<T> c = dim == 3 ? uh[2 * rows * ld_in + r * ld_in + kz] : cx<T>{0, 0};
    const cx<T> ox = cmuli(csub(cscale(c, (T)ky), cscale(b, (T)kzz)));
    const cx<T> oy = cmuli(csub(cscale(a, (T)kzz), cscale(c, (T)kx)));
    const cx<T> oz = cmuli(csub(cscale(b, (T)kx), cscale(a, (T)ky)));
    out[r * ld_out + kz] = ox;
    out[rows * ld_out + r * ld_out + kz] = oy;
    out[2 * rows * ld_out + r * ld_out + kz] = oz;
  }
}

// full spectrum F(k) of u = irfftn(u^): the Hermitian extension of the
// rfft half, with the rfft-axis DC and Nyquist planes symmetrized as the
// c2r of irfftn does.  k in integer wavenumber units (any representative).
// uq: the component's [N][N] Nyquist plane when kz = N/2 is stored apart
// from the rows (null: column N/2 of the rows)
template <typename T>
__device__ __forceinline__ cx<T> full_spectrum3(const cx<T>* u, int N, int ld, int kx, int ky, int kz,
                                               const cx<T>* uq = nullptr) {
  const int ix = ((kx % N) + N) % N, iy = ((ky % N) + N) % N, iz = ((kz % N) + N) % N;
  const int jx = ((-kx % N) + N) % N, jy = ((-ky % N) + N) % N;
  if (uq != nullptr && 2 * iz == N) {
    const cx<T> a = uq[(size_t)ix * N + iy];
    const cx<T> b = cconj(uq[(size_t)jx * N + jy]);
    return cscale(cadd(a, b), (T)0.5);
  }
  if (iz == 0 || 2 * iz == N) {
    const cx<T> a = u[((size_t)ix * N + iy) * ld + iz];
    const cx<T> b = cconj(u[((size_t)jx * N + jy) * ld + iz]);
    return cscale(cadd(a, b), (T)0.5);
  }
  if (2 * iz < N) return u[((size_t)ix * N + iy) * ld + iz];
  return cconj(u[((size_t)jx * N + jy) * ld + (N - iz)]);
}

template <typename T>
__device__ __forceinline__ cx<T> full_spectrum2(const cx<T>* u, int N, int ld, int kx, int ky) {
  const int ix = ((kx % N) + N) % N, iy = ((ky % N) + N) % N;
  const int jx = ((-kx % N) + N) % N;
  if (iy == 0 || 2 * iy == N) {
    const cx<T> a = u[(size_t)ix * ld + iy];
    const cx<T> b = cconj(u[(size_t)jx * ld + iy]);
    return cscale(cadd(a, b), (T)0.5);
  }
  if (2 * iy < N) return u[(size_t)ix * ld + iy];
  return cconj(u[(size_t)jx * ld + (N - iy)]);
}

// n-grid rfft spectrum of spectral_filtering(irfftn(u^), n) (utils.py:100-138):
// G(k) = (F(k) + F(k*)) / 2 for k in [-n/2, n/2)^dim, k* = k with every -n/2
// component replaced by +n/2 (the .real projection of the asymmetric crop).
// Output layout [ncomp][n(...)][ldo] with n/2+1 rfft entries per row.
template <typename T>
__global__ void k_window(const cx<T>* __restrict__ uh, int N, int dim, int ld, cx<T>* __restrict__ out, int n,
                         int ldo) {
  const int nh = n / 2 + 1;
  const size_t rows = dim == 3 ? (size_t)n * n : (size_t)n;
  const size_t total = rows * nh;
  const int ncomp = dim == 3 ? 3 : 2;
  const size_t cs_in = (dim == 3 ? (size_t)N * N : (size_t)N) * ld;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total * ncomp;
       m += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(m / total);
    const size_t mm = m % total;
    const int iz = (int)(mm % nh);
    const size_t r = mm / nh;
    const cx<T>* u = uh + c * cs_in;
    cx<T> g;
    if (dim == 3) {
      const int kx = kfreq((int)(r / n), n), ky = kfreq((int)(r % n), n);
      const int kz = (2 * iz == n) ? -n / 2 : iz;
      const int sx = kx == -n / 2 ? n / 2 : kx, sy = ky == -n / 2 ? n / 2 : ky, sz = kz == -n / 2 ? n / 2 : kz;
      const cx<T> f1 = full_spectrum3(u, N, ld, kx, ky, kz);
      const cx<T> f2 = full_spectrum3(u, N, ld, sx, sy, sz);
      g = cscale(cadd(f1, f2), (T)0.5);
    } else {
      const int kx = kfreq((int)r, n);
      const int ky = (2 * iz == n) ? -n / 2 : iz;
      const int sx = kx == -n / 2 ? n / 2 : kx, sy = ky == -n / 2 ? n / 2 : ky;
      const cx<T> f1 = full_spectrum2(u, N, ld, kx, ky);
      const cx<T> f2 = full_spectrum2(u, N, ld, sx, sy);
      g = cscale(cadd(f1, f2), (T)0.5);
    }
    out[(size_t)c * rows * ldo + r * ldo + iz] = g;
  }
}

// Shell-binned energy spectrum (utils.py:35-97) of an rfft-layout spectrum
// vh ([ncomp][rows][N/2+1], rows = N*N in 3-D, N in 2-D; the halved axis is
// the last one): per mode, e = fact * mul_fac * sum_c |vh_c|^2 / N^(2 dim)
// with fact = 2 inside the halved axis and 1 on its 0 / N/2 planes, binned
// by shell = int(|k| + 0.5) < nbins together with the sample count fact.
// Each block accumulates its modes in LDS (FP64 atomics: the order inside a
// block is not fixed) and writes its histogram; k_shell_final sums the
// blocks in a fixed order.
template <typename T>
__global__ void __launch_bounds__(256) k_shell_bin(const cx<T>* __restrict__ vh, int N, int dim, int ncomp,
                                                   double scale2, double mul_fac, int nbins,
                                                   double* __restrict__ partials) {
  extern __shared__ double s_h[];  // [2][nbins]
  for (int i = threadIdx.x; i < 2 * nbins; i += blockDim.x) s_h[i] = 0.0;
  __syncthreads();
  const int Nh = N / 2 + 1;
  const size_t rows = dim == 3 ? (size_t)N * N : (size_t)N;
  const size_t total = rows * Nh, cs = total;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const int kz = (int)(m % Nh);
    const size_t r = m / Nh;
    int k2;
    if (dim == 3) {
      const int kx = kfreq((int)(r / N), N), ky = kfreq((int)(r % N), N);
      k2 = kx * kx + ky * ky + kz * kz;
    } else {
      const int kx = kfreq((int)r, N);
      k2 = kx * kx + kz * kz;
    }
    const int shell = (int)(sqrt((double)k2) + 0.5);
    if (shell >= nbins) continue;
    const double fact = (kz > 0 && 2 * kz < N) ? 2.0 : 1.0;
    double e = 0.0;
    for (int c = 0; c < ncomp; ++c) {
      const cx<T> v = vh[c * cs + m];
      e += (double)v.re * v.re + (double)v.im * v.im;
    }
    atomicAdd(&s_h[shell], e * scale2 * fact * mul_fac);
    atomicAdd(&s_h[nbins + shell], fact);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 2 * nbins; i += blockDim.x) partials[(size_t)blockIdx.x * 2 * nbins + i] = s_h[i];
}

static __global__ void k_shell_final(const double* __restrict__ partials, int nblocks, int nbins, double* __restrict__ ek) {
  const double eps = 2.220446049250313e-16;
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nbins; b += gridDim.x * blockDim.x) {
    double e = 0.0, n = 0.0;
    for (int k = 0; k < nblocks; ++k) {
      e += partials[(size_t)k * 2 * nbins + b];
      n += partials[(size_t)k * 2 * nbins + nbins + b];
    }
    ek[b] = e * 4.0 * 3.14159265358979323846 * (double)b * (double)b / (n + eps);
  }
}

}  // namespace qk
