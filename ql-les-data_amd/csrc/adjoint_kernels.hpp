// Pointwise kernels of the tangent-linear (JVP) and adjoint (VJP) RHS.
//
// RHS (spectral_solver.py:50-60, explicit viscous term, no forcing):
//   R(y) = P D rfftn(u x w) - nu k^2 y,   u = irfftn(y), w = irfftn(K y),  K = i k x
// with D the 2/3 mask and P the projection I - k k^T / k^2.
//   JVP:  R'(y) d  = P D rfftn(du x w + u x dw) - nu k^2 d
//   VJP:  R'(y)^T l = c (rfftn(w x psi) + K rfftn(psi x u)) - nu k^2 l,
//         psi = irfftn(P D l / c)
// where c = 1 on the kz = 0 and kz = N/2 planes and 2 elsewhere: the adjoint
// of irfftn (whose kz sum weights the interior planes twice and takes the real
// part) is c / N^3 rfftn, and that of rfftn is N^3 irfftn(. / c), for the real
// inner products sum_x a.b and Re sum_k conj(a_k) b_k over the stored
// half-spectrum.  All spectra here are [3][N][N][N/2+1].
#pragma once
#include "ns3d.hpp"

namespace qk {

// the per-mode wavenumbers and masks used by the fused RHS kernels
struct Mode3 {
  int kx, ky, kz;  // kz in [0, N/2] (+N/2 at the rfft Nyquist), kx, ky from fftfreq
  bool keep;       // inside the 2/3 band on every axis
  int k2;
  __device__ __forceinline__ Mode3(size_t m, int N, int KZD) {
    const int Nh = N / 2 + 1;
    kz = (int)(m % Nh);
    const size_t r = m / Nh;
    const int ix = (int)(r / N), iy = (int)(r % N);
    kx = kfreq(ix, N);
    ky = kfreq(iy, N);
    keep = keep_k(ix, N, KZD) && keep_k(iy, N, KZD) && kz < KZD;
    k2 = kx * kx + ky * ky + kz * kz;
  }
};

// out = P D in (mode 0) or P D in / c (mode 1)
template <typename T>
__global__ void k_pd(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, int N, int KZD, int mode) {
  const int Nh = N / 2 + 1;
  const size_t total = (size_t)N * N * Nh, cs = total;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const Mode3 q(m, N, KZD);
    cx<T> a = in[m], b = in[cs + m], c = in[2 * cs + m];
    if (!q.keep) {
      a = b = c = {0, 0};
    } else if (q.k2 > 0) {
      const T inv = (T)1 / (T)q.k2;
      const cx<T> p = {(a.re * q.kx + b.re * q.ky + c.re * q.kz) * inv, (a.im * q.kx + b.im * q.ky + c.im * q.kz) * inv};
      a = csub(a, cscale(p, (T)q.kx));
      b = csub(b, cscale(p, (T)q.ky));
      c = csub(c, cscale(p, (T)q.kz));
    }
    if (mode == 1) {
      const T w = (q.kz == 0 || 2 * q.kz == N) ? (T)1 : (T)0.5;
      a = cscale(a, w);
      b = cscale(b, w);
      c = cscale(c, w);
    }
    out[m] = a;
    out[cs + m] = b;
    out[2 * cs + m] = c;
  }
}

// out = c (a + i k x b) - nu k^2 lam
template <typename T>
__global__ void k_vjp_combine(const cx<T>* __restrict__ a, const cx<T>* __restrict__ b, const cx<T>* __restrict__ lam,
                              cx<T>* __restrict__ out, int N, int KZD, T nu) {
  const int Nh = N / 2 + 1;
  const size_t total = (size_t)N * N * Nh, cs = total;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const Mode3 q(m, N, KZD);
    const T w = (q.kz == 0 || 2 * q.kz == N) ? (T)1 : (T)2;
    const cx<T> b0 = b[m], b1 = b[cs + m], b2 = b[2 * cs + m];
    const cx<T> kb[3] = {cmuli(csub(cscale(b2, (T)q.ky), cscale(b1, (T)q.kz))),
                         cmuli(csub(cscale(b0, (T)q.kz), cscale(b2, (T)q.kx))),
                         cmuli(csub(cscale(b1, (T)q.kx), cscale(b0, (T)q.ky)))};
    const T nk2 = nu * (T)q.k2;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      out[c * cs + m] = csub(cscale(cadd(a[c * cs + m], kb[c]), w), cscale(lam[c * cs + m], nk2));
  }
}

// out = P D g - nu k^2 d
template <typename T>
__global__ void k_jvp_combine(const cx<T>* __restrict__ g, const cx<T>* __restrict__ d, cx<T>* __restrict__ out, int N,
                              int KZD, T nu) {
  const int Nh = N / 2 + 1;
  const size_t total = (size_t)N * N * Nh, cs = total;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const Mode3 q(m, N, KZD);
    cx<T> a = g[m], b = g[cs + m], c = g[2 * cs + m];
    if (!q.keep) {
      a = b = c = {0, 0};
    } else if (q.k2 > 0) {
      const T inv = (T)1 / (T)q.k2;
      const cx<T> p = {(a.re * q.kx + b.re * q.ky + c.re * q.kz) * inv, (a.im * q.kx + b.im * q.ky + c.im * q.kz) * inv};
      a = csub(a, cscale(p, (T)q.kx));
      b = csub(b, cscale(p, (T)q.ky));
      c = csub(c, cscale(p, (T)q.kz));
    }
    const T nk2 = nu * (T)q.k2;
    out[m] = csub(a, cscale(d[m], nk2));
    out[cs + m] = csub(b, cscale(d[cs + m], nk2));
    out[2 * cs + m] = csub(c, cscale(d[2 * cs + m], nk2));
  }
}

// physical cross products: out = a x b (+ c x d when c != nullptr); component stride n
template <typename T>
__global__ void k_cross2(const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
                         const T* __restrict__ d, T* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const T a0 = a[i], a1 = a[n + i], a2 = a[2 * n + i];
    const T b0 = b[i], b1 = b[n + i], b2 = b[2 * n + i];
    T o0 = a1 * b2 - a2 * b1, o1 = a2 * b0 - a0 * b2, o2 = a0 * b1 - a1 * b0;
    if (c != nullptr) {
      const T c0 = c[i], c1 = c[n + i], c2 = c[2 * n + i];
      const T d0 = d[i], d1 = d[n + i], d2 = d[2 * n + i];
      o0 += c1 * d2 - c2 * d1;
      o1 += c2 * d0 - c0 * d2;
      o2 += c0 * d1 - c1 * d0;
    }
    out[i] = o0;
    out[n + i] = o1;
    out[2 * n + i] = o2;
  }
}

// out = alpha x + beta y (complex arrays of n entries; y may be nullptr)
template <typename T>
__global__ void k_axpby(const cx<T>* __restrict__ x, const cx<T>* __restrict__ y, cx<T>* __restrict__ out, T alpha,
                        T beta, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    cx<T> v = cscale(x[i], alpha);
    if (y != nullptr) v = cadd(v, cscale(y[i], beta));
    out[i] = v;
  }
}

// out = scale * w(kz) * in, w = c (mode 0) or 1/c (mode 1)
template <typename T>
__global__ void k_cweight(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, int N, size_t rows, int ncomp,
                          T scale, int mode) {
  const int Nh = N / 2 + 1;
  const size_t total = rows * Nh * ncomp;
  for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < total; m += (size_t)gridDim.x * blockDim.x) {
    const int kz = (int)(m % Nh);
    const bool edge = kz == 0 || 2 * kz == N;
    const T w = edge ? (T)1 : (mode == 0 ? (T)2 : (T)0.5);
    out[m] = cscale(in[m], scale * w);
  }
}

}  // namespace qk
