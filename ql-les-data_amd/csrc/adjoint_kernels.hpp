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

// FP32 occupancy target of k_adj_x_combine (waves per SIMD; 1 leaves it to the compiler)
#ifndef QL_ADJ_COMBINE_WPE32
#define QL_ADJ_COMBINE_WPE32 3
#endif

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

// k_pd (mode 1) and the inverse transform along x in one pass over
// [3][N][N*(N/2+1)]: each workgroup loads all three components of TI columns,
// forms P D l / c per mode, then transforms the components one after another
// through one LDS tile.  Same arithmetic as k_pd followed by k_c2c_strided.
template <typename T, int N>
__global__ void __launch_bounds__((StridedCfg<T, N>::THREADS))
k_adj_pd_x(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, int KZD, const cx<T>* __restrict__ tw) {
  using C = StridedCfg<T, N>;
  using F = PFFT<T, false, N, C::TP, typename Plan<N>::R>;
  constexpr int Nh = N / 2 + 1, inner = N * Nh;
  constexpr size_t cs = (size_t)N * inner;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[N * C::LS];
  const int tid = threadIdx.x, p = tid % TI, j = tid / TI;
  stage_table(s_tw, tw, N);
  __syncthreads();
  const int col = blockIdx.x * TI + p;
  const bool ok = col < inner;
  cx<T> v[3][F::MAXE];
#pragma unroll
  for (int s = 0; s < F::NIN; ++s) {
    cx<T> a = {0, 0}, b = {0, 0}, c = {0, 0};
    if (ok && F::in_valid(j, s)) {
      const size_t m = (size_t)F::in_index(j, s) * inner + col;
      a = in[m];
      b = in[cs + m];
      c = in[2 * cs + m];
      const Mode3 q(m, N, KZD);
      if (!q.keep) {
        a = b = c = {0, 0};
      } else if (q.k2 > 0) {
        const T inv = (T)1 / (T)q.k2;
        const cx<T> pr = {(a.re * q.kx + b.re * q.ky + c.re * q.kz) * inv,
                          (a.im * q.kx + b.im * q.ky + c.im * q.kz) * inv};
        a = csub(a, cscale(pr, (T)q.kx));
        b = csub(b, cscale(pr, (T)q.ky));
        c = csub(c, cscale(pr, (T)q.kz));
      }
      const T w = (q.kz == 0 || 2 * q.kz == N) ? (T)1 : (T)0.5;
      a = cscale(a, w);
      b = cscale(b, w);
      c = cscale(c, w);
    }
    v[0][s] = a;
    v[1][s] = b;
    v[2][s] = c;
  }
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    F::run(v[cc], s_buf + p, C::LS, j, s_tw, 1);  // each pass ends on a barrier: the tile is free again
#pragma unroll
    for (int s = 0; s < F::NOUT; ++s)
      if (ok && F::out_valid(j, s)) out[cc * cs + (size_t)F::out_index(j, s) * inner + col] = v[cc][s];
  }
}

// the forward transform along x of the six products [6][N][N*(N/2+1)] (a, then
// b) and k_vjp_combine in one pass: out = c (a + i k x b) - nu k^2 lam.  The
// b components are transformed first (i k x b needs all three per mode), then
// each a component is transformed and combined.  The transforms take the
// generic plan's radices in reverse order, so the last pass keeps every lane
// busy with the fewest outputs per lane (radix (16, 8) at N = 128: 8 outputs
// per lane instead of 16 on half the lanes), which keeps the three transformed
// b components in registers.
template <typename T> struct AdjCombineWpe { static constexpr int value = sizeof(T) == 4 ? QL_ADJ_COMBINE_WPE32 : 1; };
template <typename T, int N>
__global__ void __launch_bounds__((StridedCfg<T, N>::THREADS))
__attribute__((amdgpu_waves_per_eu(AdjCombineWpe<T>::value)))
k_adj_x_combine(const cx<T>* __restrict__ ab, const cx<T>* __restrict__ lam, cx<T>* __restrict__ out, int KZD, T nu,
                const cx<T>* __restrict__ tw) {
  using C = StridedCfg<T, N>;
  using F = PFFT<T, true, N, C::TP, Reverse<typename Plan<N>::R>>;
  constexpr int Nh = N / 2 + 1, inner = N * Nh;
  constexpr size_t cs = (size_t)N * inner;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[N * C::LS];
  const int tid = threadIdx.x, p = tid % TI, j = tid / TI;
  stage_table(s_tw, tw, N);
  __syncthreads();
  const int col = blockIdx.x * TI + p;
  const bool ok = col < inner;
  cx<T> kb[3][F::MAXE];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
#pragma unroll
    for (int s = 0; s < F::NIN; ++s)
      kb[cc][s] = (ok && F::in_valid(j, s)) ? ab[(3 + cc) * cs + (size_t)F::in_index(j, s) * inner + col]
                                            : cx<T>{0, 0};
    F::run(kb[cc], s_buf + p, C::LS, j, s_tw, 1);
  }
#pragma unroll
  for (int s = 0; s < F::NOUT; ++s) {  // i k x b per mode
    const Mode3 q((size_t)F::out_index(j, s) * inner + (ok ? col : 0), N, KZD);
    const cx<T> b0 = kb[0][s], b1 = kb[1][s], b2 = kb[2][s];
    kb[0][s] = cmuli(csub(cscale(b2, (T)q.ky), cscale(b1, (T)q.kz)));
    kb[1][s] = cmuli(csub(cscale(b0, (T)q.kz), cscale(b2, (T)q.kx)));
    kb[2][s] = cmuli(csub(cscale(b1, (T)q.kx), cscale(b0, (T)q.ky)));
  }
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    cx<T> v[F::MAXE];
#pragma unroll
    for (int s = 0; s < F::NIN; ++s)
      v[s] = (ok && F::in_valid(j, s)) ? ab[cc * cs + (size_t)F::in_index(j, s) * inner + col] : cx<T>{0, 0};
    F::run(v, s_buf + p, C::LS, j, s_tw, 1);
#pragma unroll
    for (int s = 0; s < F::NOUT; ++s) {
      if (ok && F::out_valid(j, s)) {
        const size_t m = (size_t)F::out_index(j, s) * inner + col;
        const Mode3 q(m, N, KZD);
        const T w = (q.kz == 0 || 2 * q.kz == N) ? (T)1 : (T)2;
        out[cc * cs + m] = csub(cscale(cadd(v[s], kb[cc][s]), w), cscale(lam[cc * cs + m], nu * (T)q.k2));
      }
    }
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

// The z-axis middle of the RHS adjoint, one pencil (x, y) per TP lanes:
//   psi   = c2r_z of the three rows of `in` (the y- and x-inverse-transformed
//           P D l / c, rows of N/2+1), without its 1/N^3 (folded into h)
//   out1  = r2c_z(w x psi), out2 = r2c_z(psi x u)     (six rows of N/2+1)
// with w, u the recorded physical omega and u of the state.  It replaces
// c2r_contig + k_cross2x2 + r2c_contig: psi and the products never go to
// HBM.  The inverse plan is the forward plan reversed, so psi's register
// order is the forward transform's input order and the products are formed
// in registers (each complex slot holds the even and odd z samples, hence
// the products are taken per part).
template <int N>
struct AdjZCfg {
  static constexpr int M = N / 2;
  using R = typename Plan<M>::R;
  static constexpr int TP = Plan<M>::TP;
  static constexpr int P = (256 / TP) > 64 ? 64 : (256 / TP);
  static constexpr int THREADS = P * TP;
  static constexpr int LSR = M + 1 + LPAD;
};
// LDS layout of the pencils: FP32 at N = 128 (8-lane radix-(8, 8) plan) in
// groups of four pencils with the exchange swizzled (fft_core.hpp LDS_QUAD8,
// no bank conflicts); otherwise one contiguous row of LSR entries per pencil
template <typename T, int N>
struct AdjZLay {
  using C = AdjZCfg<N>;
  static constexpr int LAY = (QL_ZPAIR32 && sizeof(T) == 4 && N == 128) ? LDS_QUAD8 : 0;
  static constexpr bool GROUPED = LAY == LDS_QUAD8;
  static constexpr int GS = (C::M / 8 + 1) * 32;  // elements of a group of four pencils
  static_assert(!GROUPED || (C::TP == 8 && C::P % 4 == 0), "grouped layout: four pencils of 8 lanes");
  static constexpr int BUF = GROUPED ? (C::P / 4) * GS : C::P * C::LSR;
  static __device__ __forceinline__ int pencil_off(int p) { return GROUPED ? (p >> 2) * GS + (p & 3) * 8 : p * C::LSR; }
};

template <typename T, int N>
__global__ void __launch_bounds__(AdjZCfg<N>::THREADS)
k_adj_z(const cx<T>* __restrict__ in, const T* __restrict__ w, const T* __restrict__ u, cx<T>* __restrict__ out,
        const cx<T>* __restrict__ tw, T h) {
  using C = AdjZCfg<N>;
  using L = AdjZLay<T, N>;
  constexpr int M = C::M, Nh = M + 1, TP = C::TP, LAY = L::LAY;
  using CI = PFFT<T, false, M, TP, Reverse<typename C::R>, LAY>;
  using CF = PFFT<T, true, M, TP, typename C::R, LAY>;
  static_assert(CI::NOUT == CF::NIN, "inverse output order must be the forward input order");
  constexpr int NO = CI::NOUT;
  constexpr size_t rows = (size_t)N * N;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[L::BUF];
  const int tid = threadIdx.x, p = tid / TP, j = tid % TP;
  stage_table(s_tw, tw, N);
  const size_t row = (size_t)blockIdx.x * C::P + p;  // rows is a multiple of P
  cx<T>* pb = s_buf + L::pencil_off(p);
  cx<T> psi[3][NO];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    __syncthreads();  // the previous transform's LDS scratch is free (and the table staged)
    const cx<T>* src = in + (c * rows + row) * Nh;
    for (int k = j; k <= M; k += TP) pb[lds_at<LAY>(k)] = src[k];
    __syncthreads();
    cx<T> v[CI::MAXE];
#pragma unroll
    for (int s = 0; s < CI::NIN; ++s) {
      cx<T> z = {0, 0};
      if (CI::in_valid(j, s)) {
        const int n = CI::in_index(j, s);
        cx<T> xn = pb[lds_at<LAY>(n)], xr = pb[lds_at<LAY>(M - n)];
        if (n == 0) {
          xn.im = 0;
          xr.im = 0;
        }
        xr = cconj(xr);
        z = cadd(cadd(xn, xr), cmuli(cmulc(csub(xn, xr), s_tw[n])));
      }
      v[s] = z;
    }
    __syncthreads();
    CI::run(v, pb, 1, j, s_tw, 2);
#pragma unroll
    for (int s = 0; s < NO; ++s) psi[c][s] = v[s];
  }
  // physical rows of omega / u as (even, odd) pairs in psi's register order
  auto ld = [&](const T* f, int c, int s) {
    return reinterpret_cast<const cx<T>*>(f + (c * rows + row) * N)[CI::out_index(j, s)];
  };
#pragma unroll
  for (int o = 0; o < 6; ++o) {
    const T* f = o < 3 ? w : u;
    const int a1 = (o % 3 + 1) % 3, a2 = (o % 3 + 2) % 3;  // (x, y, z) -> (y, z), (z, x), (x, y)
    cx<T> v[CF::MAXE];
#pragma unroll
    for (int s = 0; s < NO; ++s) {
      const cx<T> f1 = ld(f, a1, s), f2 = ld(f, a2, s);
      if (o < 3) {  // (w x psi)_o = w_a1 psi_a2 - w_a2 psi_a1
        v[s] = {f1.re * psi[a2][s].re - f2.re * psi[a1][s].re, f1.im * psi[a2][s].im - f2.im * psi[a1][s].im};
      } else {  // (psi x u)_o = psi_a1 u_a2 - psi_a2 u_a1
        v[s] = {psi[a1][s].re * f2.re - psi[a2][s].re * f1.re, psi[a1][s].im * f2.im - psi[a2][s].im * f1.im};
      }
    }
    __syncthreads();
    CF::run(v, pb, 1, j, s_tw, 2);
#pragma unroll
    for (int s = 0; s < CF::NOUT; ++s)
      if (CF::out_valid(j, s)) pb[lds_at<LAY>(CF::out_index(j, s))] = v[s];
    __syncthreads();
    cx<T>* dst = out + (o * rows + row) * Nh;
    for (int k = j; k <= M; k += TP) {
      const cx<T> zk = pb[lds_at<LAY>(k == M ? 0 : k)];
      const cx<T> zr = cconj(pb[lds_at<LAY>(k == 0 ? 0 : M - k)]);
      const cx<T> e = cscale(cadd(zk, zr), h);
      const cx<T> od = crot<T, true>(cscale(csub(zk, zr), h));  // (zk - zr) / (2i), times 2h
      const cx<T> wk = k == M ? cx<T>{-1, 0} : s_tw[k];
      dst[k] = cadd(e, cmul(wk, od));
    }
  }
}

// both products of the RHS adjoint in one pass: out1 = w x p, out2 = p x u
template <typename T>
__global__ void k_cross2x2(const T* __restrict__ w, const T* __restrict__ p, const T* __restrict__ u,
                           T* __restrict__ out1, T* __restrict__ out2, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const T w0 = w[i], w1 = w[n + i], w2 = w[2 * n + i];
    const T p0 = p[i], p1 = p[n + i], p2 = p[2 * n + i];
    const T u0 = u[i], u1 = u[n + i], u2 = u[2 * n + i];
    out1[i] = w1 * p2 - w2 * p1;
    out1[n + i] = w2 * p0 - w0 * p2;
    out1[2 * n + i] = w0 * p1 - w1 * p0;
    out2[i] = p1 * u2 - p2 * u1;
    out2[n + i] = p2 * u0 - p0 * u2;
    out2[2 * n + i] = p0 * u1 - p1 * u0;
  }
}

// Fixed-power forcing term F(y) = a(y) M y, a = P / (2 E_f(y) + eps),
// E_f = 0.5 sum_S w |y_k|^2 / N^6 over the shell S = {0 < k^2 <= kf2} (w = 1 on
// the kz = 0 plane, 2 elsewhere), as the fused x-pass evaluates it in FP64.
//   mode 0 (JVP, x = d):   out_S += a d - (2P / (2E+eps)^2) <g_E, d> y,   <g_E, d> = sum_S w Re(conj(y) d) / N^6
//   mode 1 (VJP, x = l):   out_S += a l - (2P / (2E+eps)^2) (sum_S Re(conj(y) l)) w y / N^6
// and, in mode 1 with dpow != nullptr, *dpow += (1 / (2E+eps)) sum_S Re(conj(l) y),
// the cotangent of the injection power P (dF/dP = M y / (2E+eps)).
// One workgroup: the shell holds at most 31 x 31 x 16 modes (kf <= 15) and
// the sums run in a fixed order.
template <typename T>
__global__ void __launch_bounds__(256) k_force_tangent(const cx<T>* __restrict__ y, const cx<T>* __restrict__ x,
                                                       cx<T>* __restrict__ out, int N, int kf2, double power,
                                                       int mode, double* dpow, int ldy) {
  __shared__ double s_red[3 * 256];
  const int Nh = N / 2 + 1;
  const size_t cs = (size_t)N * N * Nh;
  // y may have rows of ldy entries (ldy = N/2: the internal split layout, whose
  // kz < N/2 part holds every shell mode, kf < N/2)
  const size_t cy_s = (size_t)N * N * ldy;
  int kf = 0;
  while ((kf + 1) * (kf + 1) <= kf2) ++kf;
  const int nx = 2 * kf + 1, nz = kf + 1, cnt = nx * nx * nz;
  const double n6 = (double)N * N * N * (double)N * N * N;
  double acc[3] = {0.0, 0.0, 0.0};  // E (unnormalised, x0.5 later), weighted dot, plain dot
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int kx = q / (nx * nz) - kf, ky = (q / nz) % nx - kf, kz = q % nz;
    const int k2 = kx * kx + ky * ky + kz * kz;
    if (k2 == 0 || k2 > kf2) continue;
    const size_t r = (size_t)((kx + N) % N) * N + (ky + N) % N;
    const size_t m = r * Nh + kz, my = r * ldy + kz;
    const double w = kz == 0 ? 1.0 : 2.0;
    for (int c = 0; c < 3; ++c) {
      const cx<T> a = y[c * cy_s + my], b = x[c * cs + m];
      const double re = (double)a.re * b.re + (double)a.im * b.im;
      acc[0] += w * ((double)a.re * a.re + (double)a.im * a.im);
      acc[1] += w * re;
      acc[2] += re;
    }
  }
  block_sum<3>(acc, s_red);
  const double E = 0.5 * acc[0] / n6;
  const double den = 2.0 * E + 2.220446049250313e-16;
  const double a = power / den, b = 2.0 * power / (den * den);
  if (mode == 1 && dpow != nullptr && threadIdx.x == 0) *dpow += acc[2] / den;
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int kx = q / (nx * nz) - kf, ky = (q / nz) % nx - kf, kz = q % nz;
    const int k2 = kx * kx + ky * ky + kz * kz;
    if (k2 == 0 || k2 > kf2) continue;
    const size_t r = (size_t)((kx + N) % N) * N + (ky + N) % N;
    const size_t m = r * Nh + kz, my = r * ldy + kz;
    const double w = kz == 0 ? 1.0 : 2.0;
    const T cy = mode == 0 ? (T)(-b * acc[1] / n6) : (T)(-b * acc[2] * w / n6);
    for (int c = 0; c < 3; ++c)
      out[c * cs + m] = cadd(out[c * cs + m], cadd(cscale(x[c * cs + m], (T)a), cscale(y[c * cy_s + my], cy)));
  }
}

// out = alpha x + beta y (complex arrays of n entries; y may be nullptr).
// out may be x or y (in-place updates): each element is read before it is
// written, so none of the pointers is __restrict__.
template <typename T>
__global__ void k_axpby(const cx<T>* x, const cx<T>* y, cx<T>* out, T alpha, T beta, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    cx<T> v = cscale(x[i], alpha);
    if (y != nullptr) v = cadd(v, cscale(y[i], beta));
    out[i] = v;
  }
}

// two axpbys sharing y, one pass: out1 = a1 x1 + b1 y, out2 = a2 x2 + b2 y
// (an output may be its own x: each element is read before it is written)
template <typename T>
__global__ void k_axpby2(const cx<T>* x1, const cx<T>* x2, const cx<T>* __restrict__ y, cx<T>* out1, cx<T>* out2,
                         T a1, T b1, T a2, T b2, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const cx<T> yi = y[i], v1 = x1[i], v2 = x2[i];
    out1[i] = cadd(cscale(v1, a1), cscale(yi, b1));
    out2[i] = cadd(cscale(v2, a2), cscale(yi, b2));
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
