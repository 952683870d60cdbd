// Generic batched 1-D FFT passes built on the pencil engine (fft_core.hpp).
//
//   k_c2c_strided : [outer][N][inner] complex, transform along N.  A workgroup
//                   owns TI consecutive inner columns, so each row of the tile
//                   is one contiguous TI*sizeof(cx) segment in HBM.
//   k_c2c_contig  : [rows][N] complex, transform along the contiguous axis.
//   k_r2c_contig  : [rows][N] real -> [rows][ldo] complex (N/2+1 used), via an
//                   N/2-point complex FFT and the even/odd split.
//   k_c2r_contig  : [rows][ldi] complex (N/2+1 used) -> [rows][N] real; the
//                   imaginary parts of the DC and Nyquist entries are ignored,
//                   as numpy's irfft does.
// These serve the setup transforms (rfftn of the initial field), the physical
// field output, spectral filtering, the NUFFT fine grid and diagnostics; the
// time step itself runs the fused kernels in ns3d.hpp / ns2d.hpp.
#pragma once
#include "fft_core.hpp"

namespace qk {
using namespace qf;

constexpr int TI = 16;  // inner columns per tile of a strided pass
constexpr int LPAD = 1;  // LDS padding (elements) per tile row

template <typename T, int N>
struct StridedCfg {
  static constexpr int TP = Plan<N>::TP;
  static constexpr int THREADS = TI * TP;
  static constexpr int LS = TI + LPAD;
};

template <typename T, int N, bool FWD>
__global__ void __launch_bounds__((StridedCfg<T, N>::THREADS))
k_c2c_strided(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, int inner, T scale,
              const cx<T>* __restrict__ tw) {
  using C = StridedCfg<T, N>;
  using F = PFFT<T, FWD, N, C::TP, typename Plan<N>::R>;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[N * C::LS];
  const int tid = threadIdx.x, p = tid % TI, j = tid / TI;
  stage_table(s_tw, tw, N);
  __syncthreads();
  const int col = blockIdx.x * TI + p;
  const bool ok = col < inner;
  const size_t base = (size_t)blockIdx.y * N * inner + col;
  cx<T> v[F::MAXE];
#pragma unroll
  for (int s = 0; s < F::NIN; ++s) {
    v[s] = (ok && F::in_valid(j, s)) ? in[base + (size_t)F::in_index(j, s) * inner] : cx<T>{0, 0};
  }
  F::run(v, s_buf + p, C::LS, j, s_tw, 1);
#pragma unroll
  for (int s = 0; s < F::NOUT; ++s) {
    if (ok && F::out_valid(j, s)) out[base + (size_t)F::out_index(j, s) * inner] = cscale(v[s], scale);
  }
}

template <int N>
struct ContigCfg {
  static constexpr int TP = Plan<N>::TP;
  static constexpr int P = (256 / TP) > 64 ? 64 : (256 / TP);  // pencils per workgroup
  static constexpr int THREADS = P * TP;
};

template <typename T, int N, bool FWD>
__global__ void __launch_bounds__(ContigCfg<N>::THREADS)
k_c2c_contig(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, int rows, T scale,
             const cx<T>* __restrict__ tw) {
  using C = ContigCfg<N>;
  using F = PFFT<T, FWD, N, C::TP, typename Plan<N>::R>;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[C::P * (N + LPAD)];
  const int tid = threadIdx.x, p = tid / C::TP, j = tid % C::TP;
  stage_table(s_tw, tw, N);
  __syncthreads();
  const int row = blockIdx.x * C::P + p;
  const bool ok = row < rows;
  const size_t base = (size_t)row * N;
  cx<T> v[F::MAXE];
#pragma unroll
  for (int s = 0; s < F::NIN; ++s) v[s] = (ok && F::in_valid(j, s)) ? in[base + F::in_index(j, s)] : cx<T>{0, 0};
  F::run(v, s_buf + p * (N + LPAD), 1, j, s_tw, 1);
#pragma unroll
  for (int s = 0; s < F::NOUT; ++s)
    if (ok && F::out_valid(j, s)) out[base + F::out_index(j, s)] = cscale(v[s], scale);
}

// real <-> half-complex along the contiguous axis (N even, M = N/2)
template <int N>
struct RealCfg {
  static constexpr int M = N / 2;
  static constexpr int TP = Plan<M>::TP;
  static constexpr int P = (256 / TP) > 64 ? 64 : (256 / TP);
  static constexpr int THREADS = P * TP;
  static constexpr int LSR = M + 1 + LPAD;  // per-pencil LDS stride (holds N/2+1 entries)
};

// in: rows of N reals (row stride N); out: rows of N/2+1 complex (row stride ldo)
template <typename T, int N>
__global__ void __launch_bounds__(RealCfg<N>::THREADS)
k_r2c_contig(const T* __restrict__ in, cx<T>* __restrict__ out, int rows, int ldo, T scale,
             const cx<T>* __restrict__ tw) {
  using C = RealCfg<N>;
  constexpr int M = C::M;
  using F = PFFT<T, true, M, C::TP, typename Plan<M>::R>;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[C::P * C::LSR];
  const int tid = threadIdx.x, p = tid / C::TP, j = tid % C::TP;
  stage_table(s_tw, tw, N);
  __syncthreads();
  const int row = blockIdx.x * C::P + p;
  const bool ok = row < rows;
  const cx<T>* src = reinterpret_cast<const cx<T>*>(in + (size_t)row * N);
  cx<T> v[F::MAXE];
#pragma unroll
  for (int s = 0; s < F::NIN; ++s) v[s] = (ok && F::in_valid(j, s)) ? src[F::in_index(j, s)] : cx<T>{0, 0};
  cx<T>* pb = s_buf + p * C::LSR;
  F::run(v, pb, 1, j, s_tw, 2);
#pragma unroll
  for (int s = 0; s < F::NOUT; ++s)
    if (F::out_valid(j, s)) pb[F::out_index(j, s)] = v[s];
  __syncthreads();
  if (ok) {
    cx<T>* dst = out + (size_t)row * ldo;
    for (int k = j; k <= M; k += C::TP) {
      const cx<T> zk = pb[k == M ? 0 : k];
      const cx<T> zr = cconj(pb[k == 0 ? 0 : M - k]);
      const cx<T> e = cscale(cadd(zk, zr), (T)0.5);
      const cx<T> o = crot<T, true>(cscale(csub(zk, zr), (T)0.5));  // (zk - zr) / (2i)
      const cx<T> w = k == M ? cx<T>{-1, 0} : s_tw[k];
      dst[k] = cscale(cadd(e, cmul(w, o)), scale);
    }
  }
}

// in: rows of N/2+1 complex (row stride ldi); out: rows of N reals (row stride N)
template <typename T, int N>
__global__ void __launch_bounds__(RealCfg<N>::THREADS)
k_c2r_contig(const cx<T>* __restrict__ in, T* __restrict__ out, int rows, int ldi, T scale,
             const cx<T>* __restrict__ tw) {
  using C = RealCfg<N>;
  constexpr int M = C::M;
  using F = PFFT<T, false, M, C::TP, typename Plan<M>::R>;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[C::P * C::LSR];
  const int tid = threadIdx.x, p = tid / C::TP, j = tid % C::TP;
  stage_table(s_tw, tw, N);
  const int row = blockIdx.x * C::P + p;
  const bool ok = row < rows;
  cx<T>* pb = s_buf + p * C::LSR;
  if (ok) {
    const cx<T>* src = in + (size_t)row * ldi;
    for (int k = j; k <= M; k += C::TP) pb[k] = src[k];
  }
  __syncthreads();
  cx<T> v[F::MAXE];
#pragma unroll
  for (int s = 0; s < F::NIN; ++s) {
    const int n = F::in_index(j, s);
    cx<T> z = {0, 0};
    if (F::in_valid(j, s)) {
      cx<T> xn = pb[n];
      cx<T> xr = pb[M - n];
      if (n == 0) {
        xn.im = 0;
        xr.im = 0;
      }
      xr = cconj(xr);
      const cx<T> e = cadd(xn, xr);
      const cx<T> o = cmulc(csub(xn, xr), s_tw[n]);  // * W_N^{-n}
      z = cadd(e, cmuli(o));
    }
    v[s] = z;
  }
  __syncthreads();
  F::run(v, pb, 1, j, s_tw, 2);
  if (ok) {
    cx<T>* dst = reinterpret_cast<cx<T>*>(out + (size_t)row * N);
#pragma unroll
    for (int s = 0; s < F::NOUT; ++s)
      if (F::out_valid(j, s)) dst[F::out_index(j, s)] = cscale(v[s], scale);
  }
}

}  // namespace qk
