// Fused 3-D pseudo-spectral Navier-Stokes RK4 step (rotational form).
//
// Reference semantics: l3es/spectral_solver.py:27-120 (rhs_wrapper, rk4_wrapper).
// One RHS evaluation is four kernels; each reads and writes HBM once:
//
//   X  (x-pass)   forward x-FFT of the dealiased nonlinear term N^ (3 comps),
//                 dealias in kx, pressure projection, viscous term, forcing,
//                 RK4 stage update, then the inverse x-FFT of the new stage
//                 field and of the two curl components that need kx:
//                   X5 = Fx^-1 [u^, v^, w^, i(kz u^ - kx w^), i(kx v^ - ky u^)]
//   Yi (y-inv)    inverse y-FFT of the five fields plus i(ky w - kz v) -> 6
//   Z  (z-pass)   6 x c2r along z, cross product u x omega in registers,
//                 3 x r2c along z, keep kz < kmax (2/3 rule)
//   Yf (y-fwd)    forward y-FFT of the 3 components, keep |ky| < kmax
//
// Array layouts (complex, component-major, z/kz fastest):
//   state u^        [3][N][N][Nh]   (the caller's array, Nh = N/2+1)
//   stage/acc/X5/Y6 [f][N][N][LD]   (LD = Nh rounded up to 16: tile rows
//                                    start on 128-byte boundaries in FP32)
//   NZ              [3][N][N][KZP]  (KZP = KZD rounded up to 16)
//   NY              [3][N][KYD][KZP] (only the |ky| < kmax rows)
// where KZD = #{kz >= 0 : kz < kmax}, KYD = 2*KZD - 1 and
// kmax = 2/3 (N/2 + 1) exactly as spectral_solver.py:42 computes it.
#pragma once
#include "fft_kernels.hpp"

// Occupancy targets (waves per SIMD) for the fused kernels; 0 leaves the
// choice to the compiler.  These kernels are HBM-bound, so more resident
// waves hide more memory latency than extra registers buy.
#ifndef QL_XPLAN_WIDE
#define QL_XPLAN_WIDE 1
#endif
#ifndef QL_ZPLAN_WIDE
#define QL_ZPLAN_WIDE 1
#endif
#ifndef QL_WPE_X
#define QL_WPE_X 4
#endif
// z-pass: threads per workgroup, LDS padding (one slot per 2^QL_ZPAD
// entries), fields in flight in FP32
#ifndef QL_ZTHREADS
#define QL_ZTHREADS 256
#endif
#ifndef QL_ZPAD
#define QL_ZPAD 3
#endif
// the same for the FP64 3-D z-pass (16-byte elements, so another bank pattern)
#ifndef QL_ZPAD64
#define QL_ZPAD64 3
#endif
#ifndef QL_ZDEPTH32
#define QL_ZDEPTH32 2
#endif
// z-pass twiddles formed from one table read (bit 0: FFT passes, bit 1: c2r /
// r2c pre/post).  FP32 256^3 step: 455 -> 471 steps/s with both (206 -> 168
// VGPRs).  FP64: the FFT passes' half, with one field in flight and 3 waves
// per SIMD (below)
#ifndef QL_ZTWR32
#define QL_ZTWR32 3
#endif
#ifndef QL_ZTWR64
#define QL_ZTWR64 1
#endif
// FP64 z-pass at N >= 256: one field in flight and 3 waves per SIMD with the
// FFT passes' twiddles formed (168 VGPRs; no spill with the radix-(4, 4, 4, 2)
// plan at N = 256, below).  Three fields in flight at 2 waves per SIMD (206
// VGPRs) took 349.7 us per launch at 256^3; this setting with the radix-(4,
// 4, 8) plan 294.7 us (60 B per lane spilled), with the (4, 4, 4, 2) plan
// 265-266 us (two fields in flight the same)
#ifndef QL_ZDEPTH64
#define QL_ZDEPTH64 1
#endif
#ifndef QL_Z64_WPE
#define QL_Z64_WPE 3
#endif
#ifndef QL_WPE_YI
#define QL_WPE_YI 0
#endif
#ifndef QL_WPE_Z
#define QL_WPE_Z 0
#endif
#ifndef QL_WPE_YF
#define QL_WPE_YF 0
#endif
// FP64 y-forward at N >= 256: 4 waves per SIMD (128 VGPRs, 40 B per lane
// spilled) against 3 (140 VGPRs): 99.3 -> 94.0 us per launch at 256^3.  The
// y-inverse at 3 waves (168 VGPRs, 148 B spilled) took 490 against 304 us
#ifndef QL_YF64_WPE
#define QL_YF64_WPE 4
#endif
#define QL_WPE_ATTR_(n) __attribute__((amdgpu_waves_per_eu(n)))
// the x-pass target applies to FP32 at N >= 256, where it fits without
// private-memory spills; FP64 (twice the registers per element) and the
// smaller grids are left to the compiler
template <typename T, int N> struct WpeX { static constexpr int value = (sizeof(T) == 4 && N >= 256) ? QL_WPE_X : 0; };
#if QL_WPE_X > 0
#define QL_ATTR_X __attribute__((amdgpu_waves_per_eu(WpeX<T, N>::value > 0 ? WpeX<T, N>::value : 1)))
#else
#define QL_ATTR_X
#endif
#if QL_WPE_YI > 0
#define QL_ATTR_YI QL_WPE_ATTR_(QL_WPE_YI)
#else
#define QL_ATTR_YI
#endif
#if QL_WPE_Z > 0
#define QL_ATTR_Z QL_WPE_ATTR_(QL_WPE_Z)
#elif QL_Z64_WPE > 0
template <typename T, int N> struct ZWpe {
  static constexpr bool on = sizeof(T) == 8 && N >= 256;
  static constexpr int lo = on ? QL_Z64_WPE : 1, hi = on ? QL_Z64_WPE : 8;
};
#define QL_ATTR_Z __attribute__((amdgpu_waves_per_eu(ZWpe<T, N>::lo, ZWpe<T, N>::hi)))
#else
#define QL_ATTR_Z
#endif
#if QL_WPE_YF > 0
#define QL_ATTR_YF QL_WPE_ATTR_(QL_WPE_YF)
#elif QL_YF64_WPE > 0
template <typename T, int N> struct YfWpe {
  static constexpr bool on = sizeof(T) == 8 && N >= 256;
  static constexpr int lo = on ? QL_YF64_WPE : 1, hi = on ? QL_YF64_WPE : 8;
};
#define QL_ATTR_YF __attribute__((amdgpu_waves_per_eu(YfWpe<T, N>::lo, YfWpe<T, N>::hi)))
#else
#define QL_ATTR_YF
#endif

namespace qk {

// Arguments of up to K launches of one kernel, passed by value: z-slice m of
// the grid runs with entry m.  A launch for one plan fills entry 0; a batch
// group fills one entry per member (several launches when it has more than K
// members).  Plans stepped alone and groups therefore run the very same
// kernels, so a member's results are bit-identical to its solo run (two
// kernels compiled from one body need not round alike: floating-point
// contraction depends on the surrounding code).  K keeps the kernel
// arguments near 1.3 KB.  Kernels copy their entry into a local at entry, so
// its fields load together, once.
template <class A>
struct Pack {
  static constexpr int K0 = 1280 / (int)sizeof(A);
  static constexpr int K = K0 < 1 ? 1 : (K0 > 16 ? 16 : K0);
  A m[K];
};

// RK / forcing controls shared by the 3-D and 2-D x-pass kernels
template <typename T>
struct StageCtl {
  int stage;      // -1: load state and inverse-transform only; 0..3: RK stage; 4: write du (rhs only)
  int visc_if;    // 0: explicit viscous term (reference); 1: integrating factor
  T nu, dt;
  T ca, cb;       // explicit RK: a_s*dt, b_s*dt
  // fixed-power forcing on 0 < |k|^2 <= kf2: du += P/(2 E_f + eps) * u_s
  int force_power;
  int kf2;
  double power;
  // Kolmogorov forcing (2-D): du_x(kx=0, ky=kolm_k) += -i amp N^2/2 ; drag: du -= drag*u_s
  int kolm_k;
  T kolm_re, kolm_im;
  T drag;
  // ekin_tot: stage 3 writes FP64 partial sums (E, E_low) per block
  double* partials;
  // stage -1 with rescale: masked modes multiplied by *alpha and written back
  const double* alpha;
};

template <typename T>
struct NS3X {
  const cx<T>* ny;   // [3][N][KYD][KZP]
  const cx<T>* us;   // stage field u_s (row stride ld_us)
  const cx<T>* un;   // u_n (row stride ld_un)
  cx<T>* uout;       // u_{s+1} / u_{n+1} / du (row stride ld_uout)
  cx<T>* acc;        // RK accumulator (row stride ld)
  cx<T>* x5;         // [5][N][N][ld]
  cx<T>* rescale_dst;  // stage -1 + alpha: where masked modes are written back (layout of un)
  cx<T>* copy_dst;     // stage -1: optional copy of the loaded state
  // row strides and Nyquist-plane offsets (see ColAt); acc and x5 share ld
  int ld_us, ld_un, ld_uout, ld, ld_copy;
  size_t q_us, q_un, q_uout, q_acc, q_x5, q_copy;
  int KZD, KZP, KYD;
  int do_inverse;
  const cx<T>* tw;
  StageCtl<T> c;
};

// Radix plan of the strided x / y passes.  With QL_XPLAN_WIDE each pencil is
// spread over more lanes (fewer elements, hence fewer registers, per thread)
// at the price of one more LDS round trip per transform.
template <int N> struct XPlan { using R = typename Plan<N>::R; static constexpr int TP = Plan<N>::TP; };
// 1: radix (4, 4, 4, 4) on 64 lanes per column at N = 256 (1024-thread
// workgroups, 4 elements per lane).  Slower at 256^3: FP32 471.8 -> 418.8
// steps/s (x-pass 218 -> 258 us); FP64 229.7 -> 217.0 (x-pass 424 -> 492 us,
// spilling at its 128-VGPR cap; y-inverse 302 -> 290 us), so off
#ifndef QL_XPLAN256_TP64
#define QL_XPLAN256_TP64 0
#endif
#if QL_XPLAN_WIDE
#if QL_XPLAN256_TP64
template <> struct XPlan<256> { using R = Radices<4, 4, 4, 4>; static constexpr int TP = 64; };
#else
template <> struct XPlan<256> { using R = Radices<8, 4, 8>; static constexpr int TP = 32; };
#endif
template <> struct XPlan<128> { using R = Radices<4, 4, 8>; static constexpr int TP = 16; };
#endif

#ifndef QL_TIX
#define QL_TIX 16
#endif
// Non-temporal stores per output stream (st_stream): x-pass stage fields and
// X5, y-inverse Y6, z-pass NZ, y-forward NY.  They apply only where one field
// component (N^3 reals) reaches 64 MiB: smaller fields stay in the 256 MB
// last-level cache between kernels.  A cached store leaves dirty lines that
// are written back while the next kernel runs; at 256^3 FP32 the choice below
// (all streams non-temporal except X5, whose reader is the very next kernel)
// took one RK stage from 623 to 602 us.
#ifndef QL_NT_XS
#define QL_NT_XS 1
#endif
// The last step of a call writes u_{n+1} into the caller's spectrum, whose
// rows hold N/2+1 complex: a 16-column tile's 128-byte run then straddles two
// cache lines, and the neighbouring tiles that write the rest of those lines
// run on other XCDs.  Non-temporal stores send each half line to HBM on its
// own; cached stores (0) let the last-level cache merge them.
#ifndef QL_NT_CALLER
#define QL_NT_CALLER 0
#endif
#ifndef QL_NT_X5
#define QL_NT_X5 0
#endif
#ifndef QL_NT_YI
#define QL_NT_YI 1
#endif
#ifndef QL_NT_Z
#define QL_NT_Z 1
#endif
#ifndef QL_NT_YF
#define QL_NT_YF 1
#endif
// non-temporal loads (ld_stream) of the large input streams, each read once:
// NY, u_s, u_n and the accumulator in the x-pass, X5 in the y-inverse, Y6 in
// the z-pass, NZ in the y-forward (at 256^3 FP32 one RK stage 602 -> 578 us)
#ifndef QL_NTL
#define QL_NTL 1
#endif
// per stream: NY in the x-pass, NZ in the y-forward (each pairs with its
// producer's QL_NT_YF / QL_NT_Z store: a field stored and read with cached
// accesses can stay in the last-level cache between the two kernels)
#ifndef QL_NTL_NY
#define QL_NTL_NY 1
#endif
#ifndef QL_NTL_NZ
#define QL_NTL_NZ 1
#endif
// smallest field component (N^3 reals, MiB) that takes the streaming stores and
// loads; below it the last-level cache holds enough of the data between
// kernels (with no gate: 24 batched trajectories of 64^3-128^3 9440 -> 8840
// trajectory-steps/s, the 128^3 adjoint forward 20.0 -> 22.7 ms)
#ifndef QL_NT_MIN_MIB
#define QL_NT_MIN_MIB 64
#endif
template <typename T, int N>
struct NTBig {
  static constexpr bool value = (size_t)N * N * N * sizeof(T) >= ((size_t)QL_NT_MIN_MIB << 20);
};

// V selects the tile width: 0 = QL_TIX columns, 1 = half of it (more
// workgroups for small grids; chosen per plan by the autotuner)
template <int N, int V = 0>
struct XCfg {
  using R = typename XPlan<N>::R;
  static constexpr int TP = XPlan<N>::TP;
  static constexpr int TIX = V == 0 ? QL_TIX : QL_TIX / 2;  // columns (contiguous kz / ky) per workgroup tile
  static constexpr int THREADS = TIX * TP;
  static constexpr int LS = TIX + LPAD;
};
// x / y-pass twiddles formed from one table read per butterfly: no gain at
// 256^3 (FP32 470.1-471.0 against 470.8-471.1 steps/s; FP64 211.5-211.9
// against 212.3-212.5), so off
#ifndef QL_XTWR32
#define QL_XTWR32 0
#endif
#ifndef QL_XTWR64
#define QL_XTWR64 0
#endif
template <typename T>
struct XTwr {
  static constexpr bool value = sizeof(T) == 4 ? QL_XTWR32 : QL_XTWR64;
};
template <typename T, int N>
using XFwd = PFFT<T, true, N, XCfg<N>::TP, typename XCfg<N>::R, 0, SYNC_BLOCK, XTwr<T>::value>;
template <typename T, int N>
using XInv = PFFT<T, false, N, XCfg<N>::TP, Reverse<typename XCfg<N>::R>, 0, SYNC_BLOCK, XTwr<T>::value>;

__device__ __forceinline__ bool keep_k(int idx, int N, int KZD) { return idx < KZD || idx > N - KZD; }

// Column kz of an [nf][N][N] (3-D) rfft-layout spectrum: element (c, row) is
// at off + c*cs + row*rs.  Rows hold ld complex.  The kz = N/2 column is
// either column N/2 of those rows (q == 0: the caller's layout, ld >= N/2+1)
// or a separate [nf][N][N] plane starting at element q (the internal layout,
// ld = N/2: rows start on cache-line boundaries and the Nyquist column is
// contiguous across rows instead of one element in a line of its own).
struct ColAt {
  size_t off, rs, cs;
  template <int N>
  __device__ __forceinline__ static ColAt make(int ld, size_t q, int kz) {
    if (q != 0 && kz == N / 2) return {q, 1, (size_t)N * N};
    return {(size_t)kz, (size_t)ld, (size_t)N * N * ld};
  }
  __device__ __forceinline__ size_t at(int c, size_t row) const { return off + c * cs + row * rs; }
};

// deterministic block sum of NV doubles per thread (fixed tree order)
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* s_red) {
  const int tid = threadIdx.x, nt = blockDim.x;
#pragma unroll
  for (int q = 0; q < NV; ++q) s_red[q * nt + tid] = v[q];
  __syncthreads();
  for (int h = nt / 2; h > 0; h >>= 1) {
    if (tid < h) {
#pragma unroll
      for (int q = 0; q < NV; ++q) s_red[q * nt + tid] += s_red[q * nt + tid + h];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = s_red[q * nt];
  __syncthreads();
}

// E_f = 0.5 * sum_{0<|k|^2<=kf2} w |u^|^2 (unnormalized; the caller divides by N^6)
template <typename T, int N>
__device__ double forced_shell_energy3(const cx<T>* us, int ld, int kf2, double* s_red) {
  constexpr int Nh = N / 2 + 1;
  int kf = 0;
  while ((kf + 1) * (kf + 1) <= kf2) ++kf;
  const int nx = 2 * kf + 1, nz = kf + 1;
  const int cnt = nx * nx * nz;
  double acc[1] = {0.0};
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int kx = q / (nx * nz) - kf, ky = (q / nz) % nx - kf, kz = q % nz;
    const int k2 = kx * kx + ky * ky + kz * kz;
    if (k2 > 0 && k2 <= kf2) {
      const int ix = (kx + N) % N, iy = (ky + N) % N;
      const double w = (kz == 0 || kz == N / 2) ? 1.0 : 2.0;
      double e = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const cx<T> a = us[((size_t)(c * N + ix) * N + iy) * ld + kz];
        e += (double)a.re * a.re + (double)a.im * a.im;
      }
      acc[0] += w * e;
    }
  }
  (void)Nh;
  block_sum<1>(acc, s_red);
  return 0.5 * acc[0];
}

// exp in the working precision (integrating-factor weights)
__device__ __forceinline__ float tpexp(float x) { return expf(x); }
__device__ __forceinline__ double tpexp(double x) { return exp(x); }

// which RK4 stages read / write the accumulator array (explicit viscous term:
// stage 1 writes, stage 3 reads; integrating factor: stages 0-2 write, 1-3 read)
template <int SG, bool VIF> struct AccIn {
  static constexpr bool value = VIF ? (SG >= 1 && SG <= 3) : SG == 3;
};
template <int SG, bool VIF> struct AccOut {
  static constexpr bool value = VIF ? (SG >= 0 && SG <= 2) : SG == 1;
};

// SG: -1 = load the state and inverse-transform only, 0..3 = RK4 stage,
// 4 = write du (rhs); VIF: integrating-factor viscous term.  Both are
// compile-time so the epilogue is straight-line code: all of a slot's loads
// are issued unconditionally (clamped addresses, values selected) and only
// the stores are predicated, which lets loads of later slots overlap.
template <typename T, int N, int SG, bool VIF, int V = 0>
__global__ void __launch_bounds__((XCfg<N, V>::THREADS)) QL_ATTR_X
k_ns3_x(Pack<NS3X<T>> pk) {
  // the entry is copied once, so its fields load together at the top (read
  // in place at a run-time index, they were reloaded near each use: 109
  // scalar loads instead of 19 in the 256^3 stage-3 kernel); the body is
  // included, not called, so that it compiles as the kernel's own code (as a
  // called function it spilled 60-80 more registers per lane)
  const NS3X<T> a = pk.m[blockIdx.z];
#include "ns3x_body.inc"
}

// ---------------------------------------------------------------------------
// inverse y-pass: X5 -> Y6 (or the first 3 fields only)
// ---------------------------------------------------------------------------
template <typename T>
struct NS3Y {
  const cx<T>* in;  // [5][N][N] rows (ld_in, Nyquist plane at q_in)
  cx<T>* out;       // [6][N][N] rows (ld_out, Nyquist plane at q_out)
  int ld_in, ld_out;
  size_t q_in, q_out;
  int nfields;      // 6 (full RHS) or 3 (velocity only)
  const cx<T>* tw;
};

// LDS of the y passes, declared by each kernel that runs the shared body (an
// LDS variable declared in the body would be one object used by two kernels)
#define QL_YI_LDS                                                              \
  __shared__ cx<T> s_tw[XInv<T, N>::TWL > 0 ? XInv<T, N>::TWL : 1];          \
  __shared__ cx<T> s_buf[N * XCfg<N, V>::LS]

template <typename T, int N, int V = 0>
__device__ __forceinline__ void ns3_yinv_run(const NS3Y<T>& a, cx<T>* s_tw, cx<T>* s_buf) {
  using C = XCfg<N, V>;
  constexpr bool NTL = QL_NTL && NTBig<T, N>::value;
  constexpr int Nh = N / 2 + 1;
  using F = XInv<T, N>;
  const int tid = threadIdx.x, p = tid % C::TIX, j = tid / C::TIX;
  // tiles 0..NTR-1: TIX consecutive kz < N/2 of one x row; tile NTR: the
  // kz = N/2 column of TIX consecutive x rows
  constexpr int NTR = (N / 2) / C::TIX;
  const bool nyq = NTR > 0 && (int)blockIdx.x == NTR;
  if (nyq && (int)blockIdx.y >= N / C::TIX) return;  // whole workgroup, before any barrier
  const int kz = nyq ? N / 2 : (int)blockIdx.x * C::TIX + p;
  const bool ok = kz < Nh;
  const int xi = nyq ? (int)blockIdx.y * C::TIX + p : (int)blockIdx.y;
  const int kzv = kz;
  F::fill_tw(s_tw, a.tw, 1);
  __syncthreads();
  const ColAt li = ColAt::make<N>(a.ld_in, a.q_in, kz), lo = ColAt::make<N>(a.ld_out, a.q_out, kz);
  const size_t row0 = (size_t)xi * N;
  auto load = [&](int f, cx<T>* v) {
#pragma unroll
    for (int s = 0; s < F::NIN; ++s)
      v[s] = (ok && F::in_valid(j, s)) ? ld_stream<NTL>(&a.in[li.at(f, row0 + F::in_index(j, s))])
                                       : cx<T>{0, 0};
  };
  auto store = [&](int f, const cx<T>* v) {
    if (!ok) return;
#pragma unroll
    for (int s = 0; s < F::NOUT; ++s)
      if (F::out_valid(j, s))
        st_stream<QL_NT_YI && NTBig<T, N>::value>(&a.out[lo.at(f, row0 + F::out_index(j, s))], v[s]);
  };
  // loads run ahead of the transforms: each field is requested one or two
  // transforms before it is consumed (software pipelining)
  cx<T> v[F::MAXE];
  auto take = [&](const cx<T>* src) {
#pragma unroll
    for (int s = 0; s < F::NIN; ++s) v[s] = src[s];
  };
  if (a.nfields == 6) {
    cx<T> A[F::NIN], B[F::NIN], U[F::NIN], D[F::NIN], E[F::NIN];
    load(1, A);
    load(2, B);
    load(0, U);
#pragma unroll
    for (int s = 0; s < F::NIN; ++s) {  // omega_x^ = i (ky w^ - kz v^)
      const int ky = kfreq(F::in_index(j, s), N);
      v[s] = cmuli(csub(cscale(B[s], (T)ky), cscale(A[s], (T)kzv)));
    }
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(3, v);
    load(3, D);
    take(A);
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(1, v);
    load(4, E);
    take(B);
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(2, v);
    take(U);
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(0, v);
    take(D);
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(4, v);
    take(E);
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(5, v);
  } else {
    cx<T> A[F::NIN], B[F::NIN], U[F::NIN];
    load(0, U);
    load(1, A);
    load(2, B);
    take(U);
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(0, v);
    take(A);
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(1, v);
    take(B);
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    store(2, v);
  }
}

template <typename T, int N, int V = 0>
__global__ void __launch_bounds__((XCfg<N, V>::THREADS)) QL_ATTR_YI
k_ns3_yinv(Pack<NS3Y<T>> pk) {
  QL_YI_LDS;
  const NS3Y<T> a = pk.m[blockIdx.z];
  ns3_yinv_run<T, N, V>(a, s_tw, s_buf);
}

// ---------------------------------------------------------------------------
// z-pass: Y6 -> (6 x c2r) -> u x omega -> (3 x r2c, kz < kmax) -> NZ
// ---------------------------------------------------------------------------
template <typename T>
struct NS3Z {
  const cx<T>* in;   // [6][N][N][N/2] rows + [6][N][N] Nyquist plane at element q
  cx<T>* nz;         // [3][N][N][KZP]
  T* u_out;          // optional physical velocity [3][N][N][N]
  const T* u_in;     // optional physical velocity replacing the c2r of fields 0..2
  size_t q;
  int KZD, KZP;
  const cx<T>* tw;
  T* w_out = nullptr;  // optional physical vorticity [3][N][N][N] (fields 3..5 after the c2r)
};

// Radix plan (forward order) of the half-length FFTs in the z-pass; the c2r
// runs it reversed so its last pass has exactly TP butterflies.
template <int M> struct ZPlan { using R = typename Plan<M>::R; };
#if QL_ZPLAN_WIDE == 1
template <> struct ZPlan<128> { using R = Radices<4, 4, 8>; };
template <> struct ZPlan<64> { using R = Radices<4, 4, 4>; };
#elif QL_ZPLAN_WIDE == 2
template <> struct ZPlan<128> { using R = Radices<4, 4, 4, 2>; };
template <> struct ZPlan<64> { using R = Radices<4, 4, 4>; };
#endif

// V selects the radix plan of the half-length FFTs: 0 = ZPlan, 1 = the
// generic Plan (fewer passes, more elements per lane)
template <int M, int V> struct ZPlanV { using R = typename ZPlan<M>::R; };
template <int M> struct ZPlanV<M, 1> { using R = typename Plan<M>::R; };

// The 3-D z-pass in FP32 at N = 256 (M = 128) runs radices 8, 4, 4 forward
// (4, 4, 8 inverse) on 16 lanes per pencil: every pass keeps all lanes busy,
// where the 32-lane plan leaves half of them idle in its radix-8 pass.  That
// pass is bound by vector-ALU issue, so the busier lanes pay for the lower
// occupancy (2 waves per SIMD at 206 registers against 4): 146.6 -> 137.8 us
// per launch.  FP64 (256 registers, 1 wave per SIMD) keeps the 32-lane plan.
// FP64 (SEL 2) at N = 256: radices 4, 4, 4, 2 on 32 lanes.  Its four-element
// register tile (against eight for 4, 4, 8) keeps the 3-wave build free of
// spills, whose scratch reloads had waited on the in-flight row prefetch:
// 292.9 -> 265.3-266.2 us per launch.
template <int M, int V, int SEL> struct ZPlanS { using R = typename ZPlanV<M, V>::R; };
template <> struct ZPlanS<128, 0, 1> { using R = Radices<8, 4, 4>; };
template <> struct ZPlanS<128, 0, 2> { using R = Radices<4, 4, 4, 2>; };

template <int N, int V = 0, int SEL = 0, int PADV = QL_ZPAD>
struct ZCfg {
  static constexpr int M = N / 2;
  using R = typename ZPlanS<M, V, SEL>::R;
  static constexpr int TP = M / R::at(0);  // lanes per pencil = last-pass butterflies of the c2r
  static constexpr int P0 = QL_ZTHREADS / TP;
  static constexpr int P = (N <= 64) ? (P0 > 8 ? 8 : P0) : (P0 > 16 ? 16 : P0);
  static constexpr int THREADS = P * TP;
  // one LDS padding slot per 2^ZPAD elements of a pencil, or a grouped /
  // swizzled layout (fft_core.hpp LDS_PAIR16, LDS_QUAD8, LDS_XOR32)
  static constexpr int ZPAD = PADV;
  // grouped layouts: G pencils of TP lanes side by side in rows of 32
  // elements (TP of each pencil), element M in the last row
  static constexpr bool GROUPED = ZPAD == LDS_PAIR16 || ZPAD == LDS_QUAD8;
  static constexpr int G = ZPAD == LDS_PAIR16 ? 2 : 4;
  static constexpr int LSR = M + (ZPAD > 0 ? (M >> ZPAD) : 0) + 1 + LPAD;
  static constexpr int GSTRIDE = (M / (32 / G) + 1) * 32;  // elements of a group of G pencils
  static_assert(!GROUPED || (TP * G == 32 && P % G == 0), "grouped LDS layout: G pencils of 32 / G lanes");
  static constexpr int BUF = GROUPED ? (P / G) * GSTRIDE : P * LSR;  // LDS elements of the block's pencils
  // first element of pencil p of the block
  static constexpr int GL = G == 2 ? 1 : 2;  // log2 G
  static __device__ __forceinline__ int pencil_off(int p) {
    return GROUPED ? (p >> GL) * GSTRIDE + (p & (G - 1)) * TP : p * LSR;
  }
};

// W_N^{B0 r} a with N = 2M, r < R: a compile-time constant for each r (r is a
// constant wherever the caller's loop is unrolled)
template <typename T, int M, int B0, int R, int RR = 0>
__device__ __forceinline__ cx<T> rot_const(cx<T> a, int r) {
  if constexpr (RR + 1 >= R) {
    return twc<T, true, B0 * RR, 2 * M>(a);
  } else {
    return r == RR ? twc<T, true, B0 * RR, 2 * M>(a) : rot_const<T, M, B0, R, RR + 1>(a, r);
  }
}

// c2r pre-processing of one pencil held in LDS (pb[0..M]) into the pass-0
// input slots of the M-point inverse FFT: Z'[n] = (X_n + X*_{M-n}) +
// i W_N^{-n} (X_n - X*_{M-n}); Im of X_0 and X_M is ignored (numpy's irfft).
// With TWR, W_N^n for n = j + TP t + B0 r is W_N^{j + TP t} (one table read
// per t) times the constant W_N^{B0 r}.
template <typename T, int M, class CI, int PAD = 0, bool TWR = false>
__device__ __forceinline__ void c2r_prep(cx<T>* v, const cx<T>* pb, const cx<T>* s_tw, int j) {
  constexpr int R0 = CI::R(0), B0 = CI::B0;
#pragma unroll
  for (int s = 0; s < CI::NIN; ++s) {
    cx<T> z = {0, 0};
    if (CI::in_valid(j, s)) {
      const int n = CI::in_index(j, s);
      cx<T> xn = pb[lds_at<PAD>(n)], xr = pb[lds_at<PAD>(M - n)];
      if (n == 0) {
        xn.im = 0;
        xr.im = 0;
      }
      xr = cconj(xr);
      cx<T> w;
      if constexpr (TWR) {
        const cx<T> wb = s_tw[n - B0 * (s % R0)];  // n - B0 r = j + TP t
        w = rot_const<T, M, B0, R0>(wb, s % R0);
      } else {
        w = s_tw[n];
      }
      z = cadd(cadd(xn, xr), cmuli(cmulc(csub(xn, xr), w)));
    }
    v[s] = z;
  }
}

// r2c post-processing from the M-point forward FFT held in LDS (pb[0..M)):
// X_k = h (Z_k + Z*_{M-k}) + W_N^k h (Z_k - Z*_{M-k})/i, k < KZD, where h is
// 1/2 times whatever scale the caller folds into the output.  With TWR,
// W_N^k for k = j, j + TP, ... is read once and advanced by the constant W_N^TP.
template <typename T, int M, int PAD = 0, bool TWR = false, int TPC = 1>
__device__ __forceinline__ void r2c_post_store(cx<T>* dst, const cx<T>* pb, const cx<T>* s_tw, int j, int TP, int KZD,
                                               T h) {
  cx<T> w = s_tw[j];
  for (int k = j; k < KZD; k += TP) {
    const cx<T> zk = pb[lds_at<PAD>(k)];
    const cx<T> zr = cconj(pb[lds_at<PAD>(k == 0 ? 0 : M - k)]);
    const cx<T> e = cscale(cadd(zk, zr), h);
    const cx<T> o = crot<T, true>(cscale(csub(zk, zr), h));
    const cx<T> wk = TWR ? w : s_tw[k];
    st_stream<QL_NT_Z && NTBig<T, 2 * M>::value>(&dst[k], cadd(e, cmul(wk, o)));
    if constexpr (TWR) w = twc<T, true, TPC, 2 * M>(w);
  }
}

// The six input fields are streamed: while field f is transformed, the rows
// of field f+1 are already in flight into registers (software pipelining),
// and the cross product is accumulated field by field so only u, v, w and
// the three products stay live.
// state of one z-pass workgroup; the prefetch ring is two named buffers so
// every index is a compile-time constant (no private-memory arrays)
// plan selector of the 3-D z-pass (SEL of ZCfg): 1 in FP32, 2 in FP64
template <typename T> struct ZSel { static constexpr int value = sizeof(T) == 4 ? 1 : 2; };
// FP32 at N = 256 and 128 (16-lane radix-(8, 4, 4) and (4, 4, 4) plans): the
// paired, per-exchange swizzled LDS layout (fft_core.hpp LDS_PAIR16); the
// 8-lane radix-(8, 8) plan of N = 128's second variant: groups of four
// pencils (LDS_QUAD8).  No bank conflicts in the pencil exchanges
#ifndef QL_ZPAIR32
#define QL_ZPAIR32 1
#endif
// FP64 at N = 256 (32-lane radix-(4, 4, 4, 2) plan): unpadded pencils with XOR
// swizzles in the FFT exchanges (fft_core.hpp LDS_XOR32)
#ifndef QL_ZXOR64
#define QL_ZXOR64 1
#endif
template <typename T, int N, int V>
struct ZPadSel {
  static constexpr int value = sizeof(T) == 8 ? ((QL_ZXOR64 && N == 256 && V == 0) ? LDS_XOR32 : QL_ZPAD64)
                                              : !QL_ZPAIR32 ? QL_ZPAD
                                              : ((N == 256 || N == 128) && V == 0) ? LDS_PAIR16
                                              : (N == 128 && V == 1) ? LDS_QUAD8 : QL_ZPAD;
};
template <typename T, int N, int V>
using Z3Cfg = ZCfg<N, V, ZSel<T>::value, ZPadSel<T, N, V>::value>;

template <typename T, int N, int V = 0>
struct ZState {
  using C = Z3Cfg<T, N, V>;
  static constexpr int M = C::M, TP = C::TP, P = C::P, Nh = M + 1, NT = C::THREADS;
  static constexpr int ZP = C::ZPAD;
  // every pencil's TP lanes lie in one wavefront, and each wavefront loads
  // and parks its own pencils: no workgroup barrier after the table set-up
  static constexpr int WL = NT < 64 ? NT : 64;  // lanes of one wavefront of the block
  static constexpr int PWV = WL / TP;           // pencils per wavefront
  static_assert(WL % TP == 0 && P % PWV == 0, "z-pass pencils must not straddle wavefronts");
  // twiddles formed from one table read (QL_ZTWR32 / QL_ZTWR64): bit 0 in the
  // FFT passes (per butterfly), bit 1 in the c2r / r2c pre- and post-processing
  static constexpr int TWM = sizeof(T) == 4 ? QL_ZTWR32 : QL_ZTWR64;
  static constexpr bool TWR = (TWM & 2) != 0;
  using CI = PFFT<T, false, M, TP, Reverse<typename C::R>, ZP, SYNC_WAVE, (TWM & 1) != 0>;  // last pass: TP butterflies
  using CF = PFFT<T, true, M, TP, typename C::R, ZP, SYNC_WAVE, (TWM & 1) != 0>;  // pass 0 takes CI's output order
  static constexpr int NO = CI::NOUT;
  // prefetch slots per lane per field: the wavefront's PWV rows of M
  // contiguous entries, plus their Nyquist entries (one per pencil, taken by
  // lanes 0..PWV-1 from the contiguous Nyquist plane)
  static constexpr int NEM = (PWV * M + WL - 1) / WL;
  static constexpr int NE = NEM + 1;
  // fields in flight: two in FP32; QL_ZDEPTH64 in FP64 at N >= 256, one at
  // the smaller grids, where more would spill
  static constexpr int DEPTH = sizeof(T) == 4 ? QL_ZDEPTH32 : (N >= 256 ? QL_ZDEPTH64 : 1);
  static constexpr bool NTL = QL_NTL && NTBig<T, N>::value;
  const NS3Z<T>& a;
  cx<T>* s_tw;   // flat N-point table (c2r / r2c pre- and post-processing)
  cx<T>* s_twi;  // per-plan table of the inverse half-length FFT
  cx<T>* s_twf;  // per-plan table of the forward half-length FFT
  cx<T>* s_buf;
  int tid, p, j, row0, row, ln, wrow;
  size_t fs;
  cx<T> pf[DEPTH][NE];  // prefetch ring, indexed by compile-time field numbers only
  cx<T> uu[3][NO];
  cx<T> nl[3][NO];

  __device__ __forceinline__ ZState(const NS3Z<T>& a_, cx<T>* tw, cx<T>* twi, cx<T>* twf, cx<T>* buf)
      : a(a_), s_tw(tw), s_twi(twi), s_twf(twf), s_buf(buf) {
    tid = threadIdx.x;
    p = tid / TP;
    j = tid % TP;
    set_block(blockIdx.x);
    ln = tid % WL;
    wrow = (tid / WL) * PWV;  // first pencil of this wavefront within the block
    fs = (size_t)N * N * M;
  }
  // the block's P pencils are rows b*P .. b*P+P-1 of the [N*N] pencil index
  __device__ __forceinline__ void set_block(int b) {
    row0 = b * P;
    row = row0 + p;
  }
  // request field f of the pencils whose block starts at row r0
  __device__ __forceinline__ void fetch(int f, cx<T>* dst, int r0) {
    const cx<T>* src = a.in + f * fs + (size_t)(r0 + wrow) * M;
    if constexpr (C::GROUPED) {
      // TP consecutive entries of each pencil per load: lanes TP q .. TP q +
      // TP - 1 take entries TP e .. TP e + TP - 1 of pencil q
      static_assert(NEM * TP == M, "grouped fetch covers each row");
#pragma unroll
      for (int e = 0; e < NEM; ++e) dst[e] = ld_stream<NTL>(&src[(ln / TP) * M + TP * e + ln % TP]);
    } else {
#pragma unroll
      for (int e = 0; e < NEM; ++e) {
        const int idx = ln + e * WL;
        if (idx < PWV * M) dst[e] = ld_stream<NTL>(&src[idx]);
      }
    }
    if (ln < PWV) dst[NEM] = a.in[a.q + (size_t)f * N * N + r0 + wrow + ln];
  }
  __device__ __forceinline__ void park(const cx<T>* src) {
    if constexpr (C::GROUPED) {
#pragma unroll
      for (int e = 0; e < NEM; ++e) s_buf[C::pencil_off(wrow + ln / TP) + lds_at<ZP>(TP * e + ln % TP)] = src[e];
    } else {
#pragma unroll
      for (int e = 0; e < NEM; ++e) {
        const int idx = ln + e * WL;
        if (idx < PWV * M) {
          const int q = idx / M, k = idx - q * M;
          s_buf[C::pencil_off(wrow + q) + lds_at<ZP>(k)] = src[e];
        }
      }
    }
    if (ln < PWV) s_buf[C::pencil_off(wrow + ln) + lds_at<ZP>(M)] = src[NEM];
  }
  static __device__ __forceinline__ T inv_n3() { return (T)(1.0 / ((double)N * N * N)); }
  template <int F>
  __device__ __forceinline__ cx<T>* slot() {
    return pf[F % DEPTH];
  }
  // c2r of field F from its parked rows; field F+DEPTH is requested meanwhile
  template <int F>
  __device__ __forceinline__ void field() {
    cx<T>* pb = s_buf + C::pencil_off(p);
    lds_sync<SYNC_WAVE>();  // previous field's FFT scratch is free
    park(slot<F>());
    if constexpr (F + DEPTH < 6) fetch(F + DEPTH, slot<F>(), row0);
    lds_sync<SYNC_WAVE>();
    cx<T> v[CI::MAXE];
    c2r_prep<T, M, CI, ZP, TWR>(v, pb, s_tw, j);
    lds_sync<SYNC_WAVE>();
    CI::run(v, pb, 1, j, s_twi, 0);
    // (re, im) hold the (even, odd) z samples, still without the inverse
    // transform's 1/N^3, which the r2c post-processing applies to u x omega
    // (and the stores of u and omega apply to themselves)
    if constexpr (F >= 3) {
      if (a.w_out != nullptr) {
        cx<T>* dst = reinterpret_cast<cx<T>*>(a.w_out + ((size_t)(F - 3) * N * N + row) * N);
#pragma unroll
        for (int s = 0; s < NO; ++s) dst[CI::out_index(j, s)] = cscale(v[s], inv_n3());
      }
    }
    if constexpr (F < 3) {
#pragma unroll
      for (int s = 0; s < NO; ++s) uu[F][s] = v[s];
    } else if constexpr (F == 3) {  // omega_x: n_y += w ox, n_z -= v ox
#pragma unroll
      for (int s = 0; s < NO; ++s) {
        nl[1][s] = {nl[1][s].re + uu[2][s].re * v[s].re, nl[1][s].im + uu[2][s].im * v[s].im};
        nl[2][s] = {nl[2][s].re - uu[1][s].re * v[s].re, nl[2][s].im - uu[1][s].im * v[s].im};
      }
    } else if constexpr (F == 4) {  // omega_y: n_x -= w oy, n_z += u oy
#pragma unroll
      for (int s = 0; s < NO; ++s) {
        nl[0][s] = {nl[0][s].re - uu[2][s].re * v[s].re, nl[0][s].im - uu[2][s].im * v[s].im};
        nl[2][s] = {nl[2][s].re + uu[0][s].re * v[s].re, nl[2][s].im + uu[0][s].im * v[s].im};
      }
    } else {  // omega_z: n_x += v oz, n_y -= u oz
#pragma unroll
      for (int s = 0; s < NO; ++s) {
        nl[0][s] = {nl[0][s].re + uu[1][s].re * v[s].re, nl[0][s].im + uu[1][s].im * v[s].im};
        nl[1][s] = {nl[1][s].re - uu[0][s].re * v[s].re, nl[1][s].im - uu[0][s].im * v[s].im};
      }
    }
  }
  template <int F>
  __device__ __forceinline__ void fields() {
    if constexpr (F < 6) {
      field<F>();
      fields<F + 1>();
    }
  }
};

template <typename T, int N, bool UIN, int V>
__device__ __forceinline__ void ns3_z_body(const NS3Z<T>& a, cx<T>* s_tw, cx<T>* s_twi, cx<T>* s_twf,
                                           cx<T>* s_buf) {
  using St = ZState<T, N, V>;
  constexpr int M = St::M, TP = St::TP, NO = St::NO;
  using CI = typename St::CI;
  using CF = typename St::CF;
  static_assert(CI::NOUT == CF::NIN, "z-pass hand-off mismatch");
  static_assert(CI::BL == TP, "z-pass lanes");
  St z(a, s_tw, s_twi, s_twf, s_buf);
  stage_table(s_tw, a.tw, N);
  CI::fill_tw(s_twi, a.tw, 2);
  CF::fill_tw(s_twf, a.tw, 2);
  __syncthreads();  // the only workgroup barrier: tables visible to every wavefront
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int s = 0; s < NO; ++s) z.nl[c][s] = {0, 0};
  constexpr int f0 = UIN ? 3 : 0;  // first field taken through c2r
  if constexpr (UIN) {
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const cx<T>* src = reinterpret_cast<const cx<T>*>(a.u_in + ((size_t)f * N * N + z.row) * N);
#pragma unroll
      for (int s = 0; s < NO; ++s) z.uu[f][s] = src[CI::out_index(z.j, s)];
    }
  }
  z.fetch(f0, z.template slot<f0>(), z.row0);
  if constexpr (St::DEPTH >= 2) z.fetch(f0 + 1, z.template slot<f0 + 1>(), z.row0);
  if constexpr (St::DEPTH >= 3) z.fetch(f0 + 2, z.template slot<f0 + 2>(), z.row0);
  z.template fields<f0>();
  if (a.u_out != nullptr) {
    const T su = UIN ? (T)1 : St::inv_n3();
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      cx<T>* dst = reinterpret_cast<cx<T>*>(a.u_out + ((size_t)f * N * N + z.row) * N);
#pragma unroll
      for (int s = 0; s < NO; ++s) dst[CI::out_index(z.j, s)] = cscale(z.uu[f][s], su);
    }
  }
  // 1/2 of the r2c post-processing times the 1/N^3 of each inverse-transformed
  // factor of u x omega (u given in physical space carries its own)
  const double n3 = (double)N * N * N;
  const T h = (T)(UIN ? 0.5 / n3 : 0.5 / (n3 * n3));
  cx<T>* pb = s_buf + Z3Cfg<T, N, V>::pencil_off(z.p);
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    cx<T> v[CF::MAXE];
#pragma unroll
    for (int s = 0; s < NO; ++s) v[s] = z.nl[cc][s];
    lds_sync<SYNC_WAVE>();
    CF::run(v, pb, 1, z.j, s_twf, 0);
#pragma unroll
    for (int s = 0; s < CF::NOUT; ++s)
      if (CF::out_valid(z.j, s)) pb[lds_at<St::ZP>(CF::out_index(z.j, s))] = v[s];
    lds_sync<SYNC_WAVE>();
    r2c_post_store<T, M, St::ZP, St::TWR, TP>(a.nz + ((size_t)cc * N * N + z.row) * a.KZP, pb, s_tw, z.j, TP, a.KZD,
                                               h);
  }
}

#define QL_Z_LDS                                                                            \
  using St = ZState<T, N, V>;                                                               \
  __shared__ cx<T> s_tw[N];                                                                 \
  __shared__ cx<T> s_twi[St::CI::TWL > 0 ? St::CI::TWL : 1];                                \
  __shared__ cx<T> s_twf[St::CF::TWL > 0 ? St::CF::TWL : 1];                                \
  __shared__ cx<T> s_buf[Z3Cfg<T, N, V>::BUF]

template <typename T, int N, int V = 0>
__device__ __forceinline__ void ns3_z_run(const NS3Z<T>& a, cx<T>* s_tw, cx<T>* s_twi, cx<T>* s_twf,
                                          cx<T>* s_buf) {
  // the register prefetch ring is indexed by compile-time field numbers only
  if (a.u_in != nullptr) ns3_z_body<T, N, true, V>(a, s_tw, s_twi, s_twf, s_buf);
  else ns3_z_body<T, N, false, V>(a, s_tw, s_twi, s_twf, s_buf);
}

template <typename T, int N, int V = 0>
__global__ void __launch_bounds__((Z3Cfg<T, N, V>::THREADS)) QL_ATTR_Z
k_ns3_z(Pack<NS3Z<T>> pk) {
  QL_Z_LDS;
  const NS3Z<T> a = pk.m[blockIdx.z];
  ns3_z_run<T, N, V>(a, s_tw, s_twi, s_twf, s_buf);
}

// ---------------------------------------------------------------------------
// forward y-pass with |ky| < kmax truncation: NZ -> NY
// ---------------------------------------------------------------------------
template <typename T>
struct NS3YF {
  const cx<T>* nz;  // [3][N][N][KZP]
  cx<T>* ny;        // [3][N][KYD][KZP]
  int KZD, KZP, KYD;
  const cx<T>* tw;
};

#define QL_YF_LDS                                                              \
  __shared__ cx<T> s_tw[XFwd<T, N>::TWL > 0 ? XFwd<T, N>::TWL : 1];          \
  __shared__ cx<T> s_buf[N * XCfg<N, V>::LS]

template <typename T, int N, int V = 0>
__device__ __forceinline__ void ns3_yfwd_run(const NS3YF<T>& a, cx<T>* s_tw, cx<T>* s_buf) {
  using C = XCfg<N, V>;
  constexpr bool NTL = QL_NTL && QL_NTL_NZ && NTBig<T, N>::value;
  using F = XFwd<T, N>;
  const int tid = threadIdx.x, p = tid % C::TIX, j = tid / C::TIX;
  const int kz = blockIdx.x * C::TIX + p;
  const bool ok = kz < a.KZD;
  const int xi = blockIdx.y;
  F::fill_tw(s_tw, a.tw, 1);
  __syncthreads();
  cx<T> pre[3][F::NIN];
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    const size_t bi = ((size_t)cc * N + xi) * N * a.KZP + kz;
#pragma unroll
    for (int s = 0; s < F::NIN; ++s)
      pre[cc][s] = (ok && F::in_valid(j, s)) ? ld_stream<NTL>(&a.nz[bi + (size_t)F::in_index(j, s) * a.KZP])
                                             : cx<T>{0, 0};
  }
#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    cx<T> v[F::MAXE];
#pragma unroll
    for (int s = 0; s < F::NIN; ++s) v[s] = pre[cc][s];
    F::run(v, s_buf + p, C::LS, j, s_tw, 0);
    if (ok) {
      const size_t bo = ((size_t)cc * N + xi) * a.KYD * a.KZP + kz;
#pragma unroll
      for (int s = 0; s < F::NOUT; ++s) {
        const int k = F::out_index(j, s);
        if (F::out_valid(j, s) && keep_k(k, N, a.KZD)) {
          const int kd = k < a.KZD ? k : k - (N - a.KYD);
          st_stream<QL_NT_YF && NTBig<T, N>::value>(&a.ny[bo + (size_t)kd * a.KZP], v[s]);
        }
      }
    }
  }
}

template <typename T, int N, int V = 0>
__global__ void __launch_bounds__((XCfg<N, V>::THREADS)) QL_ATTR_YF
k_ns3_yfwd(Pack<NS3YF<T>> pk) {
  QL_YF_LDS;
  const NS3YF<T> a = pk.m[blockIdx.z];
  ns3_yfwd_run<T, N, V>(a, s_tw, s_buf);
}

}  // namespace qk
