// Pencil FFT engine for gfx950 (CDNA4).
//
// A "pencil" is one 1-D line of an array that is transformed; a workgroup
// transforms many pencils at once, each by TP cooperating lanes.  The engine
// is a mixed-radix Stockham autosort FFT:
//   * pass 0 takes its inputs from registers (the caller loads them straight
//     from HBM in the pass-0 order, so no LDS staging on the way in),
//   * every later pass is one LDS round trip (write, barrier, read, barrier),
//   * the last pass leaves its outputs in registers in "final order", so the
//     caller can fuse pointwise work and store straight to HBM.
// The inverse plan is the forward plan's radix list reversed, which makes the
// forward transform's output register order equal to the inverse transform's
// input register order: a forward FFT, pointwise math, and an inverse FFT can
// run back to back without touching LDS in between.
//
// Conventions: forward = exp(-2 pi i n k / M), inverse = exp(+2 pi i n k / M),
// both unnormalized.  Twiddles come from a table tw[i] = exp(-2 pi i i / TN)
// (computed in FP64 on the host), usually staged in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qf {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

template <typename T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename T> __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T> __device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
template <typename T> __device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T> __device__ __forceinline__ cx<T> cneg(cx<T> a) { return {-a.re, -a.im}; }
// multiply by +i
template <typename T> __device__ __forceinline__ cx<T> cmuli(cx<T> a) { return {-a.im, a.re}; }
// Compiler barrier on a value (no instruction): the code after it sees an
// opaque register, so the products and sums it forms from v are contracted
// into FMAs the same way whatever computed v (floating-point contraction
// otherwise depends on the surrounding code of each kernel specialisation).
template <typename T> __device__ __forceinline__ void opaque(cx<T>& v) { asm volatile("" : "+v"(v.re), "+v"(v.im)); }
// store of a value that a later kernel reads back: NT selects a non-temporal
// store, which MI355X retires at a higher write rate than a cached one but
// which leaves nothing in the last-level cache for the reader
template <bool NT, typename T> __device__ __forceinline__ void st_stream(cx<T>* p, cx<T> v) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    v2 x = {v.re, v.im};
    __builtin_nontemporal_store(x, reinterpret_cast<v2*>(p));
  } else {
    *p = v;
  }
}
// load of a value read once from a stream far larger than the caches (NT:
// non-temporal load)
template <bool NT, typename T> __device__ __forceinline__ cx<T> ld_stream(const cx<T>* p) {
  if constexpr (NT) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    const v2 x = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
    return {x.x, x.y};
  } else {
    return *p;
  }
}
// multiply by -i (FWD) or +i (!FWD)
template <typename T, bool FWD> __device__ __forceinline__ cx<T> crot(cx<T> a) {
  if constexpr (FWD) return {a.im, -a.re};
  else return {-a.im, a.re};
}

// ---------------------------------------------------------------------------
// compile-time cos/sin of 2 pi k / R (Taylor series after reduction to
// [-pi/4, pi/4]); used only for the constant in-register DFT twiddles.
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double x) {
  double t = x, s = x;
  for (int n = 1; n < 14; ++n) { t *= -x * x / ((2.0 * n) * (2.0 * n + 1.0)); s += t; }
  return s;
}
constexpr double taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int n = 1; n < 14; ++n) { t *= -x * x / ((2.0 * n - 1.0) * (2.0 * n)); s += t; }
  return s;
}
// cos / sin of 2 pi k / R using the octant symmetries on 8k/R
constexpr double cos2pi(int k, int R) {
  k = ((k % R) + R) % R;
  // angle = 2 pi k / R; reflect to [0, pi/4]
  const double f = (double)k / (double)R;  // in [0,1)
  if (f <= 0.125) return taylor_cos(2 * kPi * f);
  if (f <= 0.25) return taylor_sin(2 * kPi * (0.25 - f));
  if (f <= 0.375) return -taylor_sin(2 * kPi * (f - 0.25));
  if (f <= 0.5) return -taylor_cos(2 * kPi * (0.5 - f));
  if (f <= 0.625) return -taylor_cos(2 * kPi * (f - 0.5));
  if (f <= 0.75) return -taylor_sin(2 * kPi * (0.75 - f));
  if (f <= 0.875) return taylor_sin(2 * kPi * (f - 0.75));
  return taylor_cos(2 * kPi * (1.0 - f));
}
constexpr double sin2pi(int k, int R) { return cos2pi(4 * k - R, 4 * R); }

// multiply by W_R^K = exp(-+2 pi i K/R) (sign by FWD); trivial factors folded
template <typename T, bool FWD, int K, int R>
__device__ __forceinline__ cx<T> twc(cx<T> a) {
  constexpr int k = ((K % R) + R) % R;
  if constexpr (k == 0) {
    return a;
  } else if constexpr (2 * k == R) {
    return cneg(a);
  } else if constexpr (4 * k == R) {
    return crot<T, FWD>(a);
  } else if constexpr (4 * k == 3 * R) {
    return crot<T, !FWD>(a);
  } else {
    constexpr double c = cos2pi(k, R);
    constexpr double s = sin2pi(k, R);
    const T cc = (T)c;
    const T ss = FWD ? (T)(-s) : (T)s;  // W = c + i*ss
    return {a.re * cc - a.im * ss, a.re * ss + a.im * cc};
  }
}

// ---------------------------------------------------------------------------
// in-register small DFTs on v[0..R) (stride S in the register array)
// ---------------------------------------------------------------------------
template <typename T, bool FWD, int R> struct DFT;

template <typename T, bool FWD> struct DFT<T, FWD, 1> {
  template <int S = 1> __device__ __forceinline__ static void run(cx<T>*) {}
};

template <typename T, bool FWD> struct DFT<T, FWD, 2> {
  template <int S = 1> __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> a = v[0], b = v[S];
    v[0] = cadd(a, b);
    v[S] = csub(a, b);
  }
};

template <typename T, bool FWD> struct DFT<T, FWD, 3> {
  template <int S = 1> __device__ __forceinline__ static void run(cx<T>* v) {
    constexpr double s3 = 0.86602540378443864676372317075293618;  // sin(2 pi/3)
    cx<T> a = v[0], b = v[S], c = v[2 * S];
    cx<T> t = cadd(b, c);
    cx<T> m = {a.re - (T)0.5 * t.re, a.im - (T)0.5 * t.im};
    cx<T> d = csub(b, c);
    // s = -i*s3*d (FWD) or +i*s3*d
    cx<T> s = crot<T, FWD>(cscale(d, (T)s3));
    v[0] = cadd(a, t);
    v[S] = cadd(m, s);
    v[2 * S] = csub(m, s);
  }
};

template <typename T, bool FWD> struct DFT<T, FWD, 4> {
  template <int S = 1> __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> a0 = cadd(v[0], v[2 * S]);
    cx<T> a1 = csub(v[0], v[2 * S]);
    cx<T> a2 = cadd(v[S], v[3 * S]);
    cx<T> a3 = crot<T, FWD>(csub(v[S], v[3 * S]));
    v[0] = cadd(a0, a2);
    v[2 * S] = csub(a0, a2);
    v[S] = cadd(a1, a3);
    v[3 * S] = csub(a1, a3);
  }
};

// R = 8: radix-2 DIT over two 4-point DFTs
template <typename T, bool FWD> struct DFT<T, FWD, 8> {
  template <int S = 1> __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> e[4] = {v[0], v[2 * S], v[4 * S], v[6 * S]};
    cx<T> o[4] = {v[S], v[3 * S], v[5 * S], v[7 * S]};
    DFT<T, FWD, 4>::template run<1>(e);
    DFT<T, FWD, 4>::template run<1>(o);
    o[1] = twc<T, FWD, 1, 8>(o[1]);
    o[2] = twc<T, FWD, 2, 8>(o[2]);
    o[3] = twc<T, FWD, 3, 8>(o[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k * S] = cadd(e[k], o[k]);
      v[(k + 4) * S] = csub(e[k], o[k]);
    }
  }
};

// R = 16: 4 x 4 with internal twiddles W16^(n2 k1)
template <typename T, bool FWD> struct DFT<T, FWD, 16> {
  template <int S = 1> __device__ __forceinline__ static void run(cx<T>* v) {
    cx<T> a[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cx<T> t[4] = {v[n2 * S], v[(n2 + 4) * S], v[(n2 + 8) * S], v[(n2 + 12) * S]};
      DFT<T, FWD, 4>::template run<1>(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[n2 * 4 + k1] = t[k1];
    }
    a[1 * 4 + 1] = twc<T, FWD, 1, 16>(a[1 * 4 + 1]);
    a[1 * 4 + 2] = twc<T, FWD, 2, 16>(a[1 * 4 + 2]);
    a[1 * 4 + 3] = twc<T, FWD, 3, 16>(a[1 * 4 + 3]);
    a[2 * 4 + 1] = twc<T, FWD, 2, 16>(a[2 * 4 + 1]);
    a[2 * 4 + 2] = twc<T, FWD, 4, 16>(a[2 * 4 + 2]);
    a[2 * 4 + 3] = twc<T, FWD, 6, 16>(a[2 * 4 + 3]);
    a[3 * 4 + 1] = twc<T, FWD, 3, 16>(a[3 * 4 + 1]);
    a[3 * 4 + 2] = twc<T, FWD, 6, 16>(a[3 * 4 + 2]);
    a[3 * 4 + 3] = twc<T, FWD, 9, 16>(a[3 * 4 + 3]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cx<T> t[4] = {a[0 * 4 + k1], a[1 * 4 + k1], a[2 * 4 + k1], a[3 * 4 + k1]};
      DFT<T, FWD, 4>::template run<1>(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[(k1 + 4 * k2) * S] = t[k2];
    }
  }
};

// ---------------------------------------------------------------------------
// radix lists
// ---------------------------------------------------------------------------
template <int... Rs> struct Radices {
  static constexpr int n = sizeof...(Rs);
  static constexpr int at(int p) {
    constexpr int a[] = {Rs...};
    return a[p];
  }
  static constexpr int prod_before(int p) {
    constexpr int a[] = {Rs...};
    int s = 1;
    for (int q = 0; q < p; ++q) s *= a[q];
    return s;
  }
  static constexpr int total() { return prod_before(n); }
};

template <typename A, typename B> struct RevImpl;
template <int... Acc> struct RevImpl<Radices<Acc...>, Radices<>> { using type = Radices<Acc...>; };
template <int... Acc, int R, int... Rest> struct RevImpl<Radices<Acc...>, Radices<R, Rest...>> {
  using type = typename RevImpl<Radices<R, Acc...>, Radices<Rest...>>::type;
};
template <typename L> using Reverse = typename RevImpl<Radices<>, L>::type;

// default radix plan (forward order) and lanes per pencil, by length M
template <int M> struct Plan;
template <> struct Plan<4> { using R = Radices<4>; static constexpr int TP = 1; };
template <> struct Plan<8> { using R = Radices<8>; static constexpr int TP = 1; };
template <> struct Plan<16> { using R = Radices<4, 4>; static constexpr int TP = 4; };
template <> struct Plan<32> { using R = Radices<4, 8>; static constexpr int TP = 4; };
template <> struct Plan<48> { using R = Radices<3, 16>; static constexpr int TP = 4; };
template <> struct Plan<64> { using R = Radices<8, 8>; static constexpr int TP = 8; };
template <> struct Plan<96> { using R = Radices<3, 4, 8>; static constexpr int TP = 8; };
template <> struct Plan<128> { using R = Radices<8, 16>; static constexpr int TP = 16; };
template <> struct Plan<256> { using R = Radices<16, 16>; static constexpr int TP = 16; };
template <> struct Plan<512> { using R = Radices<8, 8, 8>; static constexpr int TP = 32; };

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------
// PFFT: M-point FFT of one pencil by TP lanes (lane index j in [0,TP)).
//   in : v[t*R0 + r] = x[(j + TP*t) + (M/R0)*r]          (pass-0 order)
//   out: v[t*RL + r] = X[(j + TP*t) + (M/RL)*r]          (final order)
// Slots whose butterfly index is >= M/R of that pass are ignored.
// lds points at this pencil's element 0; element n lives at lds[n*ls].
// tw is either a flat table exp(-2 pi i q / TN) with TN = M * tws (tws > 0),
// or, with tws == 0, the per-plan layout written by fill_tw.
// Every thread of the workgroup must call run() (it contains barriers) unless
// the plan has a single pass.
// ---------------------------------------------------------------------------
// LDS element address of pencil element n: n*ls, plus one padding slot every
// 2^PAD elements when PAD > 0 (breaks the stride-R bank pattern of the
// first Stockham pass when a pencil is contiguous in LDS).
//
// PAD == LDS_PAIR16: pencils of 16 lanes stored in pairs, element n of a
// pencil in row n / 16 (32 elements per row: 16 of each pencil of the pair,
// the pencil base pointer selecting its half) at column n % 16.  On gfx950 an
// 8-byte LDS read serves 32 lanes (two such pencils) per cycle over 64 banks
// and an 8-byte write 16 lanes over 32 banks, so any 16 consecutive elements
// of the two pencils land on distinct banks.  Inside an FFT the pair layout is
// swizzled per exchange (PFFT::xat), which keeps the strided Stockham writes
// on distinct banks as well.
//
// PAD == LDS_XOR32: pencils of 32 lanes with 16-byte elements, unpadded.  A
// 16-byte read serves 16 lanes per cycle over 64 banks, the lanes {0-3, 12-15,
// 20-27} and {4-11, 16-19, 28-31} of each half wave together, and a 16-byte
// write 8 consecutive lanes over 32 banks, so consecutive runs need no
// padding; the FFT exchanges XOR low index bits with higher ones (PFFT::xat).
//
// PAD == LDS_QUAD8: pencils of 8 lanes with 8-byte elements stored in groups
// of four, element n of a pencil in row n / 8 (32 elements per row, 8 per
// pencil of the group) at column n % 8; the radix-(8, 8) plan's one exchange
// XORs the column with the row.
constexpr int LDS_PAIR16 = -1;
constexpr int LDS_XOR32 = -2;
constexpr int LDS_QUAD8 = -3;
template <int PAD>
__device__ __forceinline__ int lds_at(int n) {
  if constexpr (PAD == LDS_PAIR16) return (n >> 4) * 32 + (n & 15);
  else if constexpr (PAD == LDS_QUAD8) return (n >> 3) * 32 + (n & 7);
  else if constexpr (PAD == LDS_XOR32) return n;
  else if constexpr (PAD > 0) return n + (n >> PAD);
  else return n;
}

// LDS hand-off between passes.  SYNC_BLOCK: lanes of one pencil may sit in
// different wavefronts (workgroup barrier).  SYNC_WAVE: every pencil lies in
// one wavefront, whose LDS instructions execute in issue order, so only
// compiler reordering has to be fenced off (no s_barrier).
enum { SYNC_BLOCK = 0, SYNC_WAVE = 1 };
template <int SYNC>
__device__ __forceinline__ void lds_sync() {
  if constexpr (SYNC == SYNC_BLOCK) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// w^r for r = 1 .. R-1 from w (= w^1): squares for the even powers, one more
// product for the odd ones, so no power is more than three products deep
template <typename T, int R>
__device__ __forceinline__ void tw_powers(cx<T> (&w)[R > 1 ? R : 2]) {
#pragma unroll
  for (int r = 2; r < R; ++r) w[r] = (r % 2 == 0) ? cmul(w[r / 2], w[r / 2]) : cmul(w[r - 1], w[1]);
}

// TWR: the twiddles of a butterfly are formed from its first one (one table
// read instead of R - 1; a few multiplies more), which moves work from the
// LDS pipe to the vector ALU
template <typename T, bool FWD, int M, int TP, typename RL, int PAD = 0, int SYNC = SYNC_BLOCK, bool TWR = false>
struct PFFT {
  static constexpr int NP = RL::n;
  static_assert(RL::total() == M, "radix plan does not multiply to M");
  static constexpr int R(int p) { return RL::at(p); }
  static constexpr int B(int p) { return M / RL::at(p); }
  static constexpr int NB(int p) { return (B(p) + TP - 1) / TP; }
  static constexpr int maxe_impl(int p) { return p >= NP ? 0 : cmax(NB(p) * R(p), maxe_impl(p + 1)); }
  static constexpr int MAXE = maxe_impl(0);
  static constexpr int NIN = NB(0) * R(0);
  static constexpr int NOUT = NB(NP - 1) * R(NP - 1);
  static constexpr int B0 = B(0);
  static constexpr int BL = B(NP - 1);
  static constexpr int RF = R(NP - 1);
  static constexpr int NBL = NB(NP - 1);

  // element index held in input slot s
  __device__ __forceinline__ static int in_index(int j, int s) {
    const int t = s / R(0), r = s % R(0);
    return (j + TP * t) + B(0) * r;
  }
  __device__ __forceinline__ static bool in_valid(int j, int s) { return (j + TP * (s / R(0))) < B(0); }
  __device__ __forceinline__ static int out_index(int j, int s) {
    const int t = s / RF, r = s % RF;
    return (j + TP * t) + BL * r;
  }
  __device__ __forceinline__ static bool out_valid(int j, int s) { return (j + TP * (s / RF)) < BL; }

  // Per-plan twiddle layout: pass p >= 1 owns Ns*(R-1) entries ordered
  // [b mod Ns][r-1], so lanes with consecutive butterflies read LDS at a
  // stride of R-1 entries (a flat table is read at strides of many entries,
  // which serializes on LDS banks).  The inverse uses the conjugates.
  static constexpr int tw_off(int p) {
    int o = 0;
    for (int q = 1; q < p; ++q) o += RL::prod_before(q) * (RL::at(q) - 1);
    return o;
  }
  static constexpr int TWL = tw_off(NP);  // entries of the per-plan table
  // fill dst[0..TWL) from a table src[q] = exp(-2 pi i q / (M*tws)) (cooperative)
  __device__ __forceinline__ static void fill_tw(cx<T>* dst, const cx<T>* src, int tws) {
    fill_tw_pass<1>(dst, src, tws);
  }
  template <int p>
  __device__ __forceinline__ static void fill_tw_pass(cx<T>* dst, const cx<T>* src, int tws) {
    if constexpr (p < NP) {
      constexpr int Rp = R(p);
      constexpr int Ns = RL::prod_before(p);
      for (int i = threadIdx.x; i < Ns * (Rp - 1); i += blockDim.x) {
        const int bm = i / (Rp - 1), r = i % (Rp - 1) + 1;
        dst[tw_off(p) + i] = src[bm * r * (M / (Ns * Rp)) * tws];
      }
      fill_tw_pass<p + 1>(dst, src, tws);
    }
  }

  template <int p>
  __device__ __forceinline__ static void butterflies(cx<T>* v, int j, const cx<T>* tw, int twstride) {
    constexpr int Rp = R(p);
    constexpr int Ns = RL::prod_before(p);
#pragma unroll
    for (int t = 0; t < NB(p); ++t) {
      const int b = j + TP * t;
      if constexpr (Ns > 1) {
        const int bm = b % Ns;
        if (twstride == 0 && TWR) {
          cx<T> w[Rp > 1 ? Rp : 2];
          w[1] = tw[tw_off(p) + bm * (Rp - 1)];
          tw_powers<T, Rp>(w);
#pragma unroll
          for (int r = 1; r < Rp; ++r) v[t * Rp + r] = FWD ? cmul(v[t * Rp + r], w[r]) : cmulc(v[t * Rp + r], w[r]);
        } else if (twstride == 0) {
          // per-plan layout (fill_tw)
          const cx<T>* tp = tw + tw_off(p) + bm * (Rp - 1) - 1;
#pragma unroll
          for (int r = 1; r < Rp; ++r) {
            cx<T> w = tp[r];
            v[t * Rp + r] = FWD ? cmul(v[t * Rp + r], w) : cmulc(v[t * Rp + r], w);
          }
        } else {
          // flat table: W_{Ns*Rp}^{bm*r} = tw[bm*r*(M/(Ns*Rp)) * twstride]
          const int step = bm * (M / (Ns * Rp)) * twstride;
#pragma unroll
          for (int r = 1; r < Rp; ++r) {
            cx<T> w = tw[step * r];
            v[t * Rp + r] = FWD ? cmul(v[t * Rp + r], w) : cmulc(v[t * Rp + r], w);
          }
        }
      }
      DFT<T, FWD, Rp>::template run<1>(v + t * Rp);
    }
  }

  // LDS position of element n in the exchange read by pass e.  With
  // LDS_PAIR16 the column is XORed with (row & mask) << shift, chosen so that
  // for every (t, r) the 16 lanes of a pencil write distinct columns: the
  // inverse radix-(4, 4, 8) plan's writes after pass 0 (4b + r) and pass 1
  // ((b / 4) 16 + b % 4 + 4r), the forward radix-(8, 4, 4) plan's after
  // pass 0 (8b + r) and pass 1 ((b / 8) 32 + b % 8 + 8r), and the M = 64
  // radix-(4, 4, 4) plan's (4b + r, (b / 4) 16 + b % 4 + 4r) both ways; the
  // reads (b + B r) are consecutive runs of 16 and stay conflict-free under
  // any row swizzle.
  static constexpr int xmask(int e) { return (FWD && M == 128) ? (e == 1 ? 7 : 2) : 3; }
  static constexpr int xshift(int e) { return e == 1 ? 0 : 2; }
  template <int e>
  __device__ __forceinline__ static int xat(int n) {
    if constexpr (PAD == LDS_PAIR16) {
      static_assert(TP == 16 && NP == 3 &&
                        ((M == 128 && (FWD ? (R(0) == 8 && R(1) == 4 && R(2) == 4)
                                           : (R(0) == 4 && R(1) == 4 && R(2) == 8))) ||
                         (M == 64 && R(0) == 4 && R(1) == 4 && R(2) == 4)),
                    "the paired LDS layout's swizzles are derived for the 16-lane radix-(8, 4, 4) and "
                    "radix-(4, 4, 4) plans");
      const int row = n >> 4;
      return row * 32 + ((n & 15) ^ ((row & xmask(e)) << xshift(e)));
    } else if constexpr (PAD == LDS_QUAD8) {
      static_assert(TP == 8 && M == 64 && NP == 2 && R(0) == 8 && R(1) == 8,
                    "the grouped 8-lane LDS layout's swizzle is derived for the radix-(8, 8) plan");
      const int row = n >> 3;
      return row * 32 + ((n & 7) ^ (row & 7));
    } else if constexpr (PAD == LDS_XOR32) {
      // n ^ ((n >> a1) & m1) << s1 ^ ((n >> a2) & m2) << s2 per exchange,
      // searched over with the 16-byte bank rule: every write (8 consecutive
      // lanes) and read (16 lanes of a group) of the radix-(4, 4, 4, 2) plan
      // and its inverse (2, 4, 4, 4) on distinct banks.  Each is a bijection
      // on 0 .. 127.
      static_assert(TP == 32 && M == 128 && NP == 4 &&
                        (FWD ? (R(0) == 4 && R(1) == 4 && R(2) == 4 && R(3) == 2)
                             : (R(0) == 2 && R(1) == 4 && R(2) == 4 && R(3) == 4)),
                    "the XOR LDS layout's swizzles are derived for the 32-lane radix-(4, 4, 4, 2) plan");
      if constexpr (!FWD) {
        if constexpr (e == 1) return n ^ ((n >> 3) & 1);
        else if constexpr (e == 2) return n ^ (((n >> 3) & 1) << 2) ^ ((n >> 3) & 3);
        else return n;
      } else {
        if constexpr (e == 1) return n ^ ((n >> 3) & 3);
        else if constexpr (e == 2) return n ^ (((n >> 3) & 3) << 1) ^ (((n >> 4) & 1) << 3);
        else return n;
      }
    } else {
      return lds_at<PAD>(n);
    }
  }

  template <int p>
  __device__ __forceinline__ static void write_lds(const cx<T>* v, cx<T>* lds, int ls, int j) {
    constexpr int Rp = R(p);
    constexpr int Ns = RL::prod_before(p);
#pragma unroll
    for (int t = 0; t < NB(p); ++t) {
      const int b = j + TP * t;
      if ((B(p) % TP == 0) || b < B(p)) {
        const int base = (b / Ns) * Ns * Rp + (b % Ns);
#pragma unroll
        for (int r = 0; r < Rp; ++r) lds[xat<p + 1>(base + Ns * r) * ls] = v[t * Rp + r];
      }
    }
  }

  template <int p>
  __device__ __forceinline__ static void read_lds(cx<T>* v, const cx<T>* lds, int ls, int j) {
    constexpr int Rp = R(p);
#pragma unroll
    for (int t = 0; t < NB(p); ++t) {
      const int b = j + TP * t;
      if ((B(p) % TP == 0) || b < B(p)) {
#pragma unroll
        for (int r = 0; r < Rp; ++r) v[t * Rp + r] = lds[xat<p>(b + B(p) * r) * ls];
      }
    }
  }

  template <int p>
  __device__ __forceinline__ static void pass_from(cx<T>* v, cx<T>* lds, int ls, int j, const cx<T>* tw, int tws) {
    if constexpr (p < NP) {
      write_lds<p - 1>(v, lds, ls, j);
      lds_sync<SYNC>();
      read_lds<p>(v, lds, ls, j);
      lds_sync<SYNC>();
      butterflies<p>(v, j, tw, tws);
      pass_from<p + 1>(v, lds, ls, j, tw, tws);
    }
  }

  // v must have MAXE slots
  __device__ __forceinline__ static void run(cx<T>* v, cx<T>* lds, int ls, int j, const cx<T>* tw, int tws) {
    butterflies<0>(v, j, tw, tws);
    pass_from<1>(v, lds, ls, j, tw, tws);
  }
};

// forward / inverse plans sharing register order at the hand-off point
template <typename T, int M, int TP>
using FwdFFT = PFFT<T, true, M, TP, typename Plan<M>::R>;
template <typename T, int M, int TP>
using InvFFT = PFFT<T, false, M, TP, Reverse<typename Plan<M>::R>>;
// the mirror pair: inverse first (plan order), then forward (reversed)
template <typename T, int M, int TP>
using InvFFTp = PFFT<T, false, M, TP, typename Plan<M>::R>;
template <typename T, int M, int TP>
using FwdFFTr = PFFT<T, true, M, TP, Reverse<typename Plan<M>::R>>;

// fftfreq(N) value of index i (Nyquist index N/2 maps to -N/2)
__device__ __forceinline__ int kfreq(int i, int N) { return i < (N + 1) / 2 ? i : i - N; }

// copy a twiddle table of n entries into LDS (caller synchronizes)
template <typename T>
__device__ __forceinline__ void stage_table(cx<T>* dst, const cx<T>* src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

}  // namespace qf
