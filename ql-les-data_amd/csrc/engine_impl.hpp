// EngineT<T, DIM, N>: launch orchestration for one solver configuration.
#pragma once
#include <algorithm>
#include <memory>
#include <vector>

#include "../../include/qlles.h"
#include "adjoint_kernels.hpp"
#include "aux_kernels.hpp"
#include "engine.hpp"
#include "fftlib.hpp"
#include "ns2d.hpp"

namespace ql {
using namespace qk;

template <typename T, int DIM, int N>
class EngineT final : public Engine {
  static constexpr int Nh = N / 2 + 1;
  Desc d_;
  int KZD, KZP, KYD, LD;
  // 3-D internal spectra (P, S1a, S1b, S2, X5 and Y6 between the fused
  // kernels): [nf][N][N][LDM] rows (LDM = N/2, cache-line aligned) followed by
  // the [nf][N][N] plane of kz = N/2 entries, at element qs(nf)
  int LDM;
  size_t qs(int nf) const { return (size_t)nf * N * N * LDM; }
  size_t rows_;  // N*N (3-D) or N (2-D)
  size_t fsz_;   // rows_ * LD: one padded spectral component
  size_t bytes_ = 0;
  bool step_ready_ = false;
  std::vector<void*> allocs_;
  // P: u_n in the padded row stride LD (the caller's spectrum has rows of
  // N/2+1, which are not cache-line aligned), refreshed by the state load of
  // each multi-step call and written back to the caller's array on the last step
  cx<T>* P = nullptr;
  cx<T>*S1a = nullptr, *S1b = nullptr, *S2 = nullptr, *X5 = nullptr, *Y6 = nullptr, *NZ = nullptr, *NY = nullptr;
  // tangent / adjoint workspace, allocated on first use: spectra AS_[i]
  // ([3][N][N][Nh]) and physical fields AP_[i] ([3][N][N][N])
  static constexpr int NAS = 10, NAP = 5;
  cx<T>* AS_[NAS] = {};
  T* AP_[NAP] = {};
  // physical u and omega of the four RK stage states of a step (u0, w0, u1, w1, ...)
  T* SF_[8] = {};
  double* partials_ = nullptr;
  double* spec_part_ = nullptr;  // per-block shell histograms (energy spectrum)
  double* alpha_ = nullptr;
  double* dscr_ = nullptr;
  const cx<T>* twN_ = nullptr;
  int nxblk_ = 0;  // capacity of the per-block partials (largest x-pass grid over the variants)

  template <class P>
  P* alloc(size_t n) {
    void* p = nullptr;
    QL_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(P)));
    allocs_.push_back(p);
    bytes_ += n * sizeof(P);
    return static_cast<P*>(p);
  }

  StageCtl<T> ctl(int stage) const {
    static const double a_rk4[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};  // spectral_solver.py:70
    static const double b_rk4[3] = {0.5, 0.5, 1.0};                              // spectral_solver.py:71
    StageCtl<T> c{};
    c.stage = stage;
    c.visc_if = d_.viscous == QLLES_VISC_IF;
    c.nu = (T)d_.nu;
    c.dt = (T)d_.dt;
    if (stage >= 0 && stage < 4) {
      c.ca = (T)(a_rk4[stage] * d_.dt);
      c.cb = stage < 3 ? (T)(b_rk4[stage] * d_.dt) : (T)0;
    }
    c.force_power = d_.forcing == QLLES_FORCING_POWER;
    c.kf2 = d_.kf * d_.kf;
    c.power = d_.power;
    if (d_.forcing == QLLES_FORCING_KOLMOGOROV) {
      // rfft2 of amp sin(k y) on the node grid: -i amp N^2 / 2 at (kx=0, ky=k)
      c.kolm_k = d_.kolm_k;
      c.kolm_re = (T)0;
      c.kolm_im = (T)(-0.5 * d_.kolm_amp * (double)N * (double)N);
    }
    c.drag = (T)d_.drag;
    return c;
  }

  // ---------------- kernel variants ----------------
  // Grids up to 128 carry a second tile width for the x / y passes (xv_) and
  // a second radix plan for the z-pass / 2-D y-pass (zv_); the autotuner
  // (set_variant / tune) picks per plan.  Larger grids keep variant 0.
  static constexpr bool HAS_V = N <= 128;
  int xv_ = 0, zv_ = 0;
  template <class F>
  void with_xv(F&& f) const {
    if constexpr (HAS_V) {
      if (xv_) {
        f(std::integral_constant<int, 1>{});
        return;
      }
    }
    f(std::integral_constant<int, 0>{});
  }
  template <class F>
  void with_zv(F&& f) const {
    if constexpr (HAS_V) {
      if (zv_) {
        f(std::integral_constant<int, 1>{});
        return;
      }
    }
    f(std::integral_constant<int, 0>{});
  }

  // ---------------- 3-D launches ----------------
  // stage and viscous mode select a compile-time specialisation of the x-pass
  // every launch passes its arguments as a Pack (ns3d.hpp): entry 0 for a
  // launch of this plan alone, one entry per member for a group
  template <class A>
  static Pack<A> pack1(const A& a) {
    Pack<A> pk{};
    pk.m[0] = a;
    return pk;
  }
  template <int SG, bool VIF, int V>
  static void launch_x3_t(const Pack<NS3X<T>>& pk, dim3 g, hipStream_t s) {
    hipLaunchKernelGGL((k_ns3_x<T, N, SG, VIF, V>), g, dim3(XCfg<N, V>::THREADS), 0, s, pk);
  }
  // stage and viscous mode of entry 0 select the specialisation (all entries share them)
  template <int V>
  static void launch_x3(const Pack<NS3X<T>>& pk, dim3 g, hipStream_t s) {
    const bool vif = pk.m[0].c.visc_if != 0;
    switch (pk.m[0].c.stage) {
      case -1: launch_x3_t<-1, false, V>(pk, g, s); break;
      case 0: vif ? launch_x3_t<0, true, V>(pk, g, s) : launch_x3_t<0, false, V>(pk, g, s); break;
      case 1: vif ? launch_x3_t<1, true, V>(pk, g, s) : launch_x3_t<1, false, V>(pk, g, s); break;
      case 2: vif ? launch_x3_t<2, true, V>(pk, g, s) : launch_x3_t<2, false, V>(pk, g, s); break;
      case 3: vif ? launch_x3_t<3, true, V>(pk, g, s) : launch_x3_t<3, false, V>(pk, g, s); break;
      case 4: vif ? launch_x3_t<4, true, V>(pk, g, s) : launch_x3_t<4, false, V>(pk, g, s); break;
      default: throw std::invalid_argument("bad stage");
    }
  }
  double C() const { return (double)rows_ * Nh * 2.0 * sizeof(T); }  // one spectral component
  // x-pass / y-inverse tiles along kz: full tiles below N/2 plus one Nyquist-column tile
  template <int V>
  static int xtiles_v() {
    constexpr int tix = XCfg<N, V>::TIX;
    constexpr int ntr = (N / 2) / tix;
    return ntr > 0 ? ntr + 1 : (Nh + tix - 1) / tix;
  }
  int xtiles() const { return xv_ && HAS_V ? xtiles_v<HAS_V ? 1 : 0>() : xtiles_v<0>(); }
  // blocks of the 3-D x-pass grid (FP64 energy partials per block) / 2-D grid
  int xblocks() const {
    if (DIM == 3) return xtiles() * N;
    return xv_ && HAS_V ? (Nh + XCfg<N, HAS_V ? 1 : 0>::TIX - 1) / XCfg<N, HAS_V ? 1 : 0>::TIX
                        : (Nh + XCfg<N>::TIX - 1) / XCfg<N>::TIX;
  }
  double xpass_bytes(const StageCtl<T>& c, bool inv, int ncomp, int nx) const {
    const double cc = C();
    double b = 0;
    if (c.stage < 0) {
      b += ncomp * cc;  // state
    } else {
      const double nyv = (double)ncomp * N * (DIM == 3 ? KYD : 1) * KZD * 2.0 * sizeof(T);
      b += nyv;                                               // N^ (dealiased modes only)
      b += ncomp * cc;                                        // u_s
      // the 3-D explicit scheme writes its accumulator (an increment) at
      // stage 1 only and reads it, with u_n, at stage 3 only
      const bool full = DIM == 2 || c.visc_if != 0;
      if (c.stage == 1 || c.stage == 2 || (!full && c.stage == 3)) b += ncomp * cc;  // u_n read
      if (full ? (c.stage >= 1 && c.stage <= 3) : c.stage == 3) b += ncomp * cc;  // acc read
      if (full ? (c.stage >= 0 && c.stage <= 2) : c.stage == 1) b += ncomp * cc;  // acc write
      b += ncomp * cc;                                        // u_{s+1} / du write
    }
    if (inv) b += nx * cc;
    return b;
  }
  void x3(NS3X<T>& a, hipStream_t s) {
    dim3 g(xtiles(), N);
    hipEvent_t e0 = nullptr;
    prof_begin(s, &e0);
    with_xv([&](auto v) { launch_x3<decltype(v)::value>(pack1(a), g, s); });
    QL_LAUNCHED();
    prof_end(s, K_XPASS, xpass_bytes(a.c, a.do_inverse, 3, 5) + (a.copy_dst ? 3.0 * C() : 0.0), e0);
  }
  void yinv3(int nf, hipStream_t s) {
    // Y6 is left in the split layout for the z-pass; the velocity-only form
    // (nf = 3) feeds the generic c2r, which takes rows of LD with kz = N/2 inside
    NS3Y<T> a{X5, Y6, LDM, nf == 6 ? LDM : LD, qs(5), nf == 6 ? qs(6) : 0, nf, twN_};
    dim3 g(xtiles(), N);
    hipEvent_t e0 = nullptr;
    prof_begin(s, &e0);
    with_xv([&](auto v) {
      constexpr int V = decltype(v)::value;
      hipLaunchKernelGGL((k_ns3_yinv<T, N, V>), g, dim3(XCfg<N, V>::THREADS), 0, s, pack1(a));
    });
    QL_LAUNCHED();
    prof_end(s, K_YINV, (nf == 6 ? 11.0 : 6.0) * C(), e0);
  }
  void z3(const T* u_in, T* u_out, hipStream_t s, T* w_out = nullptr) {
    NS3Z<T> a{Y6, NZ, u_out, u_in, qs(6), KZD, KZP, twN_, w_out};
    hipEvent_t e0 = nullptr;
    prof_begin(s, &e0);
    with_zv([&](auto v) {
      constexpr int V = decltype(v)::value;
      hipLaunchKernelGGL((k_ns3_z<T, N, V>), dim3(N * N / Z3Cfg<T, N, V>::P), dim3(Z3Cfg<T, N, V>::THREADS), 0, s,
                         pack1(a));
    });
    QL_LAUNCHED();
    const double R = (double)N * N * N * sizeof(T);
    prof_end(s, K_ZPASS, (u_in ? 3.0 * C() + 3.0 * R : 6.0 * C()) + (u_out ? 3.0 * R : 0.0) +
                             (w_out ? 3.0 * R : 0.0) + 3.0 * N * N * KZD * 2.0 * sizeof(T), e0);
  }
  void yfwd3(hipStream_t s) {
    NS3YF<T> a{NZ, NY, KZD, KZP, KYD, twN_};
    hipEvent_t e0 = nullptr;
    prof_begin(s, &e0);
    with_xv([&](auto v) {
      constexpr int V = decltype(v)::value;
      dim3 g((KZD + XCfg<N, V>::TIX - 1) / XCfg<N, V>::TIX, N);
      hipLaunchKernelGGL((k_ns3_yfwd<T, N, V>), g, dim3(XCfg<N, V>::THREADS), 0, s, pack1(a));
    });
    QL_LAUNCHED();
    prof_end(s, K_YFWD, 3.0 * N * (N + KYD) * KZD * 2.0 * sizeof(T), e0);
  }
  NS3X<T> xargs3(cx<T>* un, int ld_un, size_t q_un) const {
    NS3X<T> a{};
    a.ny = NY;
    a.un = un;
    a.ld_un = ld_un;
    a.q_un = q_un;
    a.acc = S2;
    a.x5 = X5;
    a.ld = LDM;
    a.q_acc = qs(3);
    a.q_x5 = qs(5);
    a.KZD = KZD;
    a.KZP = KZP;
    a.KYD = KYD;
    a.tw = twN_;
    return a;
  }

  // ---------------- 2-D launches ----------------
  void x2(NS2X<T>& a, hipStream_t s) {
    hipEvent_t e0 = nullptr;
    prof_begin(s, &e0);
    with_xv([&](auto v) {
      constexpr int V = decltype(v)::value;
      hipLaunchKernelGGL((k_ns2_x<T, N, V>), dim3((Nh + XCfg<N, V>::TIX - 1) / XCfg<N, V>::TIX),
                         dim3(XCfg<N, V>::THREADS), 0, s, pack1(a));
    });
    QL_LAUNCHED();
    prof_end(s, K_XPASS, xpass_bytes(a.c, a.do_inverse, 2, 3), e0);
  }
  void y2(const T* u_in, T* u_out, hipStream_t s) {
    NS2Y<T> a{X5, NY, u_out, u_in, LD, KZD, KZP, twN_};
    hipEvent_t e0 = nullptr;
    prof_begin(s, &e0);
    with_zv([&](auto v) {
      constexpr int V = decltype(v)::value;
      hipLaunchKernelGGL((k_ns2_y<T, N, V>), dim3((N + ZCfg<N, V>::P - 1) / ZCfg<N, V>::P), dim3(ZCfg<N, V>::THREADS),
                         0, s, pack1(a));
    });
    QL_LAUNCHED();
    const double R = (double)N * N * sizeof(T);
    prof_end(s, K_ZPASS, (u_in ? C() + 2.0 * R : 3.0 * C()) + (u_out ? 2.0 * R : 0.0) + 2.0 * N * KZD * 2.0 * sizeof(T),
             e0);
  }
  NS2X<T> xargs2(cx<T>* S0) const {
    NS2X<T> a{};
    a.ny = NY;
    a.un = S0;
    a.ld_un = Nh;
    a.acc = S2;
    a.x3 = X5;
    a.ld = LD;
    a.KZD = KZD;
    a.KZP = KZP;
    a.tw = twN_;
    return a;
  }

  void reduce_phys(const T* u, int ncomp, double factor, double* out, int nout, hipStream_t s) {
    const size_t npts = DIM == 3 ? (size_t)N * N * N : (size_t)N * N;
    hipLaunchKernelGGL((k_reduce_phys<T>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, u, npts, ncomp, partials_);
    QL_LAUNCHED();
    hipLaunchKernelGGL(k_reduce_phys_final, dim3(1), dim3(RED_THREADS), 0, s, partials_, RED_BLOCKS, (double)npts,
                       dscr_);
    QL_LAUNCHED();
    // dscr_[0] = 0.5 sum/npts
    hipLaunchKernelGGL(k_scale_copy, dim3(1), dim3(32), 0, s, dscr_, out, nout, 2.0 * factor);
    QL_LAUNCHED();
  }

  // out-of-place inverse of a [ncomp][rows][ld_in] rfft-layout spectrum on this grid
  void irfftn_into(const cx<T>* src, int ld_in, cx<T>* tmp, T* dst, int ncomp, hipStream_t s) {
    const T scale = (T)1 / (T)(DIM == 3 ? (double)N * N * N : (double)N * N);
    if (DIM == 3) {
      c2c_strided<T>(N, src, tmp, ncomp, N * ld_in, false, (T)1, twN_, s);
      c2c_strided<T>(N, tmp, tmp, ncomp * N, ld_in, false, (T)1, twN_, s);
      c2r_contig<T>(N, tmp, dst, ncomp * N * N, ld_in, scale, twN_, s);
    } else {
      c2c_strided<T>(N, src, tmp, ncomp, ld_in, false, (T)1, twN_, s);
      c2r_contig<T>(N, tmp, dst, ncomp * N, ld_in, scale, twN_, s);
    }
  }

  // ---------------- tangent-linear / adjoint (3-D) ----------------
  size_t spec_elems() const { return (size_t)3 * rows_ * Nh; }
  size_t phys_elems() const { return (size_t)3 * N * N * (DIM == 3 ? N : 1); }
  void ensure_adj() {
    ensure_step();
    if (AS_[0]) return;
    // one block each, so consecutive scratch arrays are adjacent: the RHS
    // adjoint transforms AP_[3..4] into AS_[7..8] as one 6-component batch
    cx<T>* as = alloc<cx<T>>(NAS * spec_elems());
    for (int i = 0; i < NAS; ++i) AS_[i] = as + i * spec_elems();
    T* ap = alloc<T>(NAP * phys_elems());
    for (int i = 0; i < NAP; ++i) AP_[i] = ap + i * phys_elems();
    for (int i = 0; i < 8; ++i) SF_[i] = alloc<T>(phys_elems());
  }
  void check_adj() const {
    if (DIM != 3) throw std::invalid_argument("tangent / adjoint maps are implemented for 3-D grids");
    if (d_.forcing != QLLES_FORCING_NONE && d_.forcing != QLLES_FORCING_POWER)
      throw std::invalid_argument("tangent / adjoint maps need forcing none or power");
    if (d_.viscous != QLLES_VISC_EXPLICIT) throw std::invalid_argument("tangent / adjoint maps need the explicit viscous term");
  }
  static int nblk(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }
  void to_phys(const cx<T>* src, T* dst, hipStream_t s) { irfftn_into(src, Nh, AS_[9], dst, 3, s); }
  void to_spec(const T* src, cx<T>* dst, hipStream_t s, int ncomp = 3) {
    r2c_contig<T>(N, src, dst, ncomp * N * N, Nh, (T)1, twN_, s);
    c2c_strided<T>(N, dst, dst, ncomp * N, Nh, true, (T)1, twN_, s);
    c2c_strided<T>(N, dst, dst, ncomp, N * Nh, true, (T)1, twN_, s);
  }
  void curl(const cx<T>* src, cx<T>* dst, hipStream_t s) {
    hipLaunchKernelGGL((k_curl_hat<T>), dim3(nblk(rows_ * Nh)), dim3(256), 0, s, src, dst, N, 3, Nh, Nh);
    QL_LAUNCHED();
  }
  void axpby(const cx<T>* x, const cx<T>* y, cx<T>* out, T a, T b, hipStream_t s) {
    hipLaunchKernelGGL((k_axpby<T>), dim3(nblk(spec_elems())), dim3(256), 0, s, x, y, out, a, b, spec_elems());
    QL_LAUNCHED();
  }
  void axpby2(const cx<T>* x1, const cx<T>* x2, const cx<T>* y, cx<T>* out1, cx<T>* out2, T a1, T b1, T a2, T b2,
              hipStream_t s) {
    hipLaunchKernelGGL((k_axpby2<T>), dim3(nblk(spec_elems())), dim3(256), 0, s, x1, x2, y, out1, out2, a1, b1, a2,
                       b2, spec_elems());
    QL_LAUNCHED();
  }
  void cross2(const T* a, const T* b, const T* c, const T* d, T* out, hipStream_t s) {
    const size_t n = (size_t)N * N * N;
    hipLaunchKernelGGL((k_cross2<T>), dim3(nblk(n)), dim3(256), 0, s, a, b, c, d, out, n);
    QL_LAUNCHED();
  }
  // u = irfftn(y) -> up, w = irfftn(K y) -> wp
  void state_fields(const cx<T>* y, T* up, T* wp, hipStream_t s) {
    to_phys(y, up, s);
    curl(y, AS_[8], s);
    to_phys(AS_[8], wp, s);
  }
  // out = R'(y) d  (temporaries AS_[7..9], AP_[0..4]); up / wp: u and omega
  // of y when the caller has them, else computed here
  void rhs_jvp_at(const cx<T>* y, const cx<T>* d, cx<T>* out, hipStream_t s, const T* up = nullptr,
                  const T* wp = nullptr) {
    if (up == nullptr) {
      state_fields(y, AP_[0], AP_[1], s);
      up = AP_[0];
      wp = AP_[1];
    }
    to_phys(d, AP_[2], s);
    curl(d, AS_[8], s);
    to_phys(AS_[8], AP_[3], s);
    cross2(AP_[2], wp, up, AP_[3], AP_[4], s);  // du x w + u x dw
    to_spec(AP_[4], AS_[7], s);
    hipLaunchKernelGGL((k_jvp_combine<T>), dim3(nblk(rows_ * Nh)), dim3(256), 0, s, AS_[7], d, out, N, KZD,
                       (T)d_.nu);
    QL_LAUNCHED();
    if (d_.forcing == QLLES_FORCING_POWER) {
      hipLaunchKernelGGL((k_force_tangent<T>), dim3(1), dim3(256), 0, s, y, d, out, N, d_.kf * d_.kf, d_.power, 0,
                         nullptr, Nh);
      QL_LAUNCHED();
    }
  }
  // out = R'(y)^T l  (temporaries AS_[7..9], AP_[0..4]); with fixed-power
  // forcing, *dpow (device, optional) accumulates <l, dR/dP>
  // y has rows of ldy entries (Nh; N/2 for the split layout of a step record)
  void rhs_vjp_at(const cx<T>* y, const cx<T>* l, cx<T>* out, hipStream_t s, double* dpow = nullptr,
                  const T* up = nullptr, const T* wp = nullptr, int ldy = Nh) {
    if (up == nullptr) {
      state_fields(y, AP_[0], AP_[1], s);
      up = AP_[0];
      wp = AP_[1];
    }
    // psi = irfftn(P D l / c): P D / c fused with the x transform (k_adj_pd_x),
    // then y; along z, fused with both products (w x psi, psi x u) and their
    // forward transforms along z (k_adj_z)
    hipLaunchKernelGGL((k_adj_pd_x<T, N>), dim3((N * Nh + TI - 1) / TI), dim3(StridedCfg<T, N>::THREADS), 0, s, l,
                       AS_[9], KZD, twN_);
    QL_LAUNCHED();
    c2c_strided<T>(N, AS_[9], AS_[9], 3 * N, Nh, false, (T)1, twN_, s);
    {
      using ZC = AdjZCfg<N>;
      static_assert((N * N) % ZC::P == 0, "pencils per workgroup must divide N^2");
      const T h = (T)(0.5 / ((double)N * N * N));  // r2c's 1/2 times the inverse's 1/N^3
      hipLaunchKernelGGL((k_adj_z<T, N>), dim3(N * N / ZC::P), dim3(ZC::THREADS), 0, s, AS_[9], wp, up, AS_[7],
                         twN_, h);
      QL_LAUNCHED();
    }
    // the six products' y forward transforms (AS_[8] follows AS_[7]), then
    // their x transforms fused with the combination (k_adj_x_combine)
    c2c_strided<T>(N, AS_[7], AS_[7], 6 * N, Nh, true, (T)1, twN_, s);
    hipLaunchKernelGGL((k_adj_x_combine<T, N>), dim3((N * Nh + TI - 1) / TI), dim3(StridedCfg<T, N>::THREADS), 0, s,
                       AS_[7], l, out, KZD, (T)d_.nu, twN_);
    QL_LAUNCHED();
    if (d_.forcing == QLLES_FORCING_POWER) {
      hipLaunchKernelGGL((k_force_tangent<T>), dim3(1), dim3(256), 0, s, y, l, out, N, d_.kf * d_.kf, d_.power, 1,
                         dpow, ldy);
      QL_LAUNCHED();
    }
  }
  // fused RHS of y that also keeps u = irfftn(y) and omega = irfftn(K y), which
  // its z-pass forms anyway (3-D)
  void rhs_fields(const cx<T>* y, cx<T>* du, T* up, T* wp, hipStream_t s) {
    cx<T>* uh = const_cast<cx<T>*>(y);
    NS3X<T> a = xargs3(uh, Nh, 0);
    a.c = ctl(-1);
    a.us = uh;
    a.ld_us = Nh;
    a.do_inverse = 1;
    x3(a, s);
    yinv3(6, s);
    z3(nullptr, up, s, wp);
    yfwd3(s);
    a = xargs3(uh, Nh, 0);
    a.c = ctl(4);
    a.us = uh;
    a.ld_us = Nh;
    a.uout = du;
    a.ld_uout = Nh;
    a.do_inverse = 0;
    x3(a, s);
  }
  // RK4 stage states y1, y2, y3 of the step from y0 (spectral_solver.py:74-80)
  // -> AS_[0..2], and u, omega of y0..y3 -> SF_[0..7]
  void stage_states(const cx<T>* y0, hipStream_t s) {
    const T h = (T)d_.dt;
    rhs_fields(y0, AS_[3], SF_[0], SF_[1], s);
    axpby(y0, AS_[3], AS_[0], (T)1, h / 2, s);
    rhs_fields(AS_[0], AS_[3], SF_[2], SF_[3], s);
    axpby(y0, AS_[3], AS_[1], (T)1, h / 2, s);
    rhs_fields(AS_[1], AS_[3], SF_[4], SF_[5], s);
    axpby(y0, AS_[3], AS_[2], (T)1, h, s);
    state_fields(AS_[2], SF_[6], SF_[7], s);
  }

  // The plan allocates its large buffers on first use: Y6 (transform
  // scratch) for the inverse transform, filter and spectrum, X5 as well for
  // the dissipation, and the whole step workspace for the first step, RHS,
  // tangent or adjoint.  A plan used only for rfftn, diagnostics or energies
  // holds the reduction partials alone.
  void ensure_y6() {
    if (!Y6) Y6 = alloc<cx<T>>((DIM == 3 ? 6 : 3) * fsz_);
  }
  void ensure_x5() {
    if (!X5) X5 = alloc<cx<T>>((DIM == 3 ? 5 : 3) * fsz_);
  }
  void ensure_step() {
    if (step_ready_) return;
    ensure_x5();
    ensure_y6();
    const int nc = DIM == 3 ? 3 : 2;
    if (DIM == 3) {
      P = alloc<cx<T>>(3 * fsz_);
      NZ = alloc<cx<T>>((size_t)3 * N * N * KZP);
      NY = alloc<cx<T>>((size_t)3 * N * KYD * KZP);
    } else {
      NY = alloc<cx<T>>((size_t)2 * N * KZP);
    }
    S1a = alloc<cx<T>>(nc * fsz_);
    S1b = alloc<cx<T>>(nc * fsz_);
    S2 = alloc<cx<T>>(nc * fsz_);
    step_ready_ = true;
  }

 public:
  explicit EngineT(const Desc& d) : d_(d) {
    KZD = dealias_count(N);
    KZP = round_up(KZD, 16);
    KYD = 2 * KZD - 1;
    LD = round_up(Nh, 16);
    LDM = N / 2;
    rows_ = DIM == 3 ? (size_t)N * N : (size_t)N;
    fsz_ = rows_ * LD;
    if (DIM == 3)
      nxblk_ = std::max(xtiles_v<0>(), xtiles_v<HAS_V ? 1 : 0>()) * N;  // partials capacity over variants
    else
      nxblk_ = (Nh + XCfg<N, HAS_V ? 1 : 0>::TIX - 1) / XCfg<N, HAS_V ? 1 : 0>::TIX;
    partials_ = alloc<double>(std::max(RED_BLOCKS * 4, nxblk_ * 2));
    alpha_ = alloc<double>(8);
    dscr_ = alloc<double>(8);
    twN_ = twiddles<T>(N);
  }
  ~EngineT() override {
    prof_drain();
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
    for (void* p : allocs_) (void)hipFree(p);
  }
  size_t workspace_bytes() const override { return bytes_; }
  void prepare() override { ensure_step(); }

  void step(void* u_out_v, void* uhat, const void* u_in_v, int nsteps, double* e_inj, hipStream_t s) override {
    ensure_step();
    cx<T>* S0 = static_cast<cx<T>*>(uhat);
    T* u_out = static_cast<T*>(u_out_v);
    const T* u_in = static_cast<const T*>(u_in_v);
    const bool ekin = d_.forcing == QLLES_FORCING_EKIN_TOT;
    if (DIM == 3) {
      NS3X<T> a = xargs3(S0, Nh, 0);
      a.c = ctl(-1);
      a.us = S0;
      a.ld_us = Nh;
      a.copy_dst = P;
      a.ld_copy = LDM;
      a.q_copy = qs(3);
      a.do_inverse = 1;
      x3(a, s);
      for (int n = 0; n < nsteps; ++n) {
        const bool more = (n < nsteps - 1) || (u_out != nullptr);
        // u_{n+1} lands in P, except on the call's last step (caller's array)
        const bool last = n == nsteps - 1;
        cx<T>* nxt = last ? S0 : P;
        const int ld_nxt = last ? Nh : LDM;
        const size_t q_nxt = last ? 0 : qs(3);
        for (int st = 0; st < 4; ++st) {
          yinv3(6, s);
          z3((n == 0 && st == 0) ? u_in : nullptr, nullptr, s);
          yfwd3(s);
          a = xargs3(P, LDM, qs(3));
          a.c = ctl(st);
          a.us = st == 0 ? P : (st == 2 ? S1b : S1a);
          a.ld_us = LDM;
          a.q_us = qs(3);
          a.uout = st == 0 ? S1a : (st == 1 ? S1b : (st == 2 ? S1a : nxt));
          a.ld_uout = st == 3 ? ld_nxt : LDM;
          a.q_uout = st == 3 ? q_nxt : qs(3);
          a.do_inverse = (st < 3) || (!ekin && more);
          if (st == 3 && ekin) a.c.partials = partials_;
          x3(a, s);
        }
        if (ekin) {
          const double n6 = (double)N * N * N * (double)N * N * N;
          hipLaunchKernelGGL(k_ekin_alpha, dim3(1), dim3(RED_THREADS), 0, s, partials_, xblocks(), 1.0 / n6,
                             d_.e_kin_init, d_.dt, alpha_, e_inj ? e_inj + n : nullptr);
          QL_LAUNCHED();
          a = xargs3(nxt, ld_nxt, q_nxt);
          a.c = ctl(-1);
          a.c.alpha = alpha_;
          a.rescale_dst = nxt;
          a.us = nxt;
          a.ld_us = ld_nxt;
          a.q_us = q_nxt;
          a.do_inverse = more ? 1 : 0;
          x3(a, s);
        } else if (e_inj) {
          hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(32), 0, s, e_inj + n, 1, 0.0);
          QL_LAUNCHED();
        }
        if (after_step) after_step(s, nxt, ld_nxt, q_nxt);
      }
      if (u_out) {
        yinv3(3, s);
        c2r_contig<T>(N, Y6, u_out, 3 * N * N, LD, (T)1 / ((T)N * (T)N * (T)N), twN_, s);
      }
    } else {
      if (ekin) throw std::invalid_argument("ekin_tot forcing is defined for 3-D HIT only");
      NS2X<T> a = xargs2(S0);
      a.c = ctl(-1);
      a.us = S0;
      a.ld_us = Nh;
      a.do_inverse = 1;
      x2(a, s);
      for (int n = 0; n < nsteps; ++n) {
        const bool more = (n < nsteps - 1) || (u_out != nullptr);
        for (int st = 0; st < 4; ++st) {
          y2((n == 0 && st == 0) ? u_in : nullptr, nullptr, s);
          a = xargs2(S0);
          a.c = ctl(st);
          a.us = st == 0 ? S0 : (st == 2 ? S1b : S1a);
          a.ld_us = st == 0 ? Nh : LD;
          a.uout = st == 0 ? S1a : (st == 1 ? S1b : (st == 2 ? S1a : S0));
          a.ld_uout = st == 3 ? Nh : LD;
          a.do_inverse = (st < 3) || more;
          x2(a, s);
        }
        if (e_inj) {
          hipLaunchKernelGGL(k_fill_double, dim3(1), dim3(32), 0, s, e_inj + n, 1, 0.0);
          QL_LAUNCHED();
        }
        if (after_step) after_step(s, S0, Nh, 0);
      }
      if (u_out) {
        c2r_contig<T>(N, X5, u_out, 2 * N, LD, (T)1 / ((T)N * (T)N), twN_, s);
        hipLaunchKernelGGL((k_fill<T>), dim3(64), dim3(256), 0, s, u_out + (size_t)2 * N * N, (size_t)N * N, (T)0);
        QL_LAUNCHED();
      }
    }
  }

  void rhs(const void* u_v, const void* uhat_v, int rk, void* du_v, hipStream_t s) override {
    ensure_step();
    cx<T>* uh = const_cast<cx<T>*>(static_cast<const cx<T>*>(uhat_v));
    cx<T>* du = static_cast<cx<T>*>(du_v);
    const T* u = (rk == 0) ? static_cast<const T*>(u_v) : nullptr;
    if (DIM == 3) {
      NS3X<T> a = xargs3(uh, Nh, 0);
      a.c = ctl(-1);
      a.us = uh;
      a.ld_us = Nh;
      a.do_inverse = 1;
      x3(a, s);
      yinv3(6, s);
      z3(u, nullptr, s);
      yfwd3(s);
      a = xargs3(uh, Nh, 0);
      a.c = ctl(4);
      a.us = uh;
      a.ld_us = Nh;
      a.uout = du;
      a.ld_uout = Nh;
      a.do_inverse = 0;
      x3(a, s);
    } else {
      NS2X<T> a = xargs2(uh);
      a.c = ctl(-1);
      a.us = uh;
      a.ld_us = Nh;
      a.do_inverse = 1;
      x2(a, s);
      y2(u, nullptr, s);
      a = xargs2(uh);
      a.c = ctl(4);
      a.us = uh;
      a.ld_us = Nh;
      a.uout = du;
      a.ld_uout = Nh;
      a.do_inverse = 0;
      x2(a, s);
      hipLaunchKernelGGL((k_fill<T>), dim3(16), dim3(256), 0, s, reinterpret_cast<T*>(du + (size_t)2 * N * Nh),
                         (size_t)2 * N * Nh, (T)0);
      QL_LAUNCHED();
    }
  }

  void rfftn(const void* u_v, void* uhat_v, hipStream_t s) override {
    const T* u = static_cast<const T*>(u_v);
    cx<T>* uh = static_cast<cx<T>*>(uhat_v);
    if (DIM == 3) {
      r2c_contig<T>(N, u, uh, 3 * N * N, Nh, (T)1, twN_, s);
      c2c_strided<T>(N, uh, uh, 3 * N, Nh, true, (T)1, twN_, s);
      c2c_strided<T>(N, uh, uh, 3, N * Nh, true, (T)1, twN_, s);
    } else {
      r2c_contig<T>(N, u, uh, 3 * N, Nh, (T)1, twN_, s);
      c2c_strided<T>(N, uh, uh, 3, Nh, true, (T)1, twN_, s);
    }
  }

  void irfftn(const void* uhat_v, void* u_v, hipStream_t s) override {
    ensure_y6();
    irfftn_into(static_cast<const cx<T>*>(uhat_v), Nh, Y6, static_cast<T*>(u_v), 3, s);
  }

  void diagnostics(const void* u_v, double* out, hipStream_t s) override {
    reduce_phys(static_cast<const T*>(u_v), 3, 0.5, out, 5, s);
  }

  void dissipation(const void* uhat_v, double* out, hipStream_t s) override {
    ensure_y6();
    ensure_x5();
    const cx<T>* uh = static_cast<const cx<T>*>(uhat_v);
    const size_t total = rows_ * Nh;
    const int nb = (int)std::min<size_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL((k_curl_hat<T>), dim3(nb), dim3(256), 0, s, uh, Y6, N, DIM, Nh, Nh);
    QL_LAUNCHED();
    T* om = reinterpret_cast<T*>(X5);
    irfftn_into(Y6, Nh, Y6, om, 3, s);
    reduce_phys(om, 3, d_.nu, out, 1, s);
  }

  void energy_hat(const void* uhat_v, double* out, hipStream_t s) override {
    const double nd = DIM == 3 ? (double)N * N * N : (double)N * N;
    hipLaunchKernelGGL((k_energy_hat<T>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s,
                       static_cast<const cx<T>*>(uhat_v), N, DIM, Nh, d_.kf * d_.kf, partials_);
    QL_LAUNCHED();
    hipLaunchKernelGGL(k_energy_hat_final, dim3(1), dim3(RED_THREADS), 0, s, partials_, RED_BLOCKS, 1.0 / (nd * nd),
                       out);
    QL_LAUNCHED();
  }

  void spectrum(const void* u_v, double mul_fac, double* ek, hipStream_t s) override {
    ensure_y6();
    cx<T>* vh = Y6;  // [3][rows][Nh] scratch
    rfftn(u_v, vh, s);
    constexpr int NBLK = 256;
    if (!spec_part_) spec_part_ = alloc<double>((size_t)NBLK * 2 * N);
    const double nd = DIM == 3 ? (double)N * N * N : (double)N * N;
    hipLaunchKernelGGL((k_shell_bin<T>), dim3(NBLK), dim3(256), 2 * N * sizeof(double), s, vh, N, DIM,
                       DIM == 3 ? 3 : 2, 1.0 / (nd * nd), mul_fac, N, spec_part_);
    QL_LAUNCHED();
    hipLaunchKernelGGL(k_shell_final, dim3((N + 255) / 256), dim3(256), 0, s, spec_part_, NBLK, N, ek);
    QL_LAUNCHED();
  }

  int variants() const override { return HAS_V ? 2 : 1; }
  void set_variant(int xv, int zv) override {
    if (xv < 0 || zv < 0 || xv >= variants() || zv >= variants())
      throw std::invalid_argument("kernel variant out of range for this grid");
    xv_ = xv;
    zv_ = zv;
  }
  void get_variant(int* xv, int* zv) const override {
    if (xv) *xv = xv_;
    if (zv) *zv = zv_;
  }

  // ---------------- groups (multi-member launches) ----------------
  // launches for members [0, m) of a group, Pack<A>::K members per launch;
  // entry(i) gives member i's arguments
  template <class A, class E, class L>
  static void packs(int m, E&& entry, L&& launch) {
    for (int lo = 0; lo < m; lo += Pack<A>::K) {
      Pack<A> pk{};
      const int c = std::min(Pack<A>::K, m - lo);
      for (int i = 0; i < c; ++i) pk.m[i] = entry(lo + i);
      launch(pk, c);
    }
  }

  // The group holds, per member, the arguments step() passes each kernel:
  // x-pass arguments for the state load (stage -1), the four stages with
  // u_{n+1} written to the member's P (3-D) and stage 3 of the member's last
  // step, which writes u_{n+1} to the caller's array and skips the inverse.
  struct Group final : EngineGroup {
    std::vector<EngineT*> e;
    std::vector<cx<T>*> s0;
    std::vector<NS3X<T>> x3[6];  // load, stages 0..3, stage 3 of the last step
    std::vector<NS3Y<T>> yi;
    std::vector<NS3Z<T>> zz;
    std::vector<NS3YF<T>> yf;
    std::vector<NS2X<T>> x2[6];
    std::vector<NS2Y<T>> y2;

    int size() const override { return (int)e.size(); }
    SpecRef spec_cont(int i) const override {
      if (DIM == 3) return SpecRef{e[i]->P, e[i]->LDM, e[i]->qs(3)};
      return SpecRef{s0[i], Nh, 0};
    }
    SpecRef spec_last(int i) const override { return SpecRef{s0[i], Nh, 0}; }

    // one pass for members [0, m); those from f on take their last step
    void pass_x(int t, int m, int f, hipStream_t s) const {
      const EngineT& E = *e[0];
      if (DIM == 3) {
        packs<NS3X<T>>(m, [&](int i) -> const NS3X<T>& { return (t == 4 && i >= f) ? x3[5][i] : x3[t][i]; },
                       [&](const Pack<NS3X<T>>& pk, int c) {
                         E.with_xv([&](auto v) { launch_x3<decltype(v)::value>(pk, dim3(E.xtiles(), N, c), s); });
                         QL_LAUNCHED();
                       });
      } else {
        packs<NS2X<T>>(m, [&](int i) -> const NS2X<T>& { return (t == 4 && i >= f) ? x2[5][i] : x2[t][i]; },
                       [&](const Pack<NS2X<T>>& pk, int c) {
                         E.with_xv([&](auto v) {
                           constexpr int V = decltype(v)::value;
                           hipLaunchKernelGGL((k_ns2_x<T, N, V>), dim3((Nh + XCfg<N, V>::TIX - 1) / XCfg<N, V>::TIX, 1, c),
                                              dim3(XCfg<N, V>::THREADS), 0, s, pk);
                         });
                         QL_LAUNCHED();
                       });
      }
    }
    void pass_yz(int m, hipStream_t s) const {
      const EngineT& E = *e[0];
      if (DIM == 3) {
        packs<NS3Y<T>>(m, [&](int i) -> const NS3Y<T>& { return yi[i]; }, [&](const Pack<NS3Y<T>>& pk, int c) {
          E.with_xv([&](auto v) {
            constexpr int V = decltype(v)::value;
            hipLaunchKernelGGL((k_ns3_yinv<T, N, V>), dim3(E.xtiles(), N, c), dim3(XCfg<N, V>::THREADS), 0, s, pk);
          });
          QL_LAUNCHED();
        });
        packs<NS3Z<T>>(m, [&](int i) -> const NS3Z<T>& { return zz[i]; }, [&](const Pack<NS3Z<T>>& pk, int c) {
          E.with_zv([&](auto v) {
            constexpr int V = decltype(v)::value;
            hipLaunchKernelGGL((k_ns3_z<T, N, V>), dim3(N * N / Z3Cfg<T, N, V>::P, 1, c),
                               dim3(Z3Cfg<T, N, V>::THREADS), 0, s, pk);
          });
          QL_LAUNCHED();
        });
        packs<NS3YF<T>>(m, [&](int i) -> const NS3YF<T>& { return yf[i]; }, [&](const Pack<NS3YF<T>>& pk, int c) {
          E.with_xv([&](auto v) {
            constexpr int V = decltype(v)::value;
            hipLaunchKernelGGL((k_ns3_yfwd<T, N, V>), dim3((E.KZD + XCfg<N, V>::TIX - 1) / XCfg<N, V>::TIX, N, c),
                               dim3(XCfg<N, V>::THREADS), 0, s, pk);
          });
          QL_LAUNCHED();
        });
      } else {
        packs<NS2Y<T>>(m, [&](int i) -> const NS2Y<T>& { return y2[i]; }, [&](const Pack<NS2Y<T>>& pk, int c) {
          E.with_zv([&](auto v) {
            constexpr int V = decltype(v)::value;
            hipLaunchKernelGGL((k_ns2_y<T, N, V>), dim3((N + ZCfg<N, V>::P - 1) / ZCfg<N, V>::P, 1, c),
                               dim3(ZCfg<N, V>::THREADS), 0, s, pk);
          });
          QL_LAUNCHED();
        });
      }
    }

    void steps(const int* n, int from, int to, hipStream_t s,
               const std::function<void(hipStream_t, int, int)>& after) override {
      const int M = size();
      for (int i = 1; i < M; ++i)
        if (n[i] > n[i - 1]) throw std::invalid_argument("group step counts must not increase along the members");
      auto active = [&](int k) {
        int c = 0;
        while (c < M && n[c] > k) ++c;
        return c;
      };
      if (from == 0) {
        const int m0 = active(0);
        if (m0 > 0) pass_x(0, m0, m0, s);
      }
      for (int k = from; k < to; ++k) {
        const int m = active(k), f = active(k + 1);
        if (m == 0) break;
        for (int st = 0; st < 4; ++st) {
          pass_yz(m, s);
          pass_x(1 + st, m, f, s);
        }
        if (after) after(s, m, f);
      }
    }
  };

  bool groupable_with(const Engine* o) const override {
    const auto* x = dynamic_cast<const EngineT*>(o);
    if (x == nullptr) return false;
    if (d_.forcing == QLLES_FORCING_EKIN_TOT || x->d_.forcing == QLLES_FORCING_EKIN_TOT) return false;
    return x->d_.viscous == d_.viscous && x->xv_ == xv_ && x->zv_ == zv_;
  }

  EngineGroup* make_group(const std::vector<Engine*>& members, const std::vector<void*>& uhat) override {
    if (members.empty() || members[0] != this || members.size() != uhat.size())
      throw std::invalid_argument("a group lists this engine first and one state per member");
    for (const Engine* m : members)
      if (!groupable_with(m)) throw std::invalid_argument("engines of different kinds cannot share a group");
    const int M = (int)members.size();
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < i; ++j)
        if (members[i] == members[j]) throw std::invalid_argument("each group member needs its own plan");
    std::unique_ptr<Group> g(new Group());
    for (int i = 0; i < M; ++i) {
      auto* x = static_cast<EngineT*>(members[i]);
      x->ensure_step();
      g->e.push_back(x);
      g->s0.push_back(static_cast<cx<T>*>(uhat[i]));
    }
    for (int i = 0; i < M; ++i) {
      const EngineT& X = *g->e[i];
      cx<T>* S0 = g->s0[i];
      if (DIM == 3) {
        NS3X<T> a = X.xargs3(S0, Nh, 0);
        a.c = X.ctl(-1);
        a.us = S0;
        a.ld_us = Nh;
        a.copy_dst = X.P;
        a.ld_copy = X.LDM;
        a.q_copy = X.qs(3);
        a.do_inverse = 1;
        g->x3[0].push_back(a);
        for (int st = 0; st < 4; ++st) {
          a = X.xargs3(X.P, X.LDM, X.qs(3));
          a.c = X.ctl(st);
          a.us = st == 0 ? X.P : (st == 2 ? X.S1b : X.S1a);
          a.ld_us = X.LDM;
          a.q_us = X.qs(3);
          a.uout = st == 0 ? X.S1a : (st == 1 ? X.S1b : (st == 2 ? X.S1a : X.P));
          a.ld_uout = X.LDM;
          a.q_uout = X.qs(3);
          a.do_inverse = 1;
          g->x3[1 + st].push_back(a);
          if (st == 3) {
            a.uout = S0;
            a.ld_uout = Nh;
            a.q_uout = 0;
            a.do_inverse = 0;
            g->x3[5].push_back(a);
          }
        }
        g->yi.push_back(NS3Y<T>{X.X5, X.Y6, X.LDM, X.LDM, X.qs(5), X.qs(6), 6, X.twN_});
        g->zz.push_back(NS3Z<T>{X.Y6, X.NZ, nullptr, nullptr, X.qs(6), X.KZD, X.KZP, X.twN_, nullptr});
        g->yf.push_back(NS3YF<T>{X.NZ, X.NY, X.KZD, X.KZP, X.KYD, X.twN_});
      } else {
        NS2X<T> a = X.xargs2(S0);
        a.c = X.ctl(-1);
        a.us = S0;
        a.ld_us = Nh;
        a.do_inverse = 1;
        g->x2[0].push_back(a);
        for (int st = 0; st < 4; ++st) {
          a = X.xargs2(S0);
          a.c = X.ctl(st);
          a.us = st == 0 ? S0 : (st == 2 ? X.S1b : X.S1a);
          a.ld_us = st == 0 ? Nh : X.LD;
          a.uout = st == 0 ? X.S1a : (st == 1 ? X.S1b : (st == 2 ? X.S1a : S0));
          a.ld_uout = st == 3 ? Nh : X.LD;
          a.do_inverse = 1;
          g->x2[1 + st].push_back(a);
          if (st == 3) {
            a.do_inverse = 0;
            g->x2[5].push_back(a);
          }
        }
        g->y2.push_back(NS2Y<T>{X.X5, X.NY, nullptr, nullptr, X.LD, X.KZD, X.KZP, X.twN_});
      }
    }
    return g.release();
  }

  void rhs_jvp(const void* uhat, const void* v, void* out, hipStream_t s) override {
    check_adj();
    ensure_adj();
    rhs_jvp_at(static_cast<const cx<T>*>(uhat), static_cast<const cx<T>*>(v), static_cast<cx<T>*>(out), s);
  }
  void rhs_vjp(const void* uhat, const void* lam, void* out, double* dpow, hipStream_t s) override {
    check_adj();
    ensure_adj();
    rhs_vjp_at(static_cast<const cx<T>*>(uhat), static_cast<const cx<T>*>(lam), static_cast<cx<T>*>(out), s, dpow);
  }
  // tangent of one classical RK4 step: dy' = v + h/6 (dk1 + 2 dk2 + 2 dk3 + dk4)
  void step_jvp(const void* uhat, const void* v_in, void* out_v, hipStream_t s) override {
    check_adj();
    ensure_adj();
    const T h = (T)d_.dt;
    const cx<T>* y0 = static_cast<const cx<T>*>(uhat);
    cx<T>* out = static_cast<cx<T>*>(out_v);
    cx<T>* v = AS_[6];
    QL_HIP(hipMemcpyAsync(v, v_in, spec_elems() * sizeof(cx<T>), hipMemcpyDeviceToDevice, s));
    stage_states(y0, s);
    cx<T>*dk = AS_[3], *dy = AS_[4];
    const cx<T>* ys[4] = {y0, AS_[0], AS_[1], AS_[2]};
    const T wk[4] = {h / 6, h / 3, h / 3, h / 6};
    const T wy[3] = {h / 2, h / 2, h};
    for (int st = 0; st < 4; ++st) {
      rhs_jvp_at(ys[st], st == 0 ? v : dy, dk, s, SF_[2 * st], SF_[2 * st + 1]);
      axpby(st == 0 ? v : out, dk, out, (T)1, wk[st], s);
      if (st < 3) axpby(v, dk, dy, (T)1, wy[st], s);
    }
  }
  // adjoint of one classical RK4 step, backward through the stages
  void step_vjp(const void* uhat, const void* lam_in, void* out_v, double* dpow, hipStream_t s) override {
    check_adj();
    ensure_adj();
    const cx<T>* y0 = static_cast<const cx<T>*>(uhat);
    stage_states(y0, s);
    const cx<T>* ys[4] = {y0, AS_[0], AS_[1], AS_[2]};
    const T* fs[8] = {SF_[0], SF_[1], SF_[2], SF_[3], SF_[4], SF_[5], SF_[6], SF_[7]};
    vjp_stages(ys, fs, static_cast<const cx<T>*>(lam_in), static_cast<cx<T>*>(out_v), dpow, s, Nh);
  }
  // a step record: the stage states y1, y2, y3 (spectra in the split layout of
  // the step's internal fields: [3][N][N][N/2] rows, then the [3][N][N] plane
  // of kz = N/2, 3 N^2 (N/2+1) entries each), then u and omega of y0..y3
  // (physical), in the plan dtype
  size_t record_bytes() const override {
    return DIM == 3 ? 3 * spec_elems() * sizeof(cx<T>) + 8 * phys_elems() * sizeof(T) : 0;
  }
  void record_views(const void* rec, const cx<T>* (&ys)[3], const T* (&fs)[8]) const {
    const cx<T>* sp = static_cast<const cx<T>*>(rec);
    for (int i = 0; i < 3; ++i) ys[i] = sp + i * spec_elems();
    const T* ph = reinterpret_cast<const T*>(sp + 3 * spec_elems());
    for (int i = 0; i < 8; ++i) fs[i] = ph + i * phys_elems();
  }
  // One fused RK4 step (the kernels and arithmetic of step()) that also
  // keeps what the step adjoint reads: the stage states y1, y2, y3 go to the
  // record instead of the internal stage buffers, and each stage's z-pass
  // writes the physical u and omega of its state.  uhat_out may be uhat: the
  // state is copied into P by the first launch and read from there.
  void step_record(const void* uhat, void* uhat_out, void* rec, hipStream_t s) override {
    check_adj();
    ensure_adj();
    cx<T>* y0 = const_cast<cx<T>*>(static_cast<const cx<T>*>(uhat));
    const cx<T>* ys[3];
    const T* fs[8];
    record_views(rec, ys, fs);
    NS3X<T> a = xargs3(y0, Nh, 0);
    a.c = ctl(-1);
    a.us = y0;
    a.ld_us = Nh;
    a.copy_dst = P;
    a.ld_copy = LDM;
    a.q_copy = qs(3);
    a.do_inverse = 1;
    x3(a, s);
    for (int st = 0; st < 4; ++st) {
      yinv3(6, s);
      z3(nullptr, const_cast<T*>(fs[2 * st]), s, const_cast<T*>(fs[2 * st + 1]));
      yfwd3(s);
      a = xargs3(P, LDM, qs(3));
      a.c = ctl(st);
      a.us = st == 0 ? P : const_cast<cx<T>*>(ys[st - 1]);
      a.ld_us = LDM;
      a.q_us = qs(3);
      a.uout = st < 3 ? const_cast<cx<T>*>(ys[st]) : static_cast<cx<T>*>(uhat_out);
      a.ld_uout = st < 3 ? LDM : Nh;
      a.q_uout = st < 3 ? qs(3) : 0;
      a.do_inverse = st < 3;
      x3(a, s);
    }
  }
  void step_vjp_record(const void* uhat, const void* rec, const void* lam_in, void* out_v, double* dpow,
                       hipStream_t s) override {
    check_adj();
    ensure_adj();
    const cx<T>* yr[3];
    const T* fs[8];
    record_views(rec, yr, fs);
    const cx<T>* ys[4] = {static_cast<const cx<T>*>(uhat), yr[0], yr[1], yr[2]};
    vjp_stages(ys, fs, static_cast<const cx<T>*>(lam_in), static_cast<cx<T>*>(out_v), dpow, s, LDM);
  }
  // backward through the four stages at the stage states ys with their u, omega fs
  // ys[1..3] have rows of ld_st entries (ys[0] is the caller's state, rows of Nh)
  void vjp_stages(const cx<T>* const (&ys)[4], const T* const (&fs)[8], const cx<T>* lam_in, cx<T>* out,
                  double* dpow, hipStream_t s, int ld_st) {
    const T h = (T)d_.dt;
    // lam_in is only read, and read for the last time before out is written,
    // so out may alias it
    const cx<T>* lam = lam_in;
    cx<T>*kb = AS_[3], *yb = AS_[4], *acc = AS_[5];
    // kbar_4 = h/6 l; kbar_3 = h/3 l + h ybar_3; kbar_2 = h/3 l + h/2 ybar_2; kbar_1 = h/6 l + h/2 ybar_1
    const T wl[4] = {h / 6, h / 3, h / 3, h / 6};  // for stages 3, 2, 1, 0
    const T wb[4] = {0, h, h / 2, h / 2};
    // acc = lam + ybar_3 + ybar_2 + ... in that order; each ybar is read once,
    // for the accumulator and the next stage's kbar together
    axpby(lam, nullptr, kb, wl[0], wb[0], s);
    for (int q = 0; q < 4; ++q) {
      const int st = 3 - q;
      rhs_vjp_at(ys[st], kb, yb, s, dpow, fs[2 * st], fs[2 * st + 1], st == 0 ? Nh : ld_st);
      const cx<T>* acc_in = q == 0 ? lam : acc;
      if (st == 0) axpby(acc_in, yb, out, (T)1, (T)1, s);
      else axpby2(lam, acc_in, yb, kb, acc, wl[q + 1], wb[q + 1], (T)1, (T)1, s);
    }
  }
  void irfftn_adjoint(const void* g, void* out_v, hipStream_t s) override {
    cx<T>* out = static_cast<cx<T>*>(out_v);
    rfftn(g, out, s);
    const double nd = DIM == 3 ? (double)N * N * N : (double)N * N;
    hipLaunchKernelGGL((k_cweight<T>), dim3(nblk(spec_elems())), dim3(256), 0, s, out, out, N, rows_, 3,
                       (T)(1.0 / nd), 0);
    QL_LAUNCHED();
  }
  void rfftn_adjoint(const void* gh, void* out, hipStream_t s) override {
    ensure_adj();
    const double nd = DIM == 3 ? (double)N * N * N : (double)N * N;
    hipLaunchKernelGGL((k_cweight<T>), dim3(nblk(spec_elems())), dim3(256), 0, s, static_cast<const cx<T>*>(gh),
                       AS_[8], N, rows_, 3, (T)nd, 1);
    QL_LAUNCHED();
    irfftn_into(AS_[8], Nh, AS_[9], static_cast<T*>(out), 3, s);
  }

  void filter(const void* uhat_v, int n, void* ulres_v, hipStream_t s) override {
    if (n > N || (n % 2) != 0 || !supported_len(n)) throw std::invalid_argument("unsupported ckp_N");
    const int nh = n / 2 + 1;
    const int ncomp = DIM == 3 ? 3 : 2;
    const size_t rows = DIM == 3 ? (size_t)n * n : (size_t)n;
    const size_t total = rows * nh * ncomp;
    const int nb = (int)std::min<size_t>((total + 255) / 256, 8192);
    ensure_y6();
    cx<T>* g = Y6;  // [ncomp][rows][nh]
    hipLaunchKernelGGL((k_window<T>), dim3(nb), dim3(256), 0, s, static_cast<const cx<T>*>(uhat_v), N, DIM, Nh, g,
                       n, nh);
    QL_LAUNCHED();
    const cx<T>* twn = twiddles<T>(n);
    const T scale = (T)1 / (T)(DIM == 3 ? (double)N * N * N : (double)N * N);
    T* dst = static_cast<T*>(ulres_v);
    if (DIM == 3) {
      c2c_strided<T>(n, g, g, 3, n * nh, false, (T)1, twn, s);
      c2c_strided<T>(n, g, g, 3 * n, nh, false, (T)1, twn, s);
      c2r_contig<T>(n, g, dst, 3 * n * n, nh, scale, twn, s);
    } else {
      c2c_strided<T>(n, g, g, 2, nh, false, (T)1, twn, s);
      c2r_contig<T>(n, g, dst, 2 * n, nh, scale, twn, s);
    }
  }
};

template <typename T, int DIM, int N>
Engine* make_engine_t(const Desc& d) {
  return new EngineT<T, DIM, N>(d);
}

}  // namespace ql
