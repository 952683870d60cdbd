// Fused 2-D pseudo-spectral Navier-Stokes RK4 step.
//
// Reference: l3es/spectral_solver.py:27-120 with fft_axes=(1,2): the rfft axis
// is y (the last axis), u^ is stored as (3, N, N/2+1) with u^_z == 0 and the
// wavenumber vector padded with kz == 0 (spectral_solver.py:38-39).  The
// nonlinear term needs only the scalar vorticity omega_z, so one RHS is two
// kernels:
//   X2 (x-pass)  forward x-FFT of N^ (2 comps), dealias kx, projection,
//                viscous term, forcing, RK4 update, inverse x-FFT of
//                [u^, v^, i(kx v^ - ky u^)]               -> X3
//   Y2 (y-pass)  3 x c2r along y, (v w, -u w), 2 x r2c, keep ky < kmax -> NY2
// Layouts: state [3][N][Nh]; X3 [3][N][LD]; NY2 [2][N][KZP].
#pragma once
#include "ns3d.hpp"

namespace qk {

template <typename T>
struct NS2X {
  const cx<T>* ny;   // [2][N][KZP]
  const cx<T>* us;
  const cx<T>* un;
  cx<T>* uout;
  cx<T>* acc;
  cx<T>* x3;         // [3][N][ld]
  cx<T>* rescale_dst;
  int ld_us, ld_un, ld_uout, ld;
  int KZD, KZP;
  int do_inverse;
  const cx<T>* tw;
  StageCtl<T> c;
};

template <typename T, int N>
__device__ double forced_shell_energy2(const cx<T>* us, int ld, int kf2, double* s_red) {
  int kf = 0;
  while ((kf + 1) * (kf + 1) <= kf2) ++kf;
  const int nx = 2 * kf + 1, ny = kf + 1;
  const int cnt = nx * ny;
  double acc[1] = {0.0};
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    const int kx = q / ny - kf, ky = q % ny;
    const int k2 = kx * kx + ky * ky;
    if (k2 > 0 && k2 <= kf2) {
      const int ix = (kx + N) % N;
      const double w = (ky == 0 || ky == N / 2) ? 1.0 : 2.0;
      double e = 0;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const cx<T> a = us[(size_t)(c * N + ix) * ld + ky];
        e += (double)a.re * a.re + (double)a.im * a.im;
      }
      acc[0] += w * e;
    }
  }
  block_sum<1>(acc, s_red);
  return 0.5 * acc[0];
}

template <typename T, int N, int V = 0>
__device__ __forceinline__ void ns2_x_run(const NS2X<T>& a) {
  using C = XCfg<N, V>;
  constexpr int Nh = N / 2 + 1;
  using FF = XFwd<T, N>;
  using IF = XInv<T, N>;
  static_assert(FF::NOUT == IF::NIN, "fft hand-off mismatch");
  constexpr int NS = FF::NOUT;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[N * C::LS];
  __shared__ double s_red[C::THREADS];

  const int tid = threadIdx.x, p = tid % C::TIX, j = tid / C::TIX;
  const int ky = blockIdx.x * C::TIX + p;  // rfft column (Nyquist carries +N/2)
  const bool ok = ky < Nh;
  stage_table(s_tw, a.tw, N);
  __syncthreads();
  const StageCtl<T>& c = a.c;
  cx<T> w[2][NS] = {};
  if (c.stage >= 0) {
    const bool fwd_ok = ok && ky < a.KZD;
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
      cx<T> v[FF::MAXE];
#pragma unroll
      for (int s = 0; s < FF::NIN; ++s)
        v[s] = (fwd_ok && FF::in_valid(j, s)) ? a.ny[((size_t)cc * N + FF::in_index(j, s)) * a.KZP + ky]
                                              : cx<T>{0, 0};
      FF::run(v, s_buf + p, C::LS, j, s_tw, 1);
#pragma unroll
      for (int s = 0; s < NS; ++s) w[cc][s] = v[s];
    }
    T fscale = 0;
    const int ky0 = blockIdx.x * C::TIX;
    if (c.force_power && ky0 * ky0 <= c.kf2) {
      const double ef = forced_shell_energy2<T, N>(a.us, a.ld_us, c.kf2, s_red) / ((double)N * N * (double)N * N);
      fscale = (T)(c.power / (2.0 * ef + 2.220446049250313e-16));
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int kxi = FF::out_index(j, s);
      const int kx = kfreq(kxi, N);
      const bool valid = ok && FF::out_valid(j, s);
      const bool keep_x = keep_k(kxi, N, a.KZD);
      const cx<T> n0 = keep_x ? w[0][s] : cx<T>{0, 0};
      const cx<T> n1 = keep_x ? w[1][s] : cx<T>{0, 0};
      const int k2 = kx * kx + ky * ky;
      const T inv = k2 == 0 ? (T)0 : (T)1 / (T)k2;
      const cx<T> pr = {(n0.re * (T)kx + n1.re * (T)ky) * inv, (n0.im * (T)kx + n1.im * (T)ky) * inv};
      cx<T> du[2] = {csub(n0, cscale(pr, (T)kx)), csub(n1, cscale(pr, (T)ky))};
      if (!valid) continue;
      cx<T> us2[2];
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) us2[cc] = a.us[((size_t)cc * N + kxi) * a.ld_us + ky];
      const T nuk2 = c.nu * (T)k2;
      const bool forced = fscale != (T)0 && k2 > 0 && k2 <= c.kf2;
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        if (!c.visc_if) du[cc] = csub(du[cc], cscale(us2[cc], nuk2));
        if (forced) du[cc] = cadd(du[cc], cscale(us2[cc], fscale));
        if (c.drag != (T)0) du[cc] = csub(du[cc], cscale(us2[cc], c.drag));
      }
      if (c.kolm_k > 0 && kx == 0 && ky == c.kolm_k) du[0] = cadd(du[0], cx<T>{c.kolm_re, c.kolm_im});
      T eh = 1, e1 = 1;
      if (c.visc_if) {
        eh = (T)exp(-(double)c.nu * k2 * 0.5 * (double)c.dt);
        e1 = (T)exp(-(double)c.nu * k2 * (double)c.dt);
      }
      const size_t cs_ld = (size_t)N * a.ld, cs_un = (size_t)N * a.ld_un, cs_out = (size_t)N * a.ld_uout;
      const size_t i_ld = (size_t)kxi * a.ld + ky, i_un = (size_t)kxi * a.ld_un + ky,
                   i_out = (size_t)kxi * a.ld_uout + ky;
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        cx<T> out;
        if (c.stage == 4) {
          out = du[cc];
        } else if (!c.visc_if) {
          if (c.stage == 0) {
            a.acc[cc * cs_ld + i_ld] = cadd(us2[cc], cscale(du[cc], c.ca));
            out = cadd(us2[cc], cscale(du[cc], c.cb));
          } else if (c.stage < 3) {
            const cx<T> un = a.un[cc * cs_un + i_un];
            a.acc[cc * cs_ld + i_ld] = cadd(a.acc[cc * cs_ld + i_ld], cscale(du[cc], c.ca));
            out = cadd(un, cscale(du[cc], c.cb));
          } else {
            out = cadd(a.acc[cc * cs_ld + i_ld], cscale(du[cc], c.ca));
          }
        } else {
          const T dt = c.dt;
          if (c.stage == 0) {
            a.acc[cc * cs_ld + i_ld] = cadd(cscale(us2[cc], e1), cscale(du[cc], dt / (T)6 * e1));
            out = cscale(cadd(us2[cc], cscale(du[cc], dt / (T)2)), eh);
          } else if (c.stage == 1) {
            const cx<T> un = a.un[cc * cs_un + i_un];
            a.acc[cc * cs_ld + i_ld] = cadd(a.acc[cc * cs_ld + i_ld], cscale(du[cc], dt / (T)3 * eh));
            out = cadd(cscale(un, eh), cscale(du[cc], dt / (T)2));
          } else if (c.stage == 2) {
            const cx<T> un = a.un[cc * cs_un + i_un];
            a.acc[cc * cs_ld + i_ld] = cadd(a.acc[cc * cs_ld + i_ld], cscale(du[cc], dt / (T)3 * eh));
            out = cadd(cscale(un, e1), cscale(du[cc], dt * eh));
          } else {
            out = cadd(a.acc[cc * cs_ld + i_ld], cscale(du[cc], dt / (T)6));
          }
        }
        a.uout[cc * cs_out + i_out] = out;
        w[cc][s] = out;
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int kxi = IF::in_index(j, s);
      if (ok && IF::in_valid(j, s)) {
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) w[cc][s] = a.un[((size_t)cc * N + kxi) * a.ld_un + ky];
      }
    }
  }
  if (a.do_inverse) {
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      cx<T> v[IF::MAXE];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const int kx = kfreq(IF::in_index(j, s), N);
        if (f < 2) v[s] = w[f][s];
        else v[s] = cmuli(csub(cscale(w[1][s], (T)kx), cscale(w[0][s], (T)ky)));  // omega_z^
      }
      IF::run(v, s_buf + p, C::LS, j, s_tw, 1);
      if (ok) {
#pragma unroll
        for (int s = 0; s < IF::NOUT; ++s)
          if (IF::out_valid(j, s)) a.x3[((size_t)f * N + IF::out_index(j, s)) * a.ld + ky] = v[s];
      }
    }
  }
}

template <typename T, int N, int V = 0>
__global__ void __launch_bounds__((XCfg<N, V>::THREADS))
k_ns2_x(Pack<NS2X<T>> pk) {
  const NS2X<T> a = pk.m[blockIdx.z];  // fields loaded once, at the top
  ns2_x_run<T, N, V>(a);
}

template <typename T>
struct NS2Y {
  const cx<T>* in;   // [3][N][ld]
  cx<T>* ny;         // [2][N][KZP]
  T* u_out;          // optional [3][N][N] (component 2 untouched)
  const T* u_in;     // optional physical velocity replacing the c2r of fields 0..1
  int ld, KZD, KZP;
  const cx<T>* tw;
};

template <typename T, int N, int V = 0>
__device__ __forceinline__ void ns2_y_run(const NS2Y<T>& a) {
  using C = ZCfg<N, V>;
  constexpr int M = C::M, TP = C::TP, P = C::P;
  constexpr int Nh = M + 1;
  constexpr int ZP = C::ZPAD;
  using CI = PFFT<T, false, M, TP, Reverse<typename C::R>, ZP>;
  using CF = PFFT<T, true, M, TP, typename C::R, ZP>;
  static_assert(CI::NOUT == CF::NIN, "y-pass hand-off mismatch");
  constexpr int NO = CI::NOUT;
  __shared__ cx<T> s_tw[N];
  __shared__ cx<T> s_buf[P * C::LSR];
  const int tid = threadIdx.x, p = tid / TP, j = tid % TP;
  const int row0 = blockIdx.x * P;
  const int row = row0 + p;
  const bool ok = row < N;
  cx<T>* pb = s_buf + p * C::LSR;
  stage_table(s_tw, a.tw, N);
  const size_t fs = (size_t)N * a.ld;
  const T scale = (T)1 / ((T)N * (T)N);
  cx<T> ph[3][NO];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (f < 2 && a.u_in != nullptr) {
      const cx<T>* src = reinterpret_cast<const cx<T>*>(a.u_in + ((size_t)f * N + (ok ? row : 0)) * N);
#pragma unroll
      for (int s = 0; s < NO; ++s) ph[f][s] = src[CI::out_index(j, s)];
      continue;
    }
    __syncthreads();
    for (int e = tid; e < P * Nh; e += C::THREADS) {
      const int q = e / Nh, k = e - q * Nh;
      if (row0 + q < N) s_buf[q * C::LSR + lds_at<ZP>(k)] = a.in[f * fs + (size_t)(row0 + q) * a.ld + k];
    }
    __syncthreads();
    cx<T> v[CI::MAXE];
#pragma unroll
    for (int s = 0; s < CI::NIN; ++s) {
      cx<T> z = {0, 0};
      if (CI::in_valid(j, s)) {
        const int n = CI::in_index(j, s);
        cx<T> xn = pb[lds_at<ZP>(n)], xr = pb[lds_at<ZP>(M - n)];
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
    for (int s = 0; s < NO; ++s) ph[f][s] = cscale(v[s], scale);
  }
  if (a.u_out != nullptr && ok) {
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      cx<T>* dst = reinterpret_cast<cx<T>*>(a.u_out + ((size_t)f * N + row) * N);
#pragma unroll
      for (int s = 0; s < NO; ++s) dst[CI::out_index(j, s)] = ph[f][s];
    }
  }
  // (u x omega)_x = v w_z, (u x omega)_y = -u w_z
  cx<T> nl[2][NO];
#pragma unroll
  for (int s = 0; s < NO; ++s) {
    const cx<T> u = ph[0][s], v = ph[1][s], oz = ph[2][s];
    nl[0][s] = {v.re * oz.re, v.im * oz.im};
    nl[1][s] = {-u.re * oz.re, -u.im * oz.im};
  }
#pragma unroll
  for (int cc = 0; cc < 2; ++cc) {
    cx<T> v[CF::MAXE];
#pragma unroll
    for (int s = 0; s < NO; ++s) v[s] = nl[cc][s];
    __syncthreads();
    CF::run(v, pb, 1, j, s_tw, 2);
#pragma unroll
    for (int s = 0; s < CF::NOUT; ++s)
      if (CF::out_valid(j, s)) pb[lds_at<ZP>(CF::out_index(j, s))] = v[s];
    __syncthreads();
    if (ok) {
      cx<T>* dst = a.ny + ((size_t)cc * N + row) * a.KZP;
      for (int k = j; k < a.KZD; k += TP) {
        const cx<T> zk = pb[lds_at<ZP>(k)];
        const cx<T> zr = cconj(pb[lds_at<ZP>(k == 0 ? 0 : M - k)]);
        const cx<T> e = cscale(cadd(zk, zr), (T)0.5);
        const cx<T> o = crot<T, true>(cscale(csub(zk, zr), (T)0.5));
        dst[k] = cadd(e, cmul(s_tw[k], o));
      }
    }
  }
}

template <typename T, int N, int V = 0>
__global__ void __launch_bounds__((ZCfg<N, V>::THREADS))
k_ns2_y(Pack<NS2Y<T>> pk) {
  const NS2Y<T> a = pk.m[blockIdx.z];
  ns2_y_run<T, N, V>(a);
}

}  // namespace qk
