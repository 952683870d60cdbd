// Runtime dispatch on transform length / grid size, and the twiddle-table cache.
#include <cmath>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/qlles.h"
#include "engine.hpp"
#include "fftlib.hpp"

namespace ql {

#define QL_LEN_SWITCH(L, CALL)                    \
  switch (L) {                                    \
    case 8: { constexpr int LL = 8; CALL; break; }     \
    case 16: { constexpr int LL = 16; CALL; break; }   \
    case 32: { constexpr int LL = 32; CALL; break; }   \
    case 64: { constexpr int LL = 64; CALL; break; }   \
    case 96: { constexpr int LL = 96; CALL; break; }   \
    case 128: { constexpr int LL = 128; CALL; break; } \
    case 256: { constexpr int LL = 256; CALL; break; } \
    case 512: { constexpr int LL = 512; CALL; break; } \
    default: throw std::invalid_argument("unsupported transform length " + std::to_string(L)); \
  }

template <typename T>
void c2c_strided(int L, const cx<T>* in, cx<T>* out, int outer, int inner, bool fwd, T scale, const cx<T>* tw, hipStream_t s) {
  QL_LEN_SWITCH(L, (c2c_strided_t<T, LL>(in, out, outer, inner, fwd, scale, tw, s)));
}
template <typename T>
void c2c_contig(int L, const cx<T>* in, cx<T>* out, int rows, bool fwd, T scale, const cx<T>* tw, hipStream_t s) {
  QL_LEN_SWITCH(L, (c2c_contig_t<T, LL>(in, out, rows, fwd, scale, tw, s)));
}
template <typename T>
void r2c_contig(int L, const T* in, cx<T>* out, int rows, int ldo, T scale, const cx<T>* tw, hipStream_t s) {
  QL_LEN_SWITCH(L, (r2c_contig_t<T, LL>(in, out, rows, ldo, scale, tw, s)));
}
template <typename T>
void c2r_contig(int L, const cx<T>* in, T* out, int rows, int ldi, T scale, const cx<T>* tw, hipStream_t s) {
  QL_LEN_SWITCH(L, (c2r_contig_t<T, LL>(in, out, rows, ldi, scale, tw, s)));
}

template void c2c_strided<float>(int, const cx<float>*, cx<float>*, int, int, bool, float, const cx<float>*, hipStream_t);
template void c2c_strided<double>(int, const cx<double>*, cx<double>*, int, int, bool, double, const cx<double>*, hipStream_t);
template void c2c_contig<float>(int, const cx<float>*, cx<float>*, int, bool, float, const cx<float>*, hipStream_t);
template void c2c_contig<double>(int, const cx<double>*, cx<double>*, int, bool, double, const cx<double>*, hipStream_t);
template void r2c_contig<float>(int, const float*, cx<float>*, int, int, float, const cx<float>*, hipStream_t);
template void r2c_contig<double>(int, const double*, cx<double>*, int, int, double, const cx<double>*, hipStream_t);
template void c2r_contig<float>(int, const cx<float>*, float*, int, int, float, const cx<float>*, hipStream_t);
template void c2r_contig<double>(int, const cx<double>*, double*, int, int, double, const cx<double>*, hipStream_t);

// twiddle tables exp(-2 pi i q / L) per (device, precision, L); kept for the
// life of the process (a few KB each)
template <typename T>
const cx<T>* twiddles(int L) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, void*> cache;
  int dev = 0;
  QL_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto key = std::make_pair(dev, L);
  auto it = cache.find(key);
  if (it != cache.end()) return static_cast<const cx<T>*>(it->second);
  std::vector<cx<T>> h(L);
  for (int q = 0; q < L; ++q) {
    // exact values on the axes; elsewhere FP64 cos/sin of 2 pi q / L
    long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)q / (long double)L;
    double c = (double)cosl(ang), sn = (double)sinl(ang);
    if (4 * q == L) { c = 0; sn = -1; }
    if (2 * q == L) { c = -1; sn = 0; }
    if (4 * q == 3 * L) { c = 0; sn = 1; }
    if (q == 0) { c = 1; sn = 0; }
    h[q].re = (T)c;
    h[q].im = (T)sn;
  }
  void* d = nullptr;
  QL_HIP(hipMalloc(&d, sizeof(cx<T>) * L));
  QL_HIP(hipMemcpy(d, h.data(), sizeof(cx<T>) * L, hipMemcpyHostToDevice));
  cache[key] = d;
  return static_cast<const cx<T>*>(d);
}
template const cx<float>* twiddles<float>(int);
template const cx<double>* twiddles<double>(int);

// ---------------- per-kernel event timing ----------------
hipEvent_t Engine::take_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  QL_HIP(hipEventCreate(&e));
  return e;
}

void Engine::prof_begin(hipStream_t s, hipEvent_t* a) {
  if (!prof_) return;
  *a = take_event();
  QL_HIP(hipEventRecord(*a, s));
}

void Engine::prof_end(hipStream_t s, int kind, double bytes, hipEvent_t a) {
  if (!prof_ || a == nullptr) return;
  hipEvent_t b = take_event();
  QL_HIP(hipEventRecord(b, s));
  recs_.push_back(Rec{kind, bytes, a, b});
  if (recs_.size() > 8192) prof_drain();
}

void Engine::prof_drain() {
  for (Rec& r : recs_) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      acc_ms_[r.kind] += ms;
      acc_bytes_[r.kind] += r.bytes;
      acc_count_[r.kind] += 1;
    }
    pool_.push_back(r.a);
    pool_.push_back(r.b);
  }
  recs_.clear();
}

void Engine::profile_enable(bool on) {
  prof_drain();
  prof_ = on;
}

void Engine::profile_read(double* ms, int* count, double* bytes) {
  prof_drain();
  for (int k = 0; k < NKIND; ++k) {
    if (ms) ms[k] = acc_ms_[k];
    if (count) count[k] = acc_count_[k];
    if (bytes) bytes[k] = acc_bytes_[k];
    acc_ms_[k] = 0;
    acc_bytes_[k] = 0;
    acc_count_[k] = 0;
  }
}

template <typename T, int DIM>
static Engine* make_dim(const Desc& d) {
  if (DIM == 3) {
    switch (d.N) {
      case 8: return make_engine_t<T, 3, 8>(d);
      case 16: return make_engine_t<T, 3, 16>(d);
      case 32: return make_engine_t<T, 3, 32>(d);
      case 64: return make_engine_t<T, 3, 64>(d);
      case 96: return make_engine_t<T, 3, 96>(d);
      case 128: return make_engine_t<T, 3, 128>(d);
      case 256: return make_engine_t<T, 3, 256>(d);
      default: break;
    }
  } else {
    switch (d.N) {
      case 8: return make_engine_t<T, 2, 8>(d);
      case 16: return make_engine_t<T, 2, 16>(d);
      case 32: return make_engine_t<T, 2, 32>(d);
      case 64: return make_engine_t<T, 2, 64>(d);
      case 96: return make_engine_t<T, 2, 96>(d);
      case 128: return make_engine_t<T, 2, 128>(d);
      case 256: return make_engine_t<T, 2, 256>(d);
      case 512: return make_engine_t<T, 2, 512>(d);
      default: break;
    }
  }
  return nullptr;
}

Engine* make_engine(const Desc& d) {
  if (d.dtype == QLLES_F32) return d.dim == 3 ? make_dim<float, 3>(d) : make_dim<float, 2>(d);
  return d.dim == 3 ? make_dim<double, 3>(d) : make_dim<double, 2>(d);
}

}  // namespace ql
