// Host-side launchers for the generic FFT passes, dispatched on the runtime
// transform length.  Each supported length is instantiated in its own
// translation unit (inst_<L>.hip) so the library builds in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdexcept>
#include <string>

#include "fft_core.hpp"

namespace ql {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define QL_HIP(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      throw ::ql::HipError(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + \
                           ":" + std::to_string(__LINE__) + ")");                                 \
  } while (0)

#define QL_LAUNCHED() QL_HIP(hipGetLastError())

inline bool supported_len(int L) {
  switch (L) {
    case 8: case 16: case 32: case 64: case 96: case 128: case 256: case 512: return true;
    default: return false;
  }
}

using qf::cx;

// per-length launchers (defined in inst_<L>.hip)
template <typename T, int L> void c2c_strided_t(const cx<T>* in, cx<T>* out, int outer, int inner, bool fwd, T scale, const cx<T>* tw, hipStream_t s);
template <typename T, int L> void c2c_contig_t(const cx<T>* in, cx<T>* out, int rows, bool fwd, T scale, const cx<T>* tw, hipStream_t s);
template <typename T, int L> void r2c_contig_t(const T* in, cx<T>* out, int rows, int ldo, T scale, const cx<T>* tw, hipStream_t s);
template <typename T, int L> void c2r_contig_t(const cx<T>* in, T* out, int rows, int ldi, T scale, const cx<T>* tw, hipStream_t s);

// runtime-length dispatch (tw must be the length-L table exp(-2 pi i q / L))
template <typename T> void c2c_strided(int L, const cx<T>* in, cx<T>* out, int outer, int inner, bool fwd, T scale, const cx<T>* tw, hipStream_t s);
template <typename T> void c2c_contig(int L, const cx<T>* in, cx<T>* out, int rows, bool fwd, T scale, const cx<T>* tw, hipStream_t s);
template <typename T> void r2c_contig(int L, const T* in, cx<T>* out, int rows, int ldo, T scale, const cx<T>* tw, hipStream_t s);
template <typename T> void c2r_contig(int L, const cx<T>* in, T* out, int rows, int ldi, T scale, const cx<T>* tw, hipStream_t s);

// device twiddle table exp(-2 pi i q / L), q < L, computed in FP64 on the host
template <typename T> const cx<T>* twiddles(int L);

}  // namespace ql
