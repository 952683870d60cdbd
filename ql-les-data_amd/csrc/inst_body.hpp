// Explicit instantiations for one transform length QL_LEN (included by inst_<L>.hip).
#include "engine_impl.hpp"

namespace ql {

template <typename T, int L>
void c2c_strided_t(const cx<T>* in, cx<T>* out, int outer, int inner, bool fwd, T scale, const cx<T>* tw,
                   hipStream_t s) {
  dim3 g((inner + TI - 1) / TI, outer);
  if (fwd)
    hipLaunchKernelGGL((k_c2c_strided<T, L, true>), g, dim3(StridedCfg<T, L>::THREADS), 0, s, in, out, inner, scale, tw);
  else
    hipLaunchKernelGGL((k_c2c_strided<T, L, false>), g, dim3(StridedCfg<T, L>::THREADS), 0, s, in, out, inner, scale, tw);
  QL_LAUNCHED();
}
template <typename T, int L>
void c2c_contig_t(const cx<T>* in, cx<T>* out, int rows, bool fwd, T scale, const cx<T>* tw, hipStream_t s) {
  dim3 g((rows + ContigCfg<L>::P - 1) / ContigCfg<L>::P);
  if (fwd)
    hipLaunchKernelGGL((k_c2c_contig<T, L, true>), g, dim3(ContigCfg<L>::THREADS), 0, s, in, out, rows, scale, tw);
  else
    hipLaunchKernelGGL((k_c2c_contig<T, L, false>), g, dim3(ContigCfg<L>::THREADS), 0, s, in, out, rows, scale, tw);
  QL_LAUNCHED();
}
template <typename T, int L>
void r2c_contig_t(const T* in, cx<T>* out, int rows, int ldo, T scale, const cx<T>* tw, hipStream_t s) {
  dim3 g((rows + RealCfg<L>::P - 1) / RealCfg<L>::P);
  hipLaunchKernelGGL((k_r2c_contig<T, L>), g, dim3(RealCfg<L>::THREADS), 0, s, in, out, rows, ldo, scale, tw);
  QL_LAUNCHED();
}
template <typename T, int L>
void c2r_contig_t(const cx<T>* in, T* out, int rows, int ldi, T scale, const cx<T>* tw, hipStream_t s) {
  dim3 g((rows + RealCfg<L>::P - 1) / RealCfg<L>::P);
  hipLaunchKernelGGL((k_c2r_contig<T, L>), g, dim3(RealCfg<L>::THREADS), 0, s, in, out, rows, ldi, scale, tw);
  QL_LAUNCHED();
}

#define QL_INST(T)                                                                                          \
  template void c2c_strided_t<T, QL_LEN>(const cx<T>*, cx<T>*, int, int, bool, T, const cx<T>*, hipStream_t); \
  template void c2c_contig_t<T, QL_LEN>(const cx<T>*, cx<T>*, int, bool, T, const cx<T>*, hipStream_t);       \
  template void r2c_contig_t<T, QL_LEN>(const T*, cx<T>*, int, int, T, const cx<T>*, hipStream_t);            \
  template void c2r_contig_t<T, QL_LEN>(const cx<T>*, T*, int, int, T, const cx<T>*, hipStream_t);
QL_INST(float)
QL_INST(double)
#undef QL_INST

template Engine* make_engine_t<float, 2, QL_LEN>(const Desc&);
template Engine* make_engine_t<double, 2, QL_LEN>(const Desc&);
#if QL_LEN <= 256
template Engine* make_engine_t<float, 3, QL_LEN>(const Desc&);
template Engine* make_engine_t<double, 3, QL_LEN>(const Desc&);
#endif

}  // namespace ql
