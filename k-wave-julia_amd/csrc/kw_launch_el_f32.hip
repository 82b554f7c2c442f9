// Elastic pass launchers for float.
#include "kw_launch.hpp"

namespace kw {

template <>
hipError_t launch_el<float>(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const ElArgs<float>& a) {
  switch (n) {
#define X(n_)                                                              \
  case n_:                                                                 \
    if (contig)                                                            \
      hipLaunchKernelGGL((k_el_pass<float, n_, true>), g, b, sh, s, a);      \
    else                                                                   \
      hipLaunchKernelGGL((k_el_pass<float, n_, false>), g, b, sh, s, a);     \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <>
hipError_t launch_el_sensor_p<float>(hipStream_t s, float* p, const float* s0, const float* s1, const float* s2,
                                   const long long* idx, int n, int ndim) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_el_sensor_p<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, s0, s1, s2, idx, n,
                     ndim);
  return hipGetLastError();
}

template <>
hipError_t launch_el_inject<float>(hipStream_t s, float* f, const long long* idx, const float* vals, int n, int dirichlet) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_el_inject<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, idx, vals, n,
                     dirichlet);
  return hipGetLastError();
}

}  // namespace kw
