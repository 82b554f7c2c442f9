// Elastic pass launchers for double.
#include "kw_launch.hpp"

namespace kw {

template <>
hipError_t launch_el<double>(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const ElArgs<double>& a) {
  switch (n) {
#define X(n_)                                                              \
  case n_:                                                                 \
    if (contig)                                                            \
      hipLaunchKernelGGL((k_el_pass<double, n_, true>), g, b, sh, s, a);      \
    else                                                                   \
      hipLaunchKernelGGL((k_el_pass<double, n_, false>), g, b, sh, s, a);     \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <>
hipError_t launch_el_sensor_p<double>(hipStream_t s, double* p, const double* s0, const double* s1, const double* s2,
                                   const long long* idx, int n, int ndim) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_el_sensor_p<double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, s0, s1, s2, idx, n,
                     ndim);
  return hipGetLastError();
}

template <>
hipError_t launch_el_inject<double>(hipStream_t s, double* f, const long long* idx, const double* vals, int n, int dirichlet) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_el_inject<double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, f, idx, vals, n,
                     dirichlet);
  return hipGetLastError();
}

}  // namespace kw
