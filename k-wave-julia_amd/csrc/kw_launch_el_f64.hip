// Elastic pass launchers for double.
#include "kw_launch.hpp"

namespace kw {

template <>
hipError_t launch_el<double>(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const ElArgs<double>& a) {
  switch (n) {
#define M(n_, c_)                                                                            \
  switch (a.mode) {                                                                          \
    case EL_DERIV: hipLaunchKernelGGL((k_el_pass<double, n_, c_, EL_DERIV>), g, b, sh, s, a); break;   \
    case EL_VEL: hipLaunchKernelGGL((k_el_pass<double, n_, c_, EL_VEL>), g, b, sh, s, a); break;       \
    case EL_NORMAL: hipLaunchKernelGGL((k_el_pass<double, n_, c_, EL_NORMAL>), g, b, sh, s, a); break; \
    case EL_SHEAR: hipLaunchKernelGGL((k_el_pass<double, n_, c_, EL_SHEAR>), g, b, sh, s, a); break;   \
    default: return hipErrorInvalidValue;                                                     \
  }
#define X(n_)      \
  case n_:         \
    if (contig)    \
      M(n_, true)  \
    else           \
      M(n_, false) \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
#undef M
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
