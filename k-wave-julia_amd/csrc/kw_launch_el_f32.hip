// Elastic pass launchers for float.
#include "kw_launch.hpp"

namespace kw {

template <>
hipError_t launch_el<float>(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const ElArgs<float>& a) {
  switch (n) {
#define M(n_, c_)                                                                            \
  switch (a.mode) {                                                                          \
    case EL_DERIV: hipLaunchKernelGGL((k_el_pass<float, n_, c_, EL_DERIV>), g, b, sh, s, a); break;   \
    case EL_VEL: hipLaunchKernelGGL((k_el_pass<float, n_, c_, EL_VEL>), g, b, sh, s, a); break;       \
    case EL_NORMAL: hipLaunchKernelGGL((k_el_pass<float, n_, c_, EL_NORMAL>), g, b, sh, s, a); break; \
    case EL_SHEAR: hipLaunchKernelGGL((k_el_pass<float, n_, c_, EL_SHEAR>), g, b, sh, s, a); break;   \
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
