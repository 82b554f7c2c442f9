// Complex-to-real launchers for double.
#include "kw_launch.hpp"

namespace kw {

template <int N>
static hipError_t c2r_alt(int pv, dim3 g, dim3 b, size_t sh, hipStream_t s, const C2RArgs<double>& a) {
  switch (pv) {
    case 0: hipLaunchKernelGGL((k_c2r_x<double, N, false, 0>), g, b, sh, s, a); break;
    case 1: hipLaunchKernelGGL((k_c2r_x<double, N, false, 1>), g, b, sh, s, a); break;
    case 2: hipLaunchKernelGGL((k_c2r_x<double, N, false, 2>), g, b, sh, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <>
hipError_t launch_c2r<double>(int n, dim3 g, dim3 b, size_t sh, hipStream_t s, const C2RArgs<double>& a, int pv) {
  if (plan_alternative(n, pv)) return n == 256 ? c2r_alt<256>(pv, g, b, sh, s, a) : c2r_alt<512>(pv, g, b, sh, s, a);
  switch (n) {
#define X(n_)                                                      \
  case n_:                                                         \
    hipLaunchKernelGGL((k_c2r_x<double, n_>), g, b, sh, s, a);        \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

// The FP64 plane does not fit a workgroup's registers: no fused form.
template <>
bool plane_c2r_supported<double>(int, int, size_t) {
  return false;
}

template <>
hipError_t launch_plane_c2r<double>(int, int, unsigned, hipStream_t, const PlaneArgs<double>&) {
  return hipErrorInvalidValue;
}

}  // namespace kw
