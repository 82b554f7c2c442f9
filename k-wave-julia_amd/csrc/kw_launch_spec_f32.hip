// Spectral-pass launchers for float.
#include "kw_launch.hpp"

namespace kw {

template <int N, typename T>
static hipError_t spec_mode(int mode, dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<T>& a) {
  switch (mode) {
    case 0:
      if (a.S2)
        hipLaunchKernelGGL((k_spec_pass<T, N, 6>), g, b, sh, s, a);
      else
        hipLaunchKernelGGL((k_spec_pass<T, N, 0>), g, b, sh, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((k_spec_pass<T, N, 1>), g, b, sh, s, a);
      break;
    case 2:
      if (a.S2)
        hipLaunchKernelGGL((k_spec_pass<T, N, 4>), g, b, sh, s, a);
      else
        hipLaunchKernelGGL((k_spec_pass<T, N, 2>), g, b, sh, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((k_spec_pass<T, N, 3>), g, b, sh, s, a);
      break;
    case 5:
      hipLaunchKernelGGL((k_spec_pass<T, N, 5>), g, b, sh, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <>
hipError_t launch_spec<float>(int n, dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<float>& a) {
  switch (n) {
#define X(n_)   \
  case n_:      \
    return spec_mode<n_, float>(a.mode, g, b, sh, s, a);
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

}  // namespace kw
