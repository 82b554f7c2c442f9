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

// the 256- and 512-point spectral passes in an alternative plan family
template <int N, int PV>
static hipError_t spec_alt(dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<float>& a) {
  switch (a.mode) {
    case 0:
      if (a.S2)
        hipLaunchKernelGGL((k_spec_pass<float, N, 6, false, PV>), g, b, sh, s, a);
      else
        hipLaunchKernelGGL((k_spec_pass<float, N, 0, false, PV>), g, b, sh, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((k_spec_pass<float, N, 1, false, PV>), g, b, sh, s, a);
      break;
    case 2:
      if (a.S2)
        hipLaunchKernelGGL((k_spec_pass<float, N, 4, false, PV>), g, b, sh, s, a);
      else
        hipLaunchKernelGGL((k_spec_pass<float, N, 2, false, PV>), g, b, sh, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((k_spec_pass<float, N, 3, false, PV>), g, b, sh, s, a);
      break;
    case 5:
      hipLaunchKernelGGL((k_spec_pass<float, N, 5, false, PV>), g, b, sh, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
template <int N>
static hipError_t spec_alt(int pv, dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<float>& a) {
  switch (pv) {
    case 0: return spec_alt<N, 0>(g, b, sh, s, a);
    case 1: return spec_alt<N, 1>(g, b, sh, s, a);
    case 2: return spec_alt<N, 2>(g, b, sh, s, a);
  }
  return hipErrorInvalidValue;
}

template <>
hipError_t launch_spec<float>(int n, dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<float>& a, int pv) {
  if (plan_alternative(n, pv)) return n == 256 ? spec_alt<256>(pv, g, b, sh, s, a) : spec_alt<512>(pv, g, b, sh, s, a);
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
