// Launchers of the any-length (Bluestein) line passes for float: a line of
// length n runs on the power-of-two plan n_plan >= 2n - 1 (kw_fft.hpp line_dft).
#include "kw_launch.hpp"

namespace kw {

template <>
hipError_t launch_dir_bs<float>(int n_plan, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s,
                               const PassArgs<float>& a) {
  switch (n_plan) {
#define X(n_)                                                                          \
  case n_:                                                                             \
    if (contig && a.S2)                                                                \
      hipLaunchKernelGGL((k_dir_pass<float, n_, true, true, true>), g, b, sh, s, a);      \
    else if (contig)                                                                   \
      hipLaunchKernelGGL((k_dir_pass<float, n_, true, false, true>), g, b, sh, s, a);     \
    else                                                                               \
      hipLaunchKernelGGL((k_dir_pass<float, n_, false, false, true>), g, b, sh, s, a);    \
    return hipGetLastError();
    KW_BS_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t spec_mode_bs(int mode, dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<float>& a) {
  switch (mode) {
    case 0:
      if (a.S2)
        hipLaunchKernelGGL((k_spec_pass<float, N, 6, true>), g, b, sh, s, a);
      else
        hipLaunchKernelGGL((k_spec_pass<float, N, 0, true>), g, b, sh, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((k_spec_pass<float, N, 1, true>), g, b, sh, s, a);
      break;
    case 2:
      if (a.S2)
        hipLaunchKernelGGL((k_spec_pass<float, N, 4, true>), g, b, sh, s, a);
      else
        hipLaunchKernelGGL((k_spec_pass<float, N, 2, true>), g, b, sh, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((k_spec_pass<float, N, 3, true>), g, b, sh, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <>
hipError_t launch_spec_bs<float>(int n_plan, dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<float>& a) {
  switch (n_plan) {
#define X(n_)   \
  case n_:      \
    return spec_mode_bs<n_>(a.mode, g, b, sh, s, a);
    KW_BS_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <>
hipError_t launch_c2r_bs<float>(int n_plan, dim3 g, dim3 b, size_t sh, hipStream_t s, const C2RArgs<float>& a) {
  switch (n_plan) {
#define X(n_)                                                       \
  case n_:                                                          \
    hipLaunchKernelGGL((k_c2r_x<float, n_, true>), g, b, sh, s, a);    \
    return hipGetLastError();
    KW_BS_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

}  // namespace kw
