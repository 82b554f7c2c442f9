// Directional-pass launchers for float.
#include "kw_launch.hpp"

namespace kw {

// the 256- and 512-point passes in an alternative plan family
template <int N, int PV>
static hipError_t dir_alt(bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const PassArgs<float>& a) {
  if (contig && a.S2)
    hipLaunchKernelGGL((k_dir_pass<float, N, true, true, false, PV>), g, b, sh, s, a);
  else if (contig)
    hipLaunchKernelGGL((k_dir_pass<float, N, true, false, false, PV>), g, b, sh, s, a);
  else
    hipLaunchKernelGGL((k_dir_pass<float, N, false, false, false, PV>), g, b, sh, s, a);
  return hipGetLastError();
}
template <int N>
static hipError_t dir_alt(int pv, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const PassArgs<float>& a) {
  switch (pv) {
    case 0: return dir_alt<N, 0>(contig, g, b, sh, s, a);
    case 1: return dir_alt<N, 1>(contig, g, b, sh, s, a);
    case 2: return dir_alt<N, 2>(contig, g, b, sh, s, a);
  }
  return hipErrorInvalidValue;
}

template <>
hipError_t launch_dir<float>(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const PassArgs<float>& a,
                           int pv) {
  if (plan_alternative(n, pv)) return n == 256 ? dir_alt<256>(pv, contig, g, b, sh, s, a) : dir_alt<512>(pv, contig, g, b, sh, s, a);
  switch (n) {
#define X(n_)                                                                  \
  case n_:                                                                     \
    if (contig && a.S2)                                                        \
      hipLaunchKernelGGL((k_dir_pass<float, n_, true, true>), g, b, sh, s, a);  \
    else if (contig)                                                           \
      hipLaunchKernelGGL((k_dir_pass<float, n_, true, false>), g, b, sh, s, a); \
    else                                                                       \
      hipLaunchKernelGGL((k_dir_pass<float, n_, false>), g, b, sh, s, a);       \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

}  // namespace kw
