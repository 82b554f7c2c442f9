// Directional-pass launchers for double.
#include "kw_launch.hpp"

namespace kw {

template <>
hipError_t launch_dir<double>(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const PassArgs<double>& a) {
  switch (n) {
#define X(n_)                                                                  \
  case n_:                                                                     \
    if (contig && a.S2)                                                        \
      hipLaunchKernelGGL((k_dir_pass<double, n_, true, true>), g, b, sh, s, a);  \
    else if (contig)                                                           \
      hipLaunchKernelGGL((k_dir_pass<double, n_, true, false>), g, b, sh, s, a); \
    else                                                                       \
      hipLaunchKernelGGL((k_dir_pass<double, n_, false>), g, b, sh, s, a);       \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

}  // namespace kw
