// Batched launchers for float: one launch per pass over many simulations.
#include "kw_launch.hpp"

namespace kw {

template <>
hipError_t launch_dir_batch<float>(int n, bool contig, bool two, unsigned nwg, dim3 b, size_t sh, hipStream_t s,
                                 const PassArgs<float>* args, const int* wg_sim, const int* wg_first,
                                 const int* t_dev) {
  switch (n) {
#define X(n_)                                                                                              \
  case n_:                                                                                                 \
    if (contig && two)                                                                                     \
      hipLaunchKernelGGL((k_dir_pass_batch<float, n_, true, true>), dim3(nwg), b, sh, s, args, wg_sim, wg_first, \
                         t_dev);                                                                           \
    else if (contig)                                                                                       \
      hipLaunchKernelGGL((k_dir_pass_batch<float, n_, true, false>), dim3(nwg), b, sh, s, args, wg_sim,        \
                         wg_first, t_dev);                                                                 \
    else                                                                                                   \
      hipLaunchKernelGGL((k_dir_pass_batch<float, n_, false>), dim3(nwg), b, sh, s, args, wg_sim, wg_first,    \
                         t_dev);                                                                           \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t spec_batch_mode(int mode, bool two, unsigned nwg, dim3 b, size_t sh, hipStream_t s,
                                  const SpecArgs<float>* args, const int* wg_sim, const int* wg_first) {
  switch (mode) {
    case 0:
      if (two)
        hipLaunchKernelGGL((k_spec_pass_batch<float, N, 6>), dim3(nwg), b, sh, s, args, wg_sim, wg_first);
      else
        hipLaunchKernelGGL((k_spec_pass_batch<float, N, 0>), dim3(nwg), b, sh, s, args, wg_sim, wg_first);
      break;
    case 1:
      hipLaunchKernelGGL((k_spec_pass_batch<float, N, 1>), dim3(nwg), b, sh, s, args, wg_sim, wg_first);
      break;
    case 2:
      if (two)
        hipLaunchKernelGGL((k_spec_pass_batch<float, N, 4>), dim3(nwg), b, sh, s, args, wg_sim, wg_first);
      else
        hipLaunchKernelGGL((k_spec_pass_batch<float, N, 2>), dim3(nwg), b, sh, s, args, wg_sim, wg_first);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <>
hipError_t launch_spec_batch<float>(int n, int mode, bool two, unsigned nwg, dim3 b, size_t sh, hipStream_t s,
                                  const SpecArgs<float>* args, const int* wg_sim, const int* wg_first) {
  switch (n) {
#define X(n_) \
  case n_:    \
    return spec_batch_mode<n_>(mode, two, nwg, b, sh, s, args, wg_sim, wg_first);
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <>
hipError_t launch_c2r_batch<float>(int n, unsigned nwg, dim3 b, size_t sh, hipStream_t s, const C2RArgs<float>* args,
                                 const int* wg_sim, const int* wg_first) {
  switch (n) {
#define X(n_)                                                                               \
  case n_:                                                                                  \
    hipLaunchKernelGGL((k_c2r_x_batch<float, n_>), dim3(nwg), b, sh, s, args, wg_sim, wg_first); \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

}  // namespace kw
