// Complex-to-real launchers for float.
#include "kw_launch.hpp"

namespace kw {

template <int N>
static hipError_t c2r_alt(int pv, dim3 g, dim3 b, size_t sh, hipStream_t s, const C2RArgs<float>& a) {
  switch (pv) {
    case 0: hipLaunchKernelGGL((k_c2r_x<float, N, false, 0>), g, b, sh, s, a); break;
    case 1: hipLaunchKernelGGL((k_c2r_x<float, N, false, 1>), g, b, sh, s, a); break;
    case 2: hipLaunchKernelGGL((k_c2r_x<float, N, false, 2>), g, b, sh, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <>
hipError_t launch_c2r<float>(int n, dim3 g, dim3 b, size_t sh, hipStream_t s, const C2RArgs<float>& a, int pv) {
  if (plan_alternative(n, pv)) return n == 256 ? c2r_alt<256>(pv, g, b, sh, s, a) : c2r_alt<512>(pv, g, b, sh, s, a);
  switch (n) {
#define X(n_)                                                      \
  case n_:                                                         \
    hipLaunchKernelGGL((k_c2r_x<float, n_>), g, b, sh, s, a);        \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <>
bool plane_c2r_supported<float>(int nx, int ny, size_t lds_max) {
  return nx == 256 && ny == 256 && PlaneGeom<float, 256, 256>::LDS <= lds_max;
}

template <>
hipError_t launch_plane_c2r<float>(int nx, int ny, unsigned nz, hipStream_t s, const PlaneArgs<float>& a) {
  if (nx == 256 && ny == 256) {
    using G = PlaneGeom<float, 256, 256>;
    hipLaunchKernelGGL((k_plane_c2r<float, 256, 256>), dim3(nz), dim3(G::NT), G::LDS, s, a);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

}  // namespace kw
