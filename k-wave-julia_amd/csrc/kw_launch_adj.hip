// Launchers of the adjoint kernels (both precisions).
#include "kw_launch.hpp"

namespace kw {

template <typename T, int N, int PV>
static hipError_t adj_alt(bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const AdjArgs<T>& a) {
  if (contig)
    hipLaunchKernelGGL((k_adj_pass<T, N, true, PV>), g, b, sh, s, a);
  else
    hipLaunchKernelGGL((k_adj_pass<T, N, false, PV>), g, b, sh, s, a);
  return hipGetLastError();
}
template <typename T, int N>
static hipError_t adj_alt(int pv, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const AdjArgs<T>& a) {
  switch (pv) {
    case 0: return adj_alt<T, N, 0>(contig, g, b, sh, s, a);
    case 1: return adj_alt<T, N, 1>(contig, g, b, sh, s, a);
    case 2: return adj_alt<T, N, 2>(contig, g, b, sh, s, a);
  }
  return hipErrorInvalidValue;
}
template <typename T, int N>
static hipError_t r2c_alt(int pv, dim3 g, dim3 b, size_t sh, hipStream_t s, const T* f, Cx<T>* S, int n_lines,
                          int nhp, const Cx<T>* tw) {
  switch (pv) {
    case 0: hipLaunchKernelGGL((k_r2c_x<T, N, 0>), g, b, sh, s, f, S, n_lines, nhp, tw); break;
    case 1: hipLaunchKernelGGL((k_r2c_x<T, N, 1>), g, b, sh, s, f, S, n_lines, nhp, tw); break;
    case 2: hipLaunchKernelGGL((k_r2c_x<T, N, 2>), g, b, sh, s, f, S, n_lines, nhp, tw); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_adj_pass(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const AdjArgs<T>& a,
                           int pv) {
  if (plan_alternative(n, pv))
    return n == 256 ? adj_alt<T, 256>(pv, contig, g, b, sh, s, a) : adj_alt<T, 512>(pv, contig, g, b, sh, s, a);
  switch (n) {
#define X(n_)                                                             \
  case n_:                                                                \
    if (contig)                                                           \
      hipLaunchKernelGGL((k_adj_pass<T, n_, true>), g, b, sh, s, a);      \
    else                                                                  \
      hipLaunchKernelGGL((k_adj_pass<T, n_, false>), g, b, sh, s, a);     \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_r2c(int n, dim3 g, dim3 b, size_t sh, hipStream_t s, const T* f, Cx<T>* S, int n_lines, int nhp,
                      const Cx<T>* tw, int pv) {
  if (plan_alternative(n, pv))
    return n == 256 ? r2c_alt<T, 256>(pv, g, b, sh, s, f, S, n_lines, nhp, tw)
                    : r2c_alt<T, 512>(pv, g, b, sh, s, f, S, n_lines, nhp, tw);
  switch (n) {
#define X(n_)                                                                      \
  case n_:                                                                         \
    hipLaunchKernelGGL((k_r2c_x<T, n_>), g, b, sh, s, f, S, n_lines, nhp, tw);     \
    return hipGetLastError();
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

// which: 0 grad (c0sq, rs, Q, Qd, g, gpart; n = one firing's voxels, nf firings), 1 src (r0, r1, r2, c0sq, Q, idx, fac, sig, g),
//        2 resid (prec, obs, res, lpart), 3 scatter (Pb, idx, r), 4 final (g, p0, c0sq, r0, r1, r2; dargs[0] = gterm)
// gpart / lpart hold adj_reduce_blocks(n) per-block partial sums.
template <typename T>
hipError_t launch_adj_small(int which, hipStream_t s, const void* const* p, const double* dargs, long long n,
                            int ndim, int nf) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  const unsigned nr = adj_reduce_blocks(n);
  switch (which) {
    case 0:
      hipLaunchKernelGGL((k_adj_grad<T>), dim3(nr), dim3(256), 0, s, (const T*)p[0], (const T*)p[1], (const T*)p[2],
                         (const T*)p[3], (double*)p[4], (double*)p[5], n, nf);
      break;
    case 1:
      hipLaunchKernelGGL((k_adj_src<T>), dim3(nb), dim3(256), 0, s, (const T*)p[0], (const T*)p[1], (const T*)p[2],
                         (const T*)p[3], (const T*)p[4], (const long long*)p[5], (const double*)p[6],
                         (const double*)p[7], (double*)p[8], ndim, (int)n);
      break;
    case 2:
      hipLaunchKernelGGL((k_adj_resid<T>), dim3(nr), dim3(256), 0, s, (const T*)p[0], (const double*)p[1],
                         (double*)p[2], (double*)p[3], n);
      break;
    case 3:
      hipLaunchKernelGGL((k_adj_scatter<T>), dim3(nb), dim3(256), 0, s, (T*)p[0], (const long long*)p[1],
                         (const double*)p[2], (int)n);
      break;
    case 4:
      hipLaunchKernelGGL((k_adj_final<T>), dim3(nb), dim3(256), 0, s, (double*)p[0], dargs[0], (const T*)p[1],
                         (const T*)p[2], (const T*)p[3], (const T*)p[4], (const T*)p[5], ndim, n);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#define KW_INST(T)                                                                                                  \
  template hipError_t launch_adj_pass<T>(int, bool, dim3, dim3, size_t, hipStream_t, const AdjArgs<T>&, int);      \
  template hipError_t launch_r2c<T>(int, dim3, dim3, size_t, hipStream_t, const T*, Cx<T>*, int, int, const Cx<T>*,  \
                                    int);                                                                           \
  template hipError_t launch_adj_small<T>(int, hipStream_t, const void* const*, const double*, long long, int, int);
KW_INST(float)
KW_INST(double)
#undef KW_INST

}  // namespace kw
