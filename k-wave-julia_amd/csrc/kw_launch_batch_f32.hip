// Batched launchers for float: one launch per pass over many simulations.
#include "kw_launch.hpp"

namespace kw {

// One batched kernel: the persistent instance (P = true) when a persistent
// grid is asked for (cus > 0) and the line length has one, else one
// workgroup per tile.
template <int N, bool C, bool TWO>
static hipError_t dir_b(unsigned nwg, dim3 b, size_t sh, hipStream_t s, int cus, int cap,
                        const PassArgs<float>* args, const int* wg_sim, const int* wg_first, const int* t_dev) {
  if constexpr (persistent_length(N))
    if (cus > 0)
      return launch_tiles(k_dir_pass_batch<float, N, C, TWO, true>, nwg, b, sh, s, cus, cap, args, wg_sim, wg_first,
                          t_dev);
  return launch_tiles(k_dir_pass_batch<float, N, C, TWO, false>, nwg, b, sh, s, 0, 0, args, wg_sim, wg_first, t_dev);
}
template <int N, int MODE>
static hipError_t spec_b(unsigned nwg, dim3 b, size_t sh, hipStream_t s, int cus, int cap,
                         const SpecArgs<float>* args, const int* wg_sim, const int* wg_first) {
  if constexpr (persistent_length(N))
    if (cus > 0)
      return launch_tiles(k_spec_pass_batch<float, N, MODE, true>, nwg, b, sh, s, cus, cap, args, wg_sim, wg_first);
  return launch_tiles(k_spec_pass_batch<float, N, MODE, false>, nwg, b, sh, s, 0, 0, args, wg_sim, wg_first);
}

template <>
hipError_t launch_dir_batch<float>(int n, bool contig, bool two, unsigned nwg, dim3 b, size_t sh, hipStream_t s,
                                   const PassArgs<float>* args, const int* wg_sim, const int* wg_first,
                                   const int* t_dev, int cus, int cap) {
  switch (n) {
#define X(n_)                                                                                              \
  case n_:                                                                                                 \
    if (contig && two) return dir_b<n_, true, true>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first, t_dev); \
    if (contig) return dir_b<n_, true, false>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first, t_dev);       \
    return dir_b<n_, false, false>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first, t_dev);
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t spec_batch_mode(int mode, bool two, unsigned nwg, dim3 b, size_t sh, hipStream_t s,
                                  const SpecArgs<float>* args, const int* wg_sim, const int* wg_first, int cus,
                                  int cap) {
  switch (mode) {
    case 0:
      if (two) return spec_b<N, 6>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first);
      return spec_b<N, 0>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first);
    case 1:
      return spec_b<N, 1>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first);
    case 2:
      if (two) return spec_b<N, 4>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first);
      return spec_b<N, 2>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first);
  }
  return hipErrorInvalidValue;
}

template <>
hipError_t launch_spec_batch<float>(int n, int mode, bool two, unsigned nwg, dim3 b, size_t sh, hipStream_t s,
                                    const SpecArgs<float>* args, const int* wg_sim, const int* wg_first, int cus,
                                    int cap) {
  switch (n) {
#define X(n_) \
  case n_:    \
    return spec_batch_mode<n_>(mode, two, nwg, b, sh, s, args, wg_sim, wg_first, cus, cap);
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

template <int N>
static hipError_t c2r_b(unsigned nwg, dim3 b, size_t sh, hipStream_t s, int cus, int cap, const C2RArgs<float>* args,
                        const int* wg_sim, const int* wg_first) {
  if constexpr (persistent_length(N))
    if (cus > 0) return launch_tiles(k_c2r_x_batch<float, N, true>, nwg, b, sh, s, cus, cap, args, wg_sim, wg_first);
  return launch_tiles(k_c2r_x_batch<float, N, false>, nwg, b, sh, s, 0, 0, args, wg_sim, wg_first);
}

template <>
hipError_t launch_c2r_batch<float>(int n, unsigned nwg, dim3 b, size_t sh, hipStream_t s, const C2RArgs<float>* args,
                                   const int* wg_sim, const int* wg_first, int cus, int cap) {
  switch (n) {
#define X(n_) \
  case n_:    \
    return c2r_b<n_>(nwg, b, sh, s, cus, cap, args, wg_sim, wg_first);
    KW_SIZES(X)
#undef X
  }
  return hipErrorInvalidValue;
}

}  // namespace kw
