// Plan queries and the small per-step kernels (recording, injection, init).
#include <map>
#include <mutex>
#include <tuple>

#include "kw_launch.hpp"

namespace kw {

hipError_t occupancy_per_cu(const void* kern, int block, size_t sh, int* per_cu) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, int, size_t>, int> known;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const auto key = std::make_tuple(dev, kern, block, sh);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(key);
    if (it != known.end()) {
      *per_cu = it->second;
      return hipSuccess;
    }
  }
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, kern, block, sh);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(mu);
  known[key] = *per_cu;
  return hipSuccess;
}

bool size_supported(long long n) {
  switch (n) {
#define X(n_) case n_:
    KW_SIZES(X)
#undef X
    return true;
    default:
      return false;
  }
}

int plan_tpl(int n, int variant) {
  switch (n) {
#define X(n_) \
  case n_:    \
    return variant == 4 ? Plan<n_, 4>::TPL                                                  \
                        : (variant == 2 ? Plan<n_, 2>::TPL : (variant ? Plan<n_, 1>::TPL : Plan<n_, 0>::TPL));
    KW_SIZES(X)
#undef X
  }
  return 0;
}

int plan_pitch(int n) {
  switch (n) {
#define X(n_) \
  case n_:    \
    return LdsLayout<n_>::PITCH;
    KW_SIZES(X)
#undef X
  }
  return 0;
}

template <typename T>
hipError_t launch_record(unsigned nb, hipStream_t s, const RecArgs<T>& a) {
  hipLaunchKernelGGL((k_record<T>), dim3(nb), dim3(256), 0, s, a);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_tr_inject(unsigned nb, hipStream_t s, T* r0, T* r1, T* r2, const long long* idx, const T* v, int n) {
  hipLaunchKernelGGL((k_tr_inject<T>), dim3(nb), dim3(256), 0, s, r0, r1, r2, idx, v, n);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_init_rho(unsigned nb, hipStream_t s, const T* p0, const T* c0sq, T c0sq_s, T ndim, T* p, T* r0, T* r1,
                           T* r2, long long n) {
  hipLaunchKernelGGL((k_init_rho<T>), dim3(nb), dim3(256), 0, s, p0, c0sq, c0sq_s, ndim, p, r0, r1, r2, n);
  return hipGetLastError();
}
hipError_t launch_finalize_rms(unsigned nb, hipStream_t s, double* a, long long n, double nt) {
  hipLaunchKernelGGL((k_finalize_rms<double>), dim3(nb), dim3(256), 0, s, a, n, nt);
  return hipGetLastError();
}

// which: 0 = p -> complex z, 1 = max-abs reduction, 2 = rescale into out
template <typename T>
hipError_t launch_smooth_helpers(int which, unsigned nb, hipStream_t s, const T* p, Cx<T>* z, unsigned long long* mx,
                                 T* out, long long n) {
  if (which == 0)
    hipLaunchKernelGGL((k_real_to_cx<T>), dim3(nb), dim3(256), 0, s, p, z, n);
  else if (which == 1)
    hipLaunchKernelGGL((k_maxabs2<T>), dim3(nb), dim3(256), 0, s, p, z, mx, n);
  else
    hipLaunchKernelGGL((k_smooth_finish<T>), dim3(nb), dim3(256), 0, s, z, mx, out, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_filter(hipStream_t s, const T* p_ts, const FiltArgs& a) {
  hipLaunchKernelGGL((k_filter_fwd<T>), dim3(a.ns), dim3(256), 0, s, p_ts, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_filter_inv<T>), dim3(a.ns), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_record_batch(hipStream_t s, const RecArgs<T>* args, const int* sens_sim, const int* sens_first,
                               int n_total, const int* t_dev) {
  hipLaunchKernelGGL((k_record_batch<T>), dim3((unsigned)((n_total + 255) / 256)), dim3(256), 0, s, args, sens_sim,
                     sens_first, n_total, t_dev);
  return hipGetLastError();
}
hipError_t launch_tick(hipStream_t s, int* t_dev) {
  hipLaunchKernelGGL((k_tick<int>), dim3(1), dim3(64), 0, s, t_dev);
  return hipGetLastError();
}

#define KW_INST(T)                                                                                                    \
  template hipError_t launch_record<T>(unsigned, hipStream_t, const RecArgs<T>&);                                  \
  template hipError_t launch_tr_inject<T>(unsigned, hipStream_t, T*, T*, T*, const long long*, const T*, int);      \
  template hipError_t launch_init_rho<T>(unsigned, hipStream_t, const T*, const T*, T, T, T*, T*, T*, T*, long long); \
  template hipError_t launch_smooth_helpers<T>(int, unsigned, hipStream_t, const T*, Cx<T>*, unsigned long long*, T*, \
                                               long long);                                                          \
  template hipError_t launch_filter<T>(hipStream_t, const T*, const FiltArgs&);                                  \
  template hipError_t launch_record_batch<T>(hipStream_t, const RecArgs<T>*, const int*, const int*, int, const int*);
KW_INST(float)
KW_INST(double)
#undef KW_INST

}  // namespace kw
