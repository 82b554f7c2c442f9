// Plan queries and the small per-step kernels (recording, injection, init).
#include "kw_launch.hpp"

namespace kw {

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

int plan_tpl(int n) {
  switch (n) {
#define X(n_) \
  case n_:    \
    return Plan<n_>::TPL;
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

#define KW_INST(T)                                                                                                    \
  template hipError_t launch_record<T>(unsigned, hipStream_t, const RecArgs<T>&);                                  \
  template hipError_t launch_tr_inject<T>(unsigned, hipStream_t, T*, T*, T*, const long long*, const T*, int);      \
  template hipError_t launch_init_rho<T>(unsigned, hipStream_t, const T*, const T*, T, T, T*, T*, T*, T*, long long);
KW_INST(float)
KW_INST(double)
#undef KW_INST

}  // namespace kw
