// Kernel launchers, one translation unit per precision and kernel family so
// the many (precision x line length) instantiations build in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "kw_elastic.hpp"
#include "kw_kernels.hpp"

// Supported line-FFT lengths (kw_fft.hpp Plan<N>).
#define KW_SIZES(X) \
  X(8) X(9) X(10) X(12) X(15) X(16) X(20) X(24) X(25) X(27) X(32) X(36) X(40) X(45) X(48) X(64) X(72) X(75) \
  X(80) X(81) X(96) X(100) X(125) X(128) X(135) X(144) X(160) X(192) X(200) X(225) X(243) X(256) X(320)  \
  X(384) X(400) X(512) X(640) X(768) X(1024) X(2048) X(4096)

// Power-of-two plans that carry lines of any other length (Bluestein).
#define KW_BS_SIZES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

namespace kw {

bool size_supported(long long n);
// Plan length of a line of length n without a compiled plan: the power of two
// n_plan >= 2n - 1 (at least 8) on which its chirp-z transform runs; 0 if n is
// longer than the largest plan allows (n > 2048).
inline long long bluestein_plan(long long n) {
  if (n < 1) return 0;
  long long m = 8;
  while (m < 2 * n - 1) m <<= 1;
  return m <= 4096 ? m : 0;
}
// grid of the adjoint's reduction kernels (one partial-sum slot per block)
inline unsigned adj_reduce_blocks(long long n) {
  const long long nb = (n + 255) / 256;
  return nb < 1 ? 1u : (nb < 1024 ? (unsigned)nb : 1024u);
}
// variant 1: contiguous x-line passes, 2: complex-to-real x pass, 4: both for chirp-z lines
int plan_tpl(int n, int variant = 0);
int plan_pitch(int n);

// Return hipSuccess or the launch error; unsupported lengths give hipErrorInvalidValue.
template <typename T>
// pv: plan family (plan_alternative), -1 for the pass's default.
hipError_t launch_dir(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const PassArgs<T>& a,
                      int pv = -1);
template <typename T>
hipError_t launch_spec(int n, dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<T>& a, int pv = -1);
template <typename T>
hipError_t launch_c2r(int n, dim3 g, dim3 b, size_t sh, hipStream_t s, const C2RArgs<T>& a, int pv = -1);
// Fused y inverse + complex-to-real pass, one workgroup per z plane of nx x ny
// (plane_c2r_supported: the plane sizes and precisions it is compiled for,
// given the workgroup LDS limit of the device).
template <typename T>
bool plane_c2r_supported(int nx, int ny, size_t lds_max);
template <typename T>
hipError_t launch_plane_c2r(int nx, int ny, unsigned nz, hipStream_t s, const PlaneArgs<T>& a);
// Bluestein forms (n_plan from bluestein_plan; the args carry the line length and chirp tables).
template <typename T>
hipError_t launch_dir_bs(int n_plan, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const PassArgs<T>& a);
template <typename T>
hipError_t launch_spec_bs(int n_plan, dim3 g, dim3 b, size_t sh, hipStream_t s, const SpecArgs<T>& a);
template <typename T>
hipError_t launch_c2r_bs(int n_plan, dim3 g, dim3 b, size_t sh, hipStream_t s, const C2RArgs<T>& a);
template <typename T>
hipError_t launch_record(unsigned nb, hipStream_t s, const RecArgs<T>& a);
template <typename T>
hipError_t launch_tr_inject(unsigned nb, hipStream_t s, T* r0, T* r1, T* r2, const long long* idx, const T* v, int n);
template <typename T>
hipError_t launch_init_rho(unsigned nb, hipStream_t s, const T* p0, const T* c0sq, T c0sq_s, T ndim, T* p, T* r0, T* r1,
                           T* r2, long long n);
hipError_t launch_finalize_rms(unsigned nb, hipStream_t s, double* a, long long n, double nt);
template <typename T>
hipError_t launch_smooth_helpers(int which, unsigned nb, hipStream_t s, const T* p, Cx<T>* z, unsigned long long* mx,
                                 T* out, long long n);

template <typename T>
hipError_t launch_filter(hipStream_t s, const T* p_ts, const FiltArgs& a);

// Batched forms (one launch over the nwg workgroup tiles of many simulations).
// cus = 0: one workgroup per tile; cus > 0 (the device's compute units): a
// persistent grid of as many workgroups as fit the device at once, at most
// cap (0: no cap), rounded down to a multiple of 8, at most nwg, each
// walking the tiles.
template <typename T>
hipError_t launch_dir_batch(int n, bool contig, bool two, unsigned nwg, dim3 b, size_t sh, hipStream_t s, const PassArgs<T>* args,
                            const int* wg_sim, const int* wg_first, const int* t_dev, int cus = 0, int cap = 0);
template <typename T>
hipError_t launch_spec_batch(int n, int mode, bool two, unsigned nwg, dim3 b, size_t sh, hipStream_t s,
                             const SpecArgs<T>* args, const int* wg_sim, const int* wg_first, int cus = 0, int cap = 0);
template <typename T>
hipError_t launch_c2r_batch(int n, unsigned nwg, dim3 b, size_t sh, hipStream_t s, const C2RArgs<T>* args,
                            const int* wg_sim, const int* wg_first, int cus = 0, int cap = 0);

// Workgroups of `kern` (block threads, sh bytes of LDS) resident per CU of
// the current device, queried once per (device, kernel, block, LDS) and kept.
hipError_t occupancy_per_cu(const void* kern, int block, size_t sh, int* per_cu);

// Launch of one batched kernel over nwg tiles (grid as described above).
template <typename... P, typename... A>
hipError_t launch_tiles(void (*kern)(P...), unsigned nwg, dim3 b, size_t sh, hipStream_t s, int cus, int cap,
                        A... args) {
  unsigned g = nwg;
  if (cus > 0) {
    int per_cu = 0;
    const hipError_t e = occupancy_per_cu(reinterpret_cast<const void*>(kern), (int)b.x, sh, &per_cu);
    if (e != hipSuccess) return e;
    unsigned fit = (unsigned)(per_cu > 0 ? per_cu : 1) * (unsigned)cus;
    if (cap > 0 && (unsigned)cap < fit) fit = (unsigned)cap;
    fit = fit / 8u * 8u;
    if (fit >= 8u && fit < nwg) g = fit;
  }
  hipLaunchKernelGGL(kern, dim3(g), b, sh, s, args..., (int)nwg);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_record_batch(hipStream_t s, const RecArgs<T>* args, const int* sens_sim, const int* sens_first,
                               int n_total, const int* t_dev);
hipError_t launch_tick(hipStream_t s, int* t_dev);

// Adjoint (gradient w.r.t. c0).
template <typename T>
hipError_t launch_adj_pass(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const AdjArgs<T>& a,
                           int pv = -1);
template <typename T>
hipError_t launch_r2c(int n, dim3 g, dim3 b, size_t sh, hipStream_t s, const T* f, Cx<T>* S, int n_lines, int nhp,
                      const Cx<T>* tw, int pv = -1);
template <typename T>
hipError_t launch_adj_small(int which, hipStream_t s, const void* const* p, const double* dargs, long long n, int ndim,
                            int nf = 1);

// Elastic stress-velocity passes and the small per-step kernels around them.
template <typename T>
hipError_t launch_el(int n, bool contig, dim3 g, dim3 b, size_t sh, hipStream_t s, const ElArgs<T>& a);
template <typename T>
hipError_t launch_el_sensor_p(hipStream_t s, T* p, const T* s0, const T* s1, const T* s2, const long long* idx, int n,
                              int ndim);
template <typename T>
hipError_t launch_el_inject(hipStream_t s, T* f, const long long* idx, const T* vals, int n, int dirichlet);

}  // namespace kw
