// Operator-level entry points for parity tests: the host tables the solvers
// build (wavenumbers, PML profiles, staggered derivative multipliers, EOS
// multipliers) and one spectral derivative along one axis on the GPU.
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kw_host.hpp"
#include "kw_launch.hpp"

using namespace kw;
using namespace kwh;

// Scratch of the allocation-free spectral derivative: the line FFT's twiddles
// and the derivative multiplier of the line axis, n complex values each
// (longest line 4096, FP64: 128 KiB).  Every device holds its own copy for the
// life of the process; host code reaches it only through the address that
// hipGetSymbolAddress returns for that device.
constexpr size_t kNoAllocScratchBytes = KW_NOALLOC_SCRATCH_BYTES;
static_assert(KW_NOALLOC_SCRATCH_BYTES == 2 * 4096 * sizeof(Cx<double>), "scratch holds two 4096-point FP64 tables");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace kwh {
void* noalloc_scratch_address(int device) {
  static std::mutex mu;
  static void* addr[64] = {};
  if (device < 0 || device >= 64) raise(KW_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void* p = nullptr;
    HIPC(hipSetDevice(device));
    HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) raise(KW_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    addr[device] = p;
  }
  return addr[device];
}
}  // namespace kwh

namespace {

// The tables of axis_tables()/unshifted_derivative() computed on the device in
// FP64 (same expressions) and cast to T: tw[m] = exp(-2 pi i m/n) and the
// derivative multiplier of `sign` with 1/n folded in.
template <typename T>
__global__ void k_line_tables(Cx<T>* tw, Cx<T>* D, int n, double h, int sign) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  int mm = m;
  double sg = -1.0;
  if (2 * mm > n) {
    mm = n - mm;
    sg = 1.0;
  }
  const double a = 2.0 * M_PI * (double)mm / (double)n;
  tw[m] = Cx<T>{(T)cos(a), (T)(sg * sin(a))};
  const int q = (n % 2 == 0) ? (m < n / 2 ? m : m - n) : (m <= (n - 1) / 2 ? m : m - n);
  const double k = (double)q * (2.0 * M_PI / ((double)n * h));
  const bool real_bin = m == 0 || (n % 2 == 0 && m == n / 2);
  double re, im;
  if (sign == 0) {
    re = 0.0;
    im = (n % 2 == 0 && m == n / 2) ? 0.0 : k / (double)n;
  } else {
    const double c = cos(k * h / 2), s = sin(k * h / 2);
    re = (sign > 0 ? -k * s : k * s) / (double)n;
    im = real_bin ? 0.0 : k * c / (double)n;
  }
  D[m] = Cx<T>{(T)re, (T)im};
}

// Launch shape and arguments of one derivative pass along `axis` (shared by the
// allocating and the allocation-free entry points).
template <typename T>
ElArgs<T> deriv_args(const long long n[3], int ndim, int axis, const PassCfg& c, const T* f, T* out,
                     const Cx<T>* tw, const Cx<T>* D) {
  ElArgs<T> a{};
  a.f = f;
  a.D = D;
  a.tw = tw;
  a.Nx = (int)n[0];
  a.Ny = (int)n[1];
  a.n_lines = (int)(n[1] * n[2]);
  a.x_blocks = (int)(((n[0] + 1) / 2 + c.lpw - 1) / c.lpw);
  a.Sa = axis == 0 ? 1 : (axis == 1 ? n[0] : n[0] * n[1]);
  a.So = axis == 0 ? n[0] : (axis == 1 ? n[0] * n[1] : n[0]);
  a.axis = axis;
  a.ndim = ndim;
  a.mode = EL_DERIV;
  a.out = out;
  a.accumulate = 0;
  return a;
}

template <typename T>
PassCfg deriv_cfg(int dev, const long long n[3], int axis) {
  int lds = 0;
  HIPC(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const long long L = n[axis], lines = n[0] * n[1] * n[2] / L;
  const long long outer = axis == 0 ? 0 : (axis == 1 ? n[2] : n[1]);
  return line_pass_cfg(L, axis == 0, axis == 0 ? lines : n[0], outer, sizeof(Cx<T>), lds > 0 ? (size_t)lds : 65536);
}

// Argument checks shared by both derivative entry points.
void check_deriv_args(int32_t precision, int32_t ndim, const int64_t* N, int32_t axis, int32_t sign) {
  if (ndim < 1 || ndim > 3) raise(KW_ERR_INVALID, "ndim must be 1, 2 or 3");
  if (axis < 0 || axis >= ndim) raise(KW_ERR_INVALID, "axis out of range");
  for (int a = 0; a < ndim; ++a)
    if (N[a] < 1) raise(KW_ERR_INVALID, "grid size must be positive");
  if (precision != 4 && precision != 8) raise(KW_ERR_INVALID, "precision must be 4 or 8");
  if (!size_supported(N[axis])) raise(KW_ERR_UNSUPPORTED, "axis length has no line-FFT plan");
  long long n3[3] = {1, 1, 1};
  for (int a = 0; a < ndim; ++a) n3[a] = N[a];
  check_lane_offsets(ndim, n3, precision);
  if (sign == 0 && axis != 0 && N[axis] % 2 == 0)
    raise(KW_ERR_UNSUPPORTED,
          "an unstaggered derivative along y/z of an even axis couples the Nyquist row across x in the N-D "
          "transform; only the staggered operators (sign +-1) and x are line-separable");
}

template <typename T>
void spectral_gradient(int dev, int ndim, const int64_t* N, const double* d, int axis, int sign, const double* f,
                       double* out) {
  HIPC(hipSetDevice(dev));
  long long n[3] = {1, 1, 1};
  for (int a = 0; a < ndim; ++a) n[a] = N[a];
  const long long nvox = n[0] * n[1] * n[2];
  const long long L = n[axis];
  AxisTables A = axis_tables(L, d[axis], 0, 0.0, 1.0);
  std::vector<double> re, im;
  if (sign > 0) {
    re = A.dpr;
    im = A.dpi;
  } else if (sign < 0) {
    re = A.dmr;
    im = A.dmi;
  } else {
    unshifted_derivative(L, d[axis], re, im);
  }
  DevBuf tw, D, fin, fout;
  upload(tw, to_cx<T>(A.twr, A.twi));
  upload(D, to_cx<T>(re, im));
  upload(fin, to_t<T>(std::vector<double>(f, f + nvox)));
  fout.alloc(nvox * sizeof(T));
  const PassCfg c = deriv_cfg<T>(dev, n, axis);
  const ElArgs<T> a = deriv_args<T>(n, ndim, axis, c, fin.as<T>(), fout.as<T>(), tw.as<Cx<T>>(), D.as<Cx<T>>());
  HIPC(launch_el<T>((int)L, axis == 0, c.grid, c.block, c.shmem, nullptr, a));
  HIPC(hipDeviceSynchronize());
  std::vector<T> h(nvox);
  HIPC(hipMemcpy(h.data(), fout.p, nvox * sizeof(T), hipMemcpyDeviceToHost));
  for (long long i = 0; i < nvox; ++i) out[i] = (double)h[i];
}

// The same derivative on device arrays, with the tables generated into the
// device's scratch array: no allocation, no host copy, enqueued on `stream`.
template <typename T>
void spectral_gradient_noalloc(int dev, void* scratch_addr, int ndim, const int64_t* N, const double* d, int axis,
                               int sign, const void* f, void* out, hipStream_t stream) {
  long long n[3] = {1, 1, 1};
  for (int a = 0; a < ndim; ++a) n[a] = N[a];
  const long long L = n[axis];
  if ((size_t)(2 * L) * sizeof(Cx<T>) > kNoAllocScratchBytes)
    raise(KW_ERR_UNSUPPORTED, "line tables exceed the allocation-free scratch (" +
                                  std::to_string(kNoAllocScratchBytes) + " bytes)");
  HIPC(hipSetDevice(dev));
  Cx<T>* tw = static_cast<Cx<T>*>(scratch_addr);
  Cx<T>* D = tw + L;
  hipLaunchKernelGGL((k_line_tables<T>), dim3((unsigned)((L + 255) / 256)), dim3(256), 0, stream, tw, D, (int)L,
                     d[axis], sign);
  HIPC(hipGetLastError());
  const PassCfg c = deriv_cfg<T>(dev, n, axis);
  const ElArgs<T> a = deriv_args<T>(n, ndim, axis, c, static_cast<const T*>(f), static_cast<T*>(out), tw, D);
  HIPC(launch_el<T>((int)L, axis == 0, c.grid, c.block, c.shmem, stream, a));
}

}  // namespace

extern "C" {

int32_t kw_test_axis_table(int64_t n, double d, int32_t pml_size, double pml_alpha, double dt, int32_t which,
                           double* out) {
  KW_GUARD_BEGIN
  if (!out || n < 1) raise(KW_ERR_INVALID, "bad argument");
  if (pml_size < 0 || pml_size > n) raise(KW_ERR_INVALID, "pml_size out of range");
  AxisTables A = axis_tables(n, d, pml_size, pml_alpha, dt);
  switch (which) {
    case KW_TAB_K: {
      std::vector<double> k = wavenumbers(n, d);
      for (long long i = 0; i < n; ++i) out[i] = k[i];
      break;
    }
    case KW_TAB_PML:
      for (long long i = 0; i < n; ++i) out[i] = A.pml[i];
      break;
    case KW_TAB_PML_SG:
      for (long long i = 0; i < n; ++i) out[i] = A.pml_sg[i];
      break;
    case KW_TAB_DPLUS:
    case KW_TAB_DMINUS: {
      const std::vector<double>& re = which == KW_TAB_DPLUS ? A.dpr : A.dmr;
      const std::vector<double>& im = which == KW_TAB_DPLUS ? A.dpi : A.dmi;
      for (long long i = 0; i < n; ++i) {
        out[2 * i] = re[i] * (double)n;
        out[2 * i + 1] = im[i] * (double)n;
      }
      break;
    }
    default:
      raise(KW_ERR_INVALID, "unknown axis table");
  }
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_test_eos_table(const kw_sim_desc* desc, int32_t which, double* out) {
  KW_GUARD_BEGIN
  if (!desc || !out) raise(KW_ERR_INVALID, "null argument");
  const int nd = desc->ndim;
  if (nd < 1 || nd > 3) raise(KW_ERR_INVALID, "ndim must be 1, 2 or 3");
  long long N[3] = {1, 1, 1};
  for (int a = 0; a < nd; ++a) N[a] = desc->N[a];
  const EosTables E = eos_tables(*desc, nd, N, desc->c0 != nullptr);
  if (which == 1 && !E.stokes) raise(KW_ERR_INVALID, "no second multiplier outside the homogeneous :stokes branch");
  if (which != 0 && which != 1) raise(KW_ERR_INVALID, "unknown EOS table");
  const std::vector<double>& m = which == 0 ? E.m1 : E.m2;
  // expand the |k| octant to the rfft grid (Nx/2+1, Ny, Nz), column-major
  for (long long z = 0; z < N[2]; ++z)
    for (long long y = 0; y < N[1]; ++y) {
      const long long iy = y <= N[1] / 2 ? y : N[1] - y, iz = z <= N[2] / 2 ? z : N[2] - z;
      for (long long x = 0; x < E.hx; ++x) out[x + E.hx * (y + N[1] * z)] = m[x + E.hx * (iz + E.hz * iy)];
    }
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_test_spectral_gradient(kw_ctx* ctx, int32_t precision, int32_t ndim, const int64_t* N, const double* d,
                                  int32_t axis, int32_t sign, const double* f, double* out) {
  KW_GUARD_BEGIN
  if (!ctx || !N || !d || !f || !out) raise(KW_ERR_INVALID, "null argument");
  check_deriv_args(precision, ndim, N, axis, sign);
  if (precision == 4)
    spectral_gradient<float>(ctx->device, ndim, N, d, axis, sign, f, out);
  else if (precision == 8)
    spectral_gradient<double>(ctx->device, ndim, N, d, axis, sign, f, out);
  else
    raise(KW_ERR_INVALID, "precision must be 4 or 8");
  return KW_OK;
  KW_GUARD_END
}

int32_t kw_spectral_gradient_noalloc(kw_ctx* ctx, int32_t precision, int32_t ndim, const int64_t* N,
                                     const double* d, int32_t axis, int32_t sign, const void* f, void* out,
                                     void* stream) {
  KW_GUARD_BEGIN
  if (!ctx || !N || !d || !f || !out) raise(KW_ERR_INVALID, "null argument");
  if (!ctx->scratch) raise(KW_ERR_INVALID, "context has no scratch address (create it with kw_ctx_create)");
  check_deriv_args(precision, ndim, N, axis, sign);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == 4)
    spectral_gradient_noalloc<float>(ctx->device, ctx->scratch, ndim, N, d, axis, sign, f, out, s);
  else
    spectral_gradient_noalloc<double>(ctx->device, ctx->scratch, ndim, N, d, axis, sign, f, out, s);
  return KW_OK;
  KW_GUARD_END
}

}  // extern "C"
