// Operator-level entry points for parity tests: the host tables the solvers
// build (wavenumbers, PML profiles, staggered derivative multipliers, EOS
// multipliers) and one spectral derivative along one axis on the GPU.
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "kw_host.hpp"
#include "kw_launch.hpp"

using namespace kw;
using namespace kwh;

namespace {

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
  int lds = 0;
  HIPC(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const long long lines = nvox / L;
  const long long outer = axis == 0 ? 0 : (axis == 1 ? n[2] : n[1]);
  const PassCfg c = line_pass_cfg(L, axis == 0, axis == 0 ? lines : n[0], outer, sizeof(Cx<T>),
                                  lds > 0 ? (size_t)lds : 65536);
  ElArgs<T> a{};
  a.f = fin.as<T>();
  a.D = D.as<Cx<T>>();
  a.tw = tw.as<Cx<T>>();
  a.Nx = (int)n[0];
  a.Ny = (int)n[1];
  a.n_lines = (int)(n[1] * n[2]);
  a.x_blocks = (int)(((n[0] + 1) / 2 + c.lpw - 1) / c.lpw);
  a.Sa = axis == 0 ? 1 : (axis == 1 ? n[0] : n[0] * n[1]);
  a.So = axis == 0 ? n[0] : (axis == 1 ? n[0] * n[1] : n[0]);
  a.axis = axis;
  a.ndim = ndim;
  a.mode = EL_DERIV;
  a.out = fout.as<T>();
  a.accumulate = 0;
  HIPC(launch_el<T>((int)L, axis == 0, c.grid, c.block, c.shmem, nullptr, a));
  HIPC(hipDeviceSynchronize());
  std::vector<T> h(nvox);
  HIPC(hipMemcpy(h.data(), fout.p, nvox * sizeof(T), hipMemcpyDeviceToHost));
  for (long long i = 0; i < nvox; ++i) out[i] = (double)h[i];
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
  if (ndim < 1 || ndim > 3) raise(KW_ERR_INVALID, "ndim must be 1, 2 or 3");
  if (axis < 0 || axis >= ndim) raise(KW_ERR_INVALID, "axis out of range");
  for (int a = 0; a < ndim; ++a)
    if (N[a] < 1) raise(KW_ERR_INVALID, "grid size must be positive");
  if (!size_supported(N[axis])) raise(KW_ERR_UNSUPPORTED, "axis length has no line-FFT plan");
  if (precision == 4 || precision == 8) {
    long long n3[3] = {1, 1, 1};
    for (int a = 0; a < ndim; ++a) n3[a] = N[a];
    check_lane_offsets(ndim, n3, precision);
  }
  if (sign == 0 && axis != 0 && N[axis] % 2 == 0)
    raise(KW_ERR_UNSUPPORTED,
          "an unstaggered derivative along y/z of an even axis couples the Nyquist row across x in the N-D "
          "transform; only the staggered operators (sign +-1) and x are line-separable");
  if (precision == 4)
    spectral_gradient<float>(ctx->device, ndim, N, d, axis, sign, f, out);
  else if (precision == 8)
    spectral_gradient<double>(ctx->device, ndim, N, d, axis, sign, f, out);
  else
    raise(KW_ERR_INVALID, "precision must be 4 or 8");
  return KW_OK;
  KW_GUARD_END
}

}  // extern "C"
