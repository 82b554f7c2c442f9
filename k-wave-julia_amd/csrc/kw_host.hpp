// Host-side helpers shared by the solver translation units: error
// reporting, device buffers, FP64 grid tables (wavenumbers, PML profiles,
// twiddles, staggered derivative multipliers) and line-pass launch shapes.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <exception>
#include <string>
#include <vector>

#include "../../include/kwave_hip.h"
#include "kw_launch.hpp"

namespace kwh {

using kw::Cx;

inline thread_local std::string g_err;

struct KwError {
  int32_t code;
  std::string msg;
};

[[noreturn]] inline void raise(int32_t code, const std::string& m) { throw KwError{code, m}; }

#define HIPC(x)                                                                         \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) raise(KW_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// Device memory
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t b) {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = b;
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      raise(KW_ERR_OOM, "hipMalloc of " + std::to_string(b) + " bytes failed: " + hipGetErrorString(e));
    }
  }
  template <typename U>
  U* as() const {
    return reinterpret_cast<U*>(p);
  }
};

template <typename U>
void upload(DevBuf& b, const std::vector<U>& h) {
  b.alloc(h.size() * sizeof(U));
  if (!h.empty()) HIPC(hipMemcpy(b.p, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice));
}

// ---------------------------------------------------------------------------
// Host tables (FP64, cast once to the kernel precision)
// ---------------------------------------------------------------------------
// grid.jl:120-130
inline std::vector<double> wavenumbers(long long n, double d) {
  std::vector<double> k(n);
  const double dk = 2.0 * M_PI / ((double)n * d);
  for (long long i = 0; i < n; ++i) {
    long long m;
    if (n % 2 == 0)
      m = (i < n / 2) ? i : i - n;
    else
      m = (i <= (n - 1) / 2) ? i : i - n;
    k[i] = (double)m * dk;
  }
  return k;
}

// pml.jl:25-55
inline std::vector<double> get_pml(long long n, int size, double alpha, double dt, bool staggered) {
  std::vector<double> p(n, 1.0);
  for (int i = 1; i <= size; ++i) {
    const double pos = staggered ? ((double)(size - i) + 0.5) / size : (double)(size - i + 1) / size;
    const double v = std::exp(-(alpha * std::pow(pos, 4)) * dt);
    p[i - 1] = v;
    p[n - i] = v;
  }
  return p;
}

// exp(-2 pi i m / n) with exact symmetric reduction of the angle
inline void twiddles(long long n, std::vector<double>& re, std::vector<double>& im) {
  re.resize(n);
  im.resize(n);
  for (long long m = 0; m < n; ++m) {
    // angle = -2 pi m/n; evaluate via the smallest equivalent argument
    long long mm = m;
    double s = -1.0;
    if (2 * mm > n) {
      mm = n - mm;
      s = 1.0;
    }
    const double a = 2.0 * M_PI * (double)mm / (double)n;
    re[m] = std::cos(a);
    im[m] = s * std::sin(a);
  }
}

template <typename T>
std::vector<Cx<T>> to_cx(const std::vector<double>& re, const std::vector<double>& im) {
  std::vector<Cx<T>> o(re.size());
  for (size_t i = 0; i < re.size(); ++i) o[i] = Cx<T>{(T)re[i], (T)im[i]};
  return o;
}
template <typename T>
std::vector<T> to_t(const std::vector<double>& v) {
  std::vector<T> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = (T)v[i];
  return o;
}

// Chirp-z (Bluestein) tables of a line of length n on the power-of-two plan m
// (kw_fft.hpp line_dft), FP64: c_j = exp(-i pi (j^2 mod 2n) / n) for j < n;
// gf = FFT_m(g) / m with g_k = conj(c_k) for 0 <= k < n and g_(m-k) = conj(c_k)
// for 0 < k < n (zero elsewhere); gb the same for the filter of c (inverse
// transform).
struct ChirpTables {
  std::vector<double> cr, ci, gfr, gfi, gbr, gbi;
};
// In-place forward DFT of length m (a power of two), iterative radix 2.
inline void host_fft_pow2(std::vector<double>& re, std::vector<double>& im) {
  const size_t m = re.size();
  for (size_t i = 1, j = 0; i < m; ++i) {
    size_t bit = m >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (size_t len = 2; len <= m; len <<= 1) {
    for (size_t i = 0; i < m; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)len;
        const double wr = std::cos(a), wi = std::sin(a);
        const size_t u = i + k, v = i + k + len / 2;
        const double xr = re[v] * wr - im[v] * wi, xi = re[v] * wi + im[v] * wr;
        re[v] = re[u] - xr;
        im[v] = im[u] - xi;
        re[u] += xr;
        im[u] += xi;
      }
  }
}
inline ChirpTables chirp_tables(long long n, long long m) {
  ChirpTables C;
  C.cr.resize(n);
  C.ci.resize(n);
  for (long long j = 0; j < n; ++j) {
    const long long q = (j * j) % (2 * n);
    const double a = -M_PI * (double)q / (double)n;
    C.cr[j] = std::cos(a);
    C.ci[j] = std::sin(a);
  }
  for (int dir = 0; dir < 2; ++dir) {
    std::vector<double> gr(m, 0.0), gi(m, 0.0);
    const double sg = dir == 0 ? -1.0 : 1.0;   // forward filter conj(c), inverse filter c
    for (long long k = 0; k < n; ++k) {
      gr[k] = C.cr[k];
      gi[k] = sg * C.ci[k];
      if (k > 0) {
        gr[m - k] = C.cr[k];
        gi[m - k] = sg * C.ci[k];
      }
    }
    host_fft_pow2(gr, gi);
    for (long long k = 0; k < m; ++k) {
      gr[k] /= (double)m;
      gi[k] /= (double)m;
    }
    if (dir == 0) {
      C.gfr = gr;
      C.gfi = gi;
    } else {
      C.gbr = gr;
      C.gbi = gi;
    }
  }
  return C;
}

struct AxisTables {
  std::vector<double> twr, twi, dpr, dpi, dmr, dmi, pml, pml_sg;
};

// Staggered derivative multipliers i k exp(+-i k h/2) (grid.jl:149-153,
// fft_utils.jl:146-175) with the inverse 1/N folded in.  The reference's
// irfft keeps only the real part of the DC and Nyquist bins of the product;
// taking the real part of the multiplier there reproduces that exactly for
// the pair-of-real-lines transform.
inline AxisTables axis_tables(long long n, double h, int pml_size, double pml_alpha, double dt) {
  AxisTables A;
  twiddles(n, A.twr, A.twi);
  std::vector<double> k = wavenumbers(n, h);
  A.dpr.resize(n);
  A.dpi.resize(n);
  A.dmr.resize(n);
  A.dmi.resize(n);
  for (long long i = 0; i < n; ++i) {
    const double c = std::cos(k[i] * h / 2), s = std::sin(k[i] * h / 2);
    // i k (c + i s) = -k s + i k c ; i k (c - i s) = k s + i k c
    double pr = -k[i] * s, pi = k[i] * c;
    double mr = k[i] * s, mi = k[i] * c;
    if (i == 0 || (n % 2 == 0 && i == n / 2)) {
      pi = 0.0;
      mi = 0.0;
    }
    A.dpr[i] = pr / (double)n;
    A.dpi[i] = pi / (double)n;
    A.dmr[i] = mr / (double)n;
    A.dmi[i] = mi / (double)n;
  }
  A.pml = get_pml(n, pml_size, pml_alpha, dt, false);
  A.pml_sg = get_pml(n, pml_size, pml_alpha, dt, true);
  return A;
}

struct PassCfg {
  int n = 0;        // line length
  int lpw = 1;      // lines (pairs) per workgroup
  int req = 1;      // lines per workgroup asked for (before shrinking to small grids)
  int pv = -1;      // plan family (plan_alternative) or -1 for the pass's default
  dim3 grid, block;
  size_t shmem = 0;
};

// Lines per workgroup of a line pass: 256 threads on contiguous lines, >= 16
// line pairs up to 512 threads on strided lines, halved until the lines fit
// the workgroup's LDS and to the lines that exist.  These passes (elastic,
// operator tests) use plan family 2 on contiguous lines.
// fill_wgs > 0: on small grids, fewer lines per workgroup until the launch has
// fill_wgs workgroups (keeping one full wave per workgroup).
inline PassCfg line_pass_cfg(long long n, bool contig, long long lines, long long outer, size_t cx_bytes,
                             size_t lds_max, long long fill_wgs = 0) {
  PassCfg c;
  c.n = (int)n;
  const int tpl = kw::plan_tpl((int)n, contig ? 2 : 0);
  const size_t line_bytes = (size_t)kw::plan_pitch((int)n) * cx_bytes;
  int lpw = (contig ? 256 : std::max(256, std::min(512, 16 * tpl))) / tpl;
  while (lpw > 1 && (size_t)lpw * line_bytes > lds_max) lpw /= 2;
  const long long pairs = (lines + 1) / 2;
  while (lpw > 1 && lpw / 2 >= pairs) lpw /= 2;
  {
    const int min_lpw = std::max(1, 64 / tpl);
    const long long o = contig ? 1 : std::max<long long>(outer, 1);
    while (lpw / 2 >= min_lpw && ((pairs + lpw - 1) / lpw) * o < fill_wgs) lpw /= 2;
  }
  c.lpw = c.req = lpw;
  c.block = dim3(lpw * tpl);
  c.shmem = (size_t)lpw * line_bytes;
  if (c.shmem > lds_max)
    raise(KW_ERR_UNSUPPORTED, "line of length " + std::to_string(n) + " exceeds the LDS of a workgroup");
  const long long blocks = (pairs + lpw - 1) / lpw;
  c.grid = dim3((unsigned)(contig ? blocks : blocks * std::max<long long>(outer, 1)));
  return c;
}

// Line passes address a uniform 64-bit workgroup base plus 32-bit unsigned
// per-lane byte offsets (scalar-base + vector-offset loads).  The largest
// lane offset is the span of one strided line: stride x (length - 1) plus the
// lines of one workgroup (< 1024 elements), in the element size of the
// buffer the pass walks.  Strided field passes walk T elements with strides
// Nx (y lines) and Nx*Ny (z lines); spectral passes walk complex elements of
// the half spectrum with strides NHP (y) and NHP*Ny (z), NHP = Nx/2+1 rounded
// up to 16; the p0 smoothing pass walks complex copies of the full field
// with strides Nx and Nx*Ny.  Contiguous x-line passes stay below
// 1024 * Nx complex elements.
inline unsigned long long max_lane_offset_bytes(int nd, const long long N[3], int prec) {
  const unsigned long long s = (unsigned long long)prec, cx = 2 * s;
  const unsigned long long nx = (unsigned long long)N[0], ny = nd > 1 ? (unsigned long long)N[1] : 1ULL,
                           nz = nd > 2 ? (unsigned long long)N[2] : 1ULL;
  const unsigned long long nhp = (nx / 2 + 1 + 15) / 16 * 16, w = 1024;
  unsigned long long m = (w * nx) * cx;
  auto upd = [&](unsigned long long stride, unsigned long long len, unsigned long long elem) {
    if (len > 1) m = std::max(m, (stride * (len - 1) + w) * elem);
  };
  upd(nx, ny, cx);           // y lines: fields (T) and smoothing (complex)
  upd(nx * ny, nz, cx);      // z lines: fields (T) and smoothing (complex)
  upd(nhp, ny, cx);          // spectral y
  upd(nhp * ny, nz, cx);     // spectral z
  return m;
}
// Grids whose lane offsets fit 32 bits (every pass, both precisions' element
// sizes as given); larger grids are rejected before any allocation.
inline bool lane_offsets_fit(int nd, const long long N[3], int prec) {
  return max_lane_offset_bytes(nd, N, prec) <= 0xFFFFFFFFULL;
}
inline void check_lane_offsets(int nd, const long long N[3], int prec) {
  if (!lane_offsets_fit(nd, N, prec))
    raise(KW_ERR_UNSUPPORTED, "grid " + std::to_string(N[0]) + "x" + std::to_string(nd > 1 ? N[1] : 1) + "x" +
                                  std::to_string(nd > 2 ? N[2] : 1) + " in " + (prec == 8 ? "Float64" : "Float32") +
                                  ": a strided line spans more than 4 GiB, beyond the 32-bit lane offsets of the "
                                  "line passes");
}

// EOS multipliers on the octant of |k| (first_order.jl:13-42, 361-375,
// 581-595; first_order_steps.jl:245-289), FP64, without the 1/N of the
// inverse transform: m1 = c^2 kappa (1 + tau |k|^y), m2 = -c^2 kappa eta |k|^(y-1)
// (homogeneous :stokes only); c^2 = 1 for heterogeneous c0 (applied in real
// space).  Layout [|ky|][|kz|][kx]: index ix + hx (iz + hz iy).
struct EosTables {
  long long hx = 1, hy = 1, hz = 1;
  std::vector<double> m1, m2;
  bool absorbing = false, stokes = false;
  double tau = 0, eta = 0;
};
inline EosTables eos_tables(const kw_sim_desc& d, int nd, const long long N[3], bool het_c) {
  EosTables E;
  E.hx = N[0] / 2 + 1;
  E.hy = N[1] / 2 + 1;
  E.hz = N[2] / 2 + 1;
  std::vector<double> kx = wavenumbers(N[0], d.d[0]);
  std::vector<double> ky = nd > 1 ? wavenumbers(N[1], d.d[1]) : std::vector<double>{0.0};
  std::vector<double> kz = nd > 2 ? wavenumbers(N[2], d.d[2]) : std::vector<double>{0.0};
  const double c_ref = d.c_ref, y = d.alpha_power;
  E.absorbing = d.absorbing != 0 && d.alpha_mode != KW_NO_ABSORPTION;
  if (E.absorbing) {
    const double a_np = d.alpha_coeff * (1e-2 / (20.0 * std::log10(std::exp(1.0)))) * std::pow(2.0 * M_PI * 1e6, -y);
    E.tau = -2.0 * a_np * std::pow(c_ref, y - 1.0);
    E.eta = 2.0 * a_np * std::pow(c_ref, y - 1.0) * std::tan(M_PI * y / 2.0);
  }
  E.stokes = E.absorbing && !het_c && d.alpha_mode != KW_NO_DISPERSION;
  const double c2 = het_c ? 1.0 : d.c0_scalar * d.c0_scalar;
  E.m1.resize(E.hx * E.hy * E.hz);
  if (E.stokes) E.m2.resize(E.m1.size());
  for (long long iz = 0; iz < E.hz; ++iz)
    for (long long iy = 0; iy < E.hy; ++iy)
      for (long long ix = 0; ix < E.hx; ++ix) {
        const double kk = std::sqrt(kx[ix] * kx[ix] + ky[iy] * ky[iy] + kz[iz] * kz[iz]);
        const double arg = c_ref * kk * d.dt / 2;
        const double kap = arg == 0.0 ? 1.0 : std::sin(arg) / arg;
        double v1 = c2 * kap, v2 = 0.0;
        if (E.absorbing) {
          v1 = c2 * kap * (1.0 + E.tau * std::pow(kk, y));
          if (E.stokes) v2 = -c2 * kap * (E.eta * std::pow(kk, y - 1.0));
        }
        const long long i = ix + E.hx * (iz + E.hz * iy);
        E.m1[i] = v1;
        if (E.stokes) E.m2[i] = v2;
      }
  return E;
}

// Unstaggered derivative multiplier i k / n (real part kept at the Nyquist bin,
// as a complex-to-real inverse transform does).
inline void unshifted_derivative(long long n, double h, std::vector<double>& re, std::vector<double>& im) {
  std::vector<double> k = wavenumbers(n, h);
  re.assign(n, 0.0);
  im.resize(n);
  for (long long i = 0; i < n; ++i) im[i] = (n % 2 == 0 && i == n / 2) ? 0.0 : k[i] / (double)n;
}

}  // namespace kwh

// Device context behind kw_ctx: the device selection and the device address
// of that device's copy of the allocation-free entry point's scratch array
// (kw_ops.hip), resolved once at kw_ctx_create.
struct kw_ctx {
  int device;
  void* scratch;
};

namespace kwh {
// Device address of the allocation-free scratch array on `device` (the
// current device when called); resolved once per device, then cached.
void* noalloc_scratch_address(int device);
}  // namespace kwh

// Every C entry point runs its body inside these: KwError carries the status
// code, anything else maps to KW_ERR_INVALID; the message goes to kw_last_error().
#define KW_GUARD_BEGIN try {
#define KW_GUARD_END                     \
  }                                      \
  catch (const kwh::KwError& e) {        \
    kwh::g_err = e.msg;                  \
    return e.code;                       \
  }                                      \
  catch (const std::exception& e) {      \
    kwh::g_err = e.what();               \
    return KW_ERR_INVALID;               \
  }                                      \
  catch (...) {                          \
    kwh::g_err = "unknown error";        \
    return KW_ERR_INVALID;               \
  }
