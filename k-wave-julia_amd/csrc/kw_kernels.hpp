// Time-step kernels of the k-space first-order PSTD loop for gfx950.
//
// One time step (first_order_steps.jl:296-535) is split into passes over
// lines of the grid:
//   * a directional pass per axis d (y, z, then x): for every line along d
//     it computes  dp/dd -> u_d update -> [u source] -> du_d/dd -> rho_d
//     update -> [p source]  entirely in registers/LDS.  This is exact because
//     the staggered derivative multiplier i k_d exp(+-i k_d h/2) depends on
//     k_d only and is Hermitian, so the N-D rfft/irfft of spectral_gradient!
//     (fft_utils.jl:126-182) equals a 1-D transform along d.
//   * the x pass also forms the equation-of-state input (c0^2) rho(1 + B/A
//     rho/(2 rho0)) and does the real-to-complex stage of the EOS transform;
//   * spectral passes along y/z (3-D: y fwd, z fwd*M*inv, y inv; 2-D: y
//     fwd*M*inv) with M = (c^2) kappa (1 + nabla1) / N read from an octant
//     table; and a complex-to-real pass along x that writes p.
// Two real lines travel as one complex line (re = line a, im = line b);
// the derivative and EOS multipliers are Hermitian so the pair separates
// exactly on the way back.
#pragma once
#include <hip/hip_fp16.h>

#include <type_traits>

#include "kw_fft.hpp"

namespace kw {

// Workgroup placement.  Workgroups are dispatched round-robin over the 8 XCDs
// (hardware workgroup b runs on XCD b % 8), each with its own L2.  xcd_wg maps
// b to a logical workgroup so that consecutive logical workgroups share an
// XCD; it is a bijection on [0, n).
__device__ __forceinline__ int xcd_wg(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7, s = b >> 3;
  return x < r ? x * (q + 1) + s : r * (q + 1) + (x - r) * q + s;
}
// Strided and spectral passes whose workgroups cover less than 128 bytes of
// each row even at 512 threads (lines of 768 points and more in FP32: 8 line
// pairs) use it, so the two workgroups reading the halves of one 128-byte
// segment run on one XCD and the second read hits its L2.  Wider rows keep
// the plain order (decided at compile time: a runtime test cost the 256-point
// passes about 1 %).
template <typename T, int N>
__device__ __forceinline__ int strided_wg(int b, int n) {
  if constexpr ((512 / Plan<N, 0>::TPL) * 2 * (int)sizeof(T) < 128)
    return xcd_wg(b, n);
  else
    return b;
}
template <typename T, int N>
__device__ __forceinline__ int strided_wg(int b) {
  return strided_wg<T, N>(b, (int)gridDim.x);
}

// ---------------------------------------------------------------------------
// Arguments
// ---------------------------------------------------------------------------
struct SrcList {
  const int* off;   // per-workgroup CSR offsets (n_wg + 1) or null
  const int* pos;   // packed (line c, part, element n): n | part<<20 | c<<21
  const int* row;   // source row j
  const void* val;  // T values, time-major [col][n_src]
  int n_src;
  int ncols;
  int dirichlet;
};

template <typename T>
struct PassArgs {
  // fields of this direction
  const T* p;
  T* u;
  T* rho;
  // medium (maps in T, or FP16 offsets from rho0_off / c0_off when maps_f16)
  const void* rho0;  // map or null
  T rho0_s;
  int maps_f16;
  T rho0_off;
  T c0_off;
  // line tables (length N of this axis)
  const T* pml;
  const T* pml_sg;
  const Cx<T>* Dp;  // i k exp(+i k h/2) / N
  const Cx<T>* Dm;  // i k exp(-i k h/2) / N
  const Cx<T>* tw;  // exp(-2 pi i m / N)
  T dt;
  // geometry
  int Nx;          // line count along x pairing (strided) / line length (contig)
  long long Sa;    // element stride along the line
  long long So;    // stride of the outer index
  int n_outer;     // outer count
  int n_lines;     // contiguous: number of real lines
  int x_blocks;    // strided: workgroups along x
  // sources
  SrcList psrc;
  SrcList usrc;
  int t;     // 1-based time index
  int mode;  // 0 = step, 1 = p0 init (u only, dt/2, no PML)
  // fused EOS (x pass only)
  int eos;             // 0 none, 1 fused r2c
  const T* rho_o1;     // other split density components (null if absent)
  const T* rho_o2;
  const void* c0sq;    // c0^2 map (FP16 mode: c0 - c0_off) or null (homogeneous: folded into M)
  const T* bona;       // B/A map or null
  T bona_s;            // scalar B/A (when bona == null)
  int nonlinear;
  Cx<T>* S1;           // half spectrum (Nx/2+1 used of a row pitch nhp, Ny, Nz)
  Cx<T>* S2;           // second spectrum for the homogeneous :stokes branch
  T* rho_prev;         // previous total density (:stokes)
  T inv_dt;
  int nhp;             // half-spectrum row pitch (Nx/2+1 rounded up to 16 complex)
  T* rho_sum_out;      // if set, the x pass also stores sum_d rho_d (adjoint recompute)
  const T* inv_r;      // axisymmetric radial pass: 1/r along the line (else null)
  // lines of any length (Bluestein kernels): line length and chirp tables
  int nl;
  Chirp<T> ch;
};

// ---------------------------------------------------------------------------
// Line indexing
// ---------------------------------------------------------------------------
// Contiguous (x) lines: thread -> (t = tid % TPL, c = tid / TPL); pair pi =
// blockIdx.x * LPW + c holds real lines 2pi (re) and 2pi+1 (im).
// Strided lines: thread -> (c = tid % LPW, t = tid / LPW); the workgroup
// covers x in [2*LPW*xb, 2*LPW*(xb+1)) at outer index o; pair c holds x =
// x0 + 2c (re) and x0 + 2c + 1 (im).

// x / y in the kernel precision; Float32 uses the hardware reciprocal
// (<= 1 ulp) so the division costs two instructions instead of ten.
template <typename T>
__device__ __forceinline__ T fdiv(T x, T y) {
  return x / y;
}
template <>
__device__ __forceinline__ float fdiv<float>(float x, float y) {
  return x * __builtin_amdgcn_rcpf(y);
}

// Element `i` of `base` addressed as base + zero-extended 32-bit byte offset,
// which maps to the scalar-base + vector-offset form of global loads/stores.
template <typename U>
__device__ __forceinline__ const U& at32(const U* base, unsigned i) {
  return *reinterpret_cast<const U*>(reinterpret_cast<const char*>(base) + (unsigned)(i * (unsigned)sizeof(U)));
}
template <typename U>
__device__ __forceinline__ U& at32(U* base, unsigned i) {
  return *reinterpret_cast<U*>(reinterpret_cast<char*>(base) + (unsigned)(i * (unsigned)sizeof(U)));
}

// Argument records and workgroup maps of the batched kernels are written by
// the host before a launch and only read by kernels.  Read through the
// constant address space they become scalar loads, and the pointers they
// hold are then taken to be global (as for kernel arguments), so field
// accesses are global scalar-base + vector-offset instructions instead of
// flat ones (64-bit per-lane addresses, LDS wait counter shared with them).
// (The host pass of the compiler only checks kernel bodies; it keeps plain
// pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define KW_CONST __attribute__((address_space(4)))
#else
#define KW_CONST
#endif
template <typename U>
__device__ __forceinline__ const KW_CONST U* cmem(const U* p) {
  return (const KW_CONST U*)p;
}

// Addresses are a per-workgroup 64-bit base (uniform) plus 32-bit lane
// offsets, so every load is a scalar-base + vector-offset access.
// BS: the line has a runtime length nl < N (Bluestein lines on the N-point
// plan); register slots j >= nl hold no element and are never loaded/stored.
template <typename T, int N, int V = 0, bool BS = false>
struct LineCtx {
  static constexpr int TPL = Plan<N, V>::TPL;
  static constexpr int E = Plan<N, V>::E;
  using V2 = typename std::conditional<sizeof(T) == 4, float2, double2>::type;
  int nl = N;           // line length (elements)
  int t, c;
  bool va, vb;          // real line a / b valid
  bool vec = false;     // strided pairs are adjacent and aligned: one 2-element access per pair
  long long wgb;        // workgroup base element offset (uniform)
  unsigned oa, ob;      // lane offsets of lines a / b from wgb
  unsigned sa;          // element stride along the line
  __device__ __forceinline__ unsigned ia(int q) const { return oa + sa * (unsigned)(t + TPL * q); }
  __device__ __forceinline__ unsigned ib(int q) const { return ob + sa * (unsigned)(t + TPL * q); }
  __device__ __forceinline__ int n(int q) const { return t + TPL * q; }
  __device__ __forceinline__ bool ok(int q) const { return !BS || t + TPL * q < nl; }   // slot holds an element
  __device__ __forceinline__ unsigned nx(int q) const {  // offset of element n+1 (periodic)
    int nn = t + TPL * q + 1;
    nn = (nn == (BS ? nl : N)) ? 0 : nn;
    return sa * (unsigned)nn;
  }
  // entry n(q) of a per-line table of nl values (0 for slots without an element)
  template <typename U>
  __device__ __forceinline__ U tab(const U* p, int q) const {
    return ok(q) ? p[t + TPL * q] : U{};
  }

  // `fill` is what an absent line of the pair holds (1 for divisors like rho0)
  __device__ __forceinline__ void load(const T* __restrict__ f0, Cx<T> (&v)[E], T fill = (T)0) const {
    const T* __restrict__ f = f0 + wgb;
    if (vec) {  // va == vb here
      const V2* __restrict__ f2 = reinterpret_cast<const V2*>(f);
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const V2 r = (va && ok(q)) ? at32(f2, ia(q) >> 1) : V2{fill, fill};
        v[q] = Cx<T>{r.x, r.y};
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      v[q].x = (va && ok(q)) ? at32(f, ia(q)) : fill;
      v[q].y = (vb && ok(q)) ? at32(f, ib(q)) : fill;
    }
  }
  __device__ __forceinline__ void store(T* __restrict__ f0, const Cx<T> (&v)[E]) const {
    T* __restrict__ f = f0 + wgb;
    if (vec) {
      V2* __restrict__ f2 = reinterpret_cast<V2*>(f);
#pragma unroll
      for (int q = 0; q < E; ++q)
        if (va && ok(q)) at32(f2, ia(q) >> 1) = V2{v[q].x, v[q].y};
      return;
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if (va && ok(q)) at32(f, ia(q)) = v[q].x;
      if (vb && ok(q)) at32(f, ib(q)) = v[q].y;
    }
  }
  // pair of map values at element offset i (even) of the workgroup base
  __device__ __forceinline__ Cx<T> map_pair(const void* f, unsigned i, bool f16, T off) const {
    if (f16) {
      const __half2 h = at32(reinterpret_cast<const __half2*>(reinterpret_cast<const __half*>(f) + wgb), i >> 1);
      return Cx<T>{off + (T)__low2float(h), off + (T)__high2float(h)};
    }
    const V2 r = at32(reinterpret_cast<const V2*>(reinterpret_cast<const T*>(f) + wgb), i >> 1);
    return Cx<T>{r.x, r.y};
  }
  // static medium map in T, or FP16 offset from `off` (converted to T)
  __device__ __forceinline__ T map_at(const void* f, unsigned i, bool f16, T off) const {
    return f16 ? off + (T)__half2float(at32(reinterpret_cast<const __half*>(f) + wgb, i))
               : at32(reinterpret_cast<const T*>(f) + wgb, i);
  }
  // The storage test is hoisted out of the element loops so each loop is one
  // straight run of independent loads.
  __device__ __forceinline__ void load_map(const void* f, Cx<T> (&v)[E], bool f16, T off, T fill) const {
    if (vec) {
      if (f16) {
        load_half_pairs(f, v, off, fill, 0);
      } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = (va && ok(q)) ? map_pair(f, ia(q), false, off) : Cx<T>{fill, fill};
      }
      return;
    }
    if (f16) {
      load_half_lines(f, v, off, fill, 0);
    } else {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        v[q].x = (va && ok(q)) ? map_at(f, ia(q), false, off) : fill;
        v[q].y = (vb && ok(q)) ? map_at(f, ib(q), false, off) : fill;
      }
    }
  }
  // FP16 map pairs (strided lines): raw 32-bit loads under one lane test, then conversion
  __device__ __forceinline__ void load_half_pairs(const void* f, Cx<T> (&v)[E], T off, T fill, int next) const {
    const unsigned* __restrict__ h2 =
        reinterpret_cast<const unsigned*>(reinterpret_cast<const unsigned short*>(f) + wgb);
    unsigned r[E] = {};
    if (va) {
#pragma unroll
      for (int q = 0; q < E; ++q) r[q] = ok(q) ? at32(h2, (next ? oa + nx(q) : ia(q)) >> 1) : 0u;
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const __half lo = __ushort_as_half((unsigned short)(r[q] & 0xFFFFu));
      const __half hi = __ushort_as_half((unsigned short)(r[q] >> 16));
      v[q] = (va && ok(q)) ? Cx<T>{off + (T)__half2float(lo), off + (T)__half2float(hi)} : Cx<T>{fill, fill};
    }
  }
  // FP16 map elements of lines a and b: all raw 16-bit loads of a line are
  // issued under one lane test, then converted (a conversion inside each
  // conditional load made every load wait for the previous one).
  __device__ __forceinline__ void load_half_lines(const void* f, Cx<T> (&v)[E], T off, T fill, int next) const {
    const unsigned short* __restrict__ h = reinterpret_cast<const unsigned short*>(f) + wgb;
    unsigned short ra[E] = {}, rb[E] = {};
    if (va) {
#pragma unroll
      for (int q = 0; q < E; ++q) ra[q] = ok(q) ? at32(h, next ? oa + nx(q) : ia(q)) : (unsigned short)0;
    }
    if (vb) {
#pragma unroll
      for (int q = 0; q < E; ++q) rb[q] = ok(q) ? at32(h, next ? ob + nx(q) : ib(q)) : (unsigned short)0;
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      v[q].x = (va && ok(q)) ? off + (T)__half2float(__ushort_as_half(ra[q])) : fill;
      v[q].y = (vb && ok(q)) ? off + (T)__half2float(__ushort_as_half(rb[q])) : fill;
    }
  }
  __device__ __forceinline__ void load_map_next(const void* f, Cx<T> (&v)[E], bool f16, T off, T fill) const {
    if (vec) {
      if (f16) {
        load_half_pairs(f, v, off, fill, 1);
      } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = (va && ok(q)) ? map_pair(f, oa + nx(q), false, off) : Cx<T>{fill, fill};
      }
      return;
    }
    if (f16) {
      load_half_lines(f, v, off, fill, 1);
    } else {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        v[q].x = (va && ok(q)) ? map_at(f, oa + nx(q), false, off) : fill;
        v[q].y = (vb && ok(q)) ? map_at(f, ob + nx(q), false, off) : fill;
      }
    }
  }
};

// VV selects the plan family (default: 1 for contiguous x lines, 0 for strided).
template <typename T, int N, bool CONTIG, int VV = (CONTIG ? 1 : 0), bool BS = false>
__device__ __forceinline__ LineCtx<T, N, VV, BS> make_ctx(const PassArgs<T>& a, int wg, int& lpw_out) {
  constexpr int V = VV;
  constexpr int TPL = Plan<N, V>::TPL;
  LineCtx<T, N, V, BS> L;
  L.nl = BS ? a.nl : N;
  const int tid = threadIdx.x;
  const int lpw = blockDim.x / TPL;
  lpw_out = lpw;
  if constexpr (CONTIG) {
    L.t = tid % TPL;
    L.c = tid / TPL;
    const long long pi = (long long)wg * lpw + L.c;
    L.va = 2 * pi < a.n_lines;
    L.vb = 2 * pi + 1 < a.n_lines;
    L.wgb = 2LL * wg * lpw * L.nl;
    L.oa = (unsigned)(2 * L.c * L.nl);
    L.ob = L.oa + L.nl;
    L.sa = 1;
  } else {
    L.c = tid % lpw;
    L.t = tid / lpw;
    const int xb = wg % a.x_blocks;
    const int o = wg / a.x_blocks;
    const int x0 = 2 * xb * lpw;
    const int x = x0 + 2 * L.c;
    L.va = x < a.Nx;
    L.vb = x + 1 < a.Nx;
    L.wgb = (long long)x0 + a.So * o;
    L.oa = (unsigned)(2 * L.c);
    L.ob = L.oa + 1;
    L.sa = (unsigned)a.Sa;
    // even Nx, Sa and So: every valid pair is complete and 2-element aligned
    L.vec = ((a.Nx | a.Sa | a.So) & 1) == 0;
  }
  return L;
}

// Apply the workgroup's source entries to the pair-line registers through LDS.
// pos packs (c << 21) | (part << 20) | n.
template <typename T, int N, int V, bool BS = false>
__device__ __forceinline__ void inject(const SrcList& s, int t_index, Cx<T> (&v)[(Plan<N, V>::E)], Cx<T>* lds_all,
                                       const LineCtx<T, N, V, BS>& L, int wg) {
  if (s.off == nullptr) return;
  const int e0 = s.off[wg], e1 = s.off[wg + 1];
  if (e0 >= e1) return;  // uniform across the workgroup
  constexpr int E = Plan<N, V>::E;
  constexpr int P = LdsLayout<N>::PITCH;
  Cx<T>* lds = lds_all + L.c * P;
#pragma unroll
  for (int q = 0; q < E; ++q) lds[LdsLayout<N>::at(L.n(q))] = v[q];
  __syncthreads();
  const int col = (t_index < s.ncols ? t_index : s.ncols) - 1;
  const T* vals = reinterpret_cast<const T*>(s.val) + (long long)col * s.n_src;
  for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
    const int pk = s.pos[e];
    const int n = pk & 0xFFFFF, part = (pk >> 20) & 1, c = pk >> 21;
    const T val = vals[s.row[e]];
    Cx<T>& cell = lds_all[c * P + LdsLayout<N>::at(n)];
    T& dst = part ? cell.y : cell.x;
    dst = s.dirichlet ? val : dst + val;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = lds[LdsLayout<N>::at(L.n(q))];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Directional pass (contiguous x lines or strided y/z lines)
// ---------------------------------------------------------------------------
// TWO (x pass only): the variant that can also form the second spectrum of
// the homogeneous :stokes branch; without it the kernel carries no registers
// for that path.
// Whether the directional pass keeps each transform's address arithmetic to
// itself (OPQ in line_fft).  Compiled for gfx950, that never costs the x pass
// occupancy and raises it (or removes spills) for most lengths; strided
// passes on 16-element plans (128 ... 4096 points) run at two waves per SIMD
// either way, and there the merged addresses are faster (512-point strided
// pass of the 512^3 step: 0.94 ms merged, 1.00 ms separate).
template <typename T, int N, bool CONTIG, int V = (CONTIG ? 1 : 0)>
constexpr bool dir_opaque() {
  return CONTIG || Plan<N, V>::E != 16;
}

// Plan families a pass can be launched with besides its default (the
// autotuner's choice of radix decomposition and threads per line): at 256
// points 16 x 16 (family 0), 64 x 4 (1) and 32 x 8 (2); at 512 points 32 x 16
// (0) and 64 x 8 (1, 2).  PV = -1 in a kernel's template arguments selects the
// pass's default family.
constexpr bool plan_alternative(int n, int v) {
  return (n == 256 || n == 512) && v >= 0 && v <= 2;
}

template <typename T, int N, bool CONTIG, bool TWO = false, bool BS = false, int PV = -1>
__device__ __forceinline__ void dir_pass_body(const PassArgs<T>& a, int wg, int t_index) {
  constexpr int V = PV >= 0 ? PV : (CONTIG ? (BS ? 4 : 1) : 0);   // plan family
  constexpr bool WL = CONTIG && wave_local_lines<N, V>();   // lines inside one wavefront
  constexpr bool OPQ = dir_opaque<T, N, CONTIG, V>();
  constexpr int E = Plan<N, V>::E;
  constexpr int P = LdsLayout<N>::PITCH;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* lds_all = reinterpret_cast<Cx<T>*>(smem_raw);
  int lpw;
  const LineCtx<T, N, V, BS> L = make_ctx<T, N, CONTIG, V, BS>(a, wg, lpw);
  Cx<T>* lds = lds_all + L.c * P;
  const Twiddles<T, N, V> tw = load_twiddles<T, N, V>(a.tw, L.t);

  // --- grad p (staggered +) : rfft -> i k e^{+ikh/2} -> irfft
  Cx<T> v[E];
  L.load(a.p, v);
  line_dft<T, N, false, V, WL, BS, OPQ>(v, lds, tw, L.t, a.ch);
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = cmul(v[q], L.tab(a.Dp, q));
  line_dft<T, N, true, V, WL, BS, OPQ>(v, lds, tw, L.t, a.ch);

  // --- velocity update  u = pml_sg (pml_sg u - dt/rho0_sg dp)
  Cx<T> r0[E];
  if (a.rho0 != nullptr) {
    Cx<T> r1[E];
    L.load_map(a.rho0, r0, a.maps_f16 != 0, a.rho0_off, (T)1);
    L.load_map_next(a.rho0, r1, a.maps_f16 != 0, a.rho0_off, (T)1);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      r1[q].x = (r0[q].x + r1[q].x) * (T)0.5;
      r1[q].y = (r0[q].y + r1[q].y) * (T)0.5;
    }
    if (a.mode == 1) {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const T h = a.dt * (T)0.5;
        v[q].x = -fdiv(h, r1[q].x) * v[q].x;
        v[q].y = -fdiv(h, r1[q].y) * v[q].y;
      }
    } else {
      Cx<T> w[E];
      L.load(a.u, w);
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const T s = L.tab(a.pml_sg, q);
        v[q].x = s * (s * w[q].x - fdiv(a.dt, r1[q].x) * v[q].x);
        v[q].y = s * (s * w[q].y - fdiv(a.dt, r1[q].y) * v[q].y);
      }
    }
  } else {
    if (a.mode == 1) {
      const T f = -(a.dt * (T)0.5) / a.rho0_s;
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = cscale(v[q], f);
    } else {
      Cx<T> w[E];
      L.load(a.u, w);
      const T f = a.dt / a.rho0_s;
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const T s = L.tab(a.pml_sg, q);
        v[q].x = s * (s * w[q].x - f * v[q].x);
        v[q].y = s * (s * w[q].y - f * v[q].y);
      }
    }
  }
  if (a.mode == 1) {
    L.store(a.u, v);
    return;
  }
  inject<T, N, V, BS>(a.usrc, t_index, v, lds_all, L, wg);
  L.store(a.u, v);

  // --- div u (staggered -)
  line_dft<T, N, false, V, WL, BS, OPQ>(v, lds, tw, L.t, a.ch);
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = cmul(v[q], L.tab(a.Dm, q));
  line_dft<T, N, true, V, WL, BS, OPQ>(v, lds, tw, L.t, a.ch);
  if constexpr (!CONTIG) {
    // axisymmetric radial pass: du_r/dr + u_r / r, u_r as stored above (an L2 hit)
    if (a.inv_r != nullptr) {
      Cx<T> ur[E];
      L.load(a.u, ur);
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const T ir = L.tab(a.inv_r, q);
        v[q].x += ur[q].x * ir;
        v[q].y += ur[q].y * ir;
      }
    }
  }

  // --- density update  rho = pml (pml rho - dt rho0 du).  Strided passes keep
  // rho0 from the velocity update in registers (their register budget already
  // allows two waves per SIMD, and a re-read misses L2 there); the contiguous
  // x pass re-reads it (an L2 hit) to stay at three waves per SIMD.
  if (CONTIG && a.rho0 != nullptr) L.load_map(a.rho0, r0, a.maps_f16 != 0, a.rho0_off, (T)1);
  {
    Cx<T> w[E];
    L.load(a.rho, w);
    if (a.rho0 != nullptr) {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const T s = L.tab(a.pml, q);
        w[q].x = s * (s * w[q].x - a.dt * r0[q].x * v[q].x);
        w[q].y = s * (s * w[q].y - a.dt * r0[q].y * v[q].y);
      }
    } else {
      const T f = a.dt * a.rho0_s;
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const T s = L.tab(a.pml, q);
        w[q].x = s * (s * w[q].x - f * v[q].x);
        w[q].y = s * (s * w[q].y - f * v[q].y);
      }
    }
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = w[q];
  }
  inject<T, N, V, BS>(a.psrc, t_index, v, lds_all, L, wg);
  L.store(a.rho, v);

  if constexpr (CONTIG) {
    if (a.eos == 0) return;
    // --- EOS input f = (c0^2) rho (1 + B/A/2 rho/rho0), then r2c along x
    Cx<T> f[E];
#pragma unroll
    for (int q = 0; q < E; ++q) f[q] = v[q];
    if (a.rho_o1) {
      Cx<T> w[E];
      L.load(a.rho_o1, w);
#pragma unroll
      for (int q = 0; q < E; ++q) f[q] = cadd(f[q], w[q]);
    }
    if (a.rho_o2) {
      Cx<T> w[E];
      L.load(a.rho_o2, w);
#pragma unroll
      for (int q = 0; q < E; ++q) f[q] = cadd(f[q], w[q]);
    }
    if (a.rho_sum_out) L.store(a.rho_sum_out, f);
    Cx<T> g[TWO ? E : 1];  // second spectrum input (:stokes)
    const bool two = TWO && a.S2 != nullptr;   // uniform per workgroup
    auto eff = [&]() {      // f <- rho (1 + B/A rho / (2 rho0))
      Cx<T> bb[E];
      if (a.bona) L.load(a.bona, bb);
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const T ba = a.bona ? bb[q].x : a.bona_s, bbv = a.bona ? bb[q].y : a.bona_s;
        const T ra = a.rho0 ? r0[q].x : a.rho0_s, rb = a.rho0 ? r0[q].y : a.rho0_s;
        f[q].x = f[q].x * ((T)1 + ba / (T)2 * fdiv(f[q].x, ra));
        f[q].y = f[q].y * ((T)1 + bbv / (T)2 * fdiv(f[q].y, rb));
      }
    };
    if constexpr (TWO) {
     if (two) {
      Cx<T> pr[E];
      L.load(a.rho_prev, pr);
      L.store(a.rho_prev, f);
      if (a.nonlinear) {
        Cx<T> bb[E];
        if (a.bona) L.load(a.bona, bb);
#pragma unroll
        for (int q = 0; q < E; ++q) {
          const T ba = a.bona ? bb[q].x : a.bona_s, bbv = a.bona ? bb[q].y : a.bona_s;
          const T ra = a.rho0 ? r0[q].x : a.rho0_s, rb = a.rho0 ? r0[q].y : a.rho0_s;
          const T ia_ = fdiv((T)1, ra), ib_ = fdiv((T)1, rb);
          const T ea = f[q].x * ((T)1 + ba / (T)2 * f[q].x * ia_);
          const T eb = f[q].y * ((T)1 + bbv / (T)2 * f[q].y * ib_);
          const T pa = pr[q].x * ((T)1 + ba / (T)2 * pr[q].x * ia_);
          const T pb = pr[q].y * ((T)1 + bbv / (T)2 * pr[q].y * ib_);
          g[q].x = (ea - pa) * a.inv_dt;
          g[q].y = (eb - pb) * a.inv_dt;
          f[q].x = ea;
          f[q].y = eb;
        }
      } else {
#pragma unroll
        for (int q = 0; q < E; ++q) {
          g[q].x = (f[q].x - pr[q].x) * a.inv_dt;
          g[q].y = (f[q].y - pr[q].y) * a.inv_dt;
        }
      }
     } else if (a.nonlinear) {
      eff();
     }
    } else if (a.nonlinear) {
      eff();
    }
    if (a.c0sq) {
      Cx<T> cc[E];
      L.load_map(a.c0sq, cc, a.maps_f16 != 0, a.c0_off, (T)0);
      if (a.maps_f16) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
          cc[q].x *= cc[q].x;
          cc[q].y *= cc[q].y;
        }
      }
#pragma unroll
      for (int q = 0; q < E; ++q) {
        f[q].x *= cc[q].x;
        f[q].y *= cc[q].y;
      }
    }
    // r2c of the pair: Z = FFT(f_a + i f_b); A_k = (Z_k + conj Z_{N-k})/2,
    // B_k = (Z_k - conj Z_{N-k})/(2i), k = 0..N/2.
    const int nhl = BS ? L.nl / 2 + 1 : N / 2 + 1;   // half-spectrum length of the line
#pragma unroll 1
    for (int s = 0; s < (TWO ? 2 : 1); ++s) {
      if (s == 1 && !two) break;
      Cx<T>* Sdst = s == 0 ? a.S1 : a.S2;
      Cx<T> z[E];
#pragma unroll
      for (int q = 0; q < E; ++q) z[q] = s == 0 ? f[q] : g[TWO ? q : 0];
      line_dft<T, N, false, V, WL, BS, OPQ>(z, lds, tw, L.t, a.ch);
#pragma unroll
      for (int q = 0; q < E; ++q) lds[LdsLayout<N>::at(L.n(q))] = z[q];
      lds_sync<WL>();
      Cx<T>* __restrict__ Sw = Sdst + 2LL * wg * lpw * a.nhp;
      const unsigned oa = (unsigned)(2 * L.c * a.nhp), ob = oa + a.nhp;
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int k = L.n(q);
        if (k < nhl) {
          const int km = (k == 0) ? 0 : (BS ? L.nl : N) - k;
          const Cx<T> zk = z[q];
          const Cx<T> zm = lds[LdsLayout<N>::at(km)];
          const Cx<T> A = {(zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5};
          const Cx<T> B = {(zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5};
          if (L.va) at32(Sw, oa + k) = A;
          if (L.vb) at32(Sw, ob + k) = B;
        }
      }
      lds_sync<WL>();
    }
  }
}

template <typename T, int N, bool CONTIG, bool TWO = false, bool BS = false, int PV = -1>
__global__ void __launch_bounds__(512) k_dir_pass(PassArgs<T> a) {
  dir_pass_body<T, N, CONTIG, TWO, BS, PV>(a, CONTIG ? (int)blockIdx.x : strided_wg<T, N>(blockIdx.x), a.t);
}

// Batched form: one launch covers the nwg workgroup tiles of many
// simulations of the same line length; wg_sim maps a tile to its simulation,
// wg_first is each simulation's first tile, *t_dev the current 1-based step.
// P = false: one workgroup per tile (grid nwg).  P = true (persistent): a
// smaller grid, a multiple of 8 workgroups, walks the tiles with stride
// gridDim.x, so tile w stays on XCD w % 8 as in the one-tile order.  The
// loop costs registers: values the tile computes from the lane index alone
// are kept live across iterations (FP32: +5 to +80 VGPRs per kernel, the
// 1024-point directional passes spill), whether the tile is inlined or
// called (a called tile is compiled without the kernel's launch bounds and
// spills more).  Persistent instances exist for the line lengths of large
// batched groups.
constexpr bool persistent_length(int n) { return n == 256 || n == 384 || n == 512 || n == 768 || n == 1024; }

template <typename T, int N, bool CONTIG, bool TWO>
__device__ __forceinline__ void dir_tile(const PassArgs<T>* __restrict__ args, const int* __restrict__ wg_sim,
                                      const int* __restrict__ wg_first, int t, int w, int nwg) {
  const int b = CONTIG ? w : strided_wg<T, N>(w, nwg);
  const int s = cmem(wg_sim)[b];
  const PassArgs<T> a = cmem(args)[s];
  dir_pass_body<T, N, CONTIG, TWO>(a, b - cmem(wg_first)[s], t);
}

template <typename T, int N, bool CONTIG, bool TWO = false, bool P = false>
__global__ void __launch_bounds__(512) k_dir_pass_batch(const PassArgs<T>* __restrict__ args,
                                                        const int* __restrict__ wg_sim,
                                                        const int* __restrict__ wg_first, const int* t_dev, int nwg) {
  if constexpr (P) {
    const int t = *cmem(t_dev);
    for (int w = blockIdx.x; w < nwg; w += gridDim.x) {
      if (w != (int)blockIdx.x) __syncthreads();   // the previous tile is done with the LDS
      dir_tile<T, N, CONTIG, TWO>(args, wg_sim, wg_first, t, w, nwg);
    }
  } else {
    const int b = CONTIG ? (int)blockIdx.x : strided_wg<T, N>(blockIdx.x);
    const int s = cmem(wg_sim)[b];
    const PassArgs<T> a = cmem(args)[s];
    dir_pass_body<T, N, CONTIG, TWO>(a, b - cmem(wg_first)[s], *cmem(t_dev));
  }
}

// ---------------------------------------------------------------------------
// Spectral passes on the half spectrum (complex lines along y or z)
// ---------------------------------------------------------------------------
template <typename T>
struct SpecArgs {
  Cx<T>* S1;
  Cx<T>* S2;          // second spectrum (:stokes) or null
  const Cx<T>* tw;
  const T* M1;        // octant multiplier tables or null
  const T* M2;
  int NH;             // lines along the first index (Nx/2+1 for the EOS passes)
  long long Sc;       // stride between those lines (1 for the EOS passes)
  long long Sa;       // stride along the line (complex elements)
  long long So;       // outer stride
  int x_blocks;
  int mode;           // 0 fwd, 1 inv, 2 fwd*M*inv, 3 fwd*W*inv (1-D window W along the line)
  const T* W;         // mode 3 window (length N, 1/N folded in)
  // octant table addressing: index = kx + MH1*(|line index| + MH2*|outer index|)
  int axis;           // 1 = line along y, 2 = line along z
  int Ny, Nz;
  int MH1, MH2;       // table row pitch (>= Nx/2+1) and the octant extent of the line axis
  // lines of any length (Bluestein kernels): line length and chirp tables
  int nl;
  Chirp<T> ch;
};

// MODE: 0 forward, 6 forward of both spectra where S2 is set, 1 inverse, 2 fwd*M1*inv,
// 4 fwd*(M1 S1 + M2 S2)*inv (:stokes), 3 fwd*W*inv, 5 fwd then S1 = inv(M1 .),
// S2 = inv(M2 .) (the adjoint's K and dK/dc_ref applied to one spectrum).  Specialising on the mode
// keeps the plain transforms at a small register footprint.
template <typename T, int N, int MODE, bool BS = false, int PV = -1>
__device__ __forceinline__ void spec_pass_body(const SpecArgs<T>& a, int wg) {
  constexpr int V = PV >= 0 ? PV : 0;   // plan family
  constexpr int TPL = Plan<N, V>::TPL;
  constexpr int E = Plan<N, V>::E;
  constexpr int P = LdsLayout<N>::PITCH;
  constexpr bool OPQ = MODE != 1;   // every mode but the plain inverse has two transforms
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* lds_all = reinterpret_cast<Cx<T>*>(smem_raw);
  const int lpw = blockDim.x / TPL;
  const int c = threadIdx.x % lpw, t = threadIdx.x / lpw;
  const int xb = wg % a.x_blocks;
  const int o = wg / a.x_blocks;
  const int kx = xb * lpw + c;
  const bool valid = kx < a.NH;
  Cx<T>* lds = lds_all + c * P;
  const Twiddles<T, N, V> tw = load_twiddles<T, N, V>(a.tw, t);
  const long long wgb = (long long)(xb * lpw) * a.Sc + a.So * o;   // uniform
  const unsigned lo = (unsigned)(c * a.Sc), sa = (unsigned)a.Sa;
  auto off = [&](int q) { return lo + sa * (unsigned)(t + TPL * q); };
  const int nl = BS ? a.nl : N;                                    // line length
  auto ok = [&](int q) { return !BS || t + TPL * q < nl; };      // slot holds an element
  Cx<T>* __restrict__ S1 = a.S1 + wgb;

  Cx<T> v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = (valid && ok(q)) ? at32(S1, off(q)) : Cx<T>{0, 0};
  if constexpr (MODE == 1) {
    line_dft<T, N, true, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
  } else if constexpr (MODE == 3) {
    line_dft<T, N, false, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = cscale(v[q], ok(q) ? a.W[t + TPL * q] : (T)0);
    line_dft<T, N, true, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
  } else if constexpr (MODE == 0 || MODE == 6) {
    // 6: forward of S1 and, where the simulation has one, of S2 (the
    // second-spectrum code would otherwise cost plain forward passes
    // registers, hence occupancy)
    line_dft<T, N, false, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
    if (MODE == 6 && a.S2) {
      Cx<T>* __restrict__ S2 = a.S2 + wgb;
      if (valid) {
#pragma unroll
        for (int q = 0; q < E; ++q)
          if (ok(q)) at32(S1, off(q)) = v[q];
      }
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = (valid && ok(q)) ? at32(S2, off(q)) : Cx<T>{0, 0};
      line_dft<T, N, false, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
      if (valid) {
#pragma unroll
        for (int q = 0; q < E; ++q)
          if (ok(q)) at32(S2, off(q)) = v[q];
      }
      return;
    }
  } else if constexpr (MODE == 5) {
    line_dft<T, N, false, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
    int io = o;
    if (a.axis == 1)
      io = a.Nz > 1 ? (io > a.Nz - io ? a.Nz - io : io) : 0;   // 2-D: outer index = stacked firing
    else
      io = (io > a.Ny - io) ? a.Ny - io : io;
    Cx<T> w[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int ik = t + TPL * q;
      ik = (ik > nl - ik) ? nl - ik : ik;
      const unsigned mi = (unsigned)(kx + a.MH1 * (ik + a.MH2 * io));
      const T m1 = (valid && ok(q)) ? at32(a.M1, mi) : (T)0;
      const T m2 = (valid && ok(q)) ? at32(a.M2, mi) : (T)0;
      w[q] = cscale(v[q], m2);
      v[q] = cscale(v[q], m1);
    }
    line_dft<T, N, true, V, false, BS, OPQ>(w, lds, tw, t, a.ch);
    if (valid) {
      Cx<T>* __restrict__ S2 = a.S2 + wgb;
#pragma unroll
      for (int q = 0; q < E; ++q)
        if (ok(q)) at32(S2, off(q)) = w[q];
    }
    line_dft<T, N, true, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
  } else {
    line_dft<T, N, false, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
    Cx<T> w[MODE == 4 ? E : 1];
    // In a batch, a MODE 4 launch also covers simulations without a second
    // spectrum (S2 and M2 null): those take w = 0 and never touch S2/M2.
    const bool has2 = MODE == 4 && a.S2 != nullptr;   // uniform per workgroup
    if constexpr (MODE == 4) {
      if (has2) {
        const Cx<T>* __restrict__ S2 = a.S2 + wgb;
#pragma unroll
        for (int q = 0; q < E; ++q) w[q] = (valid && ok(q)) ? at32(S2, off(q)) : Cx<T>{0, 0};
      } else {
#pragma unroll
        for (int q = 0; q < E; ++q) w[q] = Cx<T>{0, 0};
      }
      line_dft<T, N, false, V, false, BS, OPQ>(w, lds, tw, t, a.ch);
    }
    // octant index |i| of the outer coordinate (z for y lines, y for z lines)
    int io = o;
    if (a.axis == 1)
      io = a.Nz > 1 ? (io > a.Nz - io ? a.Nz - io : io) : 0;   // 2-D: outer index = stacked firing
    else
      io = (io > a.Ny - io) ? a.Ny - io : io;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int ik = t + TPL * q;
      ik = (ik > nl - ik) ? nl - ik : ik;
      const unsigned mi = (unsigned)(kx + a.MH1 * (ik + a.MH2 * io));
      const T m1 = (valid && ok(q)) ? at32(a.M1, mi) : (T)0;
      Cx<T> r = cscale(v[q], m1);
      if constexpr (MODE == 4) {
        const T m2 = (valid && has2 && ok(q)) ? at32(a.M2, mi) : (T)0;
        r = cadd(r, cscale(w[q], m2));
      }
      v[q] = r;
    }
    line_dft<T, N, true, V, false, BS, OPQ>(v, lds, tw, t, a.ch);
  }
  if (valid) {
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (ok(q)) at32(S1, off(q)) = v[q];
  }
}

template <typename T, int N, int MODE, bool BS = false, int PV = -1>
__global__ void __launch_bounds__(512) k_spec_pass(SpecArgs<T> a) {
  spec_pass_body<T, N, MODE, BS, PV>(a, strided_wg<T, N>(blockIdx.x));
}

template <typename T, int N, int MODE>
__device__ __forceinline__ void spec_tile(const SpecArgs<T>* __restrict__ args, const int* __restrict__ wg_sim,
                                       const int* __restrict__ wg_first, int w, int nwg) {
  const int b = strided_wg<T, N>(w, nwg);
  const int s = cmem(wg_sim)[b];
  const SpecArgs<T> a = cmem(args)[s];
  spec_pass_body<T, N, MODE>(a, b - cmem(wg_first)[s]);
}

template <typename T, int N, int MODE, bool P = false>
__global__ void __launch_bounds__(512) k_spec_pass_batch(const SpecArgs<T>* __restrict__ args,
                                                         const int* __restrict__ wg_sim,
                                                         const int* __restrict__ wg_first, int nwg) {
  if constexpr (P) {   // tiles as in k_dir_pass_batch
    for (int w = blockIdx.x; w < nwg; w += gridDim.x) {
      if (w != (int)blockIdx.x) __syncthreads();
      spec_tile<T, N, MODE>(args, wg_sim, wg_first, w, nwg);
    }
  } else {
    const int b = strided_wg<T, N>(blockIdx.x);
    const int s = cmem(wg_sim)[b];
    const SpecArgs<T> a = cmem(args)[s];
    spec_pass_body<T, N, MODE>(a, b - cmem(wg_first)[s]);
  }
}

// ---------------------------------------------------------------------------
// Complex-to-real pass along x: p = irfft(S) (1/N folded into M); 1-D grids
// also apply the multiplier here.
// ---------------------------------------------------------------------------
template <typename T>
struct C2RArgs {
  const Cx<T>* S1;
  const Cx<T>* S2;   // 1-D :stokes second spectrum
  const T* M1;       // 1-D only (length Nx/2+1) or null
  const T* M2;
  const Cx<T>* tw;
  T* p;
  int n_lines;
  int nhp;           // half-spectrum row pitch
  // lines of any length (Bluestein kernels): line length and chirp tables
  int nl;
  Chirp<T> ch;
};

template <typename T, int N, bool BS = false, int PV = -1>
__device__ __forceinline__ void c2r_body(const C2RArgs<T>& a, int wg) {
  // contiguous x lines, complex-to-real plan family (4: chirp-z lines)
  constexpr int V = PV >= 0 ? PV : (BS ? 4 : 2);
  constexpr int TPL = Plan<N, V>::TPL;
  constexpr int E = Plan<N, V>::E;
  constexpr int P = LdsLayout<N>::PITCH;
  const int nl = BS ? a.nl : N;          // line length
  const int NH = nl / 2 + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* lds_all = reinterpret_cast<Cx<T>*>(smem_raw);
  const int lpw = blockDim.x / TPL;
  const int t = threadIdx.x % TPL, c = threadIdx.x / TPL;
  const long long pi = (long long)wg * lpw + c;
  const long long la = 2 * pi, lb = 2 * pi + 1;
  const bool va = la < a.n_lines, vb = lb < a.n_lines;
  Cx<T>* lds = lds_all + c * P;
  const Twiddles<T, N, V> tw = load_twiddles<T, N, V>(a.tw, t);
  Cx<T> z[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int k = t + TPL * q;
    const bool in = !BS || k < nl;
    const bool mir = k >= NH;
    const int kk = !in ? 0 : (mir ? nl - k : k);
    Cx<T> A = (va && in) ? a.S1[la * a.nhp + kk] : Cx<T>{0, 0};
    Cx<T> B = (vb && in) ? a.S1[lb * a.nhp + kk] : Cx<T>{0, 0};
    if (a.M1) {
      const T m1 = a.M1[kk];
      A = cscale(A, m1);
      B = cscale(B, m1);
      if (a.S2) {
        const T m2 = a.M2[kk];
        const Cx<T> A2 = (va && in) ? a.S2[la * a.nhp + kk] : Cx<T>{0, 0};
        const Cx<T> B2 = (vb && in) ? a.S2[lb * a.nhp + kk] : Cx<T>{0, 0};
        A = cadd(A, cscale(A2, m2));
        B = cadd(B, cscale(B2, m2));
      }
    }
    // c2r semantics: imaginary parts of the DC and (even-N) Nyquist bins are ignored
    if (kk == 0 || 2 * kk == nl) {
      A.y = 0;
      B.y = 0;
    }
    if (mir) {
      A.y = -A.y;
      B.y = -B.y;
    }
    z[q] = Cx<T>{A.x - B.y, A.y + B.x};
  }
  line_dft<T, N, true, V, wave_local_lines<N, V>(), BS>(z, lds, tw, t, a.ch);
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int n = t + TPL * q;
    if (BS && n >= nl) continue;
    if (va) a.p[la * nl + n] = z[q].x;
    if (vb) a.p[lb * nl + n] = z[q].y;
  }
}

template <typename T, int N, bool BS = false, int PV = -1>
__global__ void __launch_bounds__(256) k_c2r_x(C2RArgs<T> a) {
  c2r_body<T, N, BS, PV>(a, blockIdx.x);
}

template <typename T, int N>
__device__ __forceinline__ void c2r_tile(const C2RArgs<T>* __restrict__ args, const int* __restrict__ wg_sim,
                                      const int* __restrict__ wg_first, int w) {
  const int s = cmem(wg_sim)[w];
  const C2RArgs<T> a = cmem(args)[s];
  c2r_body<T, N>(a, w - cmem(wg_first)[s]);
}

template <typename T, int N, bool P = false>
__global__ void __launch_bounds__(256) k_c2r_x_batch(const C2RArgs<T>* __restrict__ args,
                                                     const int* __restrict__ wg_sim,
                                                     const int* __restrict__ wg_first, int nwg) {
  if constexpr (P) {   // tiles as in k_dir_pass_batch
    for (int w = blockIdx.x; w < nwg; w += gridDim.x) {
      if (w != (int)blockIdx.x) __syncthreads();
      c2r_tile<T, N>(args, wg_sim, wg_first, w);
    }
  } else {
    const int s = cmem(wg_sim)[blockIdx.x];
    const C2RArgs<T> a = cmem(args)[s];
    c2r_body<T, N>(a, blockIdx.x - cmem(wg_first)[s]);
  }
}

// ---------------------------------------------------------------------------
// y inverse + complex-to-real x pass fused per z plane.  One workgroup holds
// a whole NX x NY plane of the half spectrum: the NX/2 + 1 columns go to
// NX/2 column slots (the DC and Nyquist columns share slot 0 as real and
// imaginary part: after the y and z transforms both are Hermitian in ky, so
// their y inverses are real and one complex transform carries both), each
// thread keeps its columns' y inverses in registers, and rows move to the
// c2r transforms through LDS in chunks.  The plane is read once and p
// written once, instead of the spectrum's extra round trip of the separate
// y inverse pass.
// ---------------------------------------------------------------------------
template <typename T>
struct PlaneArgs {
  const Cx<T>* S1;    // half spectrum: rows of nhp bins, planes of nhp * NY
  const Cx<T>* tw_x;  // line twiddle tables of the x and y axes
  const Cx<T>* tw_y;
  T* p;
  int nhp;
};

template <typename T, int NX, int NY>
struct PlaneGeom {
  static constexpr int NT = 1024;                                  // threads
  static constexpr int NC = NX / 2;                                // column slots
  static constexpr int TY = Plan<NY, 5>::TPL, EY = Plan<NY, 5>::E;
  static constexpr int TX = Plan<NX, 5>::TPL, EX = Plan<NX, 5>::E;
  static constexpr int CPR = NT / TY;                              // columns per round
  static constexpr int R = NC / CPR;                               // rounds
  static constexpr int RPC = 2 * (NT / TX);                        // rows per chunk
  static constexpr int NCH = NY / RPC;                             // chunks
  static constexpr int CP = NC + 1;                                // chunk row pitch (bins)
  static constexpr int XCH = (CPR * LdsLayout<NY>::PITCH > (NT / TX) * LdsLayout<NX>::PITCH)
                                 ? CPR * LdsLayout<NY>::PITCH
                                 : (NT / TX) * LdsLayout<NX>::PITCH;   // exchange slices
  static constexpr size_t LDS = (size_t)(XCH + RPC * CP) * sizeof(Cx<T>);
  static_assert(NC % CPR == 0 && NY % RPC == 0 && NY == TY * EY && RPC % TY == 0, "plane geometry");
};

template <typename T, int NX, int NY>
__global__ void __launch_bounds__(1024) k_plane_c2r(PlaneArgs<T> a) {
  using G = PlaneGeom<T, NX, NY>;
  constexpr int NC = G::NC, TY = G::TY, EY = G::EY, TX = G::TX, EX = G::EX, CPR = G::CPR, R = G::R;
  constexpr int RPC = G::RPC, NCH = G::NCH, CP = G::CP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* xch = reinterpret_cast<Cx<T>*>(smem_raw);
  Cx<T>* chunk = xch + G::XCH;
  const int tid = threadIdx.x;
  const int c = tid % CPR, t = tid / CPR;   // column slot in a round, thread in the column
  const Cx<T>* __restrict__ S = a.S1 + (long long)blockIdx.x * NY * a.nhp;

  // every column of the plane (rows t + TY q): round 0 with the Nyquist
  // column folded into slot 0 first, then the other rounds, whose loads are
  // in flight during round 0's transform
  Cx<T> v[R][EY];
#pragma unroll
  for (int q = 0; q < EY; ++q) v[0][q] = at32(S, (unsigned)((t + TY * q) * a.nhp + c));
  if (c == 0) {   // slot 0: DC + i Nyquist
#pragma unroll
    for (int q = 0; q < EY; ++q) {
      const Cx<T> b = at32(S, (unsigned)((t + TY * q) * a.nhp + NC));
      v[0][q] = Cx<T>{v[0][q].x - b.y, v[0][q].y + b.x};
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int r = 1; r < R; ++r)
#pragma unroll
    for (int q = 0; q < EY; ++q) v[r][q] = at32(S, (unsigned)((t + TY * q) * a.nhp + r * CPR + c));
  {
    const Twiddles<T, NY, 5> tw = load_twiddles<T, NY, 5>(a.tw_y, t);
#pragma unroll
    for (int r = 0; r < R; ++r) line_fft<T, NY, true, 5, false, true>(v[r], xch + c * LdsLayout<NY>::PITCH, tw, t);
  }

  // c2r of row pairs, RPC rows at a time
  const int pr = tid / TX, tx = tid % TX;
  const Twiddles<T, NX, 5> twx = load_twiddles<T, NX, 5>(a.tw_x, tx);
  T* __restrict__ P = a.p + (long long)blockIdx.x * NY * NX;
#pragma unroll
  for (int j = 0; j < NCH; ++j) {
    // this thread's values of the chunk's rows: q = j * RPC / TY ... (rows t + TY q)
#pragma unroll
    for (int qq = 0; qq < RPC / TY; ++qq) {
      const int q = j * (RPC / TY) + qq;
      const int row = t + TY * qq;   // row within the chunk
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r == 0 && c == 0) {
          chunk[row * CP] = Cx<T>{v[0][q].x, (T)0};
          chunk[row * CP + NC] = Cx<T>{v[0][q].y, (T)0};
        } else {
          chunk[row * CP + r * CPR + c] = v[r][q];
        }
      }
    }
    __syncthreads();
    Cx<T> z[EX];
#pragma unroll
    for (int q = 0; q < EX; ++q) {
      const int k = tx + TX * q;
      const bool mir = k > NC;
      const int kk = mir ? NX - k : k;
      Cx<T> A = chunk[(2 * pr) * CP + kk], B = chunk[(2 * pr + 1) * CP + kk];
      if (mir) {
        A.y = -A.y;
        B.y = -B.y;
      }
      z[q] = Cx<T>{A.x - B.y, A.y + B.x};
    }
    line_fft<T, NX, true, 5, wave_local_lines<NX, 5>(), true>(z, xch + pr * LdsLayout<NX>::PITCH, twx, tx);
    const int ra = j * RPC + 2 * pr;
#pragma unroll
    for (int q = 0; q < EX; ++q) {
      const int n = tx + TX * q;
      at32(P, (unsigned)(ra * NX + n)) = z[q].x;
      at32(P, (unsigned)((ra + 1) * NX + n)) = z[q].y;
    }
    __syncthreads();   // the next chunk overwrites the chunk rows and exchange slices
  }
}

// ---------------------------------------------------------------------------
// Sensor recording (record_sensor_data!, first_order_steps.jl:728-794)
// ---------------------------------------------------------------------------
enum : unsigned {
  REC_P = 1u << 0,
  REC_P_MAX = 1u << 1,
  REC_P_MIN = 1u << 2,
  REC_P_RMS = 1u << 3,
  REC_P_FINAL = 1u << 4,
  REC_UX = 1u << 5,
  REC_UY = 1u << 6,
  REC_UZ = 1u << 7,
  REC_U_MAX = 1u << 8,
  REC_U_RMS = 1u << 9,
  REC_U_FINAL = 1u << 10,
  REC_I_AVG = 1u << 11,
  REC_I_MAX = 1u << 12,
};

template <typename T>
struct RecArgs {
  const T* p;
  const T* u[3];
  int ndim;
  const long long* idx;  // sensor voxel of each row, or null when the mask is the whole grid
  int n;
  int t;       // 1-based
  int Nt;
  unsigned rec;
  T* p_ts;     // [t][j]
  T* u_ts[3];  // [t][j]
  T* p_max;
  T* p_min;
  double* p_rms;
  T* p_final;
  double* u_max;
  double* u_rms;
  double* u_final;
  double* I_avg;
  double* I_max;
};

template <typename T>
__device__ __forceinline__ void record_body(const RecArgs<T>& a, int j, int t_index) {
  const long long i = a.idx ? a.idx[j] : (long long)j;   // null: every voxel, in order
  const T pv = a.p[i];
  const long long o = (long long)(t_index - 1) * a.n + j;
  if (a.rec & REC_P) a.p_ts[o] = pv;
  if (a.rec & REC_P_MAX) a.p_max[j] = a.p_max[j] > pv ? a.p_max[j] : pv;
  if (a.rec & REC_P_MIN) a.p_min[j] = a.p_min[j] < pv ? a.p_min[j] : pv;
  if (a.rec & REC_P_RMS) a.p_rms[j] += (double)pv * (double)pv;
  if ((a.rec & REC_P_FINAL) && t_index == a.Nt) a.p_final[j] = pv;
  T uv[3] = {0, 0, 0};
  // velocities only when a velocity or intensity field is recorded
  if (a.rec & (REC_UX | REC_UY | REC_UZ | REC_U_MAX | REC_U_RMS | REC_U_FINAL | REC_I_AVG | REC_I_MAX))
    for (int d = 0; d < a.ndim; ++d) uv[d] = a.u[d][i];
  if (a.rec & REC_UX) a.u_ts[0][o] = uv[0];
  if ((a.rec & REC_UY) && a.ndim > 1) a.u_ts[1][o] = uv[1];
  if ((a.rec & REC_UZ) && a.ndim > 2) a.u_ts[2][o] = uv[2];
  if (a.rec & (REC_U_MAX | REC_U_RMS | REC_U_FINAL)) {
    double m2 = 0.0;
    for (int d = 0; d < a.ndim; ++d) m2 += (double)uv[d] * (double)uv[d];
    const double mag = sqrt(m2);
    if (a.rec & REC_U_MAX) a.u_max[j] = a.u_max[j] > mag ? a.u_max[j] : mag;
    if (a.rec & REC_U_RMS) a.u_rms[j] += mag * mag;
    if ((a.rec & REC_U_FINAL) && t_index == a.Nt) a.u_final[j] = mag;
  }
  if (a.rec & (REC_I_AVG | REC_I_MAX)) {
    double s = 0.0;
    for (int d = 0; d < a.ndim; ++d) {
      const double pu = (double)pv * (double)uv[d];
      s += pu * pu;
    }
    const double inst = sqrt(s);
    if (a.rec & REC_I_AVG) a.I_avg[j] += inst / a.Nt;
    if (a.rec & REC_I_MAX) a.I_max[j] = a.I_max[j] > inst ? a.I_max[j] : inst;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_record(RecArgs<T> a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < a.n) record_body<T>(a, j, a.t);
}

// Batched recording: sensor J of the concatenated lists belongs to
// simulation sens_sim[J] at local index J - sens_first[sim].
template <typename T>
__global__ void __launch_bounds__(256) k_record_batch(const RecArgs<T>* __restrict__ args,
                                                      const int* __restrict__ sens_sim,
                                                      const int* __restrict__ sens_first, int n_total,
                                                      const int* t_dev) {
  const int J = blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= n_total) return;
  const int s = cmem(sens_sim)[J];
  const RecArgs<T> a = cmem(args)[s];
  record_body<T>(a, J - cmem(sens_first)[s], *cmem(t_dev));
}

// Advance the device-side step counter read by the batched kernels (one
// lane, ordinary vector store), so a captured step graph replays unchanged.
template <typename D>
__global__ void k_tick(D* t_dev) {
  if (threadIdx.x == 0 && blockIdx.x == 0) t_dev[0] = t_dev[0] + 1;
}

// Time-reversal Dirichlet injection before a step (first_order.jl:678-689):
// rho_d[idx_j] = data[j, Nt - t + 1] * scale_j for every component d.
template <typename T>
__global__ void __launch_bounds__(256) k_tr_inject(T* r0, T* r1, T* r2, const long long* idx, const T* vals,
                                                   int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const long long i = idx[j];
  const T v = vals[j];
  r0[i] = v;
  if (r1) r1[i] = v;
  if (r2) r2[i] = v;
}

// rho_d = p0 / (ndim c0^2) for every component; p = p0
// (first_order_steps.jl:113-118, 161-167).
template <typename T>
__global__ void __launch_bounds__(256) k_init_rho(const T* p0, const T* c0sq, T c0sq_s, T ndim, T* p, T* r0, T* r1,
                                                  T* r2, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T v = p0[i];
  const T r = v / (ndim * (c0sq ? c0sq[i] : c0sq_s));
  p[i] = v;
  r0[i] = r;
  if (r1) r1[i] = r;
  if (r2) r2[i] = r;
}

// p_rms = sqrt(sum p^2 / Nt) (finalize_sensor_data!, first_order_steps.jl:818-825)
template <typename D>
__global__ void __launch_bounds__(256) k_finalize_rms(D* a, long long n, D nt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = sqrt(a[i] / nt);
}

// ---------------------------------------------------------------------------
// smooth(p0; restore_max=true) support (filters.jl:20-60): the per-axis
// filtering runs through k_spec_pass mode 3 on a complex copy of p0.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_real_to_cx(const T* p, Cx<T>* z, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = Cx<T>{p[i], (T)0};
}

// mx[0] = max|p0|, mx[1] = max|Re z| as IEEE-754 double bit patterns
// (non-negative doubles order like their unsigned bit patterns).
template <typename T>
__global__ void __launch_bounds__(256) k_maxabs2(const T* p, const Cx<T>* z, unsigned long long* mx, long long n) {
  __shared__ double s0[256], s1[256];
  double a = 0.0, b = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double u = fabs((double)p[i]), v = fabs((double)z[i].x);
    a = u > a ? u : a;
    b = v > b ? v : b;
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s0[threadIdx.x] = s0[threadIdx.x] > s0[threadIdx.x + s] ? s0[threadIdx.x] : s0[threadIdx.x + s];
      s1[threadIdx.x] = s1[threadIdx.x] > s1[threadIdx.x + s] ? s1[threadIdx.x] : s1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMax(&mx[0], (unsigned long long)__double_as_longlong(s0[0]));
    atomicMax(&mx[1], (unsigned long long)__double_as_longlong(s1[0]));
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_smooth_finish(const Cx<T>* z, const unsigned long long* mx, T* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T om = (T)__longlong_as_double((long long)mx[0]);
  const T sm = (T)__longlong_as_double((long long)mx[1]);
  const T s = sm > (T)0 ? om / sm : (T)1;
  p[i] = sm > (T)0 ? z[i].x * s : z[i].x;
}

// ---------------------------------------------------------------------------
// Sensor post-filters (apply_sensor_directivity!, apply_frequency_response!;
// first_order_steps.jl:836-910): y = Re(ifft(fft(x) .* H)) per sensor row,
// H(f) = sinc(2 pi |f| / h_min * size * sin(angle) / 2) * exp(-(|f| - fc)^2 / (2 bw^2)).
// Direct DFT in FP64 (any record length); one workgroup per sensor row.
// cw/sw hold cos(2 pi m/Nt) and -sin(2 pi m/Nt).
// ---------------------------------------------------------------------------
struct FiltArgs {
  int ns, nt;
  const double* cw;
  const double* sw;
  const double* absf;    // |freqs| (filters.jl:158-165)
  const double* ang;     // per-sensor angle or null
  double ang_s;
  int dir;
  double dsize, kscale;  // directivity size, 2 pi / min spacing
  int fr;
  double fc, bw;
  Cx<double>* Y;         // ns x nt spectra (row-major per sensor)
  double* out;           // column-major ns x nt
};

template <typename T>
__global__ void __launch_bounds__(256) k_filter_fwd(const T* p_ts, FiltArgs a) {
  const int j = blockIdx.x;
  const double ang = a.dir ? (a.ang ? a.ang[j] : a.ang_s) : 0.0;
  const bool use_dir = a.dir && ang != 0.0;  // angle == 0 skips this sensor (first_order_steps.jl:856-858)
  for (int f = threadIdx.x; f < a.nt; f += blockDim.x) {
    double re = 0.0, im = 0.0;
    int m = 0;
    for (int n = 0; n < a.nt; ++n) {
      const double x = (double)p_ts[(long long)n * a.ns + j];
      re += x * a.cw[m];
      im += x * a.sw[m];
      m += f;
      if (m >= a.nt) m -= a.nt;
    }
    double H = 1.0;
    if (use_dir) {
      const double arg = a.kscale * a.absf[f] * a.dsize * sin(ang) / 2;
      H *= (arg == 0.0) ? 1.0 : sin(arg) / arg;
    }
    if (a.fr) {
      const double d = a.absf[f] - a.fc;
      H *= exp(-(d * d) / (2 * a.bw * a.bw));
    }
    a.Y[(long long)j * a.nt + f] = Cx<double>{re * H, im * H};
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_filter_inv(FiltArgs a) {
  const int j = blockIdx.x;
  const Cx<double>* Y = a.Y + (long long)j * a.nt;
  for (int n = threadIdx.x; n < a.nt; n += blockDim.x) {
    double acc = 0.0;
    int m = 0;
    for (int f = 0; f < a.nt; ++f) {
      const Cx<double> y = Y[f];
      acc += y.x * a.cw[m] + y.y * a.sw[m];
      m += n;
      if (m >= a.nt) m -= a.nt;
    }
    a.out[(long long)n * a.ns + j] = acc / a.nt;
  }
}

// ---------------------------------------------------------------------------
// Discrete adjoint of the time step w.r.t. the sound-speed map.
// Per reverse step n (see the oracle's grad_c0 for the derivation):
//   Pbar += M^T r_n ; Q = K Pbar, Q' = dK/dc_ref Pbar   (k_r2c_x, k_spec_pass 5, k_c2r_x)
//   g += 2 c rho_sum^n Q ; gref += <c^2 rho_sum^n, Q'>   (k_adj_grad)
//   per axis d (k_adj_pass): Rt = Rbar_d + c^2 Q ; Rbar_d <- P^2 Rt ;
//     Ut = Ubar_d + D+ (dt rho0 P Rt) ; Ubar_d <- Psg^2 Ut ;
//     Pbar (+)= D- (dt/rho0sg Psg Ut)
// ---------------------------------------------------------------------------
template <typename T>
struct AdjArgs {
  T* Rb;            // adjoint of rho_d
  T* Ub;            // adjoint of u_d
  T* Pb;            // adjoint of p (written by the first axis, accumulated by the others)
  const T* Q;       // K Pbar
  const T* c0sq;    // c0^2 map
  const void* rho0; // map (T) or null
  T rho0_s;
  const T* pml;
  const T* pml_sg;
  const Cx<T>* Dp;
  const Cx<T>* Dm;
  const Cx<T>* tw;
  T dt;
  int Nx;
  long long Sa, So;
  int n_lines, x_blocks;
  int accumulate;   // 0: Pb = contribution, 1: Pb += contribution
};

template <typename T, int N, bool CONTIG, int PV = -1>
__global__ void __launch_bounds__(512) k_adj_pass(AdjArgs<T> a) {
  constexpr int V = PV >= 0 ? PV : (CONTIG ? 1 : 0);   // plan family, as the directional pass's
  constexpr bool OPQ = dir_opaque<T, N, CONTIG, V>();
  constexpr int E = Plan<N, V>::E;
  constexpr int P = LdsLayout<N>::PITCH;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* lds_all = reinterpret_cast<Cx<T>*>(smem_raw);
  PassArgs<T> geo;
  geo.Nx = a.Nx;
  geo.Sa = a.Sa;
  geo.So = a.So;
  geo.n_lines = a.n_lines;
  geo.x_blocks = a.x_blocks;
  int lpw;
  const LineCtx<T, N, V> L =
      make_ctx<T, N, CONTIG, V>(geo, CONTIG ? (int)blockIdx.x : strided_wg<T, N>(blockIdx.x), lpw);
  Cx<T>* lds = lds_all + L.c * P;
  const Twiddles<T, N, V> tw = load_twiddles<T, N, V>(a.tw, L.t);

  Cx<T> rt[E], cc[E], v[E];
  L.load(a.Rb, rt);
  L.load(a.Q, v);
  L.load(a.c0sq, cc);
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const T s = a.pml[L.n(q)];
    // w = P (Rbar + c^2 Q) ; Rbar <- P w
    const T wx = s * (rt[q].x + cc[q].x * v[q].x), wy = s * (rt[q].y + cc[q].y * v[q].y);
    rt[q] = Cx<T>{wx, wy};
    v[q] = Cx<T>{s * wx, s * wy};
  }
  L.store(a.Rb, v);
  // D+ (dt rho0 w)
  if (a.rho0) {
    Cx<T> r0[E];
    L.load_map(a.rho0, r0, false, (T)0, (T)1);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = Cx<T>{a.dt * r0[q].x * rt[q].x, a.dt * r0[q].y * rt[q].y};
  } else {
    const T f = a.dt * a.rho0_s;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = cscale(rt[q], f);
  }
  line_fft<T, N, false, V, false, OPQ>(v, lds, tw, L.t);
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = cmul(v[q], a.Dp[L.n(q)]);
  line_fft<T, N, true, V, false, OPQ>(v, lds, tw, L.t);
  // Ut = Ubar + D+(...) ; y = Psg Ut ; Ubar <- Psg y ; z = dt/rho0sg y
  {
    Cx<T> ub[E];
    L.load(a.Ub, ub);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const T s = a.pml_sg[L.n(q)];
      v[q] = Cx<T>{s * (ub[q].x + v[q].x), s * (ub[q].y + v[q].y)};
      ub[q] = cscale(v[q], s);
    }
    L.store(a.Ub, ub);
  }
  if (a.rho0) {
    Cx<T> r0[E], r1[E];
    L.load_map(a.rho0, r0, false, (T)0, (T)1);
    L.load_map_next(a.rho0, r1, false, (T)0, (T)1);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      v[q].x *= fdiv((T)2 * a.dt, r0[q].x + r1[q].x);
      v[q].y *= fdiv((T)2 * a.dt, r0[q].y + r1[q].y);
    }
  } else {
    const T f = a.dt / a.rho0_s;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = cscale(v[q], f);
  }
  line_fft<T, N, false, V, false, OPQ>(v, lds, tw, L.t);
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = cmul(v[q], a.Dm[L.n(q)]);
  line_fft<T, N, true, V, false, OPQ>(v, lds, tw, L.t);
  if (a.accumulate) {
    Cx<T> pb[E];
    L.load(a.Pb, pb);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = cadd(v[q], pb[q]);
  }
  L.store(a.Pb, v);
}

// Real lines (pairs) of f -> half spectrum along x (the r2c stage alone).
template <typename T, int N, int PV = -1>
__global__ void __launch_bounds__(256) k_r2c_x(const T* f, Cx<T>* S, int n_lines, int nhp, const Cx<T>* twt) {
  constexpr int V = PV >= 0 ? PV : 2;  // contiguous x lines: the complex-to-real pass's plan family
  constexpr int E = Plan<N, V>::E;
  constexpr int P = LdsLayout<N>::PITCH;
  constexpr int NH = N / 2 + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* lds_all = reinterpret_cast<Cx<T>*>(smem_raw);
  PassArgs<T> geo;
  geo.n_lines = n_lines;
  int lpw;
  const LineCtx<T, N, V> L = make_ctx<T, N, true, V>(geo, blockIdx.x, lpw);
  Cx<T>* lds = lds_all + L.c * P;
  const Twiddles<T, N, V> tw = load_twiddles<T, N, V>(twt, L.t);
  Cx<T> z[E];
  L.load(f, z);
  line_fft<T, N, false, V>(z, lds, tw, L.t);
#pragma unroll
  for (int q = 0; q < E; ++q) lds[LdsLayout<N>::at(L.n(q))] = z[q];
  __syncthreads();
  Cx<T>* __restrict__ Sw = S + 2LL * blockIdx.x * lpw * nhp;
  const unsigned oa = (unsigned)(2 * L.c * nhp), ob = oa + nhp;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int k = L.n(q);
    if (k < NH) {
      const int km = (k == 0) ? 0 : N - k;
      const Cx<T> zk = z[q];
      const Cx<T> zm = lds[LdsLayout<N>::at(km)];
      if (L.va) at32(Sw, oa + k) = Cx<T>{(zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5};
      if (L.vb) at32(Sw, ob + k) = Cx<T>{(zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5};
    }
  }
}

// g += 2 c rho_sum Q (double) ; gpart[block] += sum over the block's elements of
// c^2 rho_sum Q'.  Each block owns its slot and the grid is the same on every
// call, so the sums are deterministic (no atomics).
template <typename T>
__global__ void __launch_bounds__(256) k_adj_grad(const T* __restrict__ c0sq, const T* __restrict__ rs,
                                                  const T* __restrict__ Q, const T* __restrict__ Qd,
                                                  double* __restrict__ g, double* __restrict__ gpart, long long plane,
                                                  int nf) {
  // Stacked firings share the medium: voxel j of every firing f (j + f plane)
  // has the same c0, so the firings' terms are summed here and g holds one
  // grid (each term is read once, g once per grid voxel).  The loads of
  // eight firings are issued together before they are accumulated.
  constexpr int B = 8;
  __shared__ double red[256];
  double acc = 0.0;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < plane; j += (long long)gridDim.x * blockDim.x) {
    const double c2 = (double)c0sq[j];
    double sq = 0.0, sd = 0.0;
    int f = 0;
    for (; f + B <= nf; f += B) {
      T r[B], q[B], qd[B];
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const long long i = j + (long long)(f + u) * plane;
        r[u] = rs[i];
        q[u] = Q[i];
        qd[u] = Qd[i];
      }
#pragma unroll
      for (int u = 0; u < B; ++u) {
        sq += (double)r[u] * (double)q[u];
        sd += (double)r[u] * (double)qd[u];
      }
    }
    for (; f < nf; ++f) {
      const long long i = j + (long long)f * plane;
      const double r = (double)rs[i];
      sq += r * (double)Q[i];
      sd += r * (double)Qd[i];
    }
    g[j] += 2.0 * sqrt(c2) * sq;
    acc += c2 * sd;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) gpart[blockIdx.x] += red[0];
}

// source gradient: g[idx_j] += fac_j s_n[j] sum_d (Rbar_d + c^2 Q)[idx_j], fac_j = -2/(ndim c_j^3)
template <typename T>
__global__ void __launch_bounds__(256) k_adj_src(const T* r0, const T* r1, const T* r2, const T* c0sq, const T* Q,
                                                 const long long* idx, const double* fac, const double* sig_col,
                                                 double* g, int ndim, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const long long i = idx[j];
  const double cq = (double)c0sq[i] * (double)Q[i];
  double tot = (double)r0[i] + cq;
  if (ndim > 1) tot += (double)r1[i] + cq;
  if (ndim > 2) tot += (double)r2[i] + cq;
  g[i] += fac[j] * sig_col[j] * tot;
}

// residuals r[t][j] = p_rec[t][j] - obs[t][j] and lpart[block] = sum of r^2/2 over the block's elements
template <typename T>
__global__ void __launch_bounds__(256) k_adj_resid(const T* prec, const double* obs, double* res, double* lpart,
                                                   long long n) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double r = (double)prec[i] - obs[i];
    res[i] = r;
    acc += 0.5 * r * r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

// Pbar[idx_j] += r_j  (sensor points are distinct)
template <typename T>
__global__ void __launch_bounds__(256) k_adj_scatter(T* Pb, const long long* idx, const double* r, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) Pb[idx[j]] += (T)r[j];
}

// g += gterm (= g_ref / N) and, for p0 runs, g += -2 p0 / (ndim c^3) sum_d Rbar_d
template <typename T>
__global__ void __launch_bounds__(256) k_adj_final(double* g, double gterm, const T* p0, const T* c0sq, const T* r0,
                                                   const T* r1, const T* r2, int ndim, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = g[i] + gterm;
  if (p0) {
    double s = (double)r0[i];
    if (ndim > 1) s += (double)r1[i];
    if (ndim > 2) s += (double)r2[i];
    const double c = sqrt((double)c0sq[i]);
    v += -2.0 * (double)p0[i] / (ndim * c * c * c) * s;
  }
  g[i] = v;
}

}  // namespace kw
