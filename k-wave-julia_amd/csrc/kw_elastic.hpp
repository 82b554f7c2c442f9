// Kernels of the elastic stress-velocity PSTD loop (pstd_elastic_2d /
// pstd_elastic_3d, elastic.jl:133-588) for gfx950.
//
// Every spectral derivative of a step is a 1-D staggered derivative along
// one axis, so one line-FFT pass computes it (same line machinery and
// pair-of-real-lines packing as the acoustic passes).  The pass that
// produces the last derivative a field update needs also applies that
// update, so each step is:
//   momentum, per velocity v_a: passes along the other axes b accumulate
//     D_b+ s_ab into a scratch field; the pass along a adds D_a+ s_aa and
//     writes v_a = Psg_a (Psg_a v_a + dt / rho_sg,a * sum);
//   normal stresses: D_a- v_a along every axis a but the last into scratch
//     fields, the pass along the last axis updates all s_aa together;
//   shear stresses, per pair (a, b): D_b- v_a along b into scratch, the pass
//     along a adds D_a- v_b and updates s_ab.
// PML factors of the line's own axis are indexed along the line; those of
// the other axes are fixed per line and read once.
#pragma once

#include "kw_kernels.hpp"

namespace kw {

enum : int { EL_DERIV = 0, EL_VEL = 1, EL_NORMAL = 2, EL_SHEAR = 3 };

template <typename T>
struct ElArgs {
  const T* f;          // field differentiated along the lines
  const Cx<T>* D;      // staggered derivative multiplier of the line axis, 1/N folded in
  const Cx<T>* tw;
  // line geometry (as PassArgs)
  int Nx;
  long long Sa, So;
  int n_lines, x_blocks;
  int Ny;              // grid extent along y (per-line coordinates)
  int axis;            // axis the lines run along
  int ndim;
  int mode;            // EL_*
  T dt;
  const T* pml[3];     // non-staggered PML per axis
  // EL_DERIV: out = D f, or out += D f
  T* out;
  int accumulate;
  // EL_VEL / EL_SHEAR: sum of the derivatives taken by the earlier passes (or null)
  const T* acc;
  // EL_VEL: v = Psg (Psg v + dt / rho_sg (acc + D f)); rho_sg is the mean of
  // rho at n and n+1 along the line (periodic), or rho_s
  T* v;
  const T* rho;
  T rho_s;
  const T* pml_sg;
  // EL_NORMAL (lines along axis ndim-1): dn[a] = D_a- v_a for a < ndim-1;
  //   s[a] = P_a (P_a s[a] + dt kp (lam div + 2 mu d_a))
  // EL_SHEAR (lines along axis `axis`, partner axis `axis2`):
  //   s[0] = P_axis P_axis2 (s[0] + dt ks mu (acc + D f))
  T* s[3];
  const T* dn[2];
  int axis2;
  const T* lam;
  T lam_s;
  const T* mu;
  T mu_s;
  T kp, ks;            // 1 + chi_p, 1 + chi_s (Kelvin-Voigt factors)
};

// one pair element (line a in .x, line b in .y) of f at q
template <typename T, int N, int V>
__device__ __forceinline__ Cx<T> el_ld(const LineCtx<T, N, V>& L, const T* __restrict__ f0, int q) {
  using V2 = typename LineCtx<T, N, V>::V2;
  const T* __restrict__ f = f0 + L.wgb;
  if (L.vec) {
    if (!L.va) return Cx<T>{(T)0, (T)0};
    const V2 r = at32(reinterpret_cast<const V2*>(f), L.ia(q) >> 1);
    return Cx<T>{r.x, r.y};
  }
  return Cx<T>{L.va ? at32(f, L.ia(q)) : (T)0, L.vb ? at32(f, L.ib(q)) : (T)0};
}
template <typename T, int N, int V>
__device__ __forceinline__ void el_st(const LineCtx<T, N, V>& L, T* __restrict__ f0, int q, Cx<T> v) {
  using V2 = typename LineCtx<T, N, V>::V2;
  T* __restrict__ f = f0 + L.wgb;
  if (L.vec) {
    if (L.va) at32(reinterpret_cast<V2*>(f), L.ia(q) >> 1) = V2{v.x, v.y};
    return;
  }
  if (L.va) at32(f, L.ia(q)) = v.x;
  if (L.vb) at32(f, L.ib(q)) = v.y;
}
// medium value: map pair at q, or the scalar for both lines
template <typename T, int N, int V>
__device__ __forceinline__ Cx<T> el_med(const LineCtx<T, N, V>& L, const T* m, T s, int q) {
  return m ? el_ld(L, m, q) : Cx<T>{s, s};
}

// MODE is a template parameter so that each epilogue gets its own register
// allocation (the whole-line loads of the velocity and shear epilogues would
// otherwise cost the derivative and normal-stress passes occupancy).
template <typename T, int N, bool CONTIG, int MODE>
__global__ void __launch_bounds__(512) k_el_pass(ElArgs<T> a) {
  constexpr int V = CONTIG ? 2 : 0;  // contiguous lines: plan family 2 (fewer, longer rows per thread)
  constexpr bool OPQ = dir_opaque<T, N, CONTIG>();   // as for the acoustic directional passes
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

  Cx<T> v[E];
  L.load(a.f, v);
  line_fft<T, N, false, V, false, OPQ>(v, lds, tw, L.t);
#pragma unroll
  for (int q = 0; q < E; ++q) v[q] = cmul(v[q], a.D[L.n(q)]);
  line_fft<T, N, true, V, false, OPQ>(v, lds, tw, L.t);

  if constexpr (MODE == EL_DERIV) {
    if (a.accumulate) {
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = cadd(v[q], el_ld(L, a.out, q));
    }
    L.store(a.out, v);
    return;
  }
  if constexpr (MODE == EL_VEL) {
    // whole-line loads first (more loads in flight), then the update
    Cx<T> u[E];
    L.load(a.v, u);
    if (a.acc) {
      Cx<T> g[E];
      L.load(a.acc, g);
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = cadd(v[q], g[q]);
    }
    if (a.rho) {
      Cx<T> r0[E], r1[E];
      L.load_map(a.rho, r0, false, (T)0, (T)1);
      L.load_map_next(a.rho, r1, false, (T)0, (T)1);
#pragma unroll
      for (int q = 0; q < E; ++q) {
        v[q].x *= fdiv((T)2 * a.dt, r0[q].x + r1[q].x);
        v[q].y *= fdiv((T)2 * a.dt, r0[q].y + r1[q].y);
      }
    } else {
      const T f = a.dt / a.rho_s;
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = cscale(v[q], f);
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const T sg = a.pml_sg[L.n(q)];
      u[q] = Cx<T>{sg * (sg * u[q].x + v[q].x), sg * (sg * u[q].y + v[q].y)};
    }
    L.store(a.v, u);
    return;
  }

  const T* __restrict__ pml_line = a.axis == 0 ? a.pml[0] : (a.axis == 1 ? a.pml[1] : a.pml[2]);
  // per-line PML factors of the axes other than the line axis
  T pa[3] = {(T)1, (T)1, (T)1}, pb[3] = {(T)1, (T)1, (T)1};
  {
    const long long ba = L.wgb + L.oa, bb = L.wgb + L.ob;
    const long long ra = ba / a.Nx, rb = bb / a.Nx;
    const long long ca[3] = {ba % a.Nx, ra % a.Ny, ra / a.Ny}, cb[3] = {bb % a.Nx, rb % a.Ny, rb / a.Ny};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d >= a.ndim || d == a.axis) continue;
      if (L.va) pa[d] = a.pml[d][ca[d]];
      if (L.vb) pb[d] = a.pml[d][cb[d]];
    }
  }
  if constexpr (MODE == EL_NORMAL) {
    const bool three = a.ndim == 3;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const Cx<T> d0 = el_ld(L, a.dn[0], q);
      const Cx<T> d1 = three ? el_ld(L, a.dn[1], q) : v[q];
      const Cx<T> d2 = three ? v[q] : Cx<T>{(T)0, (T)0};
      const Cx<T> div = cadd(cadd(d0, d1), d2);
      const Cx<T> lm = el_med(L, a.lam, a.lam_s, q), mu = el_med(L, a.mu, a.mu_s, q);
      const T pn = pml_line[L.n(q)];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k == 2 && !three) break;
        const Cx<T> dk = k == 0 ? d0 : (k == 1 ? d1 : d2);
        const T fa = k == a.axis ? pn : pa[k], fb = k == a.axis ? pn : pb[k];
        const Cx<T> s = el_ld(L, a.s[k], q);
        const T ix = lm.x * div.x + (T)2 * mu.x * dk.x, iy = lm.y * div.y + (T)2 * mu.y * dk.y;
        el_st(L, a.s[k], q,
              Cx<T>{fa * (fa * s.x + a.dt * a.kp * ix), fb * (fb * s.y + a.dt * a.kp * iy)});
      }
    }
    return;
  }
  // EL_SHEAR
  const T qa = a.axis2 == 0 ? pa[0] : (a.axis2 == 1 ? pa[1] : pa[2]);
  const T qb = a.axis2 == 0 ? pb[0] : (a.axis2 == 1 ? pb[1] : pb[2]);
  Cx<T> s[E];
  L.load(a.s[0], s);
  if (a.acc) {
    Cx<T> g[E];
    L.load(a.acc, g);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = cadd(v[q], g[q]);
  }
  if (a.mu) {
    Cx<T> m[E];
    L.load(a.mu, m);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = Cx<T>{m[q].x * v[q].x, m[q].y * v[q].y};
  } else {
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = cscale(v[q], a.mu_s);
  }
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const T pn = pml_line[L.n(q)];
    const T fa = pn * qa, fb = pn * qb;
    s[q] = Cx<T>{fa * (s[q].x + a.dt * a.ks * v[q].x), fb * (s[q].y + a.dt * a.ks * v[q].y)};
  }
  L.store(a.s[0], s);
}

// p = -(sum of normal stresses) / ndim at the sensor points only
template <typename T>
__global__ void __launch_bounds__(256) k_el_sensor_p(T* p, const T* s0, const T* s1, const T* s2,
                                                      const long long* idx, int n, int ndim) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const long long i = idx[j];
  T tr = s0[i] + s1[i];
  if (ndim == 3) tr += s2[i];
  p[i] = -tr / (T)ndim;
}

// source values of column col (time-major [col][n]) added to / written over f[idx]
template <typename T>
__global__ void __launch_bounds__(256) k_el_inject(T* f, const long long* idx, const T* vals, int n, int dirichlet) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const long long i = idx[j];
  f[i] = dirichlet ? vals[j] : f[i] + vals[j];
}

}  // namespace kw
