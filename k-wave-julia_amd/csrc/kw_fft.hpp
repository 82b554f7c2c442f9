// In-register / LDS line FFT engine for gfx950 (CDNA4, wave64).
//
// A "line" of N complex values is owned by TPL consecutive threads of a
// workgroup.  Thread t holds v[q] = x[t + TPL*q] for q < E = N/TPL (the
// "stride-TPL" layout).  A mixed-radix Stockham autosort transform with the
// compile-time radix list Rs... runs the butterflies of every pass in
// registers; passes exchange through the line's LDS slice.  The last pass
// writes its outputs straight back into registers, and because the Stockham
// autosort output of the final pass lands at j + r*N/R, the output layout is
// the same stride-TPL layout as the input.  Forward and inverse transforms
// therefore chain without any reordering, which is what lets the pass
// kernels do  FFT -> pointwise multiply -> IFFT  with one LDS exchange per
// transform for N = R0*R1.
//
// Forward uses W = exp(-2 pi i / N); the inverse is unnormalised (callers
// fold 1/N into their spectral multipliers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kw {

template <typename T>
struct Cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ Cx<T> cmulc(Cx<T> a, Cx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ Cx<T> cscale(Cx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ Cx<T> rot90(Cx<T> a) {
  return INV ? Cx<T>{-a.y, a.x} : Cx<T>{a.y, -a.x};
}

// ---------------------------------------------------------------------------
// Compile-time trigonometry for the in-register butterfly constants.
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ce_sin_red(double x) {  // |x| <= pi/4, Taylor to 1e-20
  double term = x, sum = x, x2 = x * x;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double ce_cos_red(double x) {
  double term = 1.0, sum = 1.0, x2 = x * x;
  for (int n = 1; n < 14; ++n) {
    term *= -x2 / ((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}
// cos/sin of 2*pi*num/den with exact octant reduction on the rational angle
constexpr void ce_sincos_frac(long num, long den, double& s, double& c) {
  num %= den;
  if (num < 0) num += den;
  // angle = 2 pi num/den; octant = floor(8 num / den)
  long o = (8 * num) / den;
  // remainder angle in [0, pi/4): r = 2pi*num/den - o*pi/4
  double r = 2.0 * kPi * (double)(8 * num - o * den) / (8.0 * (double)den);
  double sr = ce_sin_red(r), cr = ce_cos_red(r);
  // rotate by o*pi/4 using exact octant identities
  const double h = 0.70710678118654752440084436210484904;
  double c8[8] = {1, h, 0, -h, -1, -h, 0, h};
  double s8[8] = {0, h, 1, h, 0, -h, -1, -h};
  c = cr * c8[o] - sr * s8[o];
  s = sr * c8[o] + cr * s8[o];
}

template <int R>
struct RootTable {
  double c[R > 0 ? R : 1];
  double s[R > 0 ? R : 1];
  constexpr RootTable() : c(), s() {
    for (int k = 0; k < R; ++k) {
      double sv = 0, cv = 0;
      ce_sincos_frac(k, R, sv, cv);
      c[k] = cv;
      s[k] = sv;
    }
  }
};

// W_R^m = exp(-2 pi i m / R) (forward), conjugate for the inverse.
template <typename T, int R, bool INV>
__device__ __forceinline__ Cx<T> root(int m) {
  constexpr RootTable<R> tab{};
  m %= R;
  return Cx<T>{(T)tab.c[m], INV ? (T)tab.s[m] : (T)(-tab.s[m])};
}

// ---------------------------------------------------------------------------
// Small DFTs in registers (natural order in, natural order out).
// ---------------------------------------------------------------------------
template <int R>
struct SmallestFactor {
  static constexpr int f() {
    for (int p = 2; p * p <= R; ++p)
      if (R % p == 0) return (p == 2 && R % 4 == 0 && R > 4) ? 4 : p;
    return R;
  }
  static constexpr int value = f();
};

template <typename T, int R, bool INV>
struct DFT {
  static constexpr int A = SmallestFactor<R>::value;  // outer radix
  static constexpr int B = R / A;
  __device__ __forceinline__ static void run(Cx<T> (&x)[R]) {
    if constexpr (A == R) {
      // prime radix > 5: direct O(R^2) evaluation with folded constants
      Cx<T> y[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        Cx<T> acc = x[0];
#pragma unroll
        for (int n = 1; n < R; ++n) acc = cadd(acc, cmul(x[n], root<T, R, INV>(n * k)));
        y[k] = acc;
      }
#pragma unroll
      for (int k = 0; k < R; ++k) x[k] = y[k];
    } else {
      // Cooley-Tukey: n = n1 + A*n2, k = k2 + B*k1
      Cx<T> y[A][B];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) {
        Cx<T> col[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) col[n2] = x[n1 + A * n2];
        DFT<T, B, INV>::run(col);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2)
          y[n1][k2] = (n1 * k2 == 0) ? col[k2] : cmul(col[k2], root<T, R, INV>(n1 * k2));
      }
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) {
        Cx<T> row[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) row[n1] = y[n1][k2];
        DFT<T, A, INV>::run(row);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) x[k2 + B * k1] = row[k1];
      }
    }
  }
};

template <typename T, bool INV>
struct DFT<T, 1, INV> {
  __device__ __forceinline__ static void run(Cx<T> (&)[1]) {}
};
template <typename T, bool INV>
struct DFT<T, 2, INV> {
  __device__ __forceinline__ static void run(Cx<T> (&x)[2]) {
    Cx<T> a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};
template <typename T, bool INV>
struct DFT<T, 3, INV> {
  __device__ __forceinline__ static void run(Cx<T> (&x)[3]) {
    const T c1 = (T)-0.5;
    const T s1 = (T)(INV ? 0.86602540378443864676372317075294 : -0.86602540378443864676372317075294);
    Cx<T> t1 = cadd(x[1], x[2]);
    Cx<T> t2 = csub(x[1], x[2]);
    Cx<T> m = {x[0].x + c1 * t1.x, x[0].y + c1 * t1.y};
    // s1 * i * t2
    Cx<T> r = {-s1 * t2.y, s1 * t2.x};
    x[0] = cadd(x[0], t1);
    x[1] = cadd(m, r);
    x[2] = csub(m, r);
  }
};
template <typename T, bool INV>
struct DFT<T, 4, INV> {
  __device__ __forceinline__ static void run(Cx<T> (&x)[4]) {
    Cx<T> a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    Cx<T> b0 = cadd(x[1], x[3]), b1 = rot90<T, INV>(csub(x[1], x[3]));
    x[0] = cadd(a0, b0);
    x[2] = csub(a0, b0);
    x[1] = cadd(a1, b1);
    x[3] = csub(a1, b1);
  }
};
template <typename T, bool INV>
struct DFT<T, 5, INV> {
  __device__ __forceinline__ static void run(Cx<T> (&x)[5]) {
    const T c1 = (T)0.30901699437494742410229341718282;   // cos(2pi/5)
    const T c2 = (T)-0.80901699437494742410229341718282;  // cos(4pi/5)
    const T sg = INV ? (T)1 : (T)-1;
    const T s1 = sg * (T)0.95105651629515357211643933337938;  // sin(2pi/5)
    const T s2 = sg * (T)0.58778525229247312916870595463907;  // sin(4pi/5)
    Cx<T> a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
    Cx<T> a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
    Cx<T> m1 = {x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y};
    Cx<T> m2 = {x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y};
    // i*(s1*b1 + s2*b2), i*(s2*b1 - s1*b2)
    Cx<T> n1 = {-(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x};
    Cx<T> n2 = {-(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x};
    x[0] = cadd(x[0], cadd(a1, a2));
    x[1] = cadd(m1, n1);
    x[4] = csub(m1, n1);
    x[2] = cadd(m2, n2);
    x[3] = csub(m2, n2);
  }
};

// ---------------------------------------------------------------------------
// Line plans: threads per line and the radix list for each supported N.
// ---------------------------------------------------------------------------
template <int... Rs>
struct Radices {};

// V = 0: plans of the strided and spectral passes (lines spread across
// workgroup lanes, so more lines per workgroup = wider coalesced rows);
// V = 1: plans of the contiguous x-line passes, which may trade a larger
// TPL (fewer registers, more waves) for extra LDS exchanges; V = 2 (below)
// for the complex-to-real x pass.
template <int N, int V = 0>
struct Plan;
#define KW_PLAN(N_, TPL_, ...)                     \
  template <>                                      \
  struct Plan<N_, 0> {                             \
    static constexpr int N = N_;                   \
    static constexpr int TPL = TPL_;               \
    static constexpr int E = N_ / TPL_;            \
    using R = Radices<__VA_ARGS__>;                \
  };
KW_PLAN(8, 1, 8)
KW_PLAN(10, 1, 10)
KW_PLAN(12, 1, 12)
KW_PLAN(16, 1, 16)
KW_PLAN(20, 2, 10, 2)
KW_PLAN(24, 2, 12, 2)
KW_PLAN(32, 4, 8, 4)
KW_PLAN(40, 2, 20, 2)
KW_PLAN(48, 4, 12, 4)
KW_PLAN(64, 8, 8, 8)
KW_PLAN(80, 4, 20, 4)
KW_PLAN(96, 4, 24, 4)
KW_PLAN(128, 8, 16, 8)
KW_PLAN(160, 8, 20, 4, 2)
KW_PLAN(192, 8, 24, 8)
KW_PLAN(256, 16, 16, 16)
KW_PLAN(320, 16, 20, 4, 4)
KW_PLAN(384, 32, 12, 4, 4, 2)
KW_PLAN(512, 32, 16, 8, 4)
KW_PLAN(640, 32, 20, 4, 4, 2)
KW_PLAN(768, 64, 12, 4, 4, 4)
KW_PLAN(1024, 64, 16, 16, 4)
KW_PLAN(2048, 128, 16, 16, 8)
KW_PLAN(4096, 256, 16, 16, 16)
// odd and other composite lengths (grids need not be powers of two)
KW_PLAN(9, 1, 9)
KW_PLAN(15, 1, 15)
KW_PLAN(25, 5, 5, 5)
KW_PLAN(27, 3, 9, 3)
KW_PLAN(36, 3, 12, 3)
KW_PLAN(45, 3, 15, 3)
KW_PLAN(72, 6, 12, 6)
KW_PLAN(75, 5, 15, 5)
KW_PLAN(81, 9, 9, 9)
KW_PLAN(100, 10, 10, 10)
KW_PLAN(125, 25, 5, 5, 5)
KW_PLAN(135, 9, 15, 3, 3)
KW_PLAN(144, 12, 12, 12)
KW_PLAN(200, 10, 20, 10)
KW_PLAN(225, 15, 15, 15)
KW_PLAN(243, 27, 9, 9, 3)
KW_PLAN(400, 20, 20, 20)
#undef KW_PLAN
template <int N>
struct Plan<N, 1> : Plan<N, 0> {};
template <>
struct Plan<512, 1> {
  static constexpr int N = 512;
  static constexpr int TPL = 64;
  static constexpr int E = 8;
  using R = Radices<8, 8, 8>;
};
template <>
struct Plan<256, 1> {
  static constexpr int N = 256;
  static constexpr int TPL = 64;
  static constexpr int E = 4;
  using R = Radices<4, 4, 4, 4>;
};
// V = 2: the complex-to-real pass along x (and the adjoint's r2c stage),
// which carry fewer live arrays than the directional x pass and run faster
// with more elements per thread.
template <int N>
struct Plan<N, 2> : Plan<N, 1> {};
template <>
struct Plan<256, 2> {
  static constexpr int N = 256;
  static constexpr int TPL = 32;
  static constexpr int E = 8;
  using R = Radices<8, 8, 4>;
};

// V = 5: the lines of the fused y-inverse / complex-to-real plane kernel: the
// plans of family 2, with twiddles read from the table (the kernel's
// registers hold a whole plane of the spectrum).
template <int N>
struct Plan<N, 5> : Plan<N, 2> {};

// V = 4: contiguous-line passes (x pass, complex-to-real) of chirp-z lines,
// which carry the chirp and filter arrays on top of the field arrays: the long
// plans use 8 elements per thread (the 16-element plans spill registers there).
template <int N>
struct Plan<N, 4> : Plan<N, 1> {};
template <>
struct Plan<1024, 4> {
  static constexpr int N = 1024;
  static constexpr int TPL = 128;
  static constexpr int E = 8;
  using R = Radices<8, 8, 4, 4>;
};
template <>
struct Plan<2048, 4> {
  static constexpr int N = 2048;
  static constexpr int TPL = 256;
  static constexpr int E = 8;
  using R = Radices<8, 8, 8, 4>;
};

// LDS slice of one line: one complex of padding per 16 keeps the
// stride-16/32 Stockham writes off a single bank group.  The 4096 plan is
// left unpadded so a double-precision line still fits a 64 KiB workgroup.
template <int N>
struct LdsLayout {
  static constexpr bool PAD = N >= 16 && N < 4096;
  static constexpr int PITCH = PAD ? N + N / 16 + 1 : N;
  __device__ __forceinline__ static int at(int i) { return PAD ? i + (i >> 4) : i; }
};

// Twiddles of the passes after the first, W_N^((j mod Ns) r N/(Ns R)),
// depend only on the thread's position in the line.  For small plans they
// are loaded once per kernel into registers (all four transforms of a pass
// kernel reuse them, conjugated for the inverse); larger plans read the
// table, which stays L1-resident.
template <int... Rs>
struct RadixArray {
  static constexpr int n = sizeof...(Rs);
  static constexpr int r[n] = {Rs...};
};
template <int... Rs>
constexpr int RadixArray<Rs...>::r[];

template <int N, int V, typename RL>
struct TwCount;
template <int N, int V, int... Rs>
struct TwCount<N, V, Radices<Rs...>> {
  static constexpr int f() {
    constexpr int rr[] = {Rs...};
    int E = N / Plan<N, V>::TPL, Ns = 1, tot = 0;
    for (int i = 0; i < (int)sizeof...(Rs); ++i) {
      if (Ns > 1) tot += (E / rr[i]) * (rr[i] - 1);
      Ns *= rr[i];
    }
    return tot;
  }
  static constexpr int value = f();
};

template <typename T, int N, int V = 0>
struct Twiddles {
  static constexpr int TOT = TwCount<N, V, typename Plan<N, V>::R>::value;
  static constexpr bool REG = V != 5 && TOT > 0 && TOT <= 16;
  Cx<T> w[REG ? TOT : 1];
  const Cx<T>* g;
};

template <typename T, int N, int V, int TPL, int Ns, int OFF, int R, int... Rest>
struct TwLoad {
  __device__ __forceinline__ static void run(Twiddles<T, N, V>& tw, const Cx<T>* __restrict__ table, int t) {
    constexpr int NB = (N / TPL) / R;
    if constexpr (Ns > 1) {
#pragma unroll
      for (int m = 0; m < NB; ++m) {
        const int jm = (t + TPL * m) % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) tw.w[OFF + m * (R - 1) + (r - 1)] = table[jm * r * (N / (Ns * R))];
      }
    }
    if constexpr (sizeof...(Rest) > 0)
      TwLoad<T, N, V, TPL, Ns * R, (Ns > 1 ? OFF + NB * (R - 1) : OFF), Rest...>::run(tw, table, t);
  }
};
template <typename T, int N, int V, typename RL>
struct TwLoadImpl;
template <typename T, int N, int V, int... Rs>
struct TwLoadImpl<T, N, V, Radices<Rs...>> {
  __device__ __forceinline__ static void run(Twiddles<T, N, V>& tw, const Cx<T>* __restrict__ table, int t) {
    TwLoad<T, N, V, Plan<N, V>::TPL, 1, 0, Rs...>::run(tw, table, t);
  }
};

// Prepare the twiddles of a line of length N for thread t of the line.
template <typename T, int N, int V = 0>
__device__ __forceinline__ Twiddles<T, N, V> load_twiddles(const Cx<T>* __restrict__ table, int t) {
  Twiddles<T, N, V> tw;
  tw.g = table;
  if constexpr (Twiddles<T, N, V>::REG) TwLoadImpl<T, N, V, typename Plan<N, V>::R>::run(tw, table, t);
  return tw;
}

// LDS exchange synchronisation.  WAVE: every line of the workgroup lies within
// one wavefront (contiguous lines of TPL threads, TPL dividing 64), so the
// exchange needs only an ordering fence for the compiler and the wave's own
// in-order LDS queue, not a workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <typename T, int N, int V, int TPL, bool INV, int Ns, int OFF, bool WAVE, int R, int... Rest>
struct StockhamPass;
// the pass after one (void after the last)
template <typename T, int N, int V, int TPL, bool INV, int Ns, int OFF, bool WAVE, int... Rs>
struct StockhamNext {
  using type = StockhamPass<T, N, V, TPL, INV, Ns, OFF, WAVE, Rs...>;
};
template <typename T, int N, int V, int TPL, bool INV, int Ns, int OFF, bool WAVE>
struct StockhamNext<T, N, V, TPL, INV, Ns, OFF, WAVE> {
  using type = void;
};

template <typename T, int N, int V, int TPL, bool INV, int Ns, int OFF, bool WAVE, int R, int... Rest>
struct StockhamPass {
  static constexpr int E = N / TPL;
  static constexpr int NB = E / R;
  static_assert(E % R == 0, "radix must divide the per-thread element count");
  // this pass reads its twiddles from the table (plans too large to keep
  // them in registers); NW of them per thread
  static constexpr bool TAB = !Twiddles<T, N, V>::REG && Ns > 1;
  static constexpr int NW = TAB ? NB * (R - 1) : 1;
  using Next =
      typename StockhamNext<T, N, V, TPL, INV, Ns * R, (Ns > 1 ? OFF + NB * (R - 1) : OFF), WAVE, Rest...>::type;

  __device__ __forceinline__ static void fetch(Cx<T> (&w)[NW], const Twiddles<T, N, V>& tw, int t) {
    if constexpr (TAB) {
#pragma unroll
      for (int m = 0; m < NB; ++m) {
        const int jm = (t + TPL * m) % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) w[m * (R - 1) + (r - 1)] = tw.g[jm * r * (N / (Ns * R))];
      }
    }
  }
  __device__ __forceinline__ static void run(Cx<T> (&v)[E], Cx<T>* lds, const Twiddles<T, N, V>& tw, int t) {
    Cx<T> w[NW];
    fetch(w, tw, t);
    body(v, lds, tw, t, w);
  }
  // w: this pass's table twiddles, fetched by the caller
  __device__ __forceinline__ static void body(Cx<T> (&v)[E], Cx<T>* lds, const Twiddles<T, N, V>& tw, int t,
                                              const Cx<T> (&w)[NW]) {
    constexpr bool LAST = sizeof...(Rest) == 0;
    static_assert(!LAST || Ns * R == N, "radix product must equal N");
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      Cx<T> x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = v[m + r * NB];
      const int j = t + TPL * m;
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          Cx<T> wr;
          if constexpr (Twiddles<T, N, V>::REG)
            wr = tw.w[OFF + m * (R - 1) + (r - 1)];
          else
            wr = w[m * (R - 1) + (r - 1)];
          x[r] = INV ? cmulc(x[r], wr) : cmul(x[r], wr);
        }
      }
      DFT<T, R, INV>::run(x);
      if constexpr (LAST) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[m + r * NB] = x[r];
      } else {
        const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
        for (int r = 0; r < R; ++r) lds[LdsLayout<N>::at(base + r * Ns)] = x[r];
      }
    }
    if constexpr (!LAST) {
      // the next pass's table twiddles are requested before the exchange so
      // their latency overlaps the barrier and the LDS reads
      Cx<T> wn[Next::NW];
      Next::fetch(wn, tw, t);
      lds_sync<WAVE>();
#pragma unroll
      for (int q = 0; q < E; ++q) v[q] = lds[LdsLayout<N>::at(t + TPL * q)];
      lds_sync<WAVE>();
      Next::body(v, lds, tw, t, wn);
    }
  }
};

template <typename T, int N, int V, bool INV, bool WAVE, bool OPQ, typename RL>
struct LineFFTImpl;
template <typename T, int N, int V, bool INV, bool WAVE, bool OPQ, int... Rs>
struct LineFFTImpl<T, N, V, INV, WAVE, OPQ, Radices<Rs...>> {
  static constexpr int TPL = Plan<N, V>::TPL;
  static constexpr int E = Plan<N, V>::E;
  __device__ __forceinline__ static void run(Cx<T> (&v)[E], Cx<T>* lds, const Twiddles<T, N, V>& tw, int t) {
    // OPQ: an opaque copy of the thread index, so the address arithmetic of
    // this transform is not merged with that of the kernel's other
    // transforms (which keeps every LDS and twiddle address live across all
    // of them: fewer registers and more waves per SIMD for some kernels, more
    // address arithmetic for all).
    int tt = t;
    if constexpr (OPQ) asm volatile("" : "+v"(tt));
    StockhamPass<T, N, V, TPL, INV, 1, 0, WAVE, Rs...>::run(v, lds, tw, tt);
  }
};

// Lines of plan family V whose threads are consecutive (contiguous-line
// passes) lie inside one wavefront when TPL divides 64.
template <int N, int V>
constexpr bool wave_local_lines() {
  return 64 % Plan<N, V>::TPL == 0;
}

// Public entry: forward (INV=false) or unnormalised inverse transform of the
// line held in v (stride-TPL layout), using the line's LDS slice `lds`.
// Every thread of the workgroup must call it (it contains barriers).  WAVE:
// the caller's lines are wave-local (consecutive threads, TPL dividing 64),
// so the exchanges synchronise the wavefront only.
template <typename T, int N, bool INV, int V = 0, bool WAVE = false, bool OPQ = false>
__device__ __forceinline__ void line_fft(Cx<T> (&v)[(Plan<N, V>::E)], Cx<T>* lds, const Twiddles<T, N, V>& tw,
                                         int t) {
  static_assert(!WAVE || 64 % Plan<N, V>::TPL == 0, "wave-local lines need TPL dividing 64");
  LineFFTImpl<T, N, V, INV, WAVE, OPQ, typename Plan<N, V>::R>::run(v, lds, tw, t);
}

// ---------------------------------------------------------------------------
// Lines of any length n (Bluestein / chirp-z): the DFT of length n as a
// circular convolution on a compiled power-of-two plan of length N >= 2n - 1.
//   X_k = c_k sum_j (x_j c_j) d_(k-j),  c_j = exp(-i pi j^2 / n),  d = conj(c)
//       = c_k IFFT_N( G . FFT_N(c . x) ),  G = FFT_N(g) / N
// with g_m = d_m (0 <= m < n), g_(N-m) = d_m (0 < m < n), 0 elsewhere; the
// unnormalised inverse uses conj(c) and the filter of c.  The tables are
// generated in FP64 on the host.
// ---------------------------------------------------------------------------
template <typename T>
struct Chirp {
  int n;               // logical line length
  const Cx<T>* c;      // c_j, j < n
  const Cx<T>* gf;     // N values: forward filter spectrum, 1/N folded in
  const Cx<T>* gb;     // N values: inverse filter spectrum, 1/N folded in
};

// Forward (INV=false) or unnormalised inverse DFT of a line of length ch.n held
// in v (stride-TPL layout of the N-point plan; slots j >= n hold zero on exit
// and are ignored on entry).  BS=false is the plain N-point line_fft.
template <typename T, int N, bool INV, int V = 0, bool WAVE = false, bool BS = false, bool OPQ = false>
__device__ __forceinline__ void line_dft(Cx<T> (&v)[(Plan<N, V>::E)], Cx<T>* lds, const Twiddles<T, N, V>& tw, int t,
                                         const Chirp<T>& ch) {
  if constexpr (!BS) {
    line_fft<T, N, INV, V, WAVE, OPQ>(v, lds, tw, t);
  } else {
    constexpr int TPL = Plan<N, V>::TPL, E = Plan<N, V>::E;
    const int n = ch.n;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int j = t + TPL * q;
      v[q] = j < n ? (INV ? cmulc(v[q], ch.c[j]) : cmul(v[q], ch.c[j])) : Cx<T>{0, 0};
    }
    line_fft<T, N, false, V, WAVE, OPQ>(v, lds, tw, t);
    const Cx<T>* g = INV ? ch.gb : ch.gf;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = cmul(v[q], g[t + TPL * q]);
    line_fft<T, N, true, V, WAVE, OPQ>(v, lds, tw, t);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int j = t + TPL * q;
      v[q] = j < n ? (INV ? cmulc(v[q], ch.c[j]) : cmul(v[q], ch.c[j])) : Cx<T>{0, 0};
    }
  }
}

}  // namespace kw
