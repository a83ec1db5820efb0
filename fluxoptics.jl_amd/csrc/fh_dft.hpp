// Register-resident DFT building blocks for the line-FFT passes (gfx950, wave64).
//
// A complex value is float2 (c64) or double2 (c128).  Every DFT here operates on a
// small local array whose size and strides are template constants, so after full
// unrolling the data lives in VGPRs and every internal twiddle is a literal.
//
// Conventions: forward = e^{-2*pi*i*jk/N}, inverse = e^{+2*pi*i*jk/N}, both
// unnormalised; input and output in natural order.
#pragma once

#include <hip/hip_runtime.h>

namespace fh {

template <typename T> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename cplx<T>::type;

template <typename C> __device__ __forceinline__ C cmake(decltype(C{}.x) a, decltype(C{}.x) b) {
  C r; r.x = a; r.y = b; return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmake<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmake<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmake<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  return cmake<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmake<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C{}.x) s) { return cmake<C>(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename C> __device__ __forceinline__ C cmul_mi(C a) {
  return INV ? cmake<C>(-a.y, a.x) : cmake<C>(a.y, -a.x);
}

// ---------------------------------------------------------------- compile-time twiddles
// cos(2*pi*m/N) and sin(2*pi*m/N) evaluated in double at compile time: the angle is
// folded into the first octant by symmetry and expanded as a Taylor series there
// (|x| <= pi/4, 14 terms: truncation far below double rounding).
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) { term *= -x * x / ((2 * k) * (2 * k + 1)); sum += term; }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) { term *= -x * x / ((2 * k - 1) * (2 * k)); sum += term; }
  return sum;
}
// returns cos (which=0) or sin (which=1) of 2*pi*m/N for integer m, N > 0
constexpr double turn_cs(long m, long N, int which) {
  m %= N; if (m < 0) m += N;
  // angle = 2*pi*m/N = (pi/4) * (8m/N); octant o = floor(8m/N), remainder r
  long num = 8 * m;
  long o = num / N;
  long rem = num - o * N;            // angle within octant = (pi/4) * rem/N
  double x = (kPi / 4) * (double)rem / (double)N;
  double c = taylor_cos(x), s = taylor_sin(x);
  // rotate by o * 45 degrees, using exact reflections for the cos/sin of the octant
  double cx = c, sx = s;
  // angle = o*pi/4 + x; express via quadrant q = o/2 and half-octant flag
  long q = o / 2;
  if (o % 2 == 1) {
    // pi/4 + x = pi/2 - (pi/4 - x): cos = sin(pi/4 - x), sin = cos(pi/4 - x)
    double y = (kPi / 4) * (double)(N - rem) / (double)N;
    cx = taylor_sin(y); sx = taylor_cos(y);
  }
  double cr = cx, sr = sx;
  switch (q) {
    case 0: cr = cx; sr = sx; break;
    case 1: cr = -sx; sr = cx; break;
    case 2: cr = -cx; sr = -sx; break;
    default: cr = sx; sr = -cx; break;
  }
  return which == 0 ? cr : sr;
}

// W_N^m (forward: e^{-2 pi i m/N}; inverse: conjugate) as a compile-time table
template <int N> struct TwTab {
  double c[N];
  double s[N];
  constexpr TwTab() : c(), s() {
    for (int m = 0; m < N; ++m) { c[m] = turn_cs(m, N, 0); s[m] = turn_cs(m, N, 1); }
  }
};
template <int N> struct Tw { static constexpr TwTab<N> tab{}; };
template <int N> constexpr TwTab<N> Tw<N>::tab;

// multiply a by W_N^{m} (m a compile-time constant after unrolling)
template <int N, bool INV, typename C>
__device__ __forceinline__ C twmul_const(C a, int m) {
  using R = decltype(C{}.x);
  m &= (N - 1);
  if (m == 0) return a;
  if (4 * m == N) return cmul_mi<INV>(a);
  if (2 * m == N) return cmake<C>(-a.x, -a.y);
  if (4 * m == 3 * N) return cmul_mi<!INV>(a);
  const R c = (R)Tw<N>::tab.c[m];
  const R s = INV ? (R)Tw<N>::tab.s[m] : -(R)Tw<N>::tab.s[m];
  return cmake<C>(a.x * c - a.y * s, a.x * s + a.y * c);
}

// ---------------------------------------------------------------- register DFTs
// dft<N, INV>(v): in-place natural-order DFT of v[0..N-1].
template <int N, bool INV, typename C> struct Dft;

template <bool INV, typename C> struct Dft<1, INV, C> {
  static __device__ __forceinline__ void run(C*) {}
};

template <bool INV, typename C> struct Dft<2, INV, C> {
  static __device__ __forceinline__ void run(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  }
};

template <bool INV, typename C> struct Dft<4, INV, C> {
  static __device__ __forceinline__ void run(C* v) {
    C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]), t3 = cmul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
  }
};

// N = P*Q four-step in registers: DFT_P over stride-Q subsequences, twiddle W_N^{q k1},
// DFT_Q over q, output index k1 + P*k2.
template <int N> struct Split { static constexpr int P = (N >= 16) ? 4 : 2; static constexpr int Q = N / P; };

template <int N, bool INV, typename C> struct Dft {
  static __device__ __forceinline__ void run(C* v) {
    constexpr int P = Split<N>::P, Q = Split<N>::Q;
    C y[Q][P];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      C t[P];
#pragma unroll
      for (int p = 0; p < P; ++p) t[p] = v[Q * p + q];
      Dft<P, INV, C>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) y[q][k1] = twmul_const<N, INV>(t[k1], q * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      C t[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) t[q] = y[q][k1];
      Dft<Q, INV, C>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = t[k2];
    }
  }
};

template <int N, bool INV, typename C> __device__ __forceinline__ void dft(C* v) { Dft<N, INV, C>::run(v); }

// ---------------------------------------------------------------- line decomposition
// A line of n = 2^LOGN points is split n = N1*N2 with N1 <= N2 (N1 = 2^floor(LOGN/2)).
// Each line is handled by N1 threads; every thread holds N2 values in both stages:
//   stage A ("time side"): thread p owns b in {p + N1*i : i < N2/N1}, values at
//     element N2*a + b for a < N1, and does DFT_N1 over a;
//   stage B ("frequency side"): thread k1 owns elements k1 + N1*k2 for k2 < N2 and
//     does DFT_N2 over k2.
template <int LOGN> struct LineShape {
  static constexpr int n = 1 << LOGN;
  static constexpr int N1 = 1 << (LOGN / 2);
  static constexpr int N2 = n / N1;
  static constexpr int MA = N2 / N1;       // stage-A sub-DFTs per thread (1 or 2)
  static constexpr int P = N1;              // threads per line
  static constexpr int ROW = N2 + 1;        // LDS row (one k1) with one pad slot
  static constexpr int LS = N1 * ROW + 2;   // LDS stride between lines (complex units)
};

}  // namespace fh
