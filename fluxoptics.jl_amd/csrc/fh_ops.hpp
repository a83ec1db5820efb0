// Pointwise operators fused into the line passes, and the standalone elementwise
// kernels (mask multiply, mask gradient, transfer-function tables).
#pragma once

#include "fh_lines.hpp"

namespace fh {

// ---------------------------------------------------------------- frequency-side ops

// Multiply by a transfer function stored as one quadrant (the function depends on
// |fx| and |fy| only), column-major: entry (qx, qy) = (min(kx, nx-kx), min(ky, ny-ky)) at
// qx*qcol + qy, so the values one column of the y pass needs are contiguous.
// `tab` holds one quadrant per table index; `tabidx[b]` picks the table of field b.
// The 1/(nx*ny) of the inverse FFT is folded into the table.  T is the table's complex
// type (c64, or c128 for the FP64-compute slice steps).
template <typename T> struct QuadMulT {
  const T* tab;
  const int* tabidx;
  long tabsize;
  int nx, ny, qcol;
  int conj;     // Backward direction multiplies by the conjugate
  int one_tab;  // all fields share one table
  int full = 0; // full tables: every column kx; column entries ky < n/2 at ky, the rest at n/2 + n - ky
  // the value at (kx, ky) (pre) and the multiply (apply): kernels that know their positions up
  // front issue the loads early
  template <typename C>
  __device__ __forceinline__ C pre(int b, int kx, int ky) const {
    const int qxi = full ? kx : (kx <= nx - kx ? kx : nx - kx);
    const int qyi = full ? (ky < ny / 2 ? ky : ny / 2 + ny - ky) : (ky <= ny - ky ? ky : ny - ky);
    return cvt<C>(tab[tabidx[b] * tabsize + (long)qxi * qcol + qyi]);
  }
  template <typename C>
  __device__ __forceinline__ C apply(C v, C h) const { return conj ? cmulc(v, h) : cmul(v, h); }
  template <typename C>
  __device__ __forceinline__ C operator()(C v, int b, int kx, int ky) const {
    return apply(v, pre<C>(b, kx, ky));
  }
};
using QuadMul = QuadMulT<float2>;

template <typename T> struct FreqPre<QuadMulT<T>> : std::true_type {};

// Multiply by the transfer function of a 1-D field (Fields.jl:135-136): one table of n entries
// in FFT order per table index, the line index is the field (the batch runs as the lines of
// one pass); 1/n of the inverse FFT folded in.
template <typename T> struct LineTabMulT {
  const T* tab;
  const int* tabidx;
  int n;
  int conj;  // Backward direction multiplies by the conjugate
  template <typename C>
  __device__ __forceinline__ C pre(int, int field, int k) const {
    return cvt<C>(tab[(long)tabidx[field] * n + k]);
  }
  template <typename C>
  __device__ __forceinline__ C apply(C v, C h) const { return conj ? cmulc(v, h) : cmul(v, h); }
  template <typename C>
  __device__ __forceinline__ C operator()(C v, int b, int field, int k) const {
    return apply(v, pre<C>(b, field, k));
  }
};
template <typename T> struct FreqPre<LineTabMulT<T>> : std::true_type {};

// Plain multiply by a full table (FP64 comparator path); the table is [ky][kx].
struct FullMulD {
  const double2* tab;
  int nx;
  int conj;
  __device__ __forceinline__ double2 operator()(double2 v, int, int kx, int ky) const {
    const double2 h = tab[(long)ky * nx + kx];
    return conj ? cmulc(v, h) : cmul(v, h);
  }
};

// ---------------------------------------------------------------- time-side ops (x lines)

// u * cis(s*phi[y, x]), optionally saving the input first.  The line index is the row y
// and the line position is x; the phase factors cis(s*phi) are one nx*ny c64 array shared by
// every field, formed once per call (k_phase_cis): evaluating sincos per field would repeat it
// batch times and its registers would halve the pass's resident waves.
// In the fast x pass (TwoPhaseOp) the factor loads come first, for every point of a thread, so
// that they are in flight with the field's loads rather than queued behind the stores of the saved
// input, which may alias them; there the factors and the saved input use the field's own buffer
// offsets, since they share its [ny][nx] layout.
struct PhaseTimeOp {
  const float2* cis;  // [ny][nx]
  float2* saved;      // nullable: [B][ny][nx]
  long fstride;       // nx*ny
  int nx;
  __device__ __forceinline__ float2 operator()(float2 v, int b, int y, int x) const {
    if (x >= nx) return v;
    const long i = (long)y * nx + x;
    if (saved) saved[b * fstride + i] = v;
    return cmul(v, cis[i]);
  }
};

// Time-side ops that the fast x pass applies in a load phase and an apply phase (see above).
template <typename Op> struct TwoPhaseOp : std::false_type {};
template <> struct TwoPhaseOp<PhaseTimeOp> : std::true_type {};

// Separable chirp on the time side of an x pass (CollinsProp's A and D chirps,
// collins_integral.jl:5-43): u(b, y, x) *= s * cx[t][x] * cy[t][y] (conjugated when conj),
// t = idx[b] the field's chirp table.  Positions x >= nx (zero padding) are left alone.
struct ChirpOp {
  const float2* cx;  // [tables][nx]
  const float2* cy;  // [tables][ny]
  const int* idx;    // [batch]
  int nx, ny;
  int conj;
  float s;
  __device__ __forceinline__ float2 operator()(float2 v, int b, int y, int x) const {
    if (x >= nx) return v;
    const int t = idx[b];
    const float2 h = cmul(cx[(long)t * nx + x], cy[(long)t * ny + y]);
    const float2 w = conj ? cmulc(v, h) : cmul(v, h);
    return make_float2(w.x * s, w.y * s);
  }
};

// The DotProduct / PowerCoupling pullback as the prologue of the adjoint's first x pass, which reads
// the metric's reference fields v: v_b * g_b with g_b = dc[b] (DotProduct, dot_product.jl:97-106)
// or 2 c[b] dp[b] (PowerCoupling, power_coupling.jl:66-69; dp null = ones), g rounded to c64 first
// as the standalone pullback kernel does.  Padding positions hold zeros and stay zero.
struct DotGainOp {
  const double2* dc;
  const double2* c;
  const double* dp;
  __device__ __forceinline__ float2 operator()(float2 v, int b, int, int) const {
    double gr, gi;
    if (dc) {
      gr = dc[b].x;
      gi = dc[b].y;
    } else {
      const double w = 2.0 * (dp ? dp[b] : 1.0);
      gr = w * c[b].x;
      gi = w * c[b].y;
    }
    const float fr = (float)gr, fi = (float)gi;
    return make_float2(v.x * fr - v.y * fi, v.x * fi + v.y * fr);
  }
};

// Multiply by a real constant (the 1/(nx*ny) of an inverse 2-D FFT, in the last pass).
struct ScaleOp {
  float s;
  template <typename C>
  __device__ __forceinline__ C operator()(C v, int, int, int) const {
    v.x *= s;
    v.y *= s;
    return v;
  }
};

}  // namespace fh
