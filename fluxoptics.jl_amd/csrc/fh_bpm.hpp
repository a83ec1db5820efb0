// AS_BPM slice kernels.  Between slices the field is kept as its x-spectrum (rows
// transformed along x), so one slice costs two passes: this x pass (IFFT_x, slice
// phase, FFT_x) and the y pass of the propagator (FFT_y * H * IFFT_y).
//
// The slice steps store c64 in HBM but compute in FP64: hundreds of near-identity steps
// make FP32 butterfly rounding errors add up coherently (about 3e-5 after 400 slices of
// 1 um), while rounding only the stored values keeps the error near 1e-6.
#pragma once

#include "fh_ops.hpp"

namespace fh {

template <typename R> __device__ __forceinline__ void sincos_any(R a, R* s, R* c);
template <> __device__ __forceinline__ void sincos_any<float>(float a, float* s, float* c) { sincosf(a, s, c); }
template <> __device__ __forceinline__ void sincos_any<double>(double a, double* s, double* c) { sincos(a, s, c); }

// Slice phase cis(sgn * kl[b] * dn[y, x]) with kl[b] = kdz / lambda_b
// (apply_dn_slice!, bpm.jl:235-242).  Used as a prologue of a forward x FFT.
struct DnPhaseOp {
  const float* dn;   // [ny][nx] slice
  const double* kl;  // [B]
  double sgn;
  int nx;
  template <typename C>
  __device__ __forceinline__ C operator()(C v, int b, int y, int x) const {
    using R = decltype(C{}.x);
    if (x >= nx) return v;
    R sn, cs;
    sincos_any<R>((R)(sgn * kl[b] * (double)dn[(long)y * nx + x]), &sn, &cs);
    return cmul(v, cmake<C>(cs, sn));
  }
};

// Stores the time-domain value as c64 (epilogue of an inverse x FFT).
struct SaveOp {
  float2* dst;   // [B][ny][nx]
  long fstride;
  int nx;
  template <typename C>
  __device__ __forceinline__ C operator()(C v, int b, int y, int x) const {
    if (x < nx) dst[b * fstride + (long)y * nx + x] = cvt<float2>(v);
    return v;
  }
};

// Forward slice step on x-spectrum rows, in place: IFFT_x -> [save slice input] ->
// multiply by the slice phase -> FFT_x.  The inverse plan ends in the register layout
// the forward plan starts from, so the three steps share registers.  `ckspec`, when
// set, receives the loaded x-spectrum rows (a checkpoint that replays this very step).
template <typename C, typename RDF, int L>
__global__ void __launch_bounds__(L * RDF::P)
k_bpm_xstep(LineIO<float2> io, const C* __restrict__ tw, DnPhaseOp ph, float2* __restrict__ save,
            float2* __restrict__ ckspec) {
  using RI = typename RDF::Rev;
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  extern __shared__ unsigned char smem_raw[];
  C* smem = reinterpret_cast<C*>(smem_raw);
  int l, t;
  thread_map<L, P, false>(threadIdx.x, l, t);
  const int b = blockIdx.y;
  const int li = blockIdx.x * L + l;
  const bool live = li < io.nlines;
  C* s = smem + l * LdsLine<n, L, sizeof(C)>::size;
  const long roff = b * io.out_f + (long)li * io.out_l;
  float2* row = io.out + roff;
  C v[V];
#pragma unroll
  for (int m = 0; m < V / RI::R0; ++m)
#pragma unroll
    for (int r = 0; r < RI::R0; ++r) {
      const int pos = reg_pos<n, RI::R0, P>(t, m, r);
      float2 x = make_float2(0.f, 0.f);
      if (live) {
        x = row[pos];
        if (ckspec) ckspec[roff + pos] = x;
      }
      v[m * RI::R0 + r] = cvt<C>(x);
    }
  line_fft<C, RI, true>(v, s, t, tw);
#pragma unroll
  for (int m = 0; m < V / RDF::R0; ++m)
#pragma unroll
    for (int r = 0; r < RDF::R0; ++r) {
      const int pos = reg_pos<n, RDF::R0, P>(t, m, r);
      C x = v[m * RDF::R0 + r];
      if (live && pos < ph.nx) {
        if (save) save[b * io.in_f + (long)li * ph.nx + pos] = cvt<float2>(x);
        x = ph(x, b, li, pos);
      }
      v[m * RDF::R0 + r] = x;
    }
  line_fft<C, RDF, false>(v, s, t, tw);
  constexpr int RL = (RDF::NST == 3) ? RDF::R2 : ((RDF::NST == 2) ? RDF::R1 : RDF::R0);
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int pos = reg_pos<n, RL, P>(t, m, r);
      if (live) row[pos] = cvt<float2>(v[m * RL + r]);
    }
}

// Adjoint slice step with the dn gradient, in place on x-spectrum rows of the
// cotangent: IFFT_x -> du = dv * cis(-sgn*kl[b]*dn) -> g += sgn*kl[b]*Im(du*conj(u_k))
// -> FFT_x.  One workgroup owns L rows and loops over the batch, so the gradient
// (compute_dn_gradient!, bpm.jl:275-287) is summed over fields in order and written once.
template <typename C, typename RDF, int L>
__global__ void __launch_bounds__(L * RDF::P)
k_bpm_xstep_grad(LineIO<float2> io, int batch, const C* __restrict__ tw, DnPhaseOp ph,
                 const float2* __restrict__ saved, float* __restrict__ ddn) {
  using R = decltype(C{}.x);
  using RI = typename RDF::Rev;
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  constexpr int RL = (RDF::NST == 3) ? RDF::R2 : ((RDF::NST == 2) ? RDF::R1 : RDF::R0);
  extern __shared__ unsigned char smem_raw[];
  C* smem = reinterpret_cast<C*>(smem_raw);
  int l, t;
  thread_map<L, P, false>(threadIdx.x, l, t);
  const int li = blockIdx.x * L + l;
  const bool live = li < io.nlines;
  C* s = smem + l * LdsLine<n, L, sizeof(C)>::size;
  const int nx = ph.nx;
  R g[V];
#pragma unroll
  for (int k = 0; k < V; ++k) g[k] = 0;
  for (int b = 0; b < batch; ++b) {
    float2* row = io.out + b * io.out_f + (long)li * io.out_l;
    const float2* sv = saved + b * io.in_f + (long)li * nx;
    const R coef = (R)(ph.sgn * ph.kl[b]);
    C v[V];
#pragma unroll
    for (int m = 0; m < V / RI::R0; ++m)
#pragma unroll
      for (int r = 0; r < RI::R0; ++r) {
        const int pos = reg_pos<n, RI::R0, P>(t, m, r);
        v[m * RI::R0 + r] = live ? cvt<C>(row[pos]) : cmake<C>(0, 0);
      }
    line_fft<C, RI, true>(v, s, t, tw);
#pragma unroll
    for (int m = 0; m < V / RDF::R0; ++m)
#pragma unroll
      for (int r = 0; r < RDF::R0; ++r) {
        const int pos = reg_pos<n, RDF::R0, P>(t, m, r);
        if (live && pos < nx) {
          R sn, cs;
          sincos_any<R>(-coef * (R)ph.dn[(long)li * nx + pos], &sn, &cs);
          const C du = cmul(v[m * RDF::R0 + r], cmake<C>(cs, sn));
          v[m * RDF::R0 + r] = du;
          g[m * RDF::R0 + r] += coef * cmulc(du, cvt<C>(sv[pos])).y;
        }
      }
    line_fft<C, RDF, false>(v, s, t, tw);
#pragma unroll
    for (int m = 0; m < V / RL; ++m)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int pos = reg_pos<n, RL, P>(t, m, r);
        if (live) row[pos] = cvt<float2>(v[m * RL + r]);
      }
  }
  if (live) {
#pragma unroll
    for (int m = 0; m < V / RDF::R0; ++m)
#pragma unroll
      for (int r = 0; r < RDF::R0; ++r) {
        const int pos = reg_pos<n, RDF::R0, P>(t, m, r);
        if (pos < nx) ddn[(long)li * nx + pos] = (float)g[m * RDF::R0 + r];
      }
  }
}

}  // namespace fh
