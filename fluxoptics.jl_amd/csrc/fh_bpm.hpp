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

// The slice phases of `nsl` consecutive slices, cis(sgn * kl[b] * dn) in FP64 with DnPhaseOp's
// expression, into out[slice][b][ny*nx]: computed ahead of the slice loop (every element at once,
// not one wave's dependent chain per row), so the slice steps read their phase instead of
// evaluating it, with the same values.
__global__ void k_bpm_cis(const float* __restrict__ dn, const double* __restrict__ kl, double sgn,
                          double2* __restrict__ out, long npix, int batch, int nsl) {
  const long total = npix * batch * nsl;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long p = i % npix;
    const long jb = i / npix;
    const int b = (int)(jb % batch);
    const long j = jb / batch;
    double sn, cs;
    sincos(sgn * kl[b] * (double)dn[j * npix + p], &sn, &cs);
    out[i] = make_double2(cs, sn);
  }
}

// Slice phase read from a k_bpm_cis table ([B][ny][nx] of one slice).
struct TabPhaseOp {
  const double2* cis;
  long fstride;  // ny * nx
  int nx;
  template <typename C>
  __device__ __forceinline__ C operator()(C v, int b, int y, int x) const {
    if (x >= nx) return v;
    const double2 c = cis[b * fstride + (long)y * nx + x];
    return cmul(v, cmake<C>(c.x, c.y));
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
template <typename C, typename RDF, int L, typename PH = DnPhaseOp>
__global__ void __launch_bounds__(L * RDF::P)
k_bpm_xstep(LineIO<float2> io, const C* __restrict__ tw, PH ph, float2* __restrict__ save,
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
// TAB: cis(-sgn*kl[b]*dn) is the conjugate of the forward's k_bpm_cis value (the arguments are
// exact negatives), read from `cis` ([B][ny][nx] of this slice) instead of evaluated.
template <typename C, typename RDF, int L, bool TAB = false>
__global__ void __launch_bounds__(L * RDF::P)
k_bpm_xstep_grad(LineIO<float2> io, int batch, const C* __restrict__ tw, DnPhaseOp ph,
                 const float2* __restrict__ saved, float* __restrict__ ddn, const double2* __restrict__ cis) {
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
          if constexpr (TAB) {
            const double2 c = cis[b * io.in_f + (long)li * nx + pos];
            cs = (R)c.x;
            sn = -(R)c.y;
          } else {
            sincos_any<R>(-coef * (R)ph.dn[(long)li * nx + pos], &sn, &cs);
          }
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

namespace fh {

// ---------------------------------------------------------------- 1-D fields
// A 1-D field (Fields.jl:135-136) runs the whole split-step loop of one field line in one
// workgroup: forward P(dz/2); per slice [save], phase, P(dz) (P(dz/2) after the last) (bpm.jl:
// 244-260), or the adjoint's reverse loop with the phase un-applied and the slice's gradient
// terms s * kl * Im(du * conj(saved)) (bpm.jl:275-312).  The line stays in FP64 registers across
// every slice (nothing is rounded between slices); each slice propagator is FFT_N -> transfer
// function -> IFFT_N on the line zero-padded to N points, cropped back to n after the inverse
// transform.  The inverse plan ends in the register layout the forward plan starts from, so a
// slice never leaves the registers except through the LDS exchanges of its transforms.
struct BpmLine {
  float2* u;             // [batch][n] fields, in place
  const float* dn;       // [ns][n] slices
  const double* kl;      // [batch] kdz / lambda_b (vacuum wavelength)
  const double2* tfull;  // [tables][N] P(dz), 1/N folded in
  const double2* thalf;  // [tables][N] P(dz/2)
  const int* tabidx;     // [batch] table of each field
  float2* save;          // forward: [ns][batch][n] slice inputs (nullable)
  const float2* saved;   // adjoint: the forward's slice inputs (nullable: no gradient)
  float* gpart;          // adjoint: [ns][batch][n] gradient terms (with saved)
  int n, batch, ns;
  int dir;               // direction of the forward propagation (+1 or -1)
};

template <typename RD>
__device__ __forceinline__ void bpm_line_prop(double2* v, double2* s, int t, const double2* __restrict__ tw,
                                              const double2* __restrict__ H, bool conj, int n) {
  using RI = typename RD::Rev;
  constexpr int N = RD::n, V = RD::V, P = RD::P, RL = RI::R0;
  line_fft<double2, RD, false>(v, s, t, tw);
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const double2 h = H[reg_pos<N, RL, P>(t, m, r)];
      v[m * RL + r] = conj ? cmulc(v[m * RL + r], h) : cmul(v[m * RL + r], h);
    }
  line_fft<double2, RI, true>(v, s, t, tw);
  if (N != n) {  // the padded line: back to zero outside the field
#pragma unroll
    for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
      for (int r = 0; r < RD::R0; ++r)
        if (reg_pos<N, RD::R0, P>(t, m, r) >= n) v[m * RD::R0 + r] = make_double2(0.0, 0.0);
  }
}

template <typename RD, bool ADJ>
__global__ void __launch_bounds__(RD::P) k_bpm_line(BpmLine a, const double2* __restrict__ tw) {
  constexpr int N = RD::n, V = RD::V, P = RD::P, R0 = RD::R0;
  extern __shared__ unsigned char smem_raw[];
  double2* s = reinterpret_cast<double2*>(smem_raw);
  const int t = threadIdx.x;
  const int b = blockIdx.x;
  const long fb = (long)b * a.n;
  double2 v[V];
#pragma unroll
  for (int m = 0; m < V / R0; ++m)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int pos = reg_pos<N, R0, P>(t, m, r);
      v[m * R0 + r] = pos < a.n ? cvt<double2>(a.u[fb + pos]) : make_double2(0.0, 0.0);
    }
  const double2* Hf = a.tfull + (long)a.tabidx[b] * N;
  const double2* Hh = a.thalf + (long)a.tabidx[b] * N;
  const double kl = a.kl[b];
  const double sg = a.dir;
  const bool conj = (ADJ ? -a.dir : a.dir) < 0;  // the adjoint runs every propagator reversed
  bpm_line_prop<RD>(v, s, t, tw, Hh, conj, a.n);
  for (int i = 0; i < a.ns; ++i) {
    const int k = ADJ ? a.ns - 1 - i : i;
    const float* dk = a.dn + (long)k * a.n;
    const long sb = ((long)k * a.batch + b) * a.n;
#pragma unroll
    for (int m = 0; m < V / R0; ++m)
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const int pos = reg_pos<N, R0, P>(t, m, r);
        if (pos < a.n) {
          double2 x = v[m * R0 + r];
          if (!ADJ && a.save) a.save[sb + pos] = cvt<float2>(x);
          double sn, cs;
          sincos((ADJ ? -sg : sg) * kl * (double)dk[pos], &sn, &cs);
          x = make_double2(x.x * cs - x.y * sn, x.x * sn + x.y * cs);
          if (ADJ && a.saved) {
            const float2 sv = a.saved[sb + pos];
            a.gpart[sb + pos] = (float)(sg * kl * (x.y * sv.x - x.x * sv.y));
          }
          v[m * R0 + r] = x;
        }
      }
    const bool half = ADJ ? k == 0 : k == a.ns - 1;
    bpm_line_prop<RD>(v, s, t, tw, half ? Hh : Hf, conj, a.n);
  }
#pragma unroll
  for (int m = 0; m < V / R0; ++m)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int pos = reg_pos<N, R0, P>(t, m, r);
      if (pos < a.n) a.u[fb + pos] = cvt<float2>(v[m * R0 + r]);
    }
}

// ddn[k][x] = sum over fields of gpart[k][b][x], in field order
__global__ void k_sum_line_grads(const float* __restrict__ gpart, float* __restrict__ ddn, int n, int batch, int ns) {
  const long total = (long)ns * n;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long k = i / n, x = i - k * n;
    double g = 0.0;
    for (int b = 0; b < batch; ++b) g += gpart[(k * batch + b) * n + x];
    ddn[i] = (float)g;
  }
}

}  // namespace fh
