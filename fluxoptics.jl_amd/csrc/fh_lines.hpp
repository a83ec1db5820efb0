// Stockham line-FFT passes for gfx950.
//
// A "line" is one 1-D transform of length n = R0*R1*R2 along either the contiguous
// axis (x, rows of the column-major field) or the strided axis (y, columns).  A
// workgroup owns L lines; each line is handled by P = n/V threads that hold V complex
// values in VGPRs.  Stages exchange data through LDS; the first stage reads global
// memory and the last stage's registers are stored directly (both at positions
// j + r*n/R, so consecutive lanes touch consecutive addresses).
//
// Forward plan (R0, R1, R2); the inverse runs the reversed radices so that the
// forward's last stage and the inverse's first stage share a register layout: a
// fused FFT -> pointwise multiply -> IFFT needs no LDS round trip in the middle.
#pragma once

#include "fh_dft.hpp"

namespace fh {

constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <int A, int B, int C_> struct Radices {
  static constexpr int R0 = A, R1 = B, R2 = C_;
  static constexpr int n = A * B * C_;
  static constexpr int NST = (C_ > 1) ? 3 : ((B > 1) ? 2 : 1);
  static constexpr int V = cmax(A, cmax(B, C_));
  static constexpr int P = n / V;
  static_assert(V % A == 0 && V % B == 0 && V % C_ == 0, "V must be a multiple of every radix");
  // reversed order (inverse transform)
  using Rev = Radices<(NST == 3 ? C_ : (NST == 2 ? B : A)), (NST == 3 ? B : (NST == 2 ? A : 1)),
                      (NST == 3 ? A : 1)>;
};

// LDS slot of line position pos: one pad slot every 16 complex values spreads
// power-of-two strides over the 32 banks.
__device__ __forceinline__ int lds_pos(int pos) { return pos + (pos >> 4); }
// LDS line stride.  In strided-axis passes the lanes of a wave cover L lines x (64/L)
// positions with the line index fastest.  LDS serves 16 lanes of 8-byte accesses per cycle
// over 32 four-byte banks, and a stage's positions of consecutive t are 8 or 16 slots
// apart, so lines are spaced 16/L slots past a multiple of 16: the L lines of a
// 16-lane group then fall in distinct banks.  For 16-byte elements the lines stay 8 slots
// apart.
template <int n, int L = 4, int ESZ = 8> struct LdsLine {
  static constexpr int pad = ESZ == 8 ? (L >= 16 ? 1 : 16 / L) : 8;
  static constexpr int size = ((n + n / 16 + 15) / 16) * 16 + pad;
};

// twiddle W_n^k (forward) or its conjugate (inverse) from the global table
template <bool INV, typename C>
__device__ __forceinline__ C tw_load(const C* __restrict__ tw, int k) {
  C w = tw[k];
  if (INV) w.y = -w.y;
  return w;
}

// Stockham stage: twiddle + DFT_R on the V/R butterflies held in v.
template <typename C, int n, int R, int NS, int V, int P, bool INV>
__device__ __forceinline__ void stage_compute(C* v, int t, const C* __restrict__ tw) {
#pragma unroll
  for (int m = 0; m < V / R; ++m) {
    if (NS > 1) {
      const int j = t + P * m;
      const int jm = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C w = tw_load<INV>(tw, jm * r * (n / (NS * R)));
        v[m * R + r] = cmul(v[m * R + r], w);
      }
    }
    dft<R, INV>(&v[m * R]);
  }
}

template <typename C, int n, int R, int NS, int V, int P>
__device__ __forceinline__ void stage_write(C* __restrict__ s, const C* v, int t) {
#pragma unroll
  for (int m = 0; m < V / R; ++m) {
    const int j = t + P * m;
    const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
    for (int r = 0; r < R; ++r) s[lds_pos(base + r * NS)] = v[m * R + r];
  }
}

template <typename C, int n, int R, int V, int P>
__device__ __forceinline__ void stage_read(C* v, const C* __restrict__ s, int t) {
#pragma unroll
  for (int m = 0; m < V / R; ++m) {
    const int j = t + P * m;
#pragma unroll
    for (int r = 0; r < R; ++r) v[m * R + r] = s[lds_pos(j + r * (n / R))];
  }
}

// Full line transform on registers: v enters holding positions j + r*n/R0 (first
// stage input) and leaves holding positions j + r*n/Rlast (natural-order output).
// `s` is this line's LDS slice; the whole workgroup must call this together.
template <typename C, typename RD, bool INV>
__device__ __forceinline__ void line_fft(C* v, C* s, int t, const C* __restrict__ tw) {
  constexpr int n = RD::n, V = RD::V, P = RD::P;
  stage_compute<C, n, RD::R0, 1, V, P, INV>(v, t, tw);
  if constexpr (RD::NST >= 2) {
    __syncthreads();  // earlier readers of s are done
    stage_write<C, n, RD::R0, 1, V, P>(s, v, t);
    __syncthreads();
    stage_read<C, n, RD::R1, V, P>(v, s, t);
    stage_compute<C, n, RD::R1, RD::R0, V, P, INV>(v, t, tw);
  }
  if constexpr (RD::NST >= 3) {
    __syncthreads();
    stage_write<C, n, RD::R1, RD::R0, V, P>(s, v, t);
    __syncthreads();
    stage_read<C, n, RD::R2, V, P>(v, s, t);
    stage_compute<C, n, RD::R2, RD::R0 * RD::R1, V, P, INV>(v, t, tw);
  }
}

// Frequency-side ops with a separate load (pre) and multiply (apply); specialised where the op
// is defined.
template <typename Op> struct FreqPre : std::false_type {};

// Stage twiddles of one thread, loaded up front: kernels that run one short transform per
// launch (the AS_BPM slice steps) would otherwise wait for a global load at every stage.
template <typename C, int n, int R, int NS, int V, int P, bool INV> struct StageTw {
  static constexpr bool used = NS > 1 && R > 1;
  C w[used ? V / R : 1][used ? R - 1 : 1];
  __device__ __forceinline__ void load(const C* __restrict__ tw, int t) {
    if constexpr (used) {
#pragma unroll
      for (int m = 0; m < V / R; ++m) {
        const int jm = (t + P * m) % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) w[m][r - 1] = tw_load<INV>(tw, jm * r * (n / (NS * R)));
      }
    }
  }
  __device__ __forceinline__ void compute(C* v) const {
#pragma unroll
    for (int m = 0; m < V / R; ++m) {
      if constexpr (used) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[m * R + r] = cmul(v[m * R + r], w[m][r - 1]);
      }
      dft<R, INV>(&v[m * R]);
    }
  }
};

template <typename C, typename RD, bool INV> struct LineTw {
  StageTw<C, RD::n, RD::R1, RD::R0, RD::V, RD::P, INV> s1;
  StageTw<C, RD::n, RD::R2, RD::R0 * RD::R1, RD::V, RD::P, INV> s2;
  // registers (4 bytes) that the prefetched twiddles of one thread take
  static constexpr int regs =
      (int)(sizeof(C) / 4) * ((RD::NST >= 2 ? (RD::V / RD::R1) * (RD::R1 - 1) : 0) +
                              (RD::NST >= 3 ? (RD::V / RD::R2) * (RD::R2 - 1) : 0));
  __device__ __forceinline__ void load(const C* __restrict__ tw, int t) {
    if constexpr (RD::NST >= 2) s1.load(tw, t);
    if constexpr (RD::NST >= 3) s2.load(tw, t);
  }
};

// line_fft with the stage twiddles already in registers
template <typename C, typename RD, bool INV>
__device__ __forceinline__ void line_fft_tw(C* v, C* s, int t, const LineTw<C, RD, INV>& lt) {
  constexpr int n = RD::n, V = RD::V, P = RD::P;
#pragma unroll
  for (int m = 0; m < V / RD::R0; ++m) dft<RD::R0, INV>(&v[m * RD::R0]);
  if constexpr (RD::NST >= 2) {
    __syncthreads();
    stage_write<C, n, RD::R0, 1, V, P>(s, v, t);
    __syncthreads();
    stage_read<C, n, RD::R1, V, P>(v, s, t);
    lt.s1.compute(v);
  }
  if constexpr (RD::NST >= 3) {
    __syncthreads();
    stage_write<C, n, RD::R1, RD::R0, V, P>(s, v, t);
    __syncthreads();
    stage_read<C, n, RD::R2, V, P>(v, s, t);
    lt.s2.compute(v);
  }
}

// position held in register (m, r) when the stage with radix R is the first (input) or
// last (output) stage: j + r*n/R with j = t + P*m
template <int n, int R, int P>
__device__ __forceinline__ int reg_pos(int t, int m, int r) { return t + P * m + r * (n / R); }

// ---------------------------------------------------------------- pass descriptors

template <typename C> struct LineIO {
  const C* in;
  C* out;
  long in_f, out_f;   // element stride between fields
  long in_l, out_l;   // element stride between lines
  long in_e, out_e;   // element stride between points of a line
  int nlines;         // lines per field
  int in_valid;       // points [0, in_valid) are read, the rest are zero
  int out_count;      // points [0, out_count) are stored
  int in_blk, out_blk;  // 1: 4x4-blocked spectrum layout (see elem_addr)
};

// Element (x, y) of a field with rows of `rowlen` elements.  Natural layout: y*rowlen + x.
// Blocked layout (the intermediate spectrum between passes): 4x4 blocks of 128 bytes,
// block (y/4, x/4) at ((y/4)*(rowlen/4) + x/4)*16, element (y%4)*4 + x%4 inside.  Rows
// 4g..4g+3 occupy the same address range in both layouts, so an x pass that owns whole
// 4-row groups can convert in place, and a y pass over 4 columns reads whole lines.
template <bool LFAST>
__device__ __forceinline__ long elem_addr(int li, int pos, long lstride, long estride, int blk) {
  if (!blk) return (long)li * lstride + (long)pos * estride;
  const int x = LFAST ? li : pos, y = LFAST ? pos : li;
  const long rowlen = LFAST ? estride : lstride;
  return ((((long)(y >> 2)) * (rowlen >> 2) + (x >> 2)) << 4) + ((y & 3) << 2) + (x & 3);
}

// thread -> (line in workgroup, thread within line); LFAST puts lines on adjacent
// lanes (strided-axis passes, so a wave reads L adjacent columns of one row).
template <int L, int P, bool LFAST>
__device__ __forceinline__ void thread_map(int tid, int& l, int& t) {
  if (LFAST) { l = tid % L; t = tid / L; } else { t = tid % P; l = tid / P; }
}

struct NoOp {
  template <typename C> __device__ __forceinline__ C operator()(C v, int, int, int) const { return v; }
};

// storage <-> compute conversions (c64 storage may be transformed in FP64)
template <typename C, typename S> __device__ __forceinline__ C cvt(S v) {
  C r; r.x = v.x; r.y = v.y; return r;
}

// Tile order for strided-axis passes.  Workgroups are dispatched round-robin over the
// 8 XCDs; G neighbouring column tiles share each 128-byte row segment (G = 128 B / tile
// width), so they are placed 8 blocks apart to run on the same XCD at about the same
// time and share its L2 instead of each fetching the line from HBM.
__device__ __forceinline__ int xcd_group_tile(int b, int nb, int G) {
  const int span = 8 * G;
  if (G <= 1 || nb % span) return b;
  const int i = b % span;
  return (b - i) + (i & 7) * G + (i >> 3);
}

// Forward or inverse line FFT, global -> global, with an elementwise op fused on the
// time-domain side (prologue on load for a forward transform, epilogue before store
// for an inverse one).  S is the storage type, C the compute type.
template <typename S, typename C, typename RDF, int L, bool LFAST, bool INV, typename TimeOp>
__global__ void __launch_bounds__(L * RDF::P)
k_line_fft(LineIO<S> io, const C* __restrict__ tw, TimeOp op, int G) {
  using RD = typename std::conditional<INV, typename RDF::Rev, RDF>::type;
  constexpr int n = RD::n, V = RD::V, P = RD::P;
  extern __shared__ unsigned char smem_raw[];
  C* smem = reinterpret_cast<C*>(smem_raw);
  int l, t;
  thread_map<L, P, LFAST>(threadIdx.x, l, t);
  const int b = blockIdx.y;
  const int tile = LFAST ? xcd_group_tile(blockIdx.x, gridDim.x, G) : (int)blockIdx.x;
  const int li = tile * L + l;
  const bool live = li < io.nlines;
  C* s = smem + l * LdsLine<n, L, sizeof(C)>::size;
  const S* src = io.in + b * io.in_f;
  C v[V];
#pragma unroll
  for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
    for (int r = 0; r < RD::R0; ++r) {
      const int pos = reg_pos<n, RD::R0, P>(t, m, r);
      C x = cmake<C>(0, 0);
      if (live && pos < io.in_valid) x = cvt<C>(src[elem_addr<LFAST>(li, pos, io.in_l, io.in_e, io.in_blk)]);
      if (!INV && live) x = op(x, b, li, pos);
      v[m * RD::R0 + r] = x;
    }
  constexpr int RL = (RD::NST == 3) ? RD::R2 : ((RD::NST == 2) ? RD::R1 : RD::R0);
  line_fft<C, RD, INV>(v, s, t, tw);
  S* dst = io.out + b * io.out_f;
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int pos = reg_pos<n, RL, P>(t, m, r);
      C x = v[m * RL + r];
      if (INV && live) x = op(x, b, li, pos);
      if (live && pos < io.out_count) dst[elem_addr<LFAST>(li, pos, io.out_l, io.out_e, io.out_blk)] = cvt<S>(x);
    }
}

// FFT -> pointwise multiply (freq-domain op, given line index and frequency index)
// -> IFFT along the lines, global -> global.  Used for the strided (y) axis.
template <typename S, typename C, typename RDF, int L, bool LFAST, typename FreqOp>
__global__ void __launch_bounds__(L * RDF::P)
k_line_conv(LineIO<S> io, const C* __restrict__ tw, FreqOp op, int G) {
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  using RI = typename RDF::Rev;
  extern __shared__ unsigned char smem_raw[];
  C* smem = reinterpret_cast<C*>(smem_raw);
  int l, t;
  thread_map<L, P, LFAST>(threadIdx.x, l, t);
  const int b = blockIdx.y;
  const int tile = LFAST ? xcd_group_tile(blockIdx.x, gridDim.x, G) : (int)blockIdx.x;
  const int li = tile * L + l;
  const bool live = li < io.nlines;
  C* s = smem + l * LdsLine<n, L, sizeof(C)>::size;
  const S* src = io.in + b * io.in_f;
  C v[V];
#pragma unroll
  for (int m = 0; m < V / RDF::R0; ++m)
#pragma unroll
    for (int r = 0; r < RDF::R0; ++r) {
      const int pos = reg_pos<n, RDF::R0, P>(t, m, r);
      C x = cmake<C>(0, 0);
      if (live && pos < io.in_valid) x = cvt<C>(src[elem_addr<LFAST>(li, pos, io.in_l, io.in_e, io.in_blk)]);
      v[m * RDF::R0 + r] = x;
    }
  constexpr int RL = RI::R0;  // last forward radix == first inverse radix
  // short transforms (at most 256 threads a workgroup, e.g. one 512-point FP64 line per wave)
  // load their stage twiddles and transfer-function values up front, with the line: the launch
  // then waits for one round of global loads instead of one per stage
  constexpr bool PRE = L * P <= 256 && LineTw<C, RDF, false>::regs + LineTw<C, RI, true>::regs <= 128;
  constexpr bool HPRE = PRE && FreqPre<FreqOp>::value;
  LineTw<C, RDF, false> twf;
  LineTw<C, RI, true> twi;
  C hv[HPRE ? V : 1];
  if constexpr (PRE) {
    twf.load(tw, t);
    twi.load(tw, t);
  }
  if constexpr (HPRE) {
#pragma unroll
    for (int m = 0; m < V / RL; ++m)
#pragma unroll
      for (int r = 0; r < RL; ++r)
        hv[m * RL + r] = live ? op.template pre<C>(b, li, reg_pos<n, RL, P>(t, m, r)) : cmake<C>(1, 0);
  }
  if constexpr (PRE) line_fft_tw<C, RDF, false>(v, s, t, twf);
  else line_fft<C, RDF, false>(v, s, t, tw);
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int pos = reg_pos<n, RL, P>(t, m, r);
      if constexpr (HPRE) {
        v[m * RL + r] = op.apply(v[m * RL + r], hv[m * RL + r]);
      } else {
        if (live) v[m * RL + r] = op(v[m * RL + r], b, li, pos);
      }
    }
  if constexpr (PRE) line_fft_tw<C, RI, true>(v, s, t, twi);
  else line_fft<C, RI, true>(v, s, t, tw);
  constexpr int RE = (RI::NST == 3) ? RI::R2 : ((RI::NST == 2) ? RI::R1 : RI::R0);
  S* dst = io.out + b * io.out_f;
#pragma unroll
  for (int m = 0; m < V / RE; ++m)
#pragma unroll
    for (int r = 0; r < RE; ++r) {
      const int pos = reg_pos<n, RE, P>(t, m, r);
      if (live && pos < io.out_count)
        dst[elem_addr<LFAST>(li, pos, io.out_l, io.out_e, io.out_blk)] = cvt<S>(v[m * RE + r]);
    }
}

}  // namespace fh
