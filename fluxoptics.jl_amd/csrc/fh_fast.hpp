// Fast c64 line passes: the same Stockham transforms as fh_lines.hpp with the address
// arithmetic moved out of the vector units.
//
//  - global memory goes through buffer instructions with one descriptor per field; a
//    register holds the thread's base offset and every other offset is a compile-time or
//    workgroup-uniform constant (SGPR soffset), so a load costs no VALU address math.
//    Offsets past the field's end read 0 and drop stores (the descriptor's bound).
//  - LDS positions are linear in the register index, so each exchange computes one base
//    per thread and uses immediate offsets;
//  - stage twiddles come from per-plan tables laid out [j mod NS][r-1], one base per
//    butterfly plus immediate offsets;
//  - lines shorter than the transform (zero-padded RS input, cropped RS output) are
//    exactly half of it, so the valid/zero decision is a compile-time predicate.
#pragma once

#include "fh_lines.hpp"
#include "fh_ops.hpp"

namespace fh {

constexpr int kBufWord3 = 0x00020000;  // raw buffer, gfx9-family data format

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, long bytes) {
  const long lim = bytes < 0x7fffffffL ? bytes : 0x7fffffffL;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)lim, kBufWord3);
}

// s_waitcnt immediate for vmcnt(0) with expcnt and lgkmcnt left at their maxima (gfx9
// encoding: vmcnt[3:0] in bits 3:0, expcnt in 6:4, lgkmcnt in 11:8, vmcnt[5:4] in 15:14)
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ __forceinline__ float2 buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, int voff, int soff, float2 v) {
  using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, voff, soff, 0);
}

__device__ __forceinline__ float4 buf_load4(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ __forceinline__ void buf_store4(__amdgpu_buffer_rsrc_t r, int voff, int soff, float4 v) {
  using V4 = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, v), r, voff, soff, 0);
}

// Pass geometry of the fast kernels (offsets in elements; fields are < 2 GiB).
struct FastIO {
  const float2* in;
  float2* out;
  long in_f, out_f;      // elements between fields
  int in_l, out_l;       // elements between lines (x passes) or 1 (y passes)
  int in_e, out_e;       // 1 (x passes) or elements between rows (y passes)
  int in_blk, out_blk;   // 4x4-blocked layout (elem_addr)
  int nlines;
  long in_bytes, out_bytes;  // bytes of one field (descriptor bound)
  // zero-padded lines (IN_HALF / OUT_HALF): points read / stored per line, at most n/2 (the
  // padded length of an RS plan is a power of two >= 2 n_valid - 1)
  int in_valid = 1 << 30, out_valid = 1 << 30;
};

// Offset of (line li, position pos) and the per-position stride K such that
// offset(li, t + C) = offset(li, t) + C*K whenever C is a multiple of 4.
template <bool LFAST>
__device__ __forceinline__ int fast_off(int li, int pos, int lstride, int estride, int blk) {
  return (int)elem_addr<LFAST>(li, pos, lstride, estride, blk);
}
template <bool LFAST>
__device__ __forceinline__ int fast_k(int lstride, int estride, int blk) {
  return LFAST ? estride : (blk ? 4 : 1);
}

// ---------------------------------------------------------------- LDS with linear offsets

// Stage writes go to positions base_m + r*NS with base_m = base + m*P*R.  With R and NS
// dividing 16 or being multiples of 16 and P*R a multiple of 16, (base mod 16) +
// (offset mod 16) < 16 for every thread, so lds_pos(base + offset) =
// lds_pos(base) + offset + offset/16 and every slot is an immediate offset.
template <typename C, int n, int R, int NS, int V, int P>
__device__ __forceinline__ void fast_write(C* __restrict__ s, const C* v, int t) {
  static_assert(P % NS == 0, "write stage needs P a multiple of NS");
  static_assert(NS % 16 == 0 || 16 % NS == 0, "NS must divide or be a multiple of 16");
  static_assert(R % 16 == 0 || 16 % R == 0, "R must divide or be a multiple of 16");
  static_assert((P * R) % 16 == 0, "P*R must be a multiple of 16");
  const int base = (t / NS) * NS * R + (t % NS);
  const int lb = lds_pos(base);
#pragma unroll
  for (int m = 0; m < V / R; ++m)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      // base_m = base + m*P*R (P*R a multiple of 16), element offset r*NS
      const int c = m * P * R + r * NS;
      s[lb + c + (c >> 4)] = v[m * R + r];
    }
}

template <typename C, int n, int R, int V, int P>
__device__ __forceinline__ void fast_read(C* v, const C* __restrict__ s, int t) {
  if constexpr (P % 16 == 0 && (n / R) % 16 == 0) {
    const int lb = lds_pos(t);
#pragma unroll
    for (int m = 0; m < V / R; ++m)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int c = P * m + r * (n / R);
        v[m * R + r] = s[lb + c + (c >> 4)];
      }
  } else {
#pragma unroll
    for (int m = 0; m < V / R; ++m)
#pragma unroll
      for (int r = 0; r < R; ++r) v[m * R + r] = s[lds_pos(t + P * m + r * (n / R))];
  }
}

// twiddles of a stage: table [jm][r-1], jm = j mod NS, j = t + P*m
template <typename C, int n, int R, int NS, int V, int P>
__device__ __forceinline__ void fast_twiddle(C* v, int t, const C* __restrict__ tab) {
  if constexpr (NS > 1) {
#pragma unroll
    for (int m = 0; m < V / R; ++m) {
      const int jm = (t + P * m) & (NS - 1);
      const C* w = tab + jm * (R - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) v[m * R + r] = cmul(v[m * R + r], w[r - 1]);
    }
  }
}

template <typename C, int R, int V>
__device__ __forceinline__ void fast_dfts(C* v) {
#pragma unroll
  for (int m = 0; m < V / R; ++m) dft<R, false>(&v[m * R]);
}
template <typename C, int R, int V>
__device__ __forceinline__ void fast_idfts(C* v) {
#pragma unroll
  for (int m = 0; m < V / R; ++m) dft<R, true>(&v[m * R]);
}

// Stage tables of a radix plan, concatenated: stage 1 (NS = R0, R = R1) then stage 2
// (NS = R0*R1, R = R2).  Inverse plans get their own (conjugated) tables.
template <typename RD> struct FastTw {
  static constexpr int off1 = 0;
  static constexpr int off2 = RD::R0 * (RD::R1 - 1);
  static constexpr int size = off2 + (RD::NST == 3 ? RD::R0 * RD::R1 * (RD::R2 - 1) : 0);
};

// Full transform in registers (forward or inverse according to INV; the tables
// already carry the sign).
template <typename C, typename RD, bool INV>
__device__ __forceinline__ void fast_fft(C* v, C* s, int t, const C* __restrict__ tab) {
  constexpr int n = RD::n, V = RD::V, P = RD::P;
  if (INV) fast_idfts<C, RD::R0, V>(v); else fast_dfts<C, RD::R0, V>(v);
  if constexpr (RD::NST >= 2) {
    __syncthreads();
    fast_write<C, n, RD::R0, 1, V, P>(s, v, t);
    __syncthreads();
    fast_read<C, n, RD::R1, V, P>(v, s, t);
    fast_twiddle<C, n, RD::R1, RD::R0, V, P>(v, t, tab + FastTw<RD>::off1);
    if (INV) fast_idfts<C, RD::R1, V>(v); else fast_dfts<C, RD::R1, V>(v);
  }
  if constexpr (RD::NST >= 3) {
    __syncthreads();
    fast_write<C, n, RD::R1, RD::R0, V, P>(s, v, t);
    __syncthreads();
    fast_read<C, n, RD::R2, V, P>(v, s, t);
    fast_twiddle<C, n, RD::R2, RD::R0 * RD::R1, V, P>(v, t, tab + FastTw<RD>::off2);
    if (INV) fast_idfts<C, RD::R2, V>(v); else fast_dfts<C, RD::R2, V>(v);
  }
}

template <typename RD> struct LastRadix {
  static constexpr int value = (RD::NST == 3) ? RD::R2 : ((RD::NST == 2) ? RD::R1 : RD::R0);
};

// ---------------------------------------------------------------- kernels

// Line FFT (forward or inverse), c64 global -> global, with a time-side op (prologue of a
// forward transform, epilogue of an inverse one).  IN_HALF: only the first io.in_valid <= n/2
// points of each line are read (the rest are zero); OUT_HALF: only the first io.out_valid <=
// n/2 are stored.  The n/2 bound is a compile-time predicate; the count a runtime one.
// LANES: the workgroup's L lines sit on adjacent lanes (a wave instruction then covers whole
// 128-byte lines of the 4x4-blocked layout, and 4 rows x 128 bytes of the natural one);
// otherwise a wave covers 64 consecutive positions of one line.
template <typename RDF, int L, bool LFAST, bool INV, bool IN_HALF, bool OUT_HALF, typename TimeOp,
          bool LANES = LFAST>
__global__ void __launch_bounds__(L * RDF::P)
k_fast_fft(FastIO io, const float2* __restrict__ tab, TimeOp op) {
  using RD = typename std::conditional<INV, typename RDF::Rev, RDF>::type;
  constexpr int n = RD::n, V = RD::V, P = RD::P;
  constexpr int RL = LastRadix<RD>::value;
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  int l, t;
  thread_map<L, P, LANES>(threadIdx.x, l, t);
  const int b = blockIdx.y;
  const int li = blockIdx.x * L + l;
  const bool live = li < io.nlines;
  float2* s = smem + l * LdsLine<n, L>::size;
  const auto rin = make_rsrc(io.in + b * io.in_f, io.in_bytes);
  const int vin = fast_off<LFAST>(li, t, io.in_l, io.in_e, io.in_blk) * 8;
  const int kin = fast_k<LFAST>(io.in_l, io.in_e, io.in_blk) * 8;
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
    for (int r = 0; r < RD::R0; ++r) {
      const int c = P * m + r * (n / RD::R0);
      float2 x = make_float2(0.f, 0.f);
      if (!IN_HALF || (c < n / 2 && t + c < io.in_valid)) x = buf_load(rin, vin, c * kin);
      v[m * RD::R0 + r] = x;
    }
  if (!INV && live) {
    if constexpr (TwoPhaseOp<TimeOp>::value) {
      // the phase factors and the saved input have the field's natural layout, so their loads
      // and stores take the field's offsets (one base register, immediate offsets); the factor
      // loads are all issued before the saved stores, which may alias them
      const auto rf = make_rsrc(op.cis, io.in_bytes);
      float2 pre[V];
#pragma unroll
      for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
        for (int r = 0; r < RD::R0; ++r) pre[m * RD::R0 + r] = buf_load(rf, vin, (P * m + r * (n / RD::R0)) * kin);
      if (op.saved) {
        const auto rsv = make_rsrc(op.saved + b * op.fstride, io.in_bytes);
#pragma unroll
        for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
          for (int r = 0; r < RD::R0; ++r) {
            const int c = P * m + r * (n / RD::R0);
            if (t + c < op.nx) buf_store(rsv, vin, c * kin, v[m * RD::R0 + r]);
          }
      }
#pragma unroll
      for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
        for (int r = 0; r < RD::R0; ++r) {
          const int c = P * m + r * (n / RD::R0);
          if (t + c < op.nx) v[m * RD::R0 + r] = cmul(v[m * RD::R0 + r], pre[m * RD::R0 + r]);
        }
    } else {
#pragma unroll
      for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
        for (int r = 0; r < RD::R0; ++r)
          v[m * RD::R0 + r] = op(v[m * RD::R0 + r], b, li, t + P * m + r * (n / RD::R0));
    }
  }
  fast_fft<float2, RD, INV>(v, s, t, tab);
  const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
  const int vout = fast_off<LFAST>(li, t, io.out_l, io.out_e, io.out_blk) * 8;
  const int kout = fast_k<LFAST>(io.out_l, io.out_e, io.out_blk) * 8;
  if (live) {
#pragma unroll
    for (int m = 0; m < V / RL; ++m)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int c = P * m + r * (n / RL);
        float2 x = v[m * RL + r];
        if (INV) x = op(x, b, li, t + c);
        if (!OUT_HALF || (c < n / 2 && t + c < io.out_valid)) buf_store(rout, vout, c * kout, x);
      }
  }
}

// Line FFT of whole 4-row groups whose blocked side (the input of an inverse pass, the output of
// a forward one) moves through LDS: the workgroup's rows are one contiguous run of L*n elements in
// the 4x4-blocked layout, read or written with 16-byte accesses, a wave covering 1 KB of whole
// 128-byte lines per instruction (the direct form covers 32 bytes of each of 16 lines, the rest of
// each line coming from the three other rows' waves).  Chunk k*L*P + tid of the run is elements
// 2c, 2c+1: block (2c)/16 of the run, row (2c/4) % 4 of that block, columns 2c % 4 and + 1.  The
// two 8-byte LDS accesses of a chunk stay separate instructions: merged into one ds_read2_b64 of
// adjacent slots, the first value came back stale for a few lanes per launch (measured on MI355X,
// forward pass, 0.15% of the chunks, varying from run to run), with or without an extra barrier.
// The natural side takes the same time-side op as k_fast_fft (prologue of a forward transform,
// epilogue of an inverse one).  HALF: the natural rows hold in_valid (forward) or out_valid
// (inverse) points of each n-point line, the rest zero padding (RS plans), as IN_HALF / OUT_HALF
// of k_fast_fft.
template <typename RDF, int L, bool INV, typename TimeOp = NoOp, bool HALF = false>
__global__ void __launch_bounds__(L * RDF::P)
k_fast_fft_stg(FastIO io, const float2* __restrict__ tab, TimeOp op) {
  using RD = typename std::conditional<INV, typename RDF::Rev, RDF>::type;
  constexpr int n = RD::n, V = RD::V, P = RD::P;
  constexpr int RL = LastRadix<RD>::value;
  constexpr int LS = LdsLine<n, L>::size;
  constexpr int NCH = n / (2 * P);  // 16-byte chunks a thread
  static_assert(L % 4 == 0, "whole 4-row groups");
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  const int tid = threadIdx.x;
  const int l = tid / P, t = tid % P;
  // workgroup -> (4-row group, field); with the Phase prologue the fields of one row group run on
  // one XCD one after the other (8 workgroups apart), so its factor rows come from that XCD's L2
  // after the first field instead of once per field from HBM
  int rg = blockIdx.x, b = blockIdx.y;
  if constexpr (TwoPhaseOp<TimeOp>::value) {
    if (gridDim.x % 8 == 0) {
      const int lin = blockIdx.y * gridDim.x + blockIdx.x, s8 = lin >> 3;
      b = s8 % (int)gridDim.y;
      rg = ((s8 / (int)gridDim.y) << 3) + (lin & 7);
    }
  }
  float2* s = smem + l * LS;
  // the chunk's row and column within the workgroup's run
  int crow[NCH], ccol[NCH];
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
    const int e = 2 * (k * L * P + tid);
    const int blk = e >> 4;
    crow[k] = 4 * (blk / (n / 4)) + ((e >> 2) & 3);
    ccol[k] = 4 * (blk % (n / 4)) + (e & 3);
  }
  float2 v[V];
  if (INV) {
    const auto rin = make_rsrc(io.in + b * io.in_f, io.in_bytes);
    const int vin = (rg * L * n + 2 * tid) * 8;
    float4 q[NCH];
#pragma unroll
    for (int k = 0; k < NCH; ++k) q[k] = buf_load4(rin, vin, k * L * P * 16);
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      float2* d = smem + crow[k] * LS + lds_pos(ccol[k]);
      d[0] = make_float2(q[k].x, q[k].y);
      asm volatile("" ::: "memory");
      d[1] = make_float2(q[k].z, q[k].w);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
      for (int r = 0; r < RD::R0; ++r) v[m * RD::R0 + r] = s[lds_pos(t + P * m + r * (n / RD::R0))];
  } else {
    const auto rin = make_rsrc(io.in + b * io.in_f, io.in_bytes);
    const int li = rg * L + l;
    const int vin = (li * io.in_l + t) * 8;
#pragma unroll
    for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
      for (int r = 0; r < RD::R0; ++r) v[m * RD::R0 + r] = buf_load(rin, vin, (P * m + r * (n / RD::R0)) * 8);
    if constexpr (TwoPhaseOp<TimeOp>::value) {
      // as in k_fast_fft: the factor loads first, then the saved input, then the multiply
      const auto rf = make_rsrc(op.cis, io.in_bytes);
      float2 pre[V];
#pragma unroll
      for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
        for (int r = 0; r < RD::R0; ++r) pre[m * RD::R0 + r] = buf_load(rf, vin, (P * m + r * (n / RD::R0)) * 8);
      if (op.saved) {
        const auto rsv = make_rsrc(op.saved + b * op.fstride, io.in_bytes);
#pragma unroll
        for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
          for (int r = 0; r < RD::R0; ++r) {
            const int c = P * m + r * (n / RD::R0);
            if (t + c < op.nx) buf_store(rsv, vin, c * 8, v[m * RD::R0 + r]);
          }
      }
#pragma unroll
      for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
        for (int r = 0; r < RD::R0; ++r) {
          const int c = P * m + r * (n / RD::R0);
          if (t + c < op.nx) v[m * RD::R0 + r] = cmul(v[m * RD::R0 + r], pre[m * RD::R0 + r]);
        }
    } else {
#pragma unroll
      for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
        for (int r = 0; r < RD::R0; ++r)
          v[m * RD::R0 + r] = op(v[m * RD::R0 + r], b, li, t + P * m + r * (n / RD::R0));
    }
    if constexpr (HALF) {
      // the loads past the valid points (the next row, or past the field: zero from the descriptor)
      // are replaced by the zero padding
#pragma unroll
      for (int m = 0; m < V / RD::R0; ++m)
#pragma unroll
        for (int r = 0; r < RD::R0; ++r)
          if (t + P * m + r * (n / RD::R0) >= io.in_valid) v[m * RD::R0 + r] = make_float2(0.f, 0.f);
    }
  }
  fast_fft<float2, RD, INV>(v, s, t, tab);
  if (INV) {
    const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
    const int li = rg * L + l;
    const int vout = (li * io.out_l + t) * 8;
#pragma unroll
    for (int m = 0; m < V / RL; ++m)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int c = P * m + r * (n / RL);
        if (!HALF || t + c < io.out_valid) buf_store(rout, vout, c * 8, op(v[m * RL + r], b, li, t + c));
      }
  } else {
    // the outputs go to the slots this thread read in the last exchange (its radix is RL), which
    // no other thread reads: no barrier before the writes
#pragma unroll
    for (int m = 0; m < V / RL; ++m)
#pragma unroll
      for (int r = 0; r < RL; ++r) s[lds_pos(t + P * m + r * (n / RL))] = v[m * RL + r];
    __syncthreads();
    const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
    const int vout = (rg * L * n + 2 * tid) * 8;
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const float2* d = smem + crow[k] * LS + lds_pos(ccol[k]);
      const float2 a0 = d[0];
      asm volatile("" ::: "memory");
      const float2 a1 = d[1];
      buf_store4(rout, vout, k * L * P * 16, make_float4(a0.x, a0.y, a1.x, a1.y));
    }
  }
}

// FFT -> pointwise multiply -> IFFT along the lines, c64 global -> global.
template <typename RDF, int L, bool LFAST, bool IN_HALF, bool OUT_HALF, typename FreqOp>
__global__ void __launch_bounds__(L * RDF::P)
k_fast_conv(FastIO io, const float2* __restrict__ tabf, const float2* __restrict__ tabi, FreqOp op, int G) {
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  using RI = typename RDF::Rev;
  constexpr int RL = RI::R0;
  constexpr int RE = LastRadix<RI>::value;
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  int l, t;
  thread_map<L, P, LFAST>(threadIdx.x, l, t);
  const int b = blockIdx.y;
  const int tile = LFAST ? xcd_group_tile(blockIdx.x, gridDim.x, G) : (int)blockIdx.x;
  const int li = tile * L + l;
  const bool live = li < io.nlines;
  float2* s = smem + l * LdsLine<n, L>::size;
  const auto rin = make_rsrc(io.in + b * io.in_f, io.in_bytes);
  const int vin = fast_off<LFAST>(li, t, io.in_l, io.in_e, io.in_blk) * 8;
  const int kin = fast_k<LFAST>(io.in_l, io.in_e, io.in_blk) * 8;
  float2 v[V];
#pragma unroll
  for (int m = 0; m < V / RDF::R0; ++m)
#pragma unroll
    for (int r = 0; r < RDF::R0; ++r) {
      const int c = P * m + r * (n / RDF::R0);
      float2 x = make_float2(0.f, 0.f);
      if (!IN_HALF || (c < n / 2 && t + c < io.in_valid)) x = buf_load(rin, vin, c * kin);
      v[m * RDF::R0 + r] = x;
    }
  fast_fft<float2, RDF, false>(v, s, t, tabf);
  if (live) {
#pragma unroll
    for (int m = 0; m < V / RL; ++m)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int c = P * m + r * (n / RL);
        v[m * RL + r] = op(v[m * RL + r], b, li, t + c);
      }
  }
  fast_fft<float2, RI, true>(v, s, t, tabi);
  const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
  const int vout = fast_off<LFAST>(li, t, io.out_l, io.out_e, io.out_blk) * 8;
  const int kout = fast_k<LFAST>(io.out_l, io.out_e, io.out_blk) * 8;
  if (live) {
#pragma unroll
    for (int m = 0; m < V / RE; ++m)
#pragma unroll
      for (int r = 0; r < RE; ++r) {
        const int c = P * m + r * (n / RE);
        if (!OUT_HALF || (c < n / 2 && t + c < io.out_valid)) buf_store(rout, vout, c * kout, v[m * RE + r]);
      }
  }
}

}  // namespace fh

namespace fh {

// Quadrant transfer-function tables for the batched y pass (column-major, see QuadMulT).
struct QuadTab {
  const float2* tab;
  const int* tabidx;
  long tabsize;  // entries per table
  int nx;        // transform length along x (columns of the spectrum)
  int qcol;      // entries per table column (n/2 + 1)
  int conj;
  int one_tab;   // every field uses table tabidx[0]
  int full;      // tables hold every column (no x mirror) and both column halves
  int moff;      // full tables: offset of the mirrored column half (n/2), else 0
};

// One field of the batched y pass: FFT, transfer-function multiply, IFFT, store.
template <typename RDF, int L, bool OUT_HALF>
__device__ __forceinline__ void yconv_field(float2* v, float2* s, int t, int li, bool live, int b,
                                            const FastIO& io, const float2* __restrict__ tabf,
                                            const float2* __restrict__ tabi, const QuadTab& qt,
                                            int vout, int kout, int qoff_d, int qoff_m) {
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  using RI = typename RDF::Rev;
  constexpr int RL = RI::R0;
  constexpr int RE = LastRadix<RI>::value;
  fast_fft<float2, RDF, false>(v, s, t, tabf);
  // H(kx, ky): ky = t + c; c < n/2 reads column entry t + c, c >= n/2 entry moff + n - t - c
  // (the mirror for even tables, the second half for full ones), written as
  // moff + (P - t) + (n - c - P) so both offsets stay non-negative.
  const auto rt = make_rsrc(qt.tab + (long)qt.tabidx[b] * qt.tabsize, qt.tabsize * 8);
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int c = P * m + r * (n / RL);
      const float2 h = (c < n / 2) ? buf_load(rt, qoff_d, c * 8) : buf_load(rt, qoff_m, (n - c - P) * 8);
      v[m * RL + r] = qt.conj ? cmulc(v[m * RL + r], h) : cmul(v[m * RL + r], h);
    }
  fast_fft<float2, RI, true>(v, s, t, tabi);
  if (live) {
    const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
#pragma unroll
    for (int m = 0; m < V / RE; ++m)
#pragma unroll
      for (int r = 0; r < RE; ++r) {
        const int c = P * m + r * (n / RE);
        if (!OUT_HALF || (c < n / 2 && t + c < io.out_valid)) buf_store(rout, vout, c * kout, v[m * RE + r]);
      }
  }
}

template <int V, int R0, int n, int P, bool IN_HALF>
__device__ __forceinline__ void yconv_load(float2* v, const FastIO& io, int b, int vin, int kin, int t) {
  const auto rin = make_rsrc(io.in + b * io.in_f, io.in_bytes);
#pragma unroll
  for (int m = 0; m < V / R0; ++m)
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int c = P * m + r * (n / R0);
      v[m * R0 + r] = (!IN_HALF || (c < n / 2 && t + c < io.in_valid)) ? buf_load(rin, vin, c * kin) : make_float2(0.f, 0.f);
    }
}

// Batched y pass: a workgroup owns L spectrum columns and walks `fpw` fields of the batch
// (load, FFT, multiply, IFFT, store); the grid's y dimension covers the batch in steps of fpw.  The transfer-function columns it needs are read
// from HBM for the first field and from L2 for the others; two or more resident
// workgroups per CU overlap one's memory phase with another's arithmetic.
// LANES_T: adjacent lanes take adjacent positions of one column (conflict-free LDS
// exchanges; the four columns of a 128-byte block row are read by sibling waves and meet
// in L2) instead of adjacent columns of one row.
template <typename RDF, int L, bool IN_HALF, bool OUT_HALF, bool LANES_T>
__global__ void __launch_bounds__(L * RDF::P)
k_fast_yconv(FastIO io, int batch, int fpw, const float2* __restrict__ tabf,
             const float2* __restrict__ tabi, QuadTab qt) {
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  int l, t;
  thread_map<L, P, !LANES_T>(threadIdx.x, l, t);
  // fewer than 4 columns per workgroup: the tiles sharing a 128-byte line run on one XCD
  const int li = (L < 4 ? xcd_group_tile(blockIdx.x, gridDim.x, 4 / L) : (int)blockIdx.x) * L + l;
  const bool live = li < io.nlines;
  float2* s = smem + l * LdsLine<n, L>::size;
  const int vin = fast_off<true>(li, t, io.in_l, io.in_e, io.in_blk) * 8;
  const int kin = fast_k<true>(io.in_l, io.in_e, io.in_blk) * 8;
  const int vout = fast_off<true>(li, t, io.out_l, io.out_e, io.out_blk) * 8;
  const int kout = fast_k<true>(io.out_l, io.out_e, io.out_blk) * 8;
  const int lic = live ? li : 0;
  const int qxi = qt.full ? lic : (lic <= qt.nx - lic ? lic : qt.nx - lic);
  const int qoff_d = (qxi * qt.qcol + t) * 8;
  const int qoff_m = (qxi * qt.qcol + qt.moff + P - t) * 8;
  const int b0 = blockIdx.y * fpw;
  const int b1 = b0 + fpw < batch ? b0 + fpw : batch;
  for (int b = b0; b < b1; ++b) {
    // compiler-only barrier: keeps the per-field twiddle loads inside the loop (hoisting
    // them would hold every twiddle in registers and halve the resident workgroups)
    asm volatile("" ::: "memory");
    float2 v[V];
    yconv_load<V, RDF::R0, n, P, IN_HALF>(v, io, b, vin, kin, t);
    yconv_field<RDF, L, OUT_HALF>(v, s, t, li, live, b, io, tabf, tabi, qt, vout, kout, qoff_d, qoff_m);
  }
}

}  // namespace fh

namespace fh {

// ---------------------------------------------------------------- software-pipelined y pass
// Stage twiddles from base powers: for a butterfly with j mod NS = jm the twiddles are
// w^r, r < R, with w = W_n^{jm*n/(NS*R)}; the table entries r = 1, 2, 4, 8, 16 are loaded once
// per workgroup (they depend on the thread only) and w^r is built with one complex product per
// r from w^(r - lowbit(r)) * w^lowbit(r) (at most popcount(r)-1 products: <= 3 roundings).
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }

template <typename C, int R, int NS, int V, int P>
struct TwBases {
  static constexpr int NB = ilog2c(R);   // bases per butterfly
  static constexpr int NBF = V / R;      // butterflies per thread
  C b[NBF][NB > 0 ? NB : 1];
  __device__ __forceinline__ void load(const C* __restrict__ tab, int t) {
    if constexpr (NS > 1) {
#pragma unroll
      for (int m = 0; m < NBF; ++m) {
        const int jm = (t + P * m) & (NS - 1);
#pragma unroll
        for (int k = 0; k < NB; ++k) b[m][k] = tab[jm * (R - 1) + ((1 << k) - 1)];
      }
    }
  }
  __device__ __forceinline__ void apply(C* v) const {
    if constexpr (NS > 1) {
#pragma unroll
      for (int m = 0; m < NBF; ++m) {
        C w[R];
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int low = r & (-r);
          const int k = ilog2c(low);
          w[r] = (r == low) ? b[m][k] : cmul(w[r - low], b[m][k]);
          v[m * R + r] = cmul(v[m * R + r], w[r]);
        }
      }
    }
  }
};

template <typename RD> struct PlanBases {
  using C = float2;
  static constexpr int n = RD::n, V = RD::V, P = RD::P;
  TwBases<C, RD::R1, RD::R0, V, P> s1;
  TwBases<C, RD::R2, RD::R0 * RD::R1, V, P> s2;
  __device__ __forceinline__ void load(const C* __restrict__ tab, int t) {
    if constexpr (RD::NST >= 2) s1.load(tab + FastTw<RD>::off1, t);
    if constexpr (RD::NST >= 3) s2.load(tab + FastTw<RD>::off2, t);
  }
};

template <typename RD, bool INV>
__device__ __forceinline__ void fast_fft_b(float2* v, float2* s, int t, const PlanBases<RD>& pb) {
  using C = float2;
  constexpr int n = RD::n, V = RD::V, P = RD::P;
  if (INV) fast_idfts<C, RD::R0, V>(v); else fast_dfts<C, RD::R0, V>(v);
  if constexpr (RD::NST >= 2) {
    __syncthreads();
    fast_write<C, n, RD::R0, 1, V, P>(s, v, t);
    __syncthreads();
    fast_read<C, n, RD::R1, V, P>(v, s, t);
    pb.s1.apply(v);
    if (INV) fast_idfts<C, RD::R1, V>(v); else fast_dfts<C, RD::R1, V>(v);
  }
  if constexpr (RD::NST >= 3) {
    __syncthreads();
    fast_write<C, n, RD::R1, RD::R0, V, P>(s, v, t);
    __syncthreads();
    fast_read<C, n, RD::R2, V, P>(v, s, t);
    pb.s2.apply(v);
    if (INV) fast_idfts<C, RD::R2, V>(v); else fast_dfts<C, RD::R2, V>(v);
  }
}

template <typename RDF>
__device__ __forceinline__ void ypipe_loadh(float2* h, const QuadTab& qt, int b, int qoff_d, int qoff_m) {
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  constexpr int RL = RDF::Rev::R0;
  const auto rt = make_rsrc(qt.tab + (long)qt.tabidx[b] * qt.tabsize, qt.tabsize * 8);
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int c = P * m + r * (n / RL);
      h[m * RL + r] = (c < n / 2) ? buf_load(rt, qoff_d, c * 8) : buf_load(rt, qoff_m, (n - c - P) * 8);
    }
}

// One field: FFT with cur's registers, multiply by h, IFFT, store; the caller has already
// issued the loads of the next field (data and transfer function) into other registers.
template <typename RDF, bool OUT_HALF, int HS = 1>
__device__ __forceinline__ void ypipe_field(float2* v, const float2* h, float2* s, int t, bool live, int b,
                                            const FastIO& io, const PlanBases<RDF>& pf,
                                            const PlanBases<typename RDF::Rev>& pi, int conj, int vout,
                                            int kout) {
  using RI = typename RDF::Rev;
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  constexpr int RE = LastRadix<RI>::value;
  fast_fft_b<RDF, false>(v, s, t, pf);
#pragma unroll
  for (int k = 0; k < V; ++k) v[k] = conj ? cmulc(v[k], h[k * HS]) : cmul(v[k], h[k * HS]);
  fast_fft_b<RI, true>(v, s, t, pi);
  // Loads and stores share the vector-memory counter, and a wait for a load with stores
  // still pending drains everything.  Wait here, before this field's stores, while every
  // outstanding operation is old (the next field's prefetch, the previous field's stores):
  // the next field then starts with its data known present and nothing to drain.
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  if (live) {
    const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
#pragma unroll
    for (int m = 0; m < V / RE; ++m)
#pragma unroll
      for (int r = 0; r < RE; ++r) {
        const int c = P * m + r * (n / RE);
        if (!OUT_HALF || (c < n / 2 && t + c < io.out_valid)) buf_store(rout, vout, c * kout, v[m * RE + r]);
      }
  }
}

// Software-pipelined batched y pass (one resident workgroup per CU, up to 256 VGPRs): while
// field b is transformed, the column data of field b+1 is already in flight into a second
// register set (ping-pong over two unrolled iterations).  ONE_TAB: every field uses the same
// transfer function (one wavelength), so its values are loaded once and kept in registers;
// otherwise they are prefetched with the data.
template <typename RDF, int L, bool IN_HALF, bool OUT_HALF, bool ONE_TAB, bool H_LDS = false>
__global__ void __launch_bounds__(L * RDF::P)
k_fast_ypipe(FastIO io, int batch, const float2* __restrict__ tabf, const float2* __restrict__ tabi,
             QuadTab qt) {
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  using RI = typename RDF::Rev;
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  int l, t;
  thread_map<L, P, true>(threadIdx.x, l, t);
  // fewer than 4 columns per workgroup: the tiles sharing a 128-byte line run on one XCD
  const int li = (L < 4 ? xcd_group_tile(blockIdx.x, gridDim.x, 4 / L) : (int)blockIdx.x) * L + l;
  const bool live = li < io.nlines;
  float2* s = smem + l * LdsLine<n, L>::size;
  const int vin = fast_off<true>(li, t, io.in_l, io.in_e, io.in_blk) * 8;
  const int kin = fast_k<true>(io.in_l, io.in_e, io.in_blk) * 8;
  const int vout = fast_off<true>(li, t, io.out_l, io.out_e, io.out_blk) * 8;
  const int kout = fast_k<true>(io.out_l, io.out_e, io.out_blk) * 8;
  const int lic = live ? li : 0;
  const int qxi = qt.full ? lic : (lic <= qt.nx - lic ? lic : qt.nx - lic);
  const int qoff_d = (qxi * qt.qcol + t) * 8;
  const int qoff_m = (qxi * qt.qcol + qt.moff + P - t) * 8;
  PlanBases<RDF> pf;
  PlanBases<RI> pi;
  pf.load(tabf, t);
  pi.load(tabi, t);
  float2 va[V], vb[V], ha[V];
  yconv_load<V, RDF::R0, n, P, IN_HALF>(va, io, 0, vin, kin, t);
  ypipe_loadh<RDF>(ha, qt, 0, qoff_d, qoff_m);
  // The prefetches are unconditional (past the last field they re-read field batch-1, which
  // is harmless): a load under a branch makes the wait-count pass drain every outstanding
  // load at the join, which would serialise the prefetch with the arithmetic.
  if constexpr (ONE_TAB && H_LDS) {
    // the shared transfer function lives in LDS after the exchange areas, one slot per
    // (value, thread): each thread reads back only what it wrote, so no barrier is needed
    constexpr int LP = L * P;
    float2* hs = smem + L * LdsLine<n, L>::size + threadIdx.x;
#pragma unroll
    for (int k = 0; k < V; ++k) hs[k * LP] = ha[k];
    for (int b = 0; b < batch; b += 2) {
      yconv_load<V, RDF::R0, n, P, IN_HALF>(vb, io, b + 1 < batch ? b + 1 : batch - 1, vin, kin, t);
      ypipe_field<RDF, OUT_HALF, LP>(va, hs, s, t, live, b, io, pf, pi, qt.conj, vout, kout);
      if (b + 1 >= batch) break;
      yconv_load<V, RDF::R0, n, P, IN_HALF>(va, io, b + 2 < batch ? b + 2 : batch - 1, vin, kin, t);
      ypipe_field<RDF, OUT_HALF, LP>(vb, hs, s, t, live, b + 1, io, pf, pi, qt.conj, vout, kout);
    }
  } else if constexpr (ONE_TAB) {
    for (int b = 0; b < batch; b += 2) {
      yconv_load<V, RDF::R0, n, P, IN_HALF>(vb, io, b + 1 < batch ? b + 1 : batch - 1, vin, kin, t);
      ypipe_field<RDF, OUT_HALF>(va, ha, s, t, live, b, io, pf, pi, qt.conj, vout, kout);
      if (b + 1 >= batch) break;
      yconv_load<V, RDF::R0, n, P, IN_HALF>(va, io, b + 2 < batch ? b + 2 : batch - 1, vin, kin, t);
      ypipe_field<RDF, OUT_HALF>(vb, ha, s, t, live, b + 1, io, pf, pi, qt.conj, vout, kout);
    }
  } else {
    float2 hb[V];
    for (int b = 0; b < batch; b += 2) {
      const int b1 = b + 1 < batch ? b + 1 : batch - 1;
      yconv_load<V, RDF::R0, n, P, IN_HALF>(vb, io, b1, vin, kin, t);
      ypipe_loadh<RDF>(hb, qt, b1, qoff_d, qoff_m);
      ypipe_field<RDF, OUT_HALF>(va, ha, s, t, live, b, io, pf, pi, qt.conj, vout, kout);
      if (b + 1 >= batch) break;
      const int b2 = b + 2 < batch ? b + 2 : batch - 1;
      yconv_load<V, RDF::R0, n, P, IN_HALF>(va, io, b2, vin, kin, t);
      ypipe_loadh<RDF>(ha, qt, b2, qoff_d, qoff_m);
      ypipe_field<RDF, OUT_HALF>(vb, hb, s, t, live, b + 1, io, pf, pi, qt.conj, vout, kout);
    }
  }
}

}  // namespace fh

namespace fh {

// ---------------------------------------------------------------- wave-per-column y pass
// A workgroup of four waves owns a tile of four adjacent spectrum columns of 2048 points and walks
// the batch.  Per field the tile moves between HBM and LDS cooperatively, as whole 128-byte block
// rows of the 4x4-blocked layout (16-byte loads and stores, every line read and written in full),
// and each wave transforms its own column (32 points a lane, radices 32 x 8 x 8) with exchanges in
// its own LDS column, so those exchanges need no workgroup barrier: LDS operations of one wave
// complete in order, and a compiler fence keeps the accesses where they are.  Two barriers a field
// remain: after the tile is written to LDS and after the columns are written back.  The thread
// that writes a pair of tile values into LDS is the one that reads it back out for the store, so
// the next field's tile may be written without a barrier in between.  The transfer function is
// loaded per field (from L2) into the tile registers as soon as the tile is in LDS, so the loads
// are in flight during the forward transform; after the multiply the same registers receive the
// next field's tile.

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stage twiddle bases (as TwBases) kept in LDS instead of registers, one slot per (base, lane):
// base i of lane t at tw[i * 64 + t], so a wave reads 64 consecutive slots.  A stage whose
// butterflies all have the same j mod NS (P a multiple of NS) stores one set per lane.
template <int R, int NS, int V, int P>
struct TwSlots {
  static constexpr int NB = ilog2c(R);
  static constexpr bool SHARED = P % NS == 0;
  static constexpr int NBF = NS > 1 ? (SHARED ? 1 : V / R) : 0;
  static constexpr int count = NBF * NB;
};

template <int R, int NS, int V, int P, int OFF>
__device__ __forceinline__ void tw_slots_store(float2* tw, const float2* __restrict__ tab, int t) {
  using S = TwSlots<R, NS, V, P>;
#pragma unroll
  for (int m = 0; m < S::NBF; ++m) {
    const int jm = (t + P * m) & (NS - 1);
#pragma unroll
    for (int k = 0; k < S::NB; ++k) tw[(OFF + m * S::NB + k) * 64 + t] = tab[jm * (R - 1) + ((1 << k) - 1)];
  }
}

template <int R, int NS, int V, int P, int OFF>
__device__ __forceinline__ void tw_slots_apply(float2* v, const float2* tw, int t) {
  using S = TwSlots<R, NS, V, P>;
  if constexpr (NS > 1) {
#pragma unroll
    for (int m = 0; m < V / R; ++m) {
      const int mb = S::SHARED ? 0 : m;
      float2 b[S::NB];
#pragma unroll
      for (int k = 0; k < S::NB; ++k) b[k] = tw[(OFF + mb * S::NB + k) * 64 + t];
      float2 w[R];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int low = r & (-r);
        const int k = ilog2c(low);
        w[r] = (r == low) ? b[k] : cmul(w[r - low], b[k]);
        v[m * R + r] = cmul(v[m * R + r], w[r]);
      }
    }
  }
}

// the slots of a radix plan's two twiddled stages, starting at slot OFF
template <typename RD, int OFF>
struct PlanSlots {
  using S1 = TwSlots<RD::R1, RD::R0, RD::V, RD::P>;
  using S2 = TwSlots<RD::R2, RD::R0 * RD::R1, RD::V, RD::P>;
  static constexpr int off1 = OFF, off2 = OFF + (RD::NST >= 2 ? S1::count : 0);
  static constexpr int end = off2 + (RD::NST >= 3 ? S2::count : 0);
  __device__ __forceinline__ static void store(float2* tw, const float2* __restrict__ tab, int t) {
    if constexpr (RD::NST >= 2) tw_slots_store<RD::R1, RD::R0, RD::V, RD::P, off1>(tw, tab + FastTw<RD>::off1, t);
    if constexpr (RD::NST >= 3) tw_slots_store<RD::R2, RD::R0 * RD::R1, RD::V, RD::P, off2>(tw, tab + FastTw<RD>::off2, t);
  }
};

// one column's transform within one wave, exchanging through the wave's own LDS column s, with
// the stage twiddles from the LDS slots at tw; after0() runs between the first stage and the
// first exchange
struct NoHook {
  __device__ __forceinline__ void operator()() const {}
};
template <typename RD, int OFF, bool INV, typename F = NoHook>
__device__ __forceinline__ void wave_fft(float2* v, float2* s, const float2* tw, int t, F after0 = F()) {
  using C = float2;
  using PS = PlanSlots<RD, OFF>;
  constexpr int n = RD::n, V = RD::V, P = RD::P;
  if (INV) fast_idfts<C, RD::R0, V>(v); else fast_dfts<C, RD::R0, V>(v);
  after0();
  if constexpr (RD::NST >= 2) {
    wave_lds_sync();
    fast_write<C, n, RD::R0, 1, V, P>(s, v, t);
    wave_lds_sync();
    fast_read<C, n, RD::R1, V, P>(v, s, t);
    tw_slots_apply<RD::R1, RD::R0, V, P, PS::off1>(v, tw, t);
    if (INV) fast_idfts<C, RD::R1, V>(v); else fast_dfts<C, RD::R1, V>(v);
  }
  if constexpr (RD::NST >= 3) {
    wave_lds_sync();
    fast_write<C, n, RD::R1, RD::R0, V, P>(s, v, t);
    wave_lds_sync();
    fast_read<C, n, RD::R2, V, P>(v, s, t);
    tw_slots_apply<RD::R2, RD::R0 * RD::R1, V, P, PS::off2>(v, tw, t);
    if (INV) fast_idfts<C, RD::R2, V>(v); else fast_dfts<C, RD::R2, V>(v);
  }
}


constexpr int kWaveCols = 4;   // columns (waves) of a tile
constexpr int kWaveN = 2048;   // column length
constexpr int kWavePairs = kWaveN * kWaveCols / 2 / (kWaveCols * 64);  // 16-byte pairs a thread

// the thread's 16 pairs of field b's tile: pair k is block row 32 k + (T >> 3) of the tile,
// elements 2 (T & 7) and 2 (T & 7) + 1
__device__ __forceinline__ void ywave_tile_load(float2* q, const FastIO& io, int b, int voff, int kstep) {
  const auto r = make_rsrc(io.in + b * io.in_f, io.in_bytes);
#pragma unroll
  for (int k = 0; k < kWavePairs; ++k) {
    const float4 x = buf_load4(r, voff, k * kstep);
    q[2 * k] = make_float2(x.x, x.y);
    q[2 * k + 1] = make_float2(x.z, x.w);
  }
}

// the transfer-function column values of field b at the positions the forward transform leaves
// in registers (as ypipe_loadh); the table index is read once per wave into a scalar register,
// since a descriptor the compiler cannot prove uniform turns every load into a loop over lanes
template <typename RDF>
__device__ __forceinline__ void ywave_loadh(float2* h, const QuadTab& qt, int b, int qoff_d, int qoff_m) {
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  constexpr int RL = RDF::Rev::R0;
  const int ti = __builtin_amdgcn_readfirstlane(qt.tabidx[b]);
  const auto rt = make_rsrc(qt.tab + (long)ti * qt.tabsize, qt.tabsize * 8);
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int c = P * m + r * (n / RL);
      h[m * RL + r] = (c < n / 2) ? buf_load(rt, qoff_d, c * 8) : buf_load(rt, qoff_m, (n - c - P) * 8);
    }
}

using WaveRDF = Radices<8, 8, 32>;  // forward radices (the inverse runs them reversed)
constexpr int kWaveSlots = PlanSlots<WaveRDF::Rev, PlanSlots<WaveRDF, 0>::end>::end;
// LDS of a workgroup: four columns and the twiddle slots (two workgroups fit one CU's 160 KiB)
constexpr size_t kWaveLds = ((size_t)kWaveCols * LdsLine<kWaveN, kWaveCols>::size + (size_t)kWaveSlots * 64) * 8;
static_assert(2 * kWaveLds <= 160 * 1024, "two workgroups per CU");

// Two waves per SIMD (two workgroups per CU): at most 256 registers a lane.
__global__ void __launch_bounds__(kWaveCols * 64) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_fast_ywave(FastIO io, int batch, const float2* __restrict__ tabf, const float2* __restrict__ tabi, QuadTab qt) {
  using RDF = WaveRDF;
  using RI = typename RDF::Rev;
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  constexpr int RE = LastRadix<RI>::value;
  static_assert(n == kWaveN && P == 64 && V == 32, "one wave per 2048-point column");
  static_assert((n / RDF::R0) % 16 == 0 && (n / RE) % 16 == 0, "register positions 16 slots apart");
  constexpr int CS = LdsLine<n, kWaveCols>::size;  // one column's LDS slots
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  const int T = threadIdx.x;
  const int t = T & 63;
  // tiles j and n/4 - 1 - j (whose columns kx and n - kx read 3 of the same 4 quadrant-table rows)
  // go to workgroups 8 apart, which the dispatcher places on the same XCD one after the other, so
  // the second finds the shared rows in that XCD's L2 (a bijection when the tile count is a
  // multiple of 16; otherwise tiles follow the workgroups)
  int tile = blockIdx.x;
  if (gridDim.x % 16 == 0) {
    const int b = blockIdx.x, s8 = b >> 3, pair = ((s8 >> 1) << 3) + (b & 7);
    tile = (s8 & 1) ? (int)gridDim.x - 1 - pair : pair;
  }
  const int li = tile * kWaveCols + (T >> 6);  // this wave's column
  float2* s = smem + (T >> 6) * CS;
  // the thread's tile pairs: rows 128 k + prow, columns pc and pc + 1 of the tile
  const int pw = T & 7;
  const int prow = 4 * (T >> 3) + (pw >> 1);
  float2* ps = smem + 2 * (pw & 1) * CS + lds_pos(prow);
  const int rowlen = io.in_e;  // spectrum columns (elements between rows)
  const int vin = ((((T >> 3) * (rowlen >> 2)) + tile) * 16 + 2 * pw) * 8;
  const int kin = rowlen * 1024;  // 32 block rows of the tile
  const int vout = ((((T >> 3) * (io.out_e >> 2)) + tile) * 16 + 2 * pw) * 8;
  const int kout = io.out_e * 1024;
  const int qxi = qt.full ? li : (li <= qt.nx - li ? li : qt.nx - li);
  const int qoff_d = (qxi * qt.qcol + t) * 8;
  const int qoff_m = (qxi * qt.qcol + qt.moff + P - t) * 8;
  // the stage twiddle bases of both transforms, after the columns (one wave writes them; the
  // first barrier below orders the writes before any use)
  using PF = PlanSlots<RDF, 0>;
  using PI = PlanSlots<RI, PF::end>;
  float2* tw = smem + kWaveCols * CS;
  if (T < 64) {
    PF::store(tw, tabf, t);
    PI::store(tw, tabi, t);
  }
  float2 q[V];  // the tile pairs, or the transfer-function values
  ywave_tile_load(q, io, 0, vin, kin);
  for (int b = 0; b < batch; ++b) {
#pragma unroll
    for (int k = 0; k < kWavePairs; ++k) {  // lds_pos(128 k + prow) = 136 k + lds_pos(prow)
      ps[136 * k] = q[2 * k];
      ps[CS + 136 * k] = q[2 * k + 1];
    }
    // the transfer function at the positions the forward transform leaves in registers, from L2,
    // in flight during the forward transform (the tile registers are free once written to LDS)
    ywave_loadh<RDF>(q, qt, b, qoff_d, qoff_m);
    __syncthreads();
    float2 v[V];
#pragma unroll
    for (int m = 0; m < V / RDF::R0; ++m)  // the first stage's input positions
#pragma unroll
      for (int r = 0; r < RDF::R0; ++r) {
        const int c = P * m + r * (n / RDF::R0);  // a multiple of 16: one pad slot per 16
        v[m * RDF::R0 + r] = s[lds_pos(t) + c + (c >> 4)];
      }
    wave_fft<RDF, PF::off1, false>(v, s, tw, t);
#pragma unroll
    for (int k = 0; k < V; ++k) v[k] = qt.conj ? cmulc(v[k], q[k]) : cmul(v[k], q[k]);
    // the next field's tile (or this one again after the last) goes out after the inverse's
    // first stage, the one with the most registers live, and is in flight during the rest
    wave_fft<RI, PI::off1, true>(v, s, tw, t, [&] {
      asm volatile("" ::: "memory");
      ywave_tile_load(q, io, b + 1 < batch ? b + 1 : b, vin, kin);
    });
    wave_lds_sync();
#pragma unroll
    for (int m = 0; m < V / RE; ++m)  // the last stage's output positions
#pragma unroll
      for (int r = 0; r < RE; ++r) {
        const int c = P * m + r * (n / RE);
        s[lds_pos(t) + c + (c >> 4)] = v[m * RE + r];
      }
    // wait for the prefetch here, while the only other outstanding operations are the previous
    // field's stores: a wait for a load issued after stores drains the stores too
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
    const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
#pragma unroll
    for (int k = 0; k < kWavePairs; ++k) {
      const float2 a = ps[136 * k], c = ps[CS + 136 * k];
      buf_store4(rout, vout, k * kout, make_float4(a.x, a.y, c.x, c.y));
    }
  }
}

// ---------------------------------------------------------------- staged-tile pipelined y pass
// The pipelined pass's eight-wave transform (L = 4 columns, one workgroup per CU) with the
// column tile moved between HBM and LDS as whole 128-byte block rows of the 4x4-blocked layout
// (16-byte loads and stores, a wave touching 8 lines per instruction) instead of 8-byte loads of
// each thread's first-stage positions (4 lines per instruction).  Per field: the tile registers
// go to the exchange area in column order, the next field's tile is loaded into the same
// registers (in flight during both transforms), the forward transform starts from LDS, and the
// inverse's outputs return through LDS to the tile registers' owners for the store.  The thread
// that stores a tile pair is the one that writes the next field's pair to the same slots, so no
// barrier separates them.  ONE_TAB: the shared transfer function sits in LDS after the exchange
// area (conjugated there once for the adjoint); otherwise each field's values are loaded into
// registers when the field starts and used after the forward transform.
template <int NQ>
__device__ __forceinline__ void ystg_tile_load(float4* q, const FastIO& io, int b, int voff, int kstep) {
  const auto r = make_rsrc(io.in + b * io.in_f, io.in_bytes);
#pragma unroll
  for (int k = 0; k < NQ; ++k) q[k] = buf_load4(r, voff, k * kstep);
}

template <typename RDF, bool ONE_TAB>
__global__ void __launch_bounds__(4 * RDF::P)
k_fast_ystg(FastIO io, int batch, const float2* __restrict__ tabf, const float2* __restrict__ tabi, QuadTab qt) {
  constexpr int L = 4;  // one column of 4x4 blocks
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  using RI = typename RDF::Rev;
  constexpr int RE = LastRadix<RI>::value;
  constexpr int LS = LdsLine<n, L>::size;
  constexpr int NQ = n / (2 * P);          // 16-byte tile pairs a thread
  constexpr int KS = 2 * P + 2 * P / 16;   // lds_pos step between a thread's pairs (2P rows)
  constexpr int LP = L * P;
  static_assert(P % 16 == 0 && (n / RDF::R0) % 16 == 0 && (n / RE) % 16 == 0, "slots 16 apart");
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  const int T = threadIdx.x;
  int l, t;
  thread_map<L, P, true>(T, l, t);
  const int tile = blockIdx.x;  // (pairing mirrored tiles as k_fast_ywave does: 0.147 against 0.140 ms)
  const int li = tile * L + l;
  float2* s = smem + l * LS;
  // tile pair k of the thread: block row (T >> 3) + k P / 2, elements 2 (T & 7) and 2 (T & 7) + 1,
  // i.e. row 4 (T >> 3) + ((T & 7) >> 1) + 2 P k of columns 2 (T & 1) and 2 (T & 1) + 1
  const int pw = T & 7;
  float2* ps = smem + 2 * (pw & 1) * LS + lds_pos(4 * (T >> 3) + (pw >> 1));
  const int vin = ((((T >> 3) * (io.in_e >> 2)) + tile) * 16 + 2 * pw) * 8;
  const int kin = io.in_e * (2 * P) * 8;
  const int vout = ((((T >> 3) * (io.out_e >> 2)) + tile) * 16 + 2 * pw) * 8;
  const int kout = io.out_e * (2 * P) * 8;
  const int qxi = qt.full ? li : (li <= qt.nx - li ? li : qt.nx - li);
  const int qoff_d = (qxi * qt.qcol + t) * 8;
  const int qoff_m = (qxi * qt.qcol + qt.moff + P - t) * 8;
  PlanBases<RDF> pf;
  PlanBases<RI> pi;
  pf.load(tabf, t);
  pi.load(tabi, t);
  float4 q[NQ];
  ystg_tile_load<NQ>(q, io, 0, vin, kin);
  float2* hs = smem + L * LS + T;
  if constexpr (ONE_TAB) {
    float2 ha[V];
    ywave_loadh<RDF>(ha, qt, 0, qoff_d, qoff_m);
#pragma unroll
    for (int k = 0; k < V; ++k) hs[k * LP] = qt.conj ? make_float2(ha[k].x, -ha[k].y) : ha[k];
  }
  const float hsg = qt.conj ? -1.f : 1.f;
  // nothing outstanding at the loop head: inside the loop the only wait is the one before the
  // stores (a wait for a load issued before pending stores would drain the stores as well)
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  for (int b = 0; b < batch; ++b) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      ps[KS * k] = make_float2(q[k].x, q[k].y);
      ps[LS + KS * k] = make_float2(q[k].z, q[k].w);
    }
    asm volatile("" ::: "memory");
    float2 h[ONE_TAB ? 1 : V];
    if constexpr (!ONE_TAB) ywave_loadh<RDF>(h, qt, b, qoff_d, qoff_m);
    // the next field's tile (this one again after the last), in flight during both transforms
    ystg_tile_load<NQ>(q, io, b + 1 < batch ? b + 1 : b, vin, kin);
    __syncthreads();
    float2 v[V];
#pragma unroll
    for (int m = 0; m < V / RDF::R0; ++m)  // the first stage's input positions
#pragma unroll
      for (int r = 0; r < RDF::R0; ++r) {
        const int c = P * m + r * (n / RDF::R0);
        v[m * RDF::R0 + r] = s[lds_pos(t) + c + (c >> 4)];
      }
    fast_fft_b<RDF, false>(v, s, t, pf);
    if constexpr (ONE_TAB) {
#pragma unroll
      for (int k = 0; k < V; ++k) v[k] = cmul(v[k], hs[k * LP]);
    } else {
#pragma unroll
      for (int k = 0; k < V; ++k) v[k] = cmul(v[k], make_float2(h[k].x, hsg * h[k].y));
    }
    fast_fft_b<RI, true>(v, s, t, pi);
    // the last stage's output positions are the slots this thread read in the last exchange (the
    // same radix on both sides), which no other thread reads: no barrier before the writes
#pragma unroll
    for (int m = 0; m < V / RE; ++m)
#pragma unroll
      for (int r = 0; r < RE; ++r) {
        const int c = P * m + r * (n / RE);
        s[lds_pos(t) + c + (c >> 4)] = v[m * RE + r];
      }
    // the prefetch has had both transforms to arrive; the previous field's stores are the only
    // other outstanding operations
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
    const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const float2 a = ps[KS * k], c = ps[LS + KS * k];
      buf_store4(rout, vout, k * kout, make_float4(a.x, a.y, c.x, c.y));
    }
  }
}

}  // namespace fh
