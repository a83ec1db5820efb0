// libfluxhip: the C ABI of include/fluxhip.h and its kernel launches (gfx950).
//
// Propagation = x-line FFT pass -> y-line (FFT, transfer-function multiply, IFFT)
// pass -> x-line IFFT pass, all in place on the caller's c64 array.  Transfer
// functions are tables computed in FP64 on the device and rounded once to c64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fluxhip.h"
#include "fh_bpm.hpp"
#include "fh_fast.hpp"

namespace fh {

// ================================================================ errors

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define FH_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) return fail(FH_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// ================================================================ launch configurations
// (radices, lines per workgroup) per precision, length and axis.  X passes read
// contiguous lines; Y passes read L adjacent columns per row.

template <typename T, int LOGN, bool Y> struct Cfg;
#define FH_CFG(T, LOGN, Y, A, B, C_, LL)            \
  template <> struct Cfg<T, LOGN, Y> {              \
    using RD = Radices<A, B, C_>;                   \
    static constexpr int L = LL;                    \
  };
// FP32
FH_CFG(float, 4, false, 4, 4, 1, 64)
FH_CFG(float, 5, false, 8, 4, 1, 64)
FH_CFG(float, 6, false, 8, 8, 1, 32)
FH_CFG(float, 7, false, 16, 8, 1, 32)
FH_CFG(float, 8, false, 16, 16, 1, 16)
FH_CFG(float, 9, false, 16, 32, 1, 16)
FH_CFG(float, 10, false, 32, 32, 1, 8)
FH_CFG(float, 11, false, 16, 16, 8, 4)
FH_CFG(float, 12, false, 16, 16, 16, 4)
FH_CFG(float, 4, true, 4, 4, 1, 16)
FH_CFG(float, 5, true, 8, 4, 1, 16)
FH_CFG(float, 6, true, 8, 8, 1, 16)
FH_CFG(float, 7, true, 16, 8, 1, 16)
FH_CFG(float, 8, true, 16, 16, 1, 16)
FH_CFG(float, 9, true, 16, 32, 1, 16)
FH_CFG(float, 10, true, 32, 32, 1, 8)
FH_CFG(float, 11, true, 16, 16, 8, 4)
FH_CFG(float, 12, true, 16, 16, 16, 4)
// FP64 (tables and the plain comparator)
FH_CFG(double, 4, false, 4, 4, 1, 64)
FH_CFG(double, 5, false, 8, 4, 1, 64)
FH_CFG(double, 6, false, 8, 8, 1, 32)
FH_CFG(double, 7, false, 16, 8, 1, 16)
FH_CFG(double, 8, false, 16, 16, 1, 8)
FH_CFG(double, 9, false, 8, 8, 8, 4)
FH_CFG(double, 10, false, 16, 8, 8, 4)
FH_CFG(double, 11, false, 16, 16, 8, 2)
FH_CFG(double, 12, false, 16, 16, 16, 1)
FH_CFG(double, 4, true, 4, 4, 1, 16)
FH_CFG(double, 5, true, 8, 4, 1, 16)
FH_CFG(double, 6, true, 8, 8, 1, 16)
FH_CFG(double, 7, true, 16, 8, 1, 16)
FH_CFG(double, 8, true, 16, 16, 1, 16)
FH_CFG(double, 9, true, 8, 8, 8, 8)
FH_CFG(double, 10, true, 16, 8, 8, 8)
FH_CFG(double, 11, true, 16, 16, 8, 4)
FH_CFG(double, 12, true, 16, 16, 16, 2)
#undef FH_CFG

constexpr int kMinLog = 4, kMaxLog = 12;

template <typename C> struct RealOf;
template <> struct RealOf<float2> { using type = float; };
template <> struct RealOf<double2> { using type = double; };

template <typename K>
static hipError_t ensure_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename S, typename C, int LOGN, bool Y, bool INV, typename Op>
static hipError_t launch_fft_t(const LineIO<S>& io, int batch, const C* tw, Op op, hipStream_t st) {
  using CF = Cfg<typename RealOf<C>::type, LOGN, Y>;
  using RD = typename CF::RD;
  constexpr int L = CF::L;
  static_assert(RD::n == (1 << LOGN), "radix plan does not match the length");
  auto kern = k_line_fft<S, C, RD, L, Y, INV, Op>;
  const size_t lds = (RD::NST > 1) ? (size_t)L * LdsLine<RD::n, L, sizeof(C)>::size * sizeof(C) : 0;
  hipError_t e = ensure_lds(kern, lds);
  if (e != hipSuccess) return e;
  dim3 grid((io.nlines + L - 1) / L, batch);
  const int G = (Y && !io.in_blk) ? std::max(1, (int)(128 / (L * sizeof(S)))) : 1;
  hipLaunchKernelGGL(kern, grid, dim3(L * RD::P), lds, st, io, tw, op, G);
  return hipGetLastError();
}

// Small problems (e.g. one 512^2 field per AS_BPM slice) give too few workgroups to cover the
// 256 CUs at the configured lines per workgroup; launches then halve L until there are at
// least kFillWGs workgroups (or L = 1).
constexpr long kFillWGs = 512;

// The strided FP64 convolution (the AS_BPM slice propagator) keeps two columns per workgroup down
// to kFillWGs / 2 workgroups: a 512 x 512 slice then runs in 6.6 us with 256 two-column workgroups
// against 7.2 us with 512 one-column ones (a wave's column loads touch a 128-byte line per lane;
// two columns per workgroup share those lines), and 8.8 us with four columns.
template <typename S, typename C, typename RD, int L, typename Op>
static hipError_t launch_conv_l(const LineIO<S>& io, int batch, const C* tw, Op op, hipStream_t st) {
  if constexpr (L > 2) {
    if ((long)((io.nlines + L - 1) / L) * batch < kFillWGs / 2)
      return launch_conv_l<S, C, RD, L / 2, Op>(io, batch, tw, op, st);
  }
  auto kern = k_line_conv<S, C, RD, L, true, Op>;
  const size_t lds = (RD::NST > 1) ? (size_t)L * LdsLine<RD::n, L, sizeof(C)>::size * sizeof(C) : 0;
  hipError_t e = ensure_lds(kern, lds);
  if (e != hipSuccess) return e;
  dim3 grid((io.nlines + L - 1) / L, batch);
  const int G = io.in_blk ? 1 : std::max(1, (int)(128 / (L * sizeof(S))));
  hipLaunchKernelGGL(kern, grid, dim3(L * RD::P), lds, st, io, tw, op, G);
  return hipGetLastError();
}

template <typename S, typename C, int LOGN, typename Op>
static hipError_t launch_conv_t(const LineIO<S>& io, int batch, const C* tw, Op op, hipStream_t st) {
  using CF = Cfg<typename RealOf<C>::type, LOGN, true>;
  return launch_conv_l<S, C, typename CF::RD, CF::L, Op>(io, batch, tw, op, st);
}

#define FH_LOGN_SWITCH(logn, CALL)                                       \
  switch (logn) {                                                        \
    case 4: { constexpr int LG = 4; return CALL; }                       \
    case 5: { constexpr int LG = 5; return CALL; }                       \
    case 6: { constexpr int LG = 6; return CALL; }                       \
    case 7: { constexpr int LG = 7; return CALL; }                       \
    case 8: { constexpr int LG = 8; return CALL; }                       \
    case 9: { constexpr int LG = 9; return CALL; }                       \
    case 10: { constexpr int LG = 10; return CALL; }                     \
    case 11: { constexpr int LG = 11; return CALL; }                     \
    case 12: { constexpr int LG = 12; return CALL; }                     \
    default: return hipErrorInvalidValue;                                \
  }

template <typename S, typename C, bool Y, bool INV, typename Op>
static hipError_t launch_fft(int logn, const LineIO<S>& io, int batch, const C* tw, Op op,
                             hipStream_t st) {
  FH_LOGN_SWITCH(logn, (launch_fft_t<S, C, LG, Y, INV, Op>(io, batch, tw, op, st)))
}

template <typename S, typename C, typename Op>
static hipError_t launch_conv(int logn, const LineIO<S>& io, int batch, const C* tw, Op op,
                              hipStream_t st) {
  FH_LOGN_SWITCH(logn, (launch_conv_t<S, C, LG, Op>(io, batch, tw, op, st)))
}

// FFT -> frequency op -> IFFT along contiguous lines (1-D fields: the batch's lines, one pass per
// propagation), lines per workgroup halved while a launch has fewer than kFillWGs workgroups.
template <typename S, typename C, typename RD, int L, typename Op>
static hipError_t launch_line_conv_l(const LineIO<S>& io, const C* tw, Op op, hipStream_t st) {
  if constexpr (L > 1) {
    if ((long)((io.nlines + L - 1) / L) < kFillWGs) return launch_line_conv_l<S, C, RD, L / 2, Op>(io, tw, op, st);
  }
  auto kern = k_line_conv<S, C, RD, L, false, Op>;
  const size_t lds = (RD::NST > 1) ? (size_t)L * LdsLine<RD::n, L, sizeof(C)>::size * sizeof(C) : 0;
  hipError_t e = ensure_lds(kern, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((io.nlines + L - 1) / L, 1), dim3(L * RD::P), lds, st, io, tw, op, 1);
  return hipGetLastError();
}

// at most 512 threads a workgroup, so a thread may hold 256 VGPRs: FFT, table values and IFFT of a
// 4096-point line stay in registers (1024 threads would cap them at 128 and spill)
template <typename RD, int L> constexpr int line_conv_L() { return (L * RD::P > 512 && L > 1) ? (512 / RD::P > 0 ? 512 / RD::P : 1) : L; }

template <typename S, typename C, typename Op>
static hipError_t launch_line_conv(int logn, const LineIO<S>& io, const C* tw, Op op, hipStream_t st) {
  FH_LOGN_SWITCH(logn, (launch_line_conv_l<S, C, typename Cfg<typename RealOf<C>::type, LG, false>::RD,
                                           line_conv_L<typename Cfg<typename RealOf<C>::type, LG, false>::RD,
                                                       Cfg<typename RealOf<C>::type, LG, false>::L>(), Op>(io, tw, op, st)))
}

// ================================================================ table kernels (FP64)

// Transfer-function table layouts (column-major, one column per kx so the y pass reads a
// contiguous run):
//   even tables (the function depends on |fx|, |fy| only): row i = |kx| in [0, nx/2],
//     entry j = |ky| in [0, ny/2];
//   full tables (tilted fields): row i = kx in [0, nx), entry j <= ny/2 holds ky = j and
//     entry j > ny/2 holds ky = 3*ny/2 - j (the upper half in reverse).
// Signed frequency index of an FFT bin: k < n/2 ? k : k - n (fftfreq order).
__device__ inline void tab_freq_index(int i, int j, int nx, int ny, int full, int& kx, int& ky) {
  if (!full) {
    kx = i;
    ky = j;
    return;
  }
  kx = i < nx / 2 ? i : i - nx;
  const int ku = j <= ny / 2 ? j : 3 * ny / 2 - j;
  ky = ku < ny / 2 ? ku : ku - ny;
}

// Angular-spectrum transfer functions, computed in FP64 and rounded once to c64
// (double_precision_kernel=true).  lam' = lam/n0, f0 = sin(theta)/lam'.
//   mode 0, on-axis (angular_spectrum.jl:1-12): a = 1/lam'^2 - fx^2 - fy^2 (evaluated in
//     this order, no contraction, as the reference's FP64 broadcast does);
//     a >= 0: H = cis(2*pi*t), t = z*sqrt(a) in cycles, reduced to t - rint(t) before
//     sincospi so the phase keeps full FP64 accuracy at |2*pi*t| ~ 1e5 rad;
//     a < 0: H = exp(-2*pi*|z|*sqrt(-a)) (principal sqrt; the same for z < 0, where
//     H(z) = conj(H(-z))).
//   mode 1, paraxial (:27-35): t = z/lam' - lam'*z*((fx+f0x)^2 + (fy+f0y)^2)/2
//     - [tracked] (f0x*fx + f0y*fy)/pi, for either sign of z.
//   mode 2, tilted (:45-62): a = 1/lam'^2 - (fx+f0x)^2 - (fy+f0y)^2, off = [tracked]
//     (f0x*fx + f0y*fy)*lam'; a >= 0: t = z*(sqrt(a) + off); a < 0: amplitude
//     exp(-2*pi*|z|*sqrt(-a)) and t = z*off (conj(H(-z)) for z < 0 gives the same).
//   mode 3, geometric shift (shift_kernel.jl:1-12): t = -z*(tan(theta_x)*fx + tan(theta_y)*fy).
// The inverse FFT's 1/(nx*ny) is folded in (scale).  T2 = float2 rounds the FP64 value
// once to c64; T2 = double2 keeps it (the BPM slice steps).
__device__ inline void store_tab(float2* p, double re, double im) { *p = make_float2((float)re, (float)im); }
__device__ inline void store_tab(double2* p, double re, double im) { *p = make_double2(re, im); }

// The transfer function at (fx, fy) in cycles of phase (t) and amplitude (amp); see above.
__device__ inline void as_transfer(double fx, double fy, double lam, double n0, double z, double tx, double ty,
                                   int mode, int track, double& amp, double& t) {
#pragma clang fp contract(off)
  const double le = lam / n0;
  amp = 1.0;
  t = 0.0;
  if (mode == 0) {
    const double f2 = 1.0 / (le * le);
    const double a = f2 - fx * fx - fy * fy;
    if (a >= 0.0) t = z * sqrt(a);
    else amp = exp(-2.0 * M_PI * fabs(z) * sqrt(-a));
  } else if (mode == 3) {
    t = -z * (tan(tx) * fx + tan(ty) * fy);
  } else {
    const double f0x = sin(tx) / le, f0y = sin(ty) / le;
    const double gx = fx + f0x, gy = fy + f0y;
    if (mode == 1) {
      t = z / le - 0.5 * le * z * (gx * gx + gy * gy);
      if (track) t -= (f0x * fx + f0y * fy) / M_PI;
    } else {
      const double f2 = 1.0 / (le * le);
      const double a = f2 - gx * gx - gy * gy;
      const double off = track ? (f0x * fx + f0y * fy) * le : 0.0;
      if (a >= 0.0) t = z * (sqrt(a) + off);
      else {
        amp = exp(-2.0 * M_PI * fabs(z) * sqrt(-a));
        t = z * off;
      }
    }
  }
}

// ASProp's frequency filter (angular_spectrum.jl:1-76: the kernel times Complex(filter(...))),
// sampled on the host on the fftfreq grid [ny][nx] (FFT order) at the arguments the kernel
// passes it; multiplied in FP64 before the table is rounded.  conjf: the reference conjugates
// the whole kernel, filter included, when z <= 0 (z_pos = Val(false)) except in the paraxial one.
__device__ inline void apply_filter(const double2* __restrict__ filt, int conjf, long k, double& re, double& im) {
  if (!filt) return;
  const double2 f = filt[k];
  const double fi = conjf ? -f.y : f.y;
  const double r = re * f.x - im * fi;
  im = re * fi + im * f.x;
  re = r;
}

template <typename T2>
__global__ void k_as_table(T2* __restrict__ tab, int nx, int ny, int rows, int qcol, int full,
                           double dx, double dy, double lam, double n0, double z, double tx,
                           double ty, int mode, int track, double scale,
                           const double2* __restrict__ filt = nullptr, int conjf = 0) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= rows || j >= qcol) return;
  int kx, ky;
  tab_freq_index(i, j, nx, ny, full, kx, ky);
  const double fx = (double)kx / ((double)nx * dx);
  const double fy = (double)ky / ((double)ny * dy);
  double amp, t;
  as_transfer(fx, fy, lam, n0, z, tx, ty, mode, track, amp, t);
  const double r = t - rint(t);
  double sn, cs;
  sincospi(2.0 * r, &sn, &cs);
  double re = amp * cs, im = amp * sn;
  apply_filter(filt, conjf, (long)(ky < 0 ? ky + ny : ky) * nx + (kx < 0 ? kx + nx : kx), re, im);
  store_tab(tab + (long)i * qcol + j, re * scale, im * scale);
}

// Transfer function of a 1-D field on an n-point power-of-two grid (FFT order): the reference's
// 1-D kernels (angular_spectrum.jl:14-25,37-43,64-76; shift_kernel.jl:8-12) are its 2-D ones at
// fy = 0 and theta_y = 0, so as_transfer is evaluated there; filter samples and scale as in
// k_as_table.
template <typename T2>
__global__ void k_as_line_table(T2* __restrict__ tab, int n, double dx, double lam, double n0, double z, double tx,
                                int mode, int track, double scale, const double2* __restrict__ filt, int conjf) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int kx = i < n / 2 ? i : i - n;
  double amp, t;
  as_transfer((double)kx / ((double)n * dx), 0.0, lam, n0, z, tx, 0.0, mode, track, amp, t);
  const double r = t - rint(t);
  double sn, cs;
  sincospi(2.0 * r, &sn, &cs);
  double re = amp * cs, im = amp * sn;
  apply_filter(filt, conjf, i, re, im);
  store_tab(tab + i, re * scale, im * scale);
}

// one padded FP64 line (the FFT of the periodically extended 1-D kernel) scaled into a table
template <typename T2>
__global__ void k_line_extract(T2* __restrict__ tab, const double2* __restrict__ src, int n, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) store_tab(tab + i, src[i].x * scale, src[i].y * scale);
}

// ASProp on grids that are not powers of two runs as a zero-padded linear convolution: the
// n-point circular convolution with h = IDFT_n(H) equals the linear convolution with h
// extended periodically over offsets |d| < n, which the padded (RS-geometry) plan computes.
// k_as_grid_d: H on the field's own n-point frequency grid [ny][nx] (FFT order, fftfreq for
// any n), FP64, scaled by 1/(nx ny) (the inverse DFT's normalisation).
__global__ void k_as_grid_d(double2* __restrict__ out, int nx, int ny, double dx, double dy, double lam, double n0,
                            double z, double tx, double ty, int mode, int track, const double2* __restrict__ filt,
                            int conjf) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const int kx = i < (nx + 1) / 2 ? i : i - nx;
  const int ky = j < (ny + 1) / 2 ? j : j - ny;
  double amp, t;
  as_transfer((double)kx / ((double)nx * dx), (double)ky / ((double)ny * dy), lam, n0, z, tx, ty, mode, track, amp, t);
  const double r = t - rint(t);
  double sn, cs;
  sincospi(2.0 * r, &sn, &cs);
  const double sc = 1.0 / ((double)nx * (double)ny);
  double re = amp * cs, im = amp * sn;
  apply_filter(filt, conjf, (long)j * nx + i, re, im);
  out[(long)j * nx + i] = make_double2(re * sc, im * sc);
}

// Direct inverse DFT along one axis (unnormalised, sign +), FP64: out[line][m] =
// sum_k in[line][k] e^{+2 pi i k m / n}, with (k m) mod n indexing an exact twiddle table.
// Lines are `nlines` apart by `lstride`, points by `estride`.
__global__ void k_idft_lines_d(const double2* __restrict__ in, double2* __restrict__ out, int n, int nlines,
                               long lstride, long estride, const double2* __restrict__ w) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int line = blockIdx.y;
  if (m >= n || line >= nlines) return;
  const double2* src = in + (long)line * lstride;
  double re = 0.0, im = 0.0;
  long km = 0;
  for (int k = 0; k < n; ++k) {
    const double2 a = src[(long)k * estride];
    const double2 t = w[km];
    re += a.x * t.x - a.y * t.y;
    im += a.x * t.y + a.y * t.x;
    km += m;
    if (km >= n) km -= n;
  }
  out[(long)line * lstride + (long)m * estride] = make_double2(re, im);
}

// e^{+2 pi i j / n}, j < n
__global__ void k_unit_roots_d(double2* __restrict__ w, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double sn, cs;
  sincospi(2.0 * (double)j / (double)n, &sn, &cs);
  w[j] = make_double2(cs, sn);
}

// h [ny][nx] extended periodically over offsets |d| < n onto the padded [ny2][nx2] grid
// (offsets in FFT order, i <= nx2/2 ? i : i - nx2); zero elsewhere.
__global__ void k_circ_embed_d(double2* __restrict__ out, const double2* __restrict__ h, int nx, int ny, int nx2,
                               int ny2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx2) return;
  const int ox = i <= nx2 / 2 ? i : i - nx2;
  const int oy = j <= ny2 / 2 ? j : j - ny2;
  double2 v = make_double2(0.0, 0.0);
  if (ox > -nx && ox < nx && oy > -ny && oy < ny) v = h[(long)(oy < 0 ? oy + ny : oy) * nx + (ox < 0 ? ox + nx : ox)];
  out[(long)j * nx2 + i] = v;
}

// Full FP64 transfer-function table [ny][nx] for the plain comparator (same formula).
__global__ void k_as_table_full_d(double2* __restrict__ tab, int nx, int ny, double dx, double dy,
                                  double lam, double n0, double z, double scale) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const int ki = i < (nx + 1) / 2 ? i : i - nx;
  const int kj = j < (ny + 1) / 2 ? j : j - ny;
  const double fx = (double)ki / ((double)nx * dx);
  const double fy = (double)kj / ((double)ny * dy);
  const double le = lam / n0;
  const double f2 = 1.0 / (le * le);
  const double a = f2 - fx * fx - fy * fy;
  double re, im;
  if (a >= 0.0) {
    const double t = z * sqrt(a);
    const double r = t - rint(t);
    sincospi(2.0 * r, &im, &re);
  } else {
    re = exp(-2.0 * M_PI * fabs(z) * sqrt(-a));
    im = 0.0;
  }
  tab[(long)j * nx + i] = make_double2(re * scale, im * scale);
}

// Rayleigh-Sommerfeld impulse response sampled on the (2nx)x(2ny) padded grid in FP64
// (rayleigh_sommerfeld.jl:1-12, nrm = dx*dy/(2*pi)).  Offsets are i for i <= nx and
// i - 2nx above, so the samples are even in both axes (offset +-nx is never read by a
// cropped output of the linear convolution).  The phase k*r is reduced in cycles:
// r/lam - rint(r/lam).
__global__ void k_rs_sample(double2* __restrict__ d, int nx2, int ny2, double dx, double dy,
                            double lam, double z, double nrm, double tx, double ty, int tilt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx2) return;
  // offsets in FFT order on the padded grid; the cropped convolution uses |offset| < n only
  const int oi = i <= nx2 / 2 ? i : i - nx2;
  const int oj = j <= ny2 / 2 ? j : j - ny2;
  const double x = oi * dx, y = oj * dy;
  const double az = fabs(z);
  const double r = sqrt(x * x + y * y + az * az);
  const double k = 2.0 * M_PI / lam;
  const double cyc = r / lam;
  double sn, cs;
  sincospi(2.0 * (cyc - rint(cyc)), &sn, &cs);
  // nrm * (e^{ikr}/r) * (z/r) * (1/r - i k)
  const double g = nrm * az / (r * r);
  const double ar = 1.0 / r, ai = -k;
  double re = g * (cs * ar - sn * ai);
  double im = g * (cs * ai + sn * ar);
  if (tilt) {
    // tilted field, tilts not tracked (rayleigh_sommerfeld.jl:14-22): times
    // cis(-2*pi*(x*f0x + y*f0y)), f0 = sin(theta)/lam, reduced in cycles
    const double c = x * (sin(tx) / lam) + y * (sin(ty) / lam);
    double ls, lc;
    sincospi(-2.0 * (c - rint(c)), &ls, &lc);
    const double r2 = re * lc - im * ls;
    im = re * ls + im * lc;
    re = r2;
  }
  if (z < 0) im = -im;  // h(z) = conj(h(-z))
  d[(long)j * nx2 + i] = make_double2(re, im);
}

// CollinsProp's convolution kernel (collins_integral.jl:17-21, sampled as ConvolutionKernel
// does, convolution_kernel.jl:13-18, on the padded grid with offsets i <= M/2 ? i : i - M):
// h = nrm * cis(pi*(ax*x^2 + ay*y^2)/(b*lam)), phase taken in cycles and reduced in FP64;
// nrm = dx*dy/(i*b) = prod(ds ./ sqrt(i*b)) (collins_integral.jl:122-123).
__global__ void k_collins_sample(double2* __restrict__ d, int nx2, int ny2, double dx, double dy, double lam,
                                 double ax, double ay, double b, double nrm_re, double nrm_im) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx2) return;
  const int oi = i <= nx2 / 2 ? i : i - nx2;
  const int oj = j <= ny2 / 2 ? j : j - ny2;
  const double x = oi * dx, y = oj * dy;
  const double cyc = (ax * x * x + ay * y * y) / (2.0 * b * lam);
  double sn, cs;
  sincospi(2.0 * (cyc - rint(cyc)), &sn, &cs);
  d[(long)j * nx2 + i] = make_double2(nrm_re * cs - nrm_im * sn, nrm_re * sn + nrm_im * cs);
}

// One line of a separable chirp on the centred grid x_i = (i - (n-1)/2)*ds:
// out[i] = amp * cis(2*pi*(q*x^2 + f*x)), the cycles reduced in FP64 before sincospi.
__global__ void k_chirp_line(float2* __restrict__ out, int n, double ds, double q, double f, double amp) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = ((double)i - 0.5 * (double)(n - 1)) * ds;
  const double t = q * x * x + f * x;
  double sn, cs;
  sincospi(2.0 * (t - rint(t)), &sn, &cs);
  out[i] = make_float2((float)(amp * cs), (float)(amp * sn));
}

// CZT parameters of one dimension (FFTutils.jl:30-50): a = rho_a e^{i theta_a},
// w = rho_w e^{i theta_w}; an untransformed dimension has on = 0.
struct CztDim {
  int on, center;
  double la, ta, lw, tw;  // log(rho_a), theta_a, log(rho_w), theta_w
};

// exp(re) * cis(im) with the angle reduced in FP64 cycles before sincospi.
__device__ inline double2 cexp_reduced(double re, double im) {
  const double t = im / (2.0 * M_PI);
  double sn, cs;
  sincospi(2.0 * (t - rint(t)), &sn, &cs);
  const double m = exp(re);
  return make_double2(m * cs, m * sn);
}

// The Bluestein kernel h[f] = w^{-f^2/2} of prepare_czt_data (FFTutils.jl:41), separable,
// sampled on the padded grid (a power of two >= 2n - 1) with offsets f = i <= M/2 ? i : i - M
// (the (2n-1)-grid values at |f| < n; other offsets are never reached by the cropped
// convolution); an untransformed
// dimension contributes a unit impulse at f = 0.
__device__ inline double2 czt_h(const CztDim& d, int f) {
  if (!d.on) return make_double2(f == 0 ? 1.0 : 0.0, 0.0);
  const double q = -0.5 * (double)f * (double)f;
  return cexp_reduced(q * d.lw, q * d.tw);
}

__global__ void k_czt_sample(double2* __restrict__ out, int nx, int ny, int nx2, int ny2, CztDim dx, CztDim dy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx2) return;
  const int oi = i <= nx2 / 2 ? i : i - nx2;
  const int oj = j <= ny2 / 2 ? j : j - ny2;
  if (oi <= -nx || oi >= nx || oj <= -ny || oj >= ny) {  // unused; |w| != 1 could overflow there
    out[(long)j * nx2 + i] = make_double2(0.0, 0.0);
    return;
  }
  const double2 a = czt_h(dx, oi), b = czt_h(dy, oj);
  out[(long)j * nx2 + i] = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Input (w^{k^2/2} a^{-k}) or output (w^{k^2/2}) chirp line, k = i or i - (n-1)/2 when centred
// on the grid (FFTutils.jl:33-40); 1 on an untransformed dimension.
__global__ void k_czt_chirp(float2* __restrict__ out, int n, CztDim d, int input) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!d.on) {
    out[i] = make_float2(1.f, 0.f);
    return;
  }
  const double k = d.center ? (double)i - 0.5 * (double)(n - 1) : (double)i;
  double re = 0.5 * k * k * d.lw, im = 0.5 * k * k * d.tw;
  if (input) {
    re -= k * d.la;
    im -= k * d.ta;
  }
  const double2 v = cexp_reduced(re, im);
  out[i] = make_float2((float)v.x, (float)v.y);
}

// Table of an FP64 spectrum [ny2][nx2] on the padded grid, scaled and rounded: the even
// quadrant [0..nx2/2]x[0..ny2/2] or the full layout (see tab_freq_index).
template <typename T>
__global__ void k_tab_extract(T* __restrict__ tab, const double2* __restrict__ d, int nx2, int ny2,
                              int rows, int qcol, int full, double scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= rows || j >= qcol) return;
  const int ky = (full && j > ny2 / 2) ? 3 * ny2 / 2 - j : j;
  const double2 v = d[(long)ky * nx2 + i];
  T r;
  r.x = v.x * scale;
  r.y = v.y * scale;
  tab[(long)i * qcol + j] = r;
}

// ================================================================ elementwise kernels

// dst(x, y) = src(x - x0, y - y0) where that lies inside src, else fill, per field: pad! with
// x0 = offset >= 0 and the pad value as fill (pad_crop_operators.jl:1-16), crop! with x0 = -offset
// (:61-79).  One thread per destination element (coalesced stores, row-contiguous loads).
__global__ void k_window(const float2* __restrict__ src, int snx, int sny, float2* __restrict__ dst, int dnx,
                         int dny, int batch, int x0, int y0, float2 fill) {
  const long dn = (long)dnx * dny;
  const long total = dn * batch;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long b = i / dn;
    const long r = i - b * dn;
    const int y = (int)(r / dnx), x = (int)(r - (long)y * dnx);
    const int sx = x - x0, sy = y - y0;
    dst[i] = (sx >= 0 && sx < snx && sy >= 0 && sy < sny) ? src[(b * sny + sy) * (long)snx + sx] : fill;
  }
}

// out = cis(s*phi) (the phase factors of phasemask.jl:90-113, once for every field of a batch)
__global__ void k_phase_cis(const float* __restrict__ phi, float s, float2* __restrict__ out, long npix) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    float sn, cs;
    sincosf(s * phi[i], &sn, &cs);
    out[i] = make_float2(cs, sn);
  }
}

// u <- u * cis(s*phi), saving u first when saved != nullptr (phasemask.jl:90-113).
__global__ void k_phase_apply(float2* __restrict__ u, const float* __restrict__ phi,
                              float2* __restrict__ saved, long npix, int batch, float s) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix;
       i += (long)gridDim.x * blockDim.x) {
    float sn, cs;
    sincosf(s * phi[i], &sn, &cs);
    const float2 w = make_float2(cs, sn);
    for (int b = 0; b < batch; ++b) {
      const long k = b * npix + i;
      const float2 v = u[k];
      if (saved) saved[k] = v;
      u[k] = cmul(v, w);
    }
  }
}

// du <- dv * cis(-s*phi) in place; dphi (+)= s * sum_b Im(du * conj(saved)), summed in
// field order (phasemask.jl:115-150).
__global__ void k_phase_backward(float2* __restrict__ dv, const float* __restrict__ phi,
                                 const float2* __restrict__ saved, float* __restrict__ dphi,
                                 long npix, int batch, float s, int acc) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix;
       i += (long)gridDim.x * blockDim.x) {
    float sn, cs;
    sincosf(-s * phi[i], &sn, &cs);
    const float2 w = make_float2(cs, sn);
    float g = 0.f;
    for (int b = 0; b < batch; ++b) {
      const long k = b * npix + i;
      const float2 du = cmul(dv[k], w);
      dv[k] = du;
      if (dphi) g += cmulc(du, saved[k]).y;
    }
    if (dphi) dphi[i] = (acc ? dphi[i] : 0.f) + s * g;
  }
}

// offset_tilts! (Fields.jl:280-293): u_b <- u_b * cis(2*pi*(cx_b*x + cy_b*y)) with
// cx_b = (sin(theta_x,from) - sin(theta_x,to)) / lambda_b (cycles per unit length) and the
// centred grid x = (i - (nx-1)/2)*dx (GridUtils.jl:56-60).  The phase is taken in cycles in
// FP64 and reduced to [-1/2, 1/2] before sincospi, so its error does not grow with the grid
// extent.  Fields are handled kTiltChunk at a time (coefficients passed by value).
constexpr int kTiltChunk = 32;
struct TiltCoef {
  double cx[kTiltChunk], cy[kTiltChunk];
};

__global__ void k_offset_tilts(float2* __restrict__ u, int nx, int ny, double dx, double dy, TiltCoef c,
                               int nb) {
  const long npix = (long)nx * ny;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % nx), iy = (int)(i / nx);
    const double x = ((double)ix - 0.5 * (double)(nx - 1)) * dx;
    const double y = ((double)iy - 0.5 * (double)(ny - 1)) * dy;
    for (int b = 0; b < nb; ++b) {
      const double t = c.cx[b] * x + c.cy[b] * y;
      const double r = t - rint(t);
      double sn, cs;
      sincospi(2.0 * r, &sn, &cs);
      const long k = b * npix + i;
      u[k] = cmul(u[k], make_float2((float)cs, (float)sn));
    }
  }
}

__global__ void k_scale_c64(float2* __restrict__ u, long n, float s) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float2 v = u[i];
    u[i] = make_float2(v.x * s, v.y * s);
  }
}

// Mask (modulators/mask.jl:95-141): u <- u * m (Forward) or u * conj(m) (Backward), m a c64
// [ny][nx] transmission shared by the batch, saving u first when saved != nullptr.
__global__ void k_mask_apply(float2* __restrict__ u, const float2* __restrict__ m, float2* __restrict__ saved,
                             long npix, int batch, int conj) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    const float2 w = m[i];
    for (int b = 0; b < batch; ++b) {
      const long k = b * npix + i;
      const float2 v = u[k];
      if (saved) saved[k] = v;
      u[k] = conj ? cmulc(v, w) : cmul(v, w);
    }
  }
}

// dm (+)= sum_b conj_direction(dv conj(saved)) from the cotangent at the mask output, then
// du = dv * conj_direction(m, reverse(direction)) in place (mask.jl:119-141)
__global__ void k_mask_backward(float2* __restrict__ dv, const float2* __restrict__ m,
                                const float2* __restrict__ saved, float2* __restrict__ dm, long npix,
                                int batch, int forward, int acc) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    const float2 w = m[i];
    float gr = 0.f, gi = 0.f;
    for (int b = 0; b < batch; ++b) {
      const long k = b * npix + i;
      const float2 d = dv[k];
      if (dm) {
        const float2 g = cmulc(d, saved[k]);  // dv * conj(saved)
        gr += g.x;
        gi += forward ? g.y : -g.y;
      }
      dv[k] = forward ? cmulc(d, w) : cmul(d, w);
    }
    if (dm) dm[i] = acc ? make_float2(dm[i].x + gr, dm[i].y + gi) : make_float2(gr, gi);
  }
}

// TeaDOE / TeaReflector (modulators/tea_doe.jl:146-199): per field b with kl_b = 2 pi dn(l_b)/l_b
// and complex reflectivity r_b, u_b <- u_b * r_b * cis(kl_b h) (Forward) or
// conj(r_b) * cis(-kl_b h) (Backward); h is an FP32 [ny][nx] height map.
__global__ void k_tea_apply(float2* __restrict__ u, const float* __restrict__ h, float2* __restrict__ saved,
                            long npix, int batch, const double* __restrict__ kl, const double2* __restrict__ r,
                            float s) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    const float hi = h[i];
    for (int b = 0; b < batch; ++b) {
      const long k = b * npix + i;
      float sn, cs;
      sincosf(s * ((float)kl[b] * hi), &sn, &cs);
      float2 w = make_float2(cs, sn);
      if (r) w = cmul(w, make_float2((float)r[b].x, s * (float)r[b].y));
      const float2 v = u[k];
      if (saved) saved[k] = v;
      u[k] = cmul(v, w);
    }
  }
}

// du = dv * conj(t) in place, then dh (+)= sum_b s kl_b Im(du conj(saved)) (tea_doe.jl:180-199)
__global__ void k_tea_backward(float2* __restrict__ dv, const float* __restrict__ h,
                               const float2* __restrict__ saved, float* __restrict__ dh, long npix, int batch,
                               const double* __restrict__ kl, const double2* __restrict__ r, float s, int acc) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    const float hi = h[i];
    float g = 0.f;
    for (int b = 0; b < batch; ++b) {
      const long k = b * npix + i;
      float sn, cs;
      sincosf(-s * ((float)kl[b] * hi), &sn, &cs);
      float2 w = make_float2(cs, sn);
      if (r) w = cmul(w, make_float2((float)r[b].x, -s * (float)r[b].y));
      const float2 du = cmul(dv[k], w);
      dv[k] = du;
      if (dh) g += (float)kl[b] * cmulc(du, saved[k]).y;
    }
    if (dh) dh[i] = (acc ? dh[i] : 0.f) + s * g;
  }
}

// Optimiser step with its proximal operator in one pass over the parameter (n FP32 values;
// complex parameters are their interleaved real and imaginary parts).  Rules: 0 Descent
// x -= eta dx; 1 Momentum m = rho m + eta dx, x -= m; 2 Nesterov d = -rho^2 v + (1+rho) eta dx,
// v = rho v - eta dx, x -= d (Optimisers.jl 0.4.6); 3 Fista (OptimisersExt.jl:31-40) with
// eta already squared: d = eta dx - beta (x - xk), xk = x, x -= d.  Prox
// (proximal_operators/*.jl): 0 none, 1 clamp to [p0, p1], 2 soft threshold by p0 around p1,
// 3 negatives to zero.
// beta_d (nullable): Fista's beta from device memory (k_fista_coef) in place of the argument
__global__ void k_optim_step(float* __restrict__ x, const float* __restrict__ dx, float* __restrict__ st,
                             long n, int rule, float eta, float rho, float beta, int prox, float p0, float p1,
                             const double* __restrict__ beta_d) {
  if (beta_d) beta = (float)*beta_d;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    float xi = x[i];
    const float g = dx[i];
    if (rule == 0) {
      xi -= eta * g;
    } else if (rule == 1) {
      const float m = rho * st[i] + eta * g;
      st[i] = m;
      xi -= m;
    } else if (rule == 2) {
      const float v = st[i];
      const float d = -rho * rho * v + (1.f + rho) * eta * g;
      st[i] = rho * v - eta * g;
      xi -= d;
    } else {
      const float d = eta * g - beta * (xi - st[i]);
      st[i] = xi;
      xi -= d;
    }
    if (prox == 1) {
      xi = xi < p0 ? p0 : (xi > p1 ? p1 : xi);
    } else if (prox == 2) {
      const float xc = xi - p1;
      xi = fabsf(xc) <= p0 ? p1 : (xc > p0 ? xi - p0 : xi + p0);
    } else if (prox == 3) {
      xi = xi < 0.f ? 0.f : xi;
    }
    x[i] = xi;
  }
}

// Fista's coefficients in device memory (OptimisersExt.jl:12-40), so that a recorded step
// replays correctly: tb = {t_k, beta}; beta <- (t_k - 1) / t_{k+1}, t <- t_{k+1} =
// (1 + sqrt(1 + 4 t_k^2)) / 2.  One thread.
__global__ void k_fista_coef(double* tb) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double t = tb[0];
    const double tn = (1.0 + sqrt(1.0 + 4.0 * t * t)) / 2.0;
    tb[1] = (t - 1.0) / tn;
    tb[0] = tn;
  }
}

// Adam (Optimisers.jl 0.4.6, the reference's dependency): m = b1 m + (1 - b1) g,
// v = b2 v + (1 - b2) g^2 (stored in FP32), x -= eta * m / (1 - bt1) / (sqrt(v / (1 - bt2)) + eps),
// each formed in FP64 from the FP32 values as Julia's Float64-scalar broadcasts do; bt = {b1^t,
// b2^t} of the current step, in device memory so that a recorded step replays correctly
// (k_adam_advance moves it on).  state: m in [0, n), v in [n, 2n).  Then the proximal operator.
__global__ void k_adam_step(float* __restrict__ x, const float* __restrict__ dx, float* __restrict__ st, long n,
                            double eta, double b1, double b2, double eps, const double* __restrict__ bt, int prox,
                            float p0, float p1) {
  const double bt1 = bt[0], bt2 = bt[1];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double g = dx[i];
    const float m = (float)(b1 * (double)st[i] + (1.0 - b1) * g);
    const float v = (float)(b2 * (double)st[n + i] + (1.0 - b2) * (g * g));
    st[i] = m;
    st[n + i] = v;
    const double d = (double)m / (1.0 - bt1) / (sqrt((double)v / (1.0 - bt2)) + eps) * eta;
    float xi = (float)((double)x[i] - d);
    if (prox == 1) {
      xi = xi < p0 ? p0 : (xi > p1 ? p1 : xi);
    } else if (prox == 2) {
      const float xc = xi - p1;
      xi = fabsf(xc) <= p0 ? p1 : (xc > p0 ? xi - p0 : xi + p0);
    } else if (prox == 3) {
      xi = xi < 0.f ? 0.f : xi;
    }
    x[i] = xi;
  }
}

__global__ void k_adam_advance(double* bt, double b1, double b2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    bt[0] *= b1;
    bt[1] *= b2;
  }
}

// ================================================================ metrics
// SquaredIntensityDifference (squared_intensity_difference.jl:78-90): d = |u_b|^2 - I_b per
// field ("full dimensions") or d = sum_b |u_b|^2 - I ("spatial only"); loss = sum d^2 and the
// pullback du_b = 4 dc d u_b.  d and d^2 are formed in FP64 from the FP32 data; the sum is a
// fixed-shape tree (block partials, then one ordered reduction), so repeated calls agree bit
// for bit.
constexpr int kMetricThreads = 256;
constexpr int kMetricBlocks = 256;  // partials per output slot

__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  return s;  // valid on thread 0
}

__global__ void __launch_bounds__(kMetricThreads)
k_sqint_partial(const float2* __restrict__ u, const float* __restrict__ I, long npix, int batch,
                int spatial, const double* __restrict__ dc, float2* __restrict__ du,
                double* __restrict__ part) {
  __shared__ double red[kMetricThreads / 64];
  const int slot = blockIdx.y;
  const double g = dc ? 4.0 * dc[slot] : 4.0;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)kMetricThreads + threadIdx.x; i < npix;
       i += (long)gridDim.x * kMetricThreads) {
    if (spatial) {
      double s = 0.0;
      for (int b = 0; b < batch; ++b) {
        const float2 v = u[b * npix + i];
        s += (double)v.x * v.x + (double)v.y * v.y;
      }
      const double d = s - (double)I[i];
      acc += d * d;
      if (du) {
        const float f = (float)(g * d);
        for (int b = 0; b < batch; ++b) {
          const float2 v = u[b * npix + i];
          du[b * npix + i] = make_float2(f * v.x, f * v.y);
        }
      }
    } else {
      const long k = slot * npix + i;
      const float2 v = u[k];
      const double d = (double)v.x * v.x + (double)v.y * v.y - (double)I[k];
      acc += d * d;
      if (du) {
        const float f = (float)(g * d);
        du[k] = make_float2(f * v.x, f * v.y);
      }
    }
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) part[slot * (long)gridDim.x + blockIdx.x] = s;
}

// SquaredFieldDifference (squared_field_difference.jl): loss[b] = sum |u_b - v_b|^2 (formed in
// FP64), pullback du_b = 2 dc[b] (u_b - v_b); the same fixed-shape sum as k_sqint_partial.
__global__ void __launch_bounds__(kMetricThreads)
k_sqfield_partial(const float2* __restrict__ u, const float2* __restrict__ v, long npix,
                  const double* __restrict__ dc, float2* __restrict__ du, double* __restrict__ part) {
  __shared__ double red[kMetricThreads / 64];
  const int slot = blockIdx.y;
  const double g = dc ? 2.0 * dc[slot] : 2.0;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)kMetricThreads + threadIdx.x; i < npix;
       i += (long)gridDim.x * kMetricThreads) {
    const long k = slot * npix + i;
    const float2 a = u[k], b = v[k];
    const double dx = (double)a.x - b.x, dy = (double)a.y - b.y;
    acc += dx * dx + dy * dy;
    if (du) du[k] = make_float2((float)(g * dx), (float)(g * dy));
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) part[slot * (long)gridDim.x + blockIdx.x] = s;
}

// dL/dz of an on-axis ASProp with a trainable distance (ASPropZ, angular_spectrum.jl:236-300):
// partials of Re(conj(DV) * 2 pi i w~ * V) over one field's spectrum, w = sqrt((n0/lam)^2 - f^2)
// (principal root), w~ = conj(w) for z < 0 (H(z) = conj(H(-z))), all in FP64.
__global__ void __launch_bounds__(kMetricThreads)
k_as_zgrad_partial(const float2* __restrict__ V, const float2* __restrict__ DV, int nx, int ny, double dx,
                   double dy, double kk, int zneg, double* __restrict__ part) {
  __shared__ double red[kMetricThreads / 64];
  const long npix = (long)nx * ny;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)kMetricThreads + threadIdx.x; i < npix; i += (long)gridDim.x * kMetricThreads) {
    const int y = (int)(i / nx), x = (int)(i - (long)y * nx);
    const double fx = (x <= (nx - 1) / 2 ? x : x - nx) / (nx * dx);
    const double fy = (y <= (ny - 1) / 2 ? y : y - ny) / (ny * dy);
    const double a = kk - fx * fx - fy * fy;  // (n0/lam)^2 - f^2
    const float2 v = V[i], d = DV[i];
    // P = conj(DV) V
    const double pr = (double)d.x * v.x + (double)d.y * v.y, pi = (double)d.x * v.y - (double)d.y * v.x;
    if (a >= 0.0) {
      acc += -2.0 * M_PI * sqrt(a) * pi;            // Re(P * 2 pi i r), r = w real
    } else {
      const double s = zneg ? -sqrt(-a) : sqrt(-a);  // w~ = i s
      acc += -2.0 * M_PI * s * pr;                  // Re(P * 2 pi i * i s)
    }
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s / (double)npix;  // Parseval's 1/N
}

// loss[slot] = ordered sum of the slot's partials (one workgroup per slot)
__global__ void __launch_bounds__(kMetricThreads)
k_sum_partials(const double* __restrict__ part, int nper, double* __restrict__ out) {
  __shared__ double red[kMetricThreads / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nper; i += kMetricThreads) acc += part[blockIdx.x * (long)nper + i];
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

}  // namespace fh

// Block partials of the no-allocation metric entry point (fh_sq_intensity_diff_noalloc): a
// fixed device array, one copy per GPU held until the process ends, with room for
// kNoAllocSlots loss slots of kMetricBlocks partials each (2 MiB).  Host code never passes this
// name to a kernel (on the host it is only a stand-in for the symbol): it uses the address that
// hipGetSymbolAddress returned for the device, kept in g_scratch_addr.
constexpr int kNoAllocSlots = 1024;
static __device__ __align__(256) double scratch[kNoAllocSlots * fh::kMetricBlocks];

namespace fh {

// The last pass of a propagation fused with SquaredIntensityDifference
// (squared_intensity_difference.jl:78-90) and the first pass of the adjoint propagation: per
// x line of the 4x4-blocked spectrum, IFFT -> u; d = |u|^2 - I (FP64), loss partial += d^2,
// du = 4 dc d u -> FFT -> stored back in place.  The inverse transform leaves its points in
// the forward transform's input order (its last radix is the forward's first), so the two
// transforms meet in registers.  HALF (zero-padded RS plans): points x >= nx of the IFFT are
// the discarded wrap of the linear convolution; du is zero there, as the adjoint's padding
// wants.  One FP64 partial per workgroup, summed in a fixed order afterwards.
// STG: the workgroup's 4-row groups of the blocked spectrum arrive as whole 128-byte lines (16-byte
// loads of its contiguous run, through the exchange areas, as in the staged x IFFT); the output
// side keeps the direct stores.
template <typename RDF, int L, bool HALF, bool STG = false>
__global__ void __launch_bounds__(L * RDF::P)
k_fast_sqint(FastIO io, const float2* __restrict__ tabi, const float2* __restrict__ tabf,
             const float* __restrict__ I, long tstride, int nx, const double* __restrict__ dc,
             double* __restrict__ part) {
  using RI = typename RDF::Rev;
  constexpr int n = RDF::n, V = RDF::V, P = RDF::P;
  constexpr int RL = LastRadix<RI>::value;
  constexpr int RE = LastRadix<RDF>::value;
  constexpr int LS = LdsLine<n, L>::size;
  constexpr int NCH = n / (2 * P);  // 16-byte chunks a thread (STG)
  static_assert(RL == RDF::R0, "the inverse transform must end in the forward transform's input order");
  static_assert(!STG || L % 4 == 0, "whole 4-row groups");
  __shared__ double red[(L * P + 63) / 64];
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  int l, t;
  thread_map<L, P, false>(threadIdx.x, l, t);
  const int b = blockIdx.y;
  const int li = blockIdx.x * L + l;
  const bool live = li < io.nlines;
  float2* s = smem + l * LdsLine<n, L>::size;
  const auto rin = make_rsrc(io.in + b * io.in_f, io.in_bytes);
  float2 v[V];
  float4 q[STG ? NCH : 1];
  if constexpr (STG) {
    const int vstg = (blockIdx.x * L * n + 2 * (int)threadIdx.x) * 8;
#pragma unroll
    for (int k = 0; k < NCH; ++k) q[k] = buf_load4(rin, vstg, k * L * P * 16);
  } else {
    const int vin = fast_off<false>(li, t, io.in_l, io.in_e, io.in_blk) * 8;
    const int kin = fast_k<false>(io.in_l, io.in_e, io.in_blk) * 8;
#pragma unroll
    for (int m = 0; m < V / RI::R0; ++m)
#pragma unroll
      for (int r = 0; r < RI::R0; ++r)
        v[m * RI::R0 + r] = buf_load(rin, vin, (P * m + r * (n / RI::R0)) * kin);
  }
  // the target values are loaded up front, in flight during the inverse transform (the
  // inverse leaves point (m, r) at position t + P m + r n/RL, known now); the target row is
  // nx long, so positions past it (HALF) read a clamped, unused address
  const double g = dc ? 4.0 * dc[b] : 4.0;
  const float* Ib = I + (long)b * tstride + (long)(live ? li : 0) * nx;
  float tg[V];
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int x = t + P * m + r * (n / RL);
      tg[m * RL + r] = Ib[(!HALF || x < nx) ? x : 0];
    }
  if constexpr (STG) {
    // chunk k holds elements 2 (k L P + tid) and the next of the run: block row, block, then row
    // and column within the block
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const int e = 2 * (k * L * P + (int)threadIdx.x);
      const int blk = e >> 4;
      float2* d = smem + (4 * (blk / (n / 4)) + ((e >> 2) & 3)) * LS + lds_pos(4 * (blk % (n / 4)) + (e & 3));
      d[0] = make_float2(q[k].x, q[k].y);
      asm volatile("" ::: "memory");  // two accesses, not one instruction over adjacent slots
      d[1] = make_float2(q[k].z, q[k].w);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < V / RI::R0; ++m)
#pragma unroll
      for (int r = 0; r < RI::R0; ++r) v[m * RI::R0 + r] = s[lds_pos(t + P * m + r * (n / RI::R0))];
  }
  fast_fft<float2, RI, true>(v, s, t, tabi);
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int x = t + P * m + r * (n / RL);
      const float2 u = v[m * RL + r];
      float2 w = make_float2(0.f, 0.f);
      if (live && (!HALF || x < nx)) {
        const double d = (double)u.x * u.x + (double)u.y * u.y - (double)tg[m * RL + r];
        acc += d * d;
        const float f = (float)(g * d);
        w = make_float2(f * u.x, f * u.y);
      }
      v[m * RL + r] = w;
    }
  fast_fft<float2, RDF, false>(v, s, t, tabf);
  if (live) {
    const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
    const int vout = fast_off<false>(li, t, io.out_l, io.out_e, io.out_blk) * 8;
    const int kout = fast_k<false>(io.out_l, io.out_e, io.out_blk) * 8;
#pragma unroll
    for (int m = 0; m < V / RE; ++m)
#pragma unroll
      for (int r = 0; r < RE; ++r) buf_store(rout, vout, (P * m + r * (n / RE)) * kout, v[m * RE + r]);
  }
  const double sum = block_sum(acc, red);
  if (threadIdx.x == 0) part[b * (long)gridDim.x + blockIdx.x] = sum;
}

// DotProduct / PowerCoupling (metrics/dot_product.jl:60-95, power_coupling.jl:55-64), one
// value per mode: c_b = ds * sum u_b conj(v_b) in FP64 (block partials, then an ordered sum).
__global__ void __launch_bounds__(kMetricThreads)
k_dot_partial(const float2* __restrict__ u, const float2* __restrict__ v, long npix,
              double2* __restrict__ part) {
  __shared__ double red[kMetricThreads / 64];
  const int b = blockIdx.y;
  double ar = 0.0, ai = 0.0;
  for (long i = blockIdx.x * (long)kMetricThreads + threadIdx.x; i < npix; i += (long)gridDim.x * kMetricThreads) {
    const float2 x = u[b * npix + i], y = v[b * npix + i];
    ar += (double)x.x * y.x + (double)x.y * y.y;  // Re(x conj(y))
    ai += (double)x.y * y.x - (double)x.x * y.y;  // Im(x conj(y))
  }
  const double sr = block_sum(ar, red);
  __syncthreads();
  const double si = block_sum(ai, red);
  if (threadIdx.x == 0) part[b * (long)gridDim.x + blockIdx.x] = make_double2(sr, si);
}

// c[b] = ds * ordered sum of the partials; p[b] = |c[b]|^2 when p is not null
__global__ void __launch_bounds__(kMetricThreads)
k_dot_final(const double2* __restrict__ part, int nper, double ds, double2* __restrict__ c,
            double* __restrict__ p) {
  __shared__ double red[kMetricThreads / 64];
  double ar = 0.0, ai = 0.0;
  for (int i = threadIdx.x; i < nper; i += kMetricThreads) {
    const double2 x = part[blockIdx.x * (long)nper + i];
    ar += x.x;
    ai += x.y;
  }
  const double sr = block_sum(ar, red);
  __syncthreads();
  const double si = block_sum(ai, red);
  if (threadIdx.x == 0) {
    const double2 r = make_double2(sr * ds, si * ds);
    c[blockIdx.x] = r;
    if (p) p[blockIdx.x] = r.x * r.x + r.y * r.y;
  }
}

// du_b = v_b * g_b with g_b = dc_b (DotProduct, dot_product.jl:97-106) or 2 c_b dp_b
// (PowerCoupling, power_coupling.jl:66-69; dp null = ones)
__global__ void k_dot_pullback(const float2* __restrict__ v, const double2* __restrict__ dc,
                               const double2* __restrict__ c, const double* __restrict__ dp, long npix,
                               float2* __restrict__ du) {
  const int b = blockIdx.y;
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
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    const float2 y = v[b * npix + i];
    du[b * npix + i] = make_float2(y.x * fr - y.y * fi, y.x * fi + y.y * fr);
  }
}

// DotProduct with mode_selective=false (dot_product.jl:53-66): the overlap matrix
// c[i + n j] = ds * sum_x conj(v_i(x)) u_j(x) of the n modes (the reference's column-major c[i, j],
// V' U), as FP64 block partials of pixel tiles staged in LDS (each thread keeps up to four (i, j)
// sums), then an ordered sum per entry.  Tile rows are padded by one value so that the lanes that
// read one pixel of different modes hit different banks.
constexpr int kDotMaxModes = 32;
constexpr int kDotTile = 32;
__global__ void __launch_bounds__(256)
k_dotm_partial(const float2* __restrict__ u, const float2* __restrict__ v, long npix, int n,
               double2* __restrict__ part) {
  __shared__ float2 su[kDotMaxModes][kDotTile + 1], sv[kDotMaxModes][kDotTile + 1];
  const int t = threadIdx.x;
  const int npair = n * n;
  double ar[4] = {0.0, 0.0, 0.0, 0.0}, ai[4] = {0.0, 0.0, 0.0, 0.0};
  for (long base = (long)blockIdx.x * kDotTile; base < npix; base += (long)gridDim.x * kDotTile) {
    __syncthreads();  // the previous tile's readers are done
    for (int e = t; e < n * kDotTile; e += 256) {
      const int m = e / kDotTile, x = e - m * kDotTile;
      const long i = base + x;
      su[m][x] = i < npix ? u[m * npix + i] : make_float2(0.f, 0.f);
      sv[m][x] = i < npix ? v[m * npix + i] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = t + 256 * k;
      if (q < npair) {
        const int ii = q % n, jj = q / n;
        for (int x = 0; x < kDotTile; ++x) {
          const float2 a = sv[ii][x], b = su[jj][x];
          ar[k] += (double)a.x * b.x + (double)a.y * b.y;  // Re(conj(a) b)
          ai[k] += (double)a.x * b.y - (double)a.y * b.x;  // Im(conj(a) b)
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int q = t + 256 * k;
    if (q < npair) part[(long)blockIdx.x * npair + q] = make_double2(ar[k], ai[k]);
  }
}

// c[q] = ds * (ordered sum over the nblk block partials of entry q); p[q] = |c[q]|^2 when p is set
__global__ void __launch_bounds__(kMetricThreads)
k_dotm_final(const double2* __restrict__ part, int nblk, int npair, double ds, double2* __restrict__ c,
             double* __restrict__ p) {
  __shared__ double red[kMetricThreads / 64];
  const int q = blockIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kMetricThreads) {
    const double2 x = part[(long)b * npair + q];
    ar += x.x;
    ai += x.y;
  }
  const double sr = block_sum(ar, red);
  __syncthreads();
  const double si = block_sum(ai, red);
  if (threadIdx.x == 0) {
    const double2 r = make_double2(sr * ds, si * ds);
    c[q] = r;
    if (p) p[q] = r.x * r.x + r.y * r.y;
  }
}

// Pullback of the overlap matrix (dot_product.jl:85-97, power_coupling.jl:66-69): du_j = sum_i v_i
// G[i + n j] with G = dc (DotProduct) or 2 c o dp (PowerCoupling; dp null = ones), no ds factor as
// in the reference; one thread per pixel holds the n outputs.
__global__ void __launch_bounds__(256)
k_dotm_pullback(const float2* __restrict__ v, const double2* __restrict__ dc, const double2* __restrict__ c,
                const double* __restrict__ dp, long npix, int n, float2* __restrict__ du) {
  __shared__ float2 G[kDotMaxModes * kDotMaxModes];
  for (int q = threadIdx.x; q < n * n; q += 256) {
    double gr, gi;
    if (dc) {
      gr = dc[q].x;
      gi = dc[q].y;
    } else {
      const double w = 2.0 * (dp ? dp[q] : 1.0);
      gr = w * c[q].x;
      gi = w * c[q].y;
    }
    G[q] = make_float2((float)gr, (float)gi);
  }
  __syncthreads();
  for (long x = blockIdx.x * 256L + threadIdx.x; x < npix; x += (long)gridDim.x * 256) {
    float2 acc[kDotMaxModes];
#pragma unroll
    for (int j = 0; j < kDotMaxModes; ++j) acc[j] = make_float2(0.f, 0.f);
    for (int i = 0; i < n; ++i) {
      const float2 a = v[i * npix + x];
#pragma unroll
      for (int j = 0; j < kDotMaxModes; ++j)
        if (j < n) {
          const float2 g = G[i + n * j];
          acc[j].x += a.x * g.x - a.y * g.y;
          acc[j].y += a.x * g.y + a.y * g.x;
        }
    }
#pragma unroll
    for (int j = 0; j < kDotMaxModes; ++j)
      if (j < n) du[j * npix + x] = acc[j];
  }
}

__global__ void k_mul_full_d(double2* __restrict__ u, const double2* __restrict__ tab, long npix,
                             int conj) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix;
       i += (long)gridDim.x * blockDim.x) {
    const double2 h = tab[i];
    u[i] = conj ? cmulc(u[i], h) : cmul(u[i], h);
  }
}

// ================================================================ fused x-IFFT + Phase adjoint + gradient
// x IFFT of the 4x4-blocked spectrum into natural rows, u <- u * cis(-s*phi) (un-applying
// the Phase, phasemask.jl:115-150) and dphi (+)= s * sum_b Im(du_b * conj(saved_b)).  One
// workgroup owns L rows and loops over the fields of the batch, so the mask gradient is
// summed in registers, in field order, and written once; the phase factors are taken once
// per position for the whole batch and the next field's line is loaded while this one is
// transformed.  Rows of RS plans hold 2nx points of which the first nx are kept.
// STG: the workgroup's 4-row groups of the blocked spectrum arrive as whole 128-byte lines (16-byte
// loads of its contiguous run, as in the staged x passes) and go to the exchange areas before the
// transform.  store == 0: du is formed for the gradient but not written (a caller that does not need
// the cotangent at the Phase's input).
// Where they fit beside the exchange areas (160 KB per CU), the phase factors are taken once per
// position and kept in LDS, one slot per (value, thread), read back only by the thread that wrote
// them (no barrier): that frees their registers and the per-field sincos; otherwise the mask
// values stay in registers and cis(-s*phi) is re-taken per field.
template <int n, int L, int P, int V> struct PhaseGradLds {
  static constexpr size_t exch = (size_t)L * LdsLine<n, L>::size * sizeof(float2);
  static constexpr size_t cis = (size_t)L * P * V * sizeof(float2);
  static constexpr bool in_lds = exch + cis <= 160 * 1024;
  static constexpr size_t bytes = exch + (in_lds ? cis : 0);
};

template <typename RDF, int L, bool GRAD, bool STG = false>
__global__ void __launch_bounds__(L * RDF::P)
k_fast_phase_grad(FastIO io, int batch, const float2* __restrict__ tab, const float* __restrict__ phi,
                  const float2* __restrict__ saved, float* __restrict__ dphi, int nx, float s, int acc, int store) {
  using RI = typename RDF::Rev;
  constexpr int n = RI::n, V = RI::V, P = RI::P;
  constexpr int RL = LastRadix<RI>::value;
  constexpr int LS = LdsLine<n, L>::size;
  constexpr int NCH = n / (2 * P);  // 16-byte chunks a thread (STG)
  static_assert(!STG || L % 4 == 0, "whole 4-row groups");
  extern __shared__ unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  int l, t;
  thread_map<L, P, false>(threadIdx.x, l, t);
  const int li = blockIdx.x * L + l;
  const bool live = li < io.nlines;
  float2* sl = smem + l * LdsLine<n, L>::size;
  const int vin = fast_off<false>(li, t, io.in_l, io.in_e, io.in_blk) * 8;
  const int kin = fast_k<false>(io.in_l, io.in_e, io.in_blk) * 8;
  const int vout = fast_off<false>(li, t, io.out_l, io.out_e, io.out_blk) * 8;
  const int kout = fast_k<false>(io.out_l, io.out_e, io.out_blk) * 8;
  using PL = PhaseGradLds<n, L, P, V>;
  constexpr int LP = L * P;
  float2* csl = smem + L * LdsLine<n, L>::size + threadIdx.x;  // the LDS factors (PL::in_lds)
  float ph[V], g[V];
#pragma unroll
  for (int m = 0; m < V / RL; ++m)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int x = t + P * m + r * (n / RL);
      const float a = (live && x < nx) ? -s * phi[(long)li * nx + x] : 0.f;
      if constexpr (PL::in_lds) {
        float sn, cs;
        sincosf(a, &sn, &cs);
        csl[(m * RL + r) * LP] = make_float2(cs, sn);
      } else {
        ph[m * RL + r] = a;
      }
      g[m * RL + r] = 0.f;
    }
  // the next field's spectrum, and this field's saved input, are in flight while this field
  // is transformed; the load past the last field re-reads it (unconditional, so that no load
  // sits under a branch)
  float2 nxt[STG ? 1 : V];
  float4 q[STG ? NCH : 1];
  const int vstg = (blockIdx.x * L * n + 2 * (int)threadIdx.x) * 8;  // STG: the thread's first chunk
  const auto load_spec = [&](int b) {
    const auto rin = make_rsrc(io.in + b * io.in_f, io.in_bytes);
    if constexpr (STG) {
#pragma unroll
      for (int k = 0; k < NCH; ++k) q[k] = buf_load4(rin, vstg, k * L * P * 16);
    } else {
#pragma unroll
      for (int m = 0; m < V / RI::R0; ++m)
#pragma unroll
        for (int r = 0; r < RI::R0; ++r) nxt[m * RI::R0 + r] = buf_load(rin, vin, (P * m + r * (n / RI::R0)) * kin);
    }
  };
  load_spec(0);
  for (int b = 0; b < batch; ++b) {
    float2 v[V], sv[V];
    if constexpr (STG) {
      // chunk k holds elements 2 (k L P + tid) and the next of the run: block row, then block,
      // then row and column within the block (the previous field's last exchange is read first)
      __syncthreads();
#pragma unroll
      for (int k = 0; k < NCH; ++k) {
        const int e = 2 * (k * L * P + (int)threadIdx.x);
        const int blk = e >> 4;
        float2* d = smem + (4 * (blk / (n / 4)) + ((e >> 2) & 3)) * LS + lds_pos(4 * (blk % (n / 4)) + (e & 3));
        d[0] = make_float2(q[k].x, q[k].y);
        asm volatile("" ::: "memory");  // two accesses, not one instruction over adjacent slots
        d[1] = make_float2(q[k].z, q[k].w);
      }
    } else {
#pragma unroll
      for (int k = 0; k < V; ++k) v[k] = nxt[k];
    }
    if constexpr (GRAD) {
      const auto rsv = make_rsrc(saved + b * io.out_f, io.out_bytes);
#pragma unroll
      for (int m = 0; m < V / RL; ++m)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int c = P * m + r * (n / RL);
          sv[m * RL + r] = buf_load(rsv, vout, (t + c < nx ? c : 0) * kout);
        }
    }
    load_spec(b + 1 < batch ? b + 1 : b);
    if constexpr (STG) {
      __syncthreads();
#pragma unroll
      for (int m = 0; m < V / RI::R0; ++m)
#pragma unroll
        for (int r = 0; r < RI::R0; ++r) v[m * RI::R0 + r] = sl[lds_pos(t + P * m + r * (n / RI::R0))];
    }
    fast_fft<float2, RI, true>(v, sl, t, tab);
    // loads and stores share one counter: wait while only old operations are outstanding
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    if (live) {
      const auto rout = make_rsrc(io.out + b * io.out_f, io.out_bytes);
#pragma unroll
      for (int m = 0; m < V / RL; ++m)
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int c = P * m + r * (n / RL);
          if (t + c < nx) {
            float2 f;
            if constexpr (PL::in_lds) {
              f = csl[(m * RL + r) * LP];
            } else {
              float sn, cs;
              sincosf(ph[m * RL + r], &sn, &cs);
              f = make_float2(cs, sn);
            }
            const float2 du = cmul(v[m * RL + r], f);
            if constexpr (GRAD) g[m * RL + r] += cmulc(du, sv[m * RL + r]).y;
            if (!GRAD || store) buf_store(rout, vout, c * kout, du);
          }
        }
    }
  }
  if (GRAD && live) {
#pragma unroll
    for (int m = 0; m < V / RL; ++m)
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const int x = t + P * m + r * (n / RL);
        if (x < nx) {
          const long k = (long)li * nx + x;
          dphi[k] = (acc ? dphi[k] : 0.f) + s * g[m * RL + r];
        }
      }
  }
}

// ================================================================ host objects

static int ilog2_exact(int n) {
  if (n <= 0 || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// smallest l with 2^l >= n (n >= 1)
static int ilog2_ceil(int n) {
  int l = 0;
  while ((1 << l) < n && l < 30) ++l;
  return l;
}

}  // namespace fh

using namespace fh;

struct fh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<std::pair<int, int>, void*> twiddles;  // (n, is_double) -> W_n^k table
  std::map<std::vector<int>, void*> stage_tw;      // (n, R0, R1, R2, inverse) -> stage tables
  std::mutex mu;
  double* metric_part = nullptr;  // metric block partials
  size_t metric_part_n = 0;
  // partials buffers outgrown by a later call: graphs recorded earlier still hold their
  // addresses, so they stay allocated until the context is destroyed
  std::vector<double*> metric_part_retired;
  std::map<std::vector<int>, fh_plan*> dft_plans;  // (nx, ny, batch) -> DFT-as-CZT plan (fh_fft2)
};

struct fh_plan {
  fh_ctx* ctx = nullptr;
  int kind = 0;  // 0 = AS, 1 = RS, 2 = Collins (RS geometry: zero padding to 2n)
  int nx = 0, ny = 0, batch = 0;
  int lx = 0, ly = 0;   // log2 of the transform lengths (padded for RS)
  double dx = 0, dy = 0, z = 0, n0 = 1;
  std::vector<double> lam_tables;  // per table: wavelength ...
  std::vector<double> tab_z;       // ... distance ...
  std::vector<double> tab_tx, tab_ty;  // ... and tilt angles (radians)
  std::vector<int> field_table;    // table index per field
  int mode = 0;                    // AS kernel: 0 on-axis, 1 paraxial, 2 tilted
  int track = 0;                   // tilts tracked
  int full = 0;                    // full (non-even) table layout
  int filtered = 0;                // ASProp tables include a frequency filter
  int oned = 0;                    // 1-D fields (ny == 1): lines of nx points, one pass per propagation
  float2* tab = nullptr;           // transfer-function tables, one per distinct key
  long tabsize = 0;
  int qx = 0, qy = 0;
  int* d_tabidx = nullptr;
  float2* scratch = nullptr;       // RS: (2nx) x ny x batch
  double2* plain_tab = nullptr;    // plain FP64 comparator (AS): full tables
  double2* tab_d = nullptr;        // padded AS, on request: the tables before rounding to c64
  float2* cis = nullptr;           // Phase factors cis(s*phi) of the fused Phase pass, [ny][nx]
  int conv_var = -1;               // y-pass variant (-1 = built-in default)
  int xf_var = -1, xi_var = -1;    // x-pass variants (forward, inverse)
  float2* chirp = nullptr;         // Collins: A_x, A_y, D_x, D_y chirp lines per table
  double alpha_x = 1, alpha_y = 1; // Collins: output / input sampling
};

namespace fh {

static int set_device(fh_ctx* c) {
  FH_HIP(hipSetDevice(c->device));
  return FH_OK;
}

// W_n^k = exp(-2*pi*i*k/n), k < n, from long-double sin/cos of the exactly reduced
// angle, rounded once to the table precision.
template <typename C>
static int get_twiddles(fh_ctx* c, int n, const C** out) {
  const int isd = sizeof(C) == sizeof(double2);
  std::lock_guard<std::mutex> lk(c->mu);
  auto key = std::make_pair(n, isd);
  auto it = c->twiddles.find(key);
  if (it != c->twiddles.end()) {
    *out = reinterpret_cast<const C*>(it->second);
    return FH_OK;
  }
  std::vector<C> h(n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n; ++k) {
    const long double a = two_pi * (long double)k / (long double)n;
    h[k].x = (decltype(h[k].x))cosl(a);
    h[k].y = (decltype(h[k].y))(-sinl(a));
  }
  void* d = nullptr;
  if (hipMalloc(&d, n * sizeof(C)) != hipSuccess) return fail(FH_E_NOMEM, "twiddle table: out of device memory");
  FH_HIP(hipMemcpy(d, h.data(), n * sizeof(C), hipMemcpyHostToDevice));
  c->twiddles[key] = d;
  *out = reinterpret_cast<const C*>(d);
  return FH_OK;
}

// Stage twiddle tables of the fast passes for radix plan (R0, R1, R2) of length n:
// stage 1 [jm < R0][r-1 < R1-1] = W_n^{s*jm*r*n/(R0*R1)}, then stage 2
// [jm < R0*R1][r-1 < R2-1] = W_n^{s*jm*r}, s = -1 forward / +1 inverse, each from a
// long-double sin/cos of the exactly reduced angle, rounded once.
static int get_stage_tw(fh_ctx* c, int n, int R0, int R1, int R2, bool inv, const float2** out) {
  std::lock_guard<std::mutex> lk(c->mu);
  std::vector<int> key{n, R0, R1, R2, inv ? 1 : 0};
  auto it = c->stage_tw.find(key);
  if (it != c->stage_tw.end()) {
    *out = reinterpret_cast<const float2*>(it->second);
    return FH_OK;
  }
  std::vector<float2> h;
  const long double two_pi = 6.283185307179586476925286766559005768L;
  auto push = [&](long k) {
    k %= n;
    const long double a = two_pi * (long double)k / (long double)n;
    h.push_back(make_float2((float)cosl(a), (float)(inv ? sinl(a) : -sinl(a))));
  };
  if (R1 > 1)
    for (int jm = 0; jm < R0; ++jm)
      for (int r = 1; r < R1; ++r) push((long)jm * r * (n / (R0 * R1)));
  if (R2 > 1)
    for (int jm = 0; jm < R0 * R1; ++jm)
      for (int r = 1; r < R2; ++r) push((long)jm * r);
  if (h.empty()) h.push_back(make_float2(1.f, 0.f));
  void* d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(float2)) != hipSuccess) return fail(FH_E_NOMEM, "stage twiddles: out of device memory");
  FH_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice));
  c->stage_tw[key] = d;
  *out = reinterpret_cast<const float2*>(d);
  return FH_OK;
}

// ---------------------------------------------------------------- fast c64 launches

// x-pass variants: 0 = lanes over positions of a line, 1 = lanes over the workgroup's lines,
// 2 = the blocked side staged through LDS (whole 4-row groups of unpadded rows; otherwise 0)
template <int LOGN, bool INV, bool IH, bool OH, typename Op, int L>
static int fast_x_l(fh_ctx* c, const FastIO& io, int batch, Op op, int xvar);

template <int LOGN, bool INV, bool IH, bool OH, typename Op>
static int fast_x_t(fh_ctx* c, const FastIO& io, int batch, Op op, int xvar) {
  return fast_x_l<LOGN, INV, IH, OH, Op, Cfg<float, LOGN, false>::L>(c, io, batch, op, xvar);
}

// Lines per workgroup halve (down to one full wave per workgroup) while a launch would have
// fewer than kFillWGs workgroups: single small fields otherwise leave most CUs idle.  An x
// pass converts between the natural and the 4x4-blocked layout in place, which is only safe
// when every workgroup owns whole 4-row groups, so L stays a multiple of 4.
// Plans with 8 points a thread for rows of 512 and 1024 points, taken when even 4 rows per
// workgroup leave fewer than kFillWGs workgroups (a 250 x 250 RS field has 63 4-row groups per
// field): the 32-point plans then run one wave per workgroup with four times the serial work.
template <int LOGN> struct FewRowsPlan { using RD = void; };
template <> struct FewRowsPlan<9> { using RD = Radices<8, 8, 8>; };
template <> struct FewRowsPlan<10> { using RD = Radices<16, 8, 8>; };

template <typename RDF, int L, bool INV, bool IH, bool OH, typename Op>
static int fast_x_launch(fh_ctx* c, const FastIO& io, int batch, Op op, int xvar);

// whether an x pass of 4-row workgroups takes the staged form: the blocked side (input of an
// inverse pass, output of a forward one) one contiguous run of whole 4-row groups of n-point rows;
// the natural side may hold fewer points a row (RS plans: IH of a forward pass, OH of an inverse one)
template <typename RDF, bool INV, bool IH, bool OH>
static bool x_staged(const FastIO& io, int xvar) {
  using RD = typename std::conditional<INV, typename RDF::Rev, RDF>::type;
  if ((IH && INV) || (OH && !INV)) return false;
  const bool blk_side = INV ? (io.in_blk && !io.out_blk && io.in_l == RD::n) : (io.out_blk && !io.in_blk && io.out_l == RD::n);
  const bool full = IH || OH || (INV ? io.out_valid : io.in_valid) >= RD::n;
  return xvar == 2 && blk_side && full && io.nlines % 4 == 0;
}

template <int LOGN, bool INV, bool IH, bool OH, typename Op, int L>
static int fast_x_l(fh_ctx* c, const FastIO& io, int batch, Op op, int xvar) {
  using RDF = typename Cfg<float, LOGN, false>::RD;
  const bool few = (long)((io.nlines + L - 1) / L) * batch < kFillWGs;
  if constexpr (L >= 8 && (L / 2) * RDF::P >= 64) {
    if (few) return fast_x_l<LOGN, INV, IH, OH, Op, L / 2>(c, io, batch, op, xvar);
  }
  using FR = typename FewRowsPlan<LOGN>::RD;
  if constexpr (!std::is_void<FR>::value && L == 4) {
    if (few) return fast_x_launch<FR, 4, INV, IH, OH, Op>(c, io, batch, op, xvar);
  }
  // Out-of-place passes (RS, Collins and CZT plans: natural rows <-> the scratch spectrum, or a
  // source read into a separate field) need no whole 4-row groups.  Rows of 4096 points then go
  // two to a workgroup: a 4-row workgroup takes 139 KB of LDS, one per CU, which then alternates
  // between loading, transforming and storing; two 2-row workgroups share a CU and overlap.
  // The staged form keeps whole 4-row groups (one 16-wave workgroup per CU).
  if constexpr (L == 4 && LOGN == 12) {
    if (static_cast<const void*>(io.in) != static_cast<const void*>(io.out) && !x_staged<RDF, INV, IH, OH>(io, xvar))
      return fast_x_launch<RDF, 2, INV, IH, OH, Op>(c, io, batch, op, xvar);
  }
  return fast_x_launch<RDF, L, INV, IH, OH, Op>(c, io, batch, op, xvar);
}

template <typename RDF, int L, bool INV, bool IH, bool OH, typename Op>
static int fast_x_launch(fh_ctx* c, const FastIO& io, int batch, Op op, int xvar) {
  using RD = typename std::conditional<INV, typename RDF::Rev, RDF>::type;
  const float2* tab = nullptr;
  int rc = get_stage_tw(c, RD::n, RD::R0, RD::R1, RD::R2, INV, &tab);
  if (rc) return rc;
  if constexpr (L % 4 == 0 && !(IH && INV) && !(OH && !INV)) {
    if (x_staged<RDF, INV, IH, OH>(io, xvar) && io.nlines % L == 0) {
      auto kern = k_fast_fft_stg<RDF, L, INV, Op, IH || OH>;
      const size_t lds = (size_t)L * LdsLine<RD::n, L>::size * sizeof(float2);
      FH_HIP(ensure_lds(kern, lds));
      hipLaunchKernelGGL(kern, dim3(io.nlines / L, batch), dim3(L * RDF::P), lds, c->stream, io, tab, op);
      FH_HIP(hipGetLastError());
      return FH_OK;
    }
  }
  const bool lanes = xvar == 1;
  auto kern = lanes ? k_fast_fft<RDF, L, false, INV, IH, OH, Op, true> : k_fast_fft<RDF, L, false, INV, IH, OH, Op, false>;
  const size_t lds = (size_t)L * LdsLine<RD::n, L>::size * sizeof(float2);
  FH_HIP(ensure_lds(kern, lds));
  hipLaunchKernelGGL(kern, dim3((io.nlines + L - 1) / L, batch), dim3(L * RDF::P), lds, c->stream, io, tab, op);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

#define FH_LOGN_SWITCH_INT(logn, CALL)                                   \
  switch (logn) {                                                        \
    case 4: { constexpr int LG = 4; return CALL; }                       \
    case 5: { constexpr int LG = 5; return CALL; }                       \
    case 6: { constexpr int LG = 6; return CALL; }                       \
    case 7: { constexpr int LG = 7; return CALL; }                       \
    case 8: { constexpr int LG = 8; return CALL; }                       \
    case 9: { constexpr int LG = 9; return CALL; }                       \
    case 10: { constexpr int LG = 10; return CALL; }                     \
    case 11: { constexpr int LG = 11; return CALL; }                     \
    case 12: { constexpr int LG = 12; return CALL; }                     \
    default: return fail(FH_E_UNSUPP, "transform length 2^%d unsupported", logn); \
  }

// forward x FFT; half_in: rows hold n/2 points (zero-padded RS input)
template <typename Op>
static int fast_xfwd(fh_ctx* c, int logn, bool half_in, const FastIO& io, int batch, Op op, int xvar) {
  if (half_in) { FH_LOGN_SWITCH_INT(logn, (fast_x_t<LG, false, true, false, Op>(c, io, batch, op, xvar))) }
  if constexpr (std::is_same<Op, ChirpOp>::value) {
    return fail(FH_E_UNSUPP, "chirp prologue needs a padded plan");
  } else {
    FH_LOGN_SWITCH_INT(logn, (fast_x_t<LG, false, false, false, Op>(c, io, batch, op, xvar)))
  }
}

// inverse x FFT; half_out: only n/2 points are stored (RS crop).  The chirp op only runs on
// padded (Collins) plans, so its unpadded variants are not instantiated.
template <typename Op>
static int fast_xinv(fh_ctx* c, int logn, bool half_out, const FastIO& io, int batch, Op op, int xvar) {
  if (half_out) { FH_LOGN_SWITCH_INT(logn, (fast_x_t<LG, true, false, true, Op>(c, io, batch, op, xvar))) }
  if constexpr (std::is_same<Op, ChirpOp>::value) {
    return fail(FH_E_UNSUPP, "chirp epilogue needs a padded plan");
  } else {
    FH_LOGN_SWITCH_INT(logn, (fast_x_t<LG, true, false, false, Op>(c, io, batch, op, xvar)))
  }
}

// The fused metric pass (k_fast_sqint): lines per workgroup halve like the x passes'; the
// partials buffer must hold batch * ceil(ny / L) values for the L the launch ends with, which
// sqint_parts() reports for the same arguments.
template <int LOGN, bool HALF, int L>
static int fast_sqint_l(fh_ctx* c, const FastIO& io, int batch, const float* I, long tstride, int nx,
                        const double* dc, double* part, long* nparts) {
  using RDF = typename Cfg<float, LOGN, false>::RD;
  using RI = typename RDF::Rev;
  if constexpr (L >= 8 && (L / 2) * RDF::P >= 64) {
    if ((long)((io.nlines + L - 1) / L) * batch < kFillWGs)
      return fast_sqint_l<LOGN, HALF, L / 2>(c, io, batch, I, tstride, nx, dc, part, nparts);
  }
  const long gx = (io.nlines + L - 1) / L;
  if (!part) {
    *nparts = gx;
    return FH_OK;
  }
  const float2 *tabi = nullptr, *tabf = nullptr;
  int rc = get_stage_tw(c, RI::n, RI::R0, RI::R1, RI::R2, true, &tabi);
  if (!rc) rc = get_stage_tw(c, RDF::n, RDF::R0, RDF::R1, RDF::R2, false, &tabf);
  if (rc) return rc;
  auto kern = k_fast_sqint<RDF, L, HALF>;
  if constexpr (LOGN == 11 && L == 4) {
    // rows of 2048 points in whole, unpadded 4-row groups: the input staged through LDS
    if (io.in_blk && io.in_l == RDF::n && io.nlines % L == 0) kern = k_fast_sqint<RDF, L, HALF, true>;
  }
  const size_t lds = (size_t)L * LdsLine<RDF::n, L>::size * sizeof(float2);
  FH_HIP(ensure_lds(kern, lds));
  hipLaunchKernelGGL(kern, dim3(gx, batch), dim3(L * RDF::P), lds, c->stream, io, tabi, tabf, I, tstride, nx, dc,
                     part);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

template <int LOGN, bool HALF>
static int fast_sqint_t(fh_ctx* c, const FastIO& io, int batch, const float* I, long tstride, int nx,
                        const double* dc, double* part, long* nparts) {
  return fast_sqint_l<LOGN, HALF, Cfg<float, LOGN, false>::L>(c, io, batch, I, tstride, nx, dc, part, nparts);
}

// part == nullptr: only report the number of partials per field in *nparts
static int fast_sqint(fh_ctx* c, int logn, bool half, const FastIO& io, int batch, const float* I, long tstride,
                      int nx, const double* dc, double* part, long* nparts) {
  if (half) { FH_LOGN_SWITCH_INT(logn, (fast_sqint_t<LG, true>(c, io, batch, I, tstride, nx, dc, part, nparts))) }
  FH_LOGN_SWITCH_INT(logn, (fast_sqint_t<LG, false>(c, io, batch, I, tstride, nx, dc, part, nparts)))
}

template <typename RDF, int L, bool STG = false>
static int fast_phase_grad_launch(fh_ctx* c, const FastIO& io, int batch, const float* phi, const float2* saved,
                                  float* dphi, int nx, float s, int acc, int store) {
  using RI = typename RDF::Rev;
  const float2* tab = nullptr;
  int rc = get_stage_tw(c, RI::n, RI::R0, RI::R1, RI::R2, true, &tab);
  if (rc) return rc;
  auto kern = dphi ? k_fast_phase_grad<RDF, L, true, STG> : k_fast_phase_grad<RDF, L, false, STG>;
  const size_t lds = PhaseGradLds<RI::n, L, RI::P, RI::V>::bytes;
  FH_HIP(ensure_lds(kern, lds));
  hipLaunchKernelGGL(kern, dim3((io.nlines + L - 1) / L), dim3(L * RDF::P), lds, c->stream, io, batch, tab, phi,
                     saved, dphi, nx, s, acc, store);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

// (FewRowsPlan as for the x passes: this pass loops over the batch, so only the 4-row groups count)
template <int LOGN, int L>
static int fast_phase_grad_l(fh_ctx* c, const FastIO& io, int batch, const float* phi, const float2* saved,
                             float* dphi, int nx, float s, int acc, int store) {
  using RDF = typename Cfg<float, LOGN, false>::RD;
  const bool few = (long)((io.nlines + L - 1) / L) < kFillWGs;
  if constexpr (L >= 8 && (L / 2) * RDF::P >= 64) {
    if (few) return fast_phase_grad_l<LOGN, L / 2>(c, io, batch, phi, saved, dphi, nx, s, acc, store);
  }
  using FR = typename FewRowsPlan<LOGN>::RD;
  if constexpr (!std::is_void<FR>::value && L == 4) {
    if (few) return fast_phase_grad_launch<FR, 4>(c, io, batch, phi, saved, dphi, nx, s, acc, store);
  }
  // rows of 2048 points: the blocked spectrum staged through LDS when the workgroups' 4-row groups
  // are whole, unpadded runs (as the staged x passes take them)
  if constexpr (LOGN == 11 && L == 4) {
    if (io.in_blk && io.in_l == RDF::n && io.nlines % L == 0)
      return fast_phase_grad_launch<RDF, L, true>(c, io, batch, phi, saved, dphi, nx, s, acc, store);
  }
  // Out of place (RS, Collins and CZT plans read the scratch spectrum), rows of 4096 points go two
  // to a workgroup: 512 threads may hold 256 VGPRs, and the phase factors then fit in LDS beside
  // the exchange areas.  In place, a workgroup must own whole 4-row groups of the blocked input.
  if constexpr (L == 4 && LOGN == 12) {
    // every caller hands rows of 4096 points over out of place (grad_split): four such rows and
    // the gradient accumulators do not fit one 1024-thread workgroup's registers
    if (static_cast<const void*>(io.in) == static_cast<const void*>(io.out))
      return fail(FH_E_UNSUPP, "the mask-gradient pass of 4096-point rows reads another buffer than it writes");
    return fast_phase_grad_launch<RDF, 2>(c, io, batch, phi, saved, dphi, nx, s, acc, store);
  } else {
    return fast_phase_grad_launch<RDF, L>(c, io, batch, phi, saved, dphi, nx, s, acc, store);
  }
}

// A workgroup owns whole 4-row groups (L >= 4): on ASProp plans this pass writes natural rows over
// the 4x4-blocked spectrum it reads, and the rows 4g..4g+3 share those bytes.
template <int LOGN>
static int fast_phase_grad_t(fh_ctx* c, const FastIO& io, int batch, const float* phi, const float2* saved,
                             float* dphi, int nx, float s, int acc, int store) {
  static_assert(Cfg<float, LOGN, false>::L >= 4, "the in-place pass needs whole 4-row groups");
  return fast_phase_grad_l<LOGN, Cfg<float, LOGN, false>::L>(c, io, batch, phi, saved, dphi, nx, s, acc, store);
}

// x IFFT of the blocked spectrum into natural rows, un-applying a Phase and accumulating its
// gradient over the batch (the adjoint end of a Phase -> propagator pair); store == 0 (with a
// gradient) leaves the un-applied field unwritten
static int fast_phase_grad(fh_ctx* c, int logn, const FastIO& io, int batch, const float* phi, const float2* saved,
                           float* dphi, int nx, float s, int acc, int store = 1) {
  FH_LOGN_SWITCH_INT(logn, (fast_phase_grad_t<LG>(c, io, batch, phi, saved, dphi, nx, s, acc, store)))
}

// x-pass variants (fast_x_l): names for the autotuner's report.  The staged form is the default for
// rows of 2048 and 4096 points of ASProp plans (forward + adjoint step: 8 x 2048^2 0.607 ms against
// 0.656 ms, 2 x 4096^2 0.885 against 0.977 ms) and not for shorter rows (32 x 512^2: 0.160 against
// 0.158 ms, 16 x 1024^2: 0.325 against 0.314 ms).  Plans padded to 2n (RS, Collins: half rows on
// the natural side, out of place) take lanes over rows there: 4096-point forward x pass of one
// field 0.0235 ms against 0.0316 staged and 0.0324 lanes over x, of 16 fields 0.355 against 0.467
// and 0.633 ms; of 2048-point rows the forward passes tie and the inverse takes the staged form
// (3 fields 0.0156 against 0.0169 ms).
static const char* const kXVariants[] = {"lanes over x", "lanes over rows", "blocked side via LDS"};
constexpr int kNumXVariants = 3;
static int default_x_variant(int logn, bool padded, bool inv) {
  if (logn < 11) return 0;
  if (!padded) return 2;
  return (logn == 11 && inv) ? 2 : 1;
}

// y-pass variants (radices, columns per workgroup) for the autotuner
using FastConvLaunch = int (*)(fh_ctx*, const FastIO&, int, QuadMul, bool);
struct FastConvVariant {
  const char* name;
  FastConvLaunch fn;
};

template <int A, int B, int C_, int L, int FPW = 1 << 20, bool LT = false>
static int fast_conv_v(fh_ctx* c, const FastIO& io, int batch, QuadMul op, bool half) {
  using RDF = Radices<A, B, C_>;
  using RI = typename RDF::Rev;
  if constexpr (L > 1 && (L / 2) * RDF::P >= 64 && !LT) {
    const int fpw0 = FPW < batch ? FPW : batch;
    if ((long)((io.nlines + L - 1) / L) * ((batch + fpw0 - 1) / fpw0) < kFillWGs)
      return fast_conv_v<A, B, C_, L / 2, FPW, LT>(c, io, batch, op, half);
  }
  const float2 *tf = nullptr, *ti = nullptr;
  int rc;
  if ((rc = get_stage_tw(c, RDF::n, RDF::R0, RDF::R1, RDF::R2, false, &tf)) ||
      (rc = get_stage_tw(c, RI::n, RI::R0, RI::R1, RI::R2, true, &ti)))
    return rc;
  const size_t lds = (size_t)L * LdsLine<RDF::n, L>::size * sizeof(float2);
  QuadTab qt{op.tab, op.tabidx, op.tabsize, op.nx, op.qcol, op.conj, op.one_tab, op.full,
             op.full ? op.ny / 2 : 0};
  const int fpw = FPW < batch ? FPW : batch;
  dim3 grid((io.nlines + L - 1) / L, (batch + fpw - 1) / fpw);
  if (half) {
    auto kern = k_fast_yconv<RDF, L, true, true, LT>;
    FH_HIP(ensure_lds(kern, lds));
    hipLaunchKernelGGL(kern, grid, dim3(L * RDF::P), lds, c->stream, io, batch, fpw, tf, ti, qt);
  } else {
    auto kern = k_fast_yconv<RDF, L, false, false, LT>;
    FH_HIP(ensure_lds(kern, lds));
    hipLaunchKernelGGL(kern, grid, dim3(L * RDF::P), lds, c->stream, io, batch, fpw, tf, ti, qt);
  }
  FH_HIP(hipGetLastError());
  return FH_OK;
}

static const FastConvVariant kFast4[] = {{"4x4 L16", fast_conv_v<4, 4, 1, 16>}};
static const FastConvVariant kFast5[] = {{"8x4 L16", fast_conv_v<8, 4, 1, 16>}};
static const FastConvVariant kFast6[] = {{"8x8 L16", fast_conv_v<8, 8, 1, 16>}};
static const FastConvVariant kFast7[] = {{"16x8 L16", fast_conv_v<16, 8, 1, 16>}};
static const FastConvVariant kFast8[] = {{"16x16 L16", fast_conv_v<16, 16, 1, 16>},
                                         {"16x16 L8", fast_conv_v<16, 16, 1, 8>}};
template <int A, int B, int C_, int L, bool H_LDS = false>
static int fast_pipe_v(fh_ctx* c, const FastIO& io, int batch, QuadMul op, bool half) {
  using RDF = Radices<A, B, C_>;
  using RI = typename RDF::Rev;
  const float2 *tf = nullptr, *ti = nullptr;
  int rc;
  if ((rc = get_stage_tw(c, RDF::n, RDF::R0, RDF::R1, RDF::R2, false, &tf)) ||
      (rc = get_stage_tw(c, RI::n, RI::R0, RI::R1, RI::R2, true, &ti)))
    return rc;
  const size_t lds0 = (size_t)L * LdsLine<RDF::n, L>::size * sizeof(float2);
  QuadTab qt{op.tab, op.tabidx, op.tabsize, op.nx, op.qcol, op.conj, op.one_tab, op.full,
             op.full ? op.ny / 2 : 0};
  dim3 grid((io.nlines + L - 1) / L);
  auto launch = [&](auto kern, size_t lds) -> int {
    FH_HIP(ensure_lds(kern, lds));
    hipLaunchKernelGGL(kern, grid, dim3(L * RDF::P), lds, c->stream, io, batch, tf, ti, qt);
    return FH_OK;
  };
  // the LDS-resident transfer function (H_LDS) applies to one-table batches only
  const size_t ldsh = lds0 + (H_LDS ? (size_t)L * RDF::P * RDF::V * sizeof(float2) : 0);
  if (half) rc = qt.one_tab ? launch(k_fast_ypipe<RDF, L, true, true, true, H_LDS>, ldsh)
                            : launch(k_fast_ypipe<RDF, L, true, true, false>, lds0);
  else rc = qt.one_tab ? launch(k_fast_ypipe<RDF, L, false, false, true, H_LDS>, ldsh)
                       : launch(k_fast_ypipe<RDF, L, false, false, false>, lds0);
  if (rc) return rc;
  FH_HIP(hipGetLastError());
  return FH_OK;
}

// Wave-per-column y pass (2048-point columns of the blocked layout, whole lines only); RS
// half lines, the natural layout and a ragged last tile take the LDS-table pipelined pass.
static int fast_wave_v(fh_ctx* c, const FastIO& io, int batch, QuadMul op, bool half) {
  if (half || !io.in_blk || !io.out_blk || io.nlines % kWaveCols || io.in_e % 4 || io.out_e % 4 ||
      io.in_valid < kWaveN || io.out_valid < kWaveN)
    return fast_pipe_v<16, 16, 8, 4, true>(c, io, batch, op, half);
  using RDF = WaveRDF;
  using RI = typename RDF::Rev;
  const float2 *tf = nullptr, *ti = nullptr;
  int rc;
  if ((rc = get_stage_tw(c, RDF::n, RDF::R0, RDF::R1, RDF::R2, false, &tf)) ||
      (rc = get_stage_tw(c, RI::n, RI::R0, RI::R1, RI::R2, true, &ti)))
    return rc;
  // the tile's offsets: 16-byte pairs k * rowlen * 1024 bytes apart within one field
  if ((long)io.in_e * 1024 * (kWavePairs - 1) + (long)io.in_e * 128 > 0x7fffffffL ||
      (long)io.out_e * 1024 * (kWavePairs - 1) + (long)io.out_e * 128 > 0x7fffffffL)
    return fast_pipe_v<16, 16, 8, 4, true>(c, io, batch, op, half);
  const size_t lds = kWaveLds;
  QuadTab qt{op.tab, op.tabidx, op.tabsize, op.nx, op.qcol, op.conj, op.one_tab, op.full,
             op.full ? op.ny / 2 : 0};
  FH_HIP(ensure_lds(k_fast_ywave, lds));
  hipLaunchKernelGGL(k_fast_ywave, dim3(io.nlines / kWaveCols), dim3(kWaveCols * 64), lds, c->stream, io,
                     batch, tf, ti, qt);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

// Staged-tile pipelined y pass (whole 4-column tiles of the blocked layout on both sides, full
// lines); otherwise the LDS-table pipelined pass.
template <int A, int B, int C_>
static int fast_ystg_v(fh_ctx* c, const FastIO& io, int batch, QuadMul op, bool half) {
  using RDF = Radices<A, B, C_>;
  using RI = typename RDF::Rev;
  constexpr int n = RDF::n, P = RDF::P, NQ = n / (2 * P);
  if (half || !io.in_blk || !io.out_blk || io.nlines % 4 || io.in_e % 4 || io.out_e % 4 ||
      io.in_valid < n || io.out_valid < n ||
      (long)io.in_e * 2 * P * 8 * NQ > 0x7fffffffL || (long)io.out_e * 2 * P * 8 * NQ > 0x7fffffffL)
    return fast_pipe_v<A, B, C_, 4, true>(c, io, batch, op, half);
  const float2 *tf = nullptr, *ti = nullptr;
  int rc;
  if ((rc = get_stage_tw(c, RDF::n, RDF::R0, RDF::R1, RDF::R2, false, &tf)) ||
      (rc = get_stage_tw(c, RI::n, RI::R0, RI::R1, RI::R2, true, &ti)))
    return rc;
  QuadTab qt{op.tab, op.tabidx, op.tabsize, op.nx, op.qcol, op.conj, op.one_tab, op.full,
             op.full ? op.ny / 2 : 0};
  const size_t lds0 = (size_t)4 * LdsLine<n, 4>::size * sizeof(float2);
  auto launch = [&](auto kern, size_t lds) -> int {
    FH_HIP(ensure_lds(kern, lds));
    hipLaunchKernelGGL(kern, dim3(io.nlines / 4), dim3(4 * P), lds, c->stream, io, batch, tf, ti, qt);
    FH_HIP(hipGetLastError());
    return FH_OK;
  };
  return qt.one_tab ? launch(k_fast_ystg<RDF, true>, lds0 + (size_t)4 * P * RDF::V * sizeof(float2))
                    : launch(k_fast_ystg<RDF, false>, lds0);
}

// Variant lists per length; default_conv_variant picks the fastest measured on MI355X.
static const FastConvVariant kFast9[] = {{"8x8x8 L4 f1", fast_conv_v<8, 8, 8, 4, 1>},
                                         {"8x8x8 L4", fast_conv_v<8, 8, 8, 4>},
                                         {"16x32 L8", fast_conv_v<16, 32, 1, 8>}};
static const FastConvVariant kFast10[] = {{"16x8x8 L4 f1", fast_conv_v<16, 8, 8, 4, 1>},
                                          {"pipe 16x8x8 L4", fast_pipe_v<16, 8, 8, 4>},
                                          {"16x8x8 L4", fast_conv_v<16, 8, 8, 4>},
                                          {"32x32 L8 f1", fast_conv_v<32, 32, 1, 8, 1>}};
static const FastConvVariant kFast11[] = {{"8x16x16 L4 f1", fast_conv_v<8, 16, 16, 4, 1>},
                                          {"pipe 8x16x16 L4", fast_pipe_v<8, 16, 16, 4>},
                                          {"pipe 16x16x8 L4", fast_pipe_v<16, 16, 8, 4>},
                                          {"16x16x8 L4 f1", fast_conv_v<16, 16, 8, 4, 1>},
                                          {"16x16x8 L4 f2", fast_conv_v<16, 16, 8, 4, 2>},
                                          {"pipe 16x16x8 L4 hlds", fast_pipe_v<16, 16, 8, 4, true>},
                                          {"wave 8x8x32 L4", fast_wave_v},
                                          {"stg 16x16x8 L4", fast_ystg_v<16, 16, 8>}};
static const FastConvVariant kFast12[] = {{"16x16x16 L4 f1", fast_conv_v<16, 16, 16, 4, 1>},
                                          {"16x16x16 L4", fast_conv_v<16, 16, 16, 4>},
                                          {"16x16x16 L2 f1", fast_conv_v<16, 16, 16, 2, 1>},
                                          {"pipe 16x16x16 L2", fast_pipe_v<16, 16, 16, 2>}};

// default variant index for a y length and whether the batch shares one transfer function
static int default_conv_variant(int logn, bool one_tab, bool half, int batch) {
  // (times: autotuner, 10 reps, per call)
  if (logn == 11 && !half) {
    if (!one_tab) return 6;  // wave per column (a table per field)
    // one table: wave per column for one or two fields (1 field 0.0197 ms against 0.0286 staged),
    // pipe 16x16x8 L4 with the table in LDS up to four (3 fields 0.0542 against 0.0568), then the
    // staged-tile pipe 16x16x8 L4 (8 fields 0.138 ms, the fastest)
    return batch <= 2 ? 6 : (batch <= 4 ? 5 : 7);
  }
  if (one_tab && logn == 11) return 5;           // pipe 16x16x8 L4, transfer function in LDS
  if (logn == 11) return 3;                      // 16x16x8 L4 f1 (RS, a table per field)
  // RS, one field: 16x16x16 L2 f1 (0.0758 against 0.0836 ms pipelined); more fields: pipe L2
  if (logn == 12 && half && !one_tab) return batch == 1 ? 2 : 3;
  // ASProp 4096: one table over 3 or more fields pipe 16x16x16 L2 (8 fields 0.790 against 0.933 ms
  // for 16x16x16 L4); a table per field over 8 or more 16x16x16 L4 f1 (0.961 against 0.993)
  if (logn == 12 && !half) return one_tab ? (batch >= 3 ? 3 : 1) : (batch >= 8 ? 0 : 1);
  if (logn == 12) return 1;                      // 16x16x16 L4
  // RS, one field of 1024-point columns: pipe 16x8x8 L4 (0.0078 against 0.0089 ms; ASProp plans of
  // that length measure within noise of one another)
  if (logn == 10 && half && batch == 1) return 1;
  return 0;
}

template <size_t N>
static const FastConvVariant* variant_list(const FastConvVariant (&v)[N], int* count) {
  *count = (int)N;
  return v;
}

static const FastConvVariant* fast_conv_variants(int logn, int* count) {
  switch (logn) {
    case 4: return variant_list(kFast4, count);
    case 5: return variant_list(kFast5, count);
    case 6: return variant_list(kFast6, count);
    case 7: return variant_list(kFast7, count);
    case 8: return variant_list(kFast8, count);
    case 9: return variant_list(kFast9, count);
    case 10: return variant_list(kFast10, count);
    case 11: return variant_list(kFast11, count);
    case 12: return variant_list(kFast12, count);
    default: *count = 0; return nullptr;
  }
}

static int check_dir(int direction) {
  if (direction != 1 && direction != -1) return fail(FH_E_ARG, "direction must be +1 (Forward) or -1 (Backward), got %d", direction);
  return FH_OK;
}

// Validates the common plan parameters and assigns one table per distinct
// (wavelength, z, tilt x, tilt y) of the batch.  zs / tilts may be null (z = 0, no tilt).
static int plan_common(fh_ctx* ctx, int nx, int ny, int batch, double dx, double dy,
                       const double* lambdas, int nlambdas, fh_plan* p, const double* zs = nullptr,
                       int nz = 0, const double* tx = nullptr, const double* ty = nullptr, int nt = 0) {
  if (!ctx) return fail(FH_E_ARG, "null context");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive (nx=%d ny=%d batch=%d)", nx, ny, batch);
  if (!(dx > 0) || !(dy > 0)) return fail(FH_E_ARG, "sampling must be positive");
  if (!lambdas || (nlambdas != 1 && nlambdas != batch))
    return fail(FH_E_ARG, "need 1 or batch (=%d) wavelengths, got %d", batch, nlambdas);
  for (int k = 0; k < nlambdas; ++k)
    if (!(lambdas[k] > 0)) return fail(FH_E_ARG, "wavelengths must be positive");
  if (zs && nz != 1 && nz != batch) return fail(FH_E_ARG, "need 1 or batch (=%d) distances, got %d", batch, nz);
  for (int k = 0; zs && k < nz; ++k)
    if (!std::isfinite(zs[k])) return fail(FH_E_ARG, "distances must be finite");
  if (nt != 0 && ((nt != 1 && nt != batch) || !tx || !ty))
    return fail(FH_E_ARG, "need 0, 1 or batch (=%d) tilt pairs, got %d", batch, nt);
  for (int k = 0; k < nt; ++k)
    if (!std::isfinite(tx[k]) || !std::isfinite(ty[k]) || fabs(tx[k]) >= M_PI / 2 || fabs(ty[k]) >= M_PI / 2)
      return fail(FH_E_ARG, "tilt angles must be finite radians in (-pi/2, pi/2)");
  p->ctx = ctx;
  p->nx = nx; p->ny = ny; p->batch = batch; p->dx = dx; p->dy = dy;
  p->field_table.resize(batch);
  for (int b = 0; b < batch; ++b) {
    const double lam = lambdas[nlambdas == 1 ? 0 : b];
    const double z = zs ? zs[nz == 1 ? 0 : b] : 0.0;
    const double ax = nt ? tx[nt == 1 ? 0 : b] : 0.0, ay = nt ? ty[nt == 1 ? 0 : b] : 0.0;
    int idx = -1;
    for (size_t k = 0; k < p->lam_tables.size(); ++k)
      if (p->lam_tables[k] == lam && p->tab_z[k] == z && p->tab_tx[k] == ax && p->tab_ty[k] == ay) {
        idx = (int)k;
        break;
      }
    if (idx < 0) {
      idx = (int)p->lam_tables.size();
      p->lam_tables.push_back(lam);
      p->tab_z.push_back(z);
      p->tab_tx.push_back(ax);
      p->tab_ty.push_back(ay);
    }
    p->field_table[b] = idx;
  }
  return FH_OK;
}

static bool plan_tilted(const fh_plan* p) {
  for (size_t k = 0; k < p->tab_tx.size(); ++k)
    if (p->tab_tx[k] != 0.0 || p->tab_ty[k] != 0.0) return true;
  return false;
}

static void plan_free(fh_plan* p) {
  if (!p) return;
  if (p->tab) (void)hipFree(p->tab);
  if (p->d_tabidx) (void)hipFree(p->d_tabidx);
  if (p->scratch) (void)hipFree(p->scratch);
  if (p->plain_tab) (void)hipFree(p->plain_tab);
  if (p->tab_d) (void)hipFree(p->tab_d);
  if (p->cis) (void)hipFree(p->cis);
  if (p->chirp) (void)hipFree(p->chirp);
  delete p;
}

static int upload_tabidx(fh_plan* p) {
  if (hipMalloc(&p->d_tabidx, p->batch * sizeof(int)) != hipSuccess) return fail(FH_E_NOMEM, "table index: out of device memory");
  FH_HIP(hipMemcpy(p->d_tabidx, p->field_table.data(), p->batch * sizeof(int), hipMemcpyHostToDevice));
  return FH_OK;
}

// In-place FP64 2-D FFT of a [ny2][nx2] double2 array (table generation).
static int fft2_d(fh_ctx* c, double2* d, int nx2, int ny2, int lx, int ly) {
  const double2 *twx = nullptr, *twy = nullptr;
  int rc;
  if ((rc = get_twiddles(c, nx2, &twx)) || (rc = get_twiddles(c, ny2, &twy))) return rc;
  LineIO<double2> io{};
  io.in = d; io.out = d; io.in_f = io.out_f = (long)nx2 * ny2;
  io.in_l = io.out_l = nx2; io.in_e = io.out_e = 1;
  io.nlines = ny2; io.in_valid = io.out_count = nx2;
  FH_HIP((launch_fft<double2, double2, false, false>(lx, io, 1, twx, NoOp{}, c->stream)));
  io.in_l = io.out_l = 1; io.in_e = io.out_e = nx2;
  io.nlines = nx2; io.in_valid = io.out_count = ny2;
  FH_HIP((launch_fft<double2, double2, true, false>(ly, io, 1, twy, NoOp{}, c->stream)));
  return FH_OK;
}

// The passes of one propagation of fields u (or, fused, of Phase then propagation):
// pass 0 (x FFT) and pass 1 (y FFT * kernel * IFFT); `only` >= 0 runs just that pass.
// Between passes the spectrum sits in the 4x4-blocked layout (elem_addr); for padded plans
// (RS, Collins, CZT) it sits in the nxl x ny workspace (rows rounded up to whole 4-row
// groups) and only its first ny rows are ever stored.
static long mid_rows(const fh_plan* p) { return p->kind == 0 ? p->ny : ((p->ny + 3) / 4) * 4; }

// ASProp plans with rows of 4096 points end an adjoint-with-gradient in the out-of-place
// mask-gradient pass (two rows per workgroup, no spill): their y pass then writes the spectrum into
// the plan's scratch (fields-sized, allocated at the first such call) instead of in place.
static bool grad_split(const fh_plan* p) { return p->kind == 0 && !p->oned && p->lx == 12; }

static int grad_scratch(fh_plan* p) {
  if (p->scratch) return FH_OK;
  if (hipMalloc(&p->scratch, (size_t)p->nx * p->ny * p->batch * sizeof(float2)) != hipSuccess)
    return fail(FH_E_NOMEM, "mask-gradient spectrum: out of device memory");
  return FH_OK;
}

template <typename XOp>
// src (nullable): the first pass reads the input there instead of in u (u is then only written);
// y_to_scratch: an ASProp y pass writes the plan's scratch (see grad_split)
static int run_prop(fh_plan* p, float2* u, int direction, XOp xop, int only = -1, const float2* src = nullptr,
                    bool y_to_scratch = false) {
  fh_ctx* c = p->ctx;
  const int nxl = 1 << p->lx;
  const bool rs = p->kind != 0;
  float2* mid = rs ? p->scratch : u;
  const long midf = (long)nxl * mid_rows(p);
  int rc;
  FastIO io{};
  if (only < 0 || only == 0) {
    io.in = src ? src : u; io.out = mid; io.in_f = (long)p->nx * p->ny; io.out_f = midf;
    io.in_l = p->nx; io.out_l = nxl; io.in_e = io.out_e = 1; io.in_blk = 0; io.out_blk = 1;
    io.nlines = p->ny; io.in_bytes = io.in_f * 8; io.out_bytes = midf * 8;
    io.in_valid = p->nx;
    const int xv = p->xf_var >= 0 ? p->xf_var : default_x_variant(p->lx, rs, false);
    if ((rc = fast_xfwd(c, p->lx, rs, io, p->batch, xop, xv))) return rc;
  }
  if (only == 0) return FH_OK;
  QuadMul q{p->tab, p->d_tabidx, p->tabsize, nxl, 1 << p->ly, p->qy, direction < 0,
            p->lam_tables.size() == 1, p->full};
  io.in = mid; io.out = (y_to_scratch && !rs) ? p->scratch : mid; io.in_f = io.out_f = midf;
  io.in_l = io.out_l = 1; io.in_e = io.out_e = nxl; io.in_blk = io.out_blk = 1;
  io.nlines = nxl; io.in_bytes = io.out_bytes = midf * 8;
  io.in_valid = io.out_valid = p->ny;
  int nvar = 0;
  const FastConvVariant* vars = fast_conv_variants(p->ly, &nvar);
  if (!vars) return fail(FH_E_UNSUPP, "y length 2^%d unsupported", p->ly);
  const int v = (p->conv_var >= 0 && p->conv_var < nvar) ? p->conv_var
                                                          : default_conv_variant(p->ly, q.one_tab, rs, p->batch);
  return vars[v].fn(c, io, p->batch, q, rs);
}

template <typename XOp>
static int run_xifft_op(fh_plan* p, float2* u, XOp xop) {
  const int nxl = 1 << p->lx;
  const bool rs = p->kind != 0;
  float2* mid = rs ? p->scratch : u;
  FastIO io{};
  io.in = mid; io.out = u; io.in_f = (long)nxl * mid_rows(p); io.out_f = (long)p->nx * p->ny;
  io.in_l = nxl; io.out_l = p->nx; io.in_e = io.out_e = 1; io.in_blk = 1; io.out_blk = 0;
  io.nlines = p->ny; io.in_bytes = io.in_f * 8; io.out_bytes = io.out_f * 8;
  io.out_valid = p->nx;
  const int xv = p->xi_var >= 0 ? p->xi_var : default_x_variant(p->lx, rs, true);
  return fast_xinv(p->ctx, p->lx, rs, io, p->batch, xop, xv);
}

// CollinsProp's chirps for one direction (collins_integral.jl:155-189, reversed for
// Backward): Forward applies A before the convolution and D after it; Backward applies
// conj(D) before and conj(A) after, times alpha_x*alpha_y = the ratio of the reference's
// backward normalisation prod(ds'/sqrt(-i b)) to the conjugate of the forward one folded
// into the convolution table.
static ChirpOp collins_op(const fh_plan* p, int direction, bool pre) {
  const long nt = (long)p->lam_tables.size();
  const float2* ax = p->chirp;
  const float2* ay = ax + nt * p->nx;
  const float2* dx = ay + nt * p->ny;
  const float2* dy = dx + nt * p->nx;
  const bool a = (direction > 0) == pre;
  const float s = (direction < 0 && !pre) ? (float)(p->alpha_x * p->alpha_y) : 1.f;
  return ChirpOp{a ? ax : dx, a ? ay : dy, p->d_tabidx, p->nx, p->ny, direction < 0, s};
}

// The passes of a plan's own propagation (chirps included for Collins).
static int run_prop_plan(fh_plan* p, float2* u, int direction, int only = -1) {
  if (p->kind == 2) return run_prop(p, u, direction, collins_op(p, direction, true), only);
  return run_prop(p, u, direction, NoOp{}, only);
}

static int run_xifft(fh_plan* p, float2* u, int direction) {
  if (p->kind == 2) return run_xifft_op(p, u, collins_op(p, direction, false));
  return run_xifft_op(p, u, NoOp{});
}

// The adjoint's last pass for a Phase -> propagator pair: x IFFT of the blocked spectrum,
// the Phase un-applied, its gradient summed over the batch (fh_propagate_phase_backward).
static int phase_grad_pass(fh_plan* p, float2* u, const float* phi, const void* saved, float* dphi, int direction,
                           int accumulate, int store = 1) {
  const int nxl = 1 << p->lx;
  FastIO io{};
  io.in = (p->kind == 0 && !grad_split(p)) ? u : p->scratch; io.out = u;
  io.in_f = (long)nxl * mid_rows(p); io.out_f = (long)p->nx * p->ny;
  io.in_l = nxl; io.out_l = p->nx; io.in_e = io.out_e = 1; io.in_blk = 1; io.out_blk = 0;
  io.nlines = p->ny; io.in_bytes = io.in_f * 8; io.out_bytes = io.out_f * 8;
  return fast_phase_grad(p->ctx, p->lx, io, p->batch, phi, reinterpret_cast<const float2*>(saved), dphi, p->nx,
                         (float)direction, accumulate, store);
}

}  // namespace fh

// ================================================================ C ABI

// Device address of `scratch` per device, looked up once per device (with that device
// current) when the first context on it is created.
static std::mutex g_scratch_mu;
static std::map<int, double*> g_scratch_addr;

static int scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_addr.count(device)) return FH_OK;
  int prev = 0;
  FH_HIP(hipGetDevice(&prev));
  FH_HIP(hipSetDevice(device));
  void* addr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
  (void)hipSetDevice(prev);
  if (e != hipSuccess || !addr)
    return fail(FH_E_HIP, "device %d: address of the no-allocation scratch array: %s", device, hipGetErrorString(e));
  g_scratch_addr[device] = static_cast<double*>(addr);
  return FH_OK;
}

static double* scratch_addr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  return it == g_scratch_addr.end() ? nullptr : it->second;
}

extern "C" {

int fh_version(void) { return 100; }

const char* fh_last_error(void) { return g_err.c_str(); }

int fh_max_length(void) { return 1 << kMaxLog; }

int fh_ctx_create(int device, void* stream, fh_ctx** out) {
  if (!out) return fail(FH_E_ARG, "null output pointer");
  int ndev = 0;
  FH_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(FH_E_ARG, "device %d out of range (%d devices)", device, ndev);
  int rc;
  if ((rc = scratch_setup(device))) return rc;
  fh_ctx* c = new fh_ctx();
  c->device = device;
  c->stream = reinterpret_cast<hipStream_t>(stream);
  *out = c;
  return FH_OK;
}

int fh_ctx_set_stream(fh_ctx* ctx, void* stream) {
  if (!ctx) return fail(FH_E_ARG, "null context");
  ctx->stream = reinterpret_cast<hipStream_t>(stream);
  return FH_OK;
}

int fh_ctx_synchronize(fh_ctx* ctx) {
  if (!ctx) return fail(FH_E_ARG, "null context");
  FH_HIP(hipStreamSynchronize(ctx->stream));
  return FH_OK;
}

int fh_ctx_destroy(fh_ctx* ctx) {
  if (!ctx) return FH_OK;
  (void)hipSetDevice(ctx->device);
  for (auto& kv : ctx->twiddles) (void)hipFree(kv.second);
  for (auto& kv : ctx->stage_tw) (void)hipFree(kv.second);
  for (auto& kv : ctx->dft_plans) plan_free(kv.second);
  if (ctx->metric_part) (void)hipFree(ctx->metric_part);
  for (double* d : ctx->metric_part_retired) (void)hipFree(d);
  delete ctx;
  return FH_OK;
}

// ---------------------------------------------------------------- 1-D fields
// A 1-D ScalarField (Fields.jl:135-136: electric [nx, batch...], Nd = 1) is a batch of lines of
// nx points; a propagation is one pass of FFT -> transfer function -> IFFT per line.  Power-of-two
// nx transforms at length nx; any other nx as the n-point circular convolution written as a
// zero-padded linear one at the power of two N >= 2 nx - 1 (as as_padded_tables does in 2-D).

struct LineKey { double lam, z, tx; };

// Tables of 1-D propagators, [keys][N] in FFT order with 1/N folded in: the n-point transfer
// function when N = n, else FFT_N of h = IDFT_n(H) extended periodically over |d| < n; FP64
// throughout, rounded once into tab (c64) and/or kept in tab_d (FP64).  filt: 1 or one per key
// grid of n filter samples (nullable).
static int line_tables(fh_ctx* ctx, int n, int lx, double dx, double n0, int mode, int track,
                       const std::vector<LineKey>& keys, const double2* filt, int nfilt, float2* tab,
                       double2* tab_d) {
  const int N = 1 << lx;
  const bool direct = N == n;
  hipStream_t st = ctx->stream;
  int rc = FH_OK;
  double2 *hn = nullptr, *tmp = nullptr, *wx = nullptr, *dbuf = nullptr;
  const double2* twN = nullptr;
  if (!direct) {
    if (hipMalloc(&hn, n * sizeof(double2)) != hipSuccess || hipMalloc(&tmp, n * sizeof(double2)) != hipSuccess ||
        hipMalloc(&wx, n * sizeof(double2)) != hipSuccess || hipMalloc(&dbuf, (size_t)N * sizeof(double2)) != hipSuccess)
      rc = fail(FH_E_NOMEM, "1-D tables: out of device memory");
    if (rc == FH_OK) rc = get_twiddles(ctx, N, &twN);
    if (rc == FH_OK) hipLaunchKernelGGL(k_unit_roots_d, dim3((n + 255) / 256), dim3(256), 0, st, wx, n);
  }
  const dim3 gN((N + 255) / 256), gn((n + 255) / 256);
  for (size_t k = 0; k < keys.size() && rc == FH_OK; ++k) {
    const double2* fk = filt ? filt + (nfilt == 1 ? 0 : k) * (size_t)n : nullptr;
    const int conjf = mode != 1 && !(keys[k].z > 0.0);
    if (direct) {
      if (tab)
        hipLaunchKernelGGL(k_as_line_table<float2>, gN, dim3(256), 0, st, tab + k * N, N, dx, keys[k].lam, n0,
                           keys[k].z, keys[k].tx, mode, track, 1.0 / N, fk, conjf);
      if (tab_d)
        hipLaunchKernelGGL(k_as_line_table<double2>, gN, dim3(256), 0, st, tab_d + k * N, N, dx, keys[k].lam, n0,
                           keys[k].z, keys[k].tx, mode, track, 1.0 / N, fk, conjf);
    } else {
      hipLaunchKernelGGL(k_as_grid_d, gn, dim3(256), 0, st, hn, n, 1, dx, 1.0, keys[k].lam, n0, keys[k].z,
                         keys[k].tx, 0.0, mode, track, fk, conjf);
      hipLaunchKernelGGL(k_idft_lines_d, gn, dim3(256), 0, st, (const double2*)hn, tmp, n, 1, (long)n, 1L,
                         (const double2*)wx);
      hipLaunchKernelGGL(k_circ_embed_d, gN, dim3(256), 0, st, dbuf, (const double2*)tmp, n, 1, N, 1);
      LineIO<double2> io{};
      io.in = dbuf; io.out = dbuf; io.in_l = io.out_l = N; io.in_e = io.out_e = 1;
      io.nlines = 1; io.in_valid = io.out_count = N;
      if (launch_fft<double2, double2, false, false>(lx, io, 1, twN, NoOp{}, st) != hipSuccess) {
        rc = fail(FH_E_HIP, "1-D table FFT failed to launch");
        break;
      }
      if (tab) hipLaunchKernelGGL(k_line_extract<float2>, gN, dim3(256), 0, st, tab + k * N, (const double2*)dbuf, N, 1.0 / N);
      if (tab_d)
        hipLaunchKernelGGL(k_line_extract<double2>, gN, dim3(256), 0, st, tab_d + k * N, (const double2*)dbuf, N, 1.0 / N);
    }
    if (hipGetLastError() != hipSuccess) rc = fail(FH_E_HIP, "1-D table kernels failed to launch");
  }
  // the buffers below (and the caller's filter samples) are read by the kernels: wait first
  if (hipStreamSynchronize(st) != hipSuccess && rc == FH_OK) rc = fail(FH_E_HIP, "1-D table generation failed");
  for (void* q : {(void*)hn, (void*)tmp, (void*)wx, (void*)dbuf})
    if (q) (void)hipFree(q);
  return rc;
}

// Uploads nfilters grids of nx*ny filter samples after checking them (nullptr when filter is).
static int upload_filter(const double* filter, int nfilters, int batch, size_t npix, double2** out) {
  *out = nullptr;
  if (!filter) return FH_OK;
  if (nfilters != 1 && nfilters != batch) return fail(FH_E_ARG, "need 1 or batch (=%d) filter grids, got %d", batch, nfilters);
  const size_t nf = (size_t)nfilters * npix;
  for (size_t k = 0; k < 2 * nf; ++k)
    if (!std::isfinite(filter[k])) return fail(FH_E_ARG, "filter samples must be finite");
  if (hipMalloc(out, nf * sizeof(double2)) != hipSuccess) return fail(FH_E_NOMEM, "filter samples: out of device memory");
  if (hipMemcpy(*out, filter, nf * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(*out);
    *out = nullptr;
    return fail(FH_E_HIP, "filter upload failed");
  }
  return FH_OK;
}

// Whether table k takes the conjugated filter: the reference's on-axis and tilted kernels for
// z <= 0 are conj(kernel(-z)) with the filter inside (angular_spectrum.jl:9-12,64-68); its
// paraxial kernel multiplies the filter as given for either sign (:27-35).
static int filter_conj(const fh_plan* p, long k) { return p->mode != 1 && !(p->tab_z[k] > 0.0); }

// Tables of a padded ASProp plan (grids that are not powers of two): per key, H on the
// n-point grid, its inverse 2-D DFT h (FP64, direct sums), h extended periodically onto the
// padded grid, the padded FFT (FP64), rounded once; the plan then runs the padded passes.
// fp64: also keep the tables unrounded in p->tab_d (FP64-compute users such as the general
// AS_BPM slice steps, whose hundreds of steps would otherwise repeat the same rounding).
static int as_padded_tables(fh_ctx* ctx, const fh_prop_desc* d, fh_plan* p, bool fp64,
                            const double2* filt = nullptr, int nfilt = 0) {
  const int nx = d->nx, ny = d->ny;
  const int nx2 = 1 << p->lx, ny2 = 1 << p->ly;
  p->kind = 1;
  p->qx = p->full ? nx2 : nx2 / 2 + 1;
  p->qy = p->full ? ny2 + 1 : ny2 / 2 + 1;
  p->tabsize = (long)p->qx * p->qy;
  const long nt = (long)p->lam_tables.size();
  if (hipMalloc(&p->tab, p->tabsize * nt * sizeof(float2)) != hipSuccess ||
      hipMalloc(&p->scratch, (size_t)nx2 * mid_rows(p) * d->batch * sizeof(float2)) != hipSuccess ||
      (fp64 && hipMalloc(&p->tab_d, p->tabsize * nt * sizeof(double2)) != hipSuccess))
    return fail(FH_E_NOMEM, "ASProp tables and workspace: out of device memory");
  double2 *hn = nullptr, *tmp = nullptr, *wx = nullptr, *wy = nullptr, *dbuf = nullptr;
  int rc = FH_OK;
  const size_t nn = (size_t)nx * ny;
  if (hipMalloc(&hn, nn * sizeof(double2)) != hipSuccess || hipMalloc(&tmp, nn * sizeof(double2)) != hipSuccess ||
      hipMalloc(&wx, nx * sizeof(double2)) != hipSuccess || hipMalloc(&wy, ny * sizeof(double2)) != hipSuccess ||
      hipMalloc(&dbuf, (size_t)nx2 * ny2 * sizeof(double2)) != hipSuccess)
    rc = fail(FH_E_NOMEM, "ASProp FP64 table buffers: out of device memory");
  if (rc == FH_OK) {
    hipLaunchKernelGGL(k_unit_roots_d, dim3((nx + 255) / 256), dim3(256), 0, ctx->stream, wx, nx);
    hipLaunchKernelGGL(k_unit_roots_d, dim3((ny + 255) / 256), dim3(256), 0, ctx->stream, wy, ny);
  }
  for (long k = 0; k < nt && rc == FH_OK; ++k) {
    hipLaunchKernelGGL(k_as_grid_d, dim3((nx + 255) / 256, ny), dim3(256), 0, ctx->stream, hn, nx, ny, d->dx, d->dy,
                       p->lam_tables[k], d->n0, p->tab_z[k], p->tab_tx[k], p->tab_ty[k], p->mode, p->track,
                       filt ? filt + (nfilt == 1 ? 0 : k) * (long)nx * ny : nullptr, filter_conj(p, k));
    hipLaunchKernelGGL(k_idft_lines_d, dim3((nx + 255) / 256, ny), dim3(256), 0, ctx->stream, (const double2*)hn, tmp,
                       nx, ny, (long)nx, 1L, (const double2*)wx);
    hipLaunchKernelGGL(k_idft_lines_d, dim3((ny + 255) / 256, nx), dim3(256), 0, ctx->stream, (const double2*)tmp, hn,
                       ny, nx, 1L, (long)nx, (const double2*)wy);
    hipLaunchKernelGGL(k_circ_embed_d, dim3((nx2 + 255) / 256, ny2), dim3(256), 0, ctx->stream, dbuf,
                       (const double2*)hn, nx, ny, nx2, ny2);
    if (hipGetLastError() != hipSuccess) { rc = fail(FH_E_HIP, "ASProp padded table kernels failed to launch"); break; }
    if ((rc = fft2_d(ctx, dbuf, nx2, ny2, p->lx, p->ly))) break;
    hipLaunchKernelGGL(k_tab_extract<float2>, dim3((p->qx + 255) / 256, p->qy), dim3(256), 0, ctx->stream,
                       p->tab + k * p->tabsize, (const double2*)dbuf, nx2, ny2, p->qx, p->qy, p->full,
                       1.0 / ((double)nx2 * (double)ny2));
    if (fp64)
      hipLaunchKernelGGL(k_tab_extract<double2>, dim3((p->qx + 255) / 256, p->qy), dim3(256), 0, ctx->stream,
                         p->tab_d + k * p->tabsize, (const double2*)dbuf, nx2, ny2, p->qx, p->qy, p->full,
                         1.0 / ((double)nx2 * (double)ny2));
    if (hipGetLastError() != hipSuccess) rc = fail(FH_E_HIP, "ASProp table kernel launch failed");
  }
  if (rc == FH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(FH_E_HIP, "ASProp table generation failed");
  for (void* q : {(void*)hn, (void*)tmp, (void*)wx, (void*)wy, (void*)dbuf})
    if (q) (void)hipFree(q);
  if (rc == FH_OK) rc = upload_tabidx(p);
  return rc;
}

// One table per field (per-field filter samples): the keys of plan_common are undone.
static void plan_table_per_field(fh_plan* p) {
  std::vector<double> l(p->batch), z(p->batch), tx(p->batch), ty(p->batch);
  for (int b = 0; b < p->batch; ++b) {
    const int k = p->field_table[b];
    l[b] = p->lam_tables[k]; z[b] = p->tab_z[k]; tx[b] = p->tab_tx[k]; ty[b] = p->tab_ty[k];
    p->field_table[b] = b;
  }
  p->lam_tables = l; p->tab_z = z; p->tab_tx = tx; p->tab_ty = ty;
}

// A 1-D ASProp / ShiftProp plan (ny == 1): the reference's kernel choice with theta_y = 0.
static int line_plan(fh_ctx* ctx, const fh_prop_desc* d, fh_plan* p, bool fp64, const double* filter, int nfilters) {
  const int nx = d->nx;
  for (double ty : p->tab_ty)
    if (ty != 0.0) return fail(FH_E_ARG, "a 1-D field has no y tilt");
  const bool pow2 = ilog2_exact(nx) >= 0;
  p->lx = pow2 ? ilog2_exact(nx) : ilog2_ceil(2 * nx - 1);
  p->ly = 0;
  p->oned = 1;
  if (nx < 2 || p->lx < kMinLog || p->lx > kMaxLog)
    return fail(FH_E_UNSUPP, "1-D ASProp needs power-of-two nx in [%d, %d], or another nx with 2 nx - 1 <= %d "
                "(got %d)", 1 << kMinLog, 1 << kMaxLog, 1 << kMaxLog, nx);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const bool tilted = plan_tilted(p);
  p->mode = (d->flags & FH_PROP_SHIFT) ? 3 : ((d->flags & FH_PROP_PARAXIAL) ? 1 : (tilted ? 2 : 0));
  p->track = (d->flags & FH_PROP_TRACK_TILTS) ? 1 : 0;
  p->full = 1;
  if (filter && p->mode == 3) return fail(FH_E_ARG, "ShiftProp takes no frequency filter");
  double2* filt = nullptr;
  if ((rc = upload_filter(filter, nfilters, d->batch, (size_t)nx, &filt))) return rc;
  if (filter) {
    if (nfilters > 1) plan_table_per_field(p);
    p->filtered = 1;
  }
  const int N = 1 << p->lx;
  p->qx = N; p->qy = 1;
  p->tabsize = N;
  const size_t ntab = p->lam_tables.size();
  if (hipMalloc(&p->tab, ntab * N * sizeof(float2)) != hipSuccess ||
      (fp64 && hipMalloc(&p->tab_d, ntab * N * sizeof(double2)) != hipSuccess))
    rc = fail(FH_E_NOMEM, "1-D ASProp tables: out of device memory");
  std::vector<LineKey> keys;
  for (size_t k = 0; k < ntab; ++k) keys.push_back({p->lam_tables[k], p->tab_z[k], p->tab_tx[k]});
  if (rc == FH_OK) rc = line_tables(ctx, nx, p->lx, d->dx, d->n0, p->mode, p->track, keys, filt, nfilters, p->tab, p->tab_d);
  if (filt) (void)hipFree(filt);
  if (rc == FH_OK) rc = upload_tabidx(p);
  return rc;
}

// One propagation of a 1-D plan: every field a line, one launch.
static int run_line_prop(fh_plan* p, float2* u, int direction) {
  const float2* tw = nullptr;
  int rc;
  if ((rc = get_twiddles(p->ctx, 1 << p->lx, &tw))) return rc;
  LineIO<float2> io{};
  io.in = u; io.out = u; io.in_f = io.out_f = 0;
  io.in_l = io.out_l = p->nx; io.in_e = io.out_e = 1;
  io.nlines = p->batch; io.in_valid = io.out_count = p->nx;
  const LineTabMulT<float2> op{p->tab, p->d_tabidx, 1 << p->lx, direction < 0};
  FH_HIP((launch_line_conv<float2, float2>(p->lx, io, tw, op, p->ctx->stream)));
  return FH_OK;
}

static int asprop_create(fh_ctx* ctx, const fh_prop_desc* d, fh_plan** out, bool fp64_tables,
                         const double* filter = nullptr, int nfilters = 0) {
  if (!out || !d) return fail(FH_E_ARG, "null descriptor or output pointer");
  *out = nullptr;
  if (d->flags & ~(FH_PROP_PARAXIAL | FH_PROP_TRACK_TILTS | FH_PROP_SHIFT)) return fail(FH_E_ARG, "unknown flags 0x%x", d->flags);
  if ((d->flags & FH_PROP_SHIFT) && (d->flags & FH_PROP_PARAXIAL))
    return fail(FH_E_ARG, "FH_PROP_SHIFT and FH_PROP_PARAXIAL exclude each other");
  if (!d->z) return fail(FH_E_ARG, "null distances");
  fh_plan* p = new fh_plan();
  int rc = plan_common(ctx, d->nx, d->ny, d->batch, d->dx, d->dy, d->lambdas, d->nlambdas, p, d->z, d->nz,
                       d->tilts_x, d->tilts_y, d->ntilts);
  if (rc) { plan_free(p); return rc; }
  const int nx = d->nx, ny = d->ny;
  if (!(d->n0 > 0)) { plan_free(p); return fail(FH_E_ARG, "n0 must be positive"); }
  p->kind = 0; p->z = p->tab_z[0]; p->n0 = d->n0;
  if (ny == 1) {
    rc = line_plan(ctx, d, p, fp64_tables, filter, nfilters);
    if (rc) { plan_free(p); return rc; }
    *out = p;
    return FH_OK;
  }
  const bool pow2 = ilog2_exact(nx) >= 0 && ilog2_exact(ny) >= 0;
  if (pow2) {
    p->lx = ilog2_exact(nx); p->ly = ilog2_exact(ny);
  } else {
    // any other grid: the padded convolution form (as_padded_tables)
    p->lx = ilog2_ceil(2 * nx - 1); p->ly = ilog2_ceil(2 * ny - 1);
  }
  if (nx < 2 || ny < 2 || p->lx < kMinLog || p->lx > kMaxLog || p->ly < kMinLog || p->ly > kMaxLog) {
    plan_free(p);
    return fail(FH_E_UNSUPP, "ASProp needs power-of-two nx, ny in [%d, %d], or other sizes with 2n - 1 <= %d "
                "(got %d x %d)", 1 << kMinLog, 1 << kMaxLog, 1 << kMaxLog, nx, ny);
  }
  if ((rc = set_device(ctx))) { plan_free(p); return rc; }
  // kernel choice of angular_spectrum.jl:126-139; tilted tables are not even in fx, fy
  const bool tilted = plan_tilted(p);
  p->mode = (d->flags & FH_PROP_SHIFT) ? 3 : ((d->flags & FH_PROP_PARAXIAL) ? 1 : (tilted ? 2 : 0));
  p->track = (d->flags & FH_PROP_TRACK_TILTS) ? 1 : 0;
  p->full = tilted ? 1 : 0;
  // frequency filter: samples are not even in fx, fy, so the tables take the full layout
  double2* filt = nullptr;
  if (filter && p->mode == 3) { plan_free(p); return fail(FH_E_ARG, "ShiftProp takes no frequency filter"); }
  if ((rc = upload_filter(filter, nfilters, d->batch, (size_t)nx * ny, &filt))) { plan_free(p); return rc; }
  if (filter) {
    if (nfilters > 1) plan_table_per_field(p);
    p->full = 1;
    p->filtered = 1;
  }
  if (!pow2) {
    rc = as_padded_tables(ctx, d, p, fp64_tables, filt, nfilters);
    if (filt) (void)hipFree(filt);
    if (rc) { plan_free(p); return rc; }
    *out = p;
    return FH_OK;
  }
  p->qx = p->full ? nx : nx / 2 + 1;
  p->qy = p->full ? ny + 1 : ny / 2 + 1;
  p->tabsize = (long)p->qx * p->qy;
  const size_t bytes = p->tabsize * p->lam_tables.size() * sizeof(float2);
  if (hipMalloc(&p->tab, bytes) != hipSuccess) {
    if (filt) (void)hipFree(filt);
    plan_free(p);
    return fail(FH_E_NOMEM, "ASProp table: out of device memory");
  }
  const double scale = 1.0 / ((double)nx * (double)ny);
  for (size_t k = 0; k < p->lam_tables.size() && rc == FH_OK; ++k) {
    dim3 grid((p->qx + 255) / 256, p->qy);
    hipLaunchKernelGGL(k_as_table<float2>, grid, dim3(256), 0, ctx->stream, p->tab + k * p->tabsize, nx, ny,
                       p->qx, p->qy, p->full, d->dx, d->dy, p->lam_tables[k], d->n0, p->tab_z[k],
                       p->tab_tx[k], p->tab_ty[k], p->mode, p->track, scale,
                       (const double2*)(filt ? filt + (nfilters == 1 ? 0 : k) * (size_t)nx * ny : nullptr),
                       filter_conj(p, (long)k));
    if (hipGetLastError() != hipSuccess) rc = fail(FH_E_HIP, "ASProp table kernel launch failed");
  }
  if (filt) {  // the table kernels read the samples: wait before freeing them
    if (rc == FH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(FH_E_HIP, "ASProp table generation failed");
    (void)hipFree(filt);
  }
  if (rc == FH_OK) rc = upload_tabidx(p);
  if (rc) { plan_free(p); return rc; }
  *out = p;
  return FH_OK;
}

int fh_asprop_create_ex(fh_ctx* ctx, const fh_prop_desc* d, fh_plan** out) {
  return asprop_create(ctx, d, out, false);
}

int fh_asprop_create_filtered(fh_ctx* ctx, const fh_prop_desc* d, const double* filter, int nfilters, fh_plan** out) {
  if (!filter) return fail(FH_E_ARG, "null filter samples");
  return asprop_create(ctx, d, out, false, filter, nfilters);
}

int fh_asprop_create(fh_ctx* ctx, int nx, int ny, int batch, double dx, double dy,
                     const double* lambdas, int nlambdas, double z, double n0, fh_plan** out) {
  fh_prop_desc d{};
  d.nx = nx; d.ny = ny; d.batch = batch; d.dx = dx; d.dy = dy;
  d.lambdas = lambdas; d.nlambdas = nlambdas; d.z = &z; d.nz = 1; d.n0 = n0;
  return fh_asprop_create_ex(ctx, &d, out);
}

int fh_rsprop_create_ex(fh_ctx* ctx, const fh_prop_desc* d, fh_plan** out) {
  if (!out || !d) return fail(FH_E_ARG, "null descriptor or output pointer");
  *out = nullptr;
  if (d->flags & ~FH_PROP_TRACK_TILTS) return fail(FH_E_ARG, "RSProp takes FH_PROP_TRACK_TILTS only (flags 0x%x)", d->flags);
  if (!d->z) return fail(FH_E_ARG, "null distances");
  fh_plan* p = new fh_plan();
  int rc = plan_common(ctx, d->nx, d->ny, d->batch, d->dx, d->dy, d->lambdas, d->nlambdas, p, d->z, d->nz,
                       d->tilts_x, d->tilts_y, d->ntilts);
  if (rc) { plan_free(p); return rc; }
  for (double z : p->tab_z)
    if (z == 0.0) { plan_free(p); return fail(FH_E_ARG, "RSProp needs a finite, nonzero z"); }
  const int nx = d->nx, ny = d->ny;
  p->kind = 1; p->z = p->tab_z[0];
  // zero padding to the power of two >= 2n - 1 (the reference pads to exactly 2n - 1,
  // rayleigh_sommerfeld.jl:135-156; any length >= 2n - 1 gives the same cropped linear
  // convolution), so any n up to half the longest transform
  p->lx = ilog2_ceil(2 * nx - 1); p->ly = ilog2_ceil(2 * ny - 1);
  if (nx < 2 || ny < 2 || p->lx < kMinLog || p->lx > kMaxLog || p->ly < kMinLog || p->ly > kMaxLog) {
    plan_free(p);
    return fail(FH_E_UNSUPP, "RSProp needs 2 <= n with 2n - 1 <= %d and a padded length >= %d (got %d x %d)",
                1 << kMaxLog, 1 << kMinLog, nx, ny);
  }
  if ((rc = set_device(ctx))) { plan_free(p); return rc; }
  const int nx2 = 1 << p->lx, ny2 = 1 << p->ly;
  // tilted kernels (rayleigh_sommerfeld.jl:90-96) carry a linear phase unless tracked: odd spectra
  const bool tilted = plan_tilted(p);
  p->track = (d->flags & FH_PROP_TRACK_TILTS) ? 1 : 0;
  p->mode = tilted ? 2 : 0;
  p->full = (tilted && !p->track) ? 1 : 0;
  p->qx = p->full ? nx2 : nx2 / 2 + 1;
  p->qy = p->full ? ny2 + 1 : ny2 / 2 + 1;
  p->tabsize = (long)p->qx * p->qy;
  if (hipMalloc(&p->tab, p->tabsize * p->lam_tables.size() * sizeof(float2)) != hipSuccess) {
    plan_free(p);
    return fail(FH_E_NOMEM, "RSProp table: out of device memory");
  }
  if (hipMalloc(&p->scratch, (size_t)nx2 * mid_rows(p) * d->batch * sizeof(float2)) != hipSuccess) {
    plan_free(p);
    return fail(FH_E_NOMEM, "RSProp workspace: out of device memory");
  }
  double2* dd = nullptr;
  if (hipMalloc(&dd, (size_t)nx2 * ny2 * sizeof(double2)) != hipSuccess) {
    plan_free(p);
    return fail(FH_E_NOMEM, "RSProp FP64 kernel buffer: out of device memory");
  }
  const double scale = 1.0 / ((double)nx2 * (double)ny2);
  const double nrm = d->dx * d->dy / (2.0 * M_PI);
  for (size_t k = 0; k < p->lam_tables.size() && rc == FH_OK; ++k) {
    const int tilt = (p->tab_tx[k] != 0.0 || p->tab_ty[k] != 0.0) && !p->track;
    hipLaunchKernelGGL(k_rs_sample, dim3((nx2 + 255) / 256, ny2), dim3(256), 0, ctx->stream, dd, nx2, ny2,
                       d->dx, d->dy, p->lam_tables[k], p->tab_z[k], nrm, p->tab_tx[k], p->tab_ty[k], tilt);
    if (hipGetLastError() != hipSuccess) { rc = fail(FH_E_HIP, "RS sample kernel launch failed"); break; }
    if ((rc = fft2_d(ctx, dd, nx2, ny2, p->lx, p->ly))) break;
    hipLaunchKernelGGL(k_tab_extract<float2>, dim3((p->qx + 255) / 256, p->qy), dim3(256), 0, ctx->stream,
                       p->tab + k * p->tabsize, dd, nx2, ny2, p->qx, p->qy, p->full, scale);
    if (hipGetLastError() != hipSuccess) { rc = fail(FH_E_HIP, "RS table kernel launch failed"); break; }
  }
  if (rc == FH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(FH_E_HIP, "RS table generation failed");
  (void)hipFree(dd);
  if (rc == FH_OK) rc = upload_tabidx(p);
  if (rc) { plan_free(p); return rc; }
  *out = p;
  return FH_OK;
}

int fh_collins_create(fh_ctx* ctx, const fh_prop_desc* d, double dx_out, double dy_out, double a, double b,
                      double dd, fh_plan** out) {
  if (!out || !d) return fail(FH_E_ARG, "null descriptor or output pointer");
  *out = nullptr;
  if (d->flags & ~FH_PROP_TRACK_TILTS) return fail(FH_E_ARG, "CollinsProp takes FH_PROP_TRACK_TILTS only (flags 0x%x)", d->flags);
  if (d->z || d->nz) return fail(FH_E_ARG, "CollinsProp takes (A, B, D), not distances (z must be null)");
  if (!std::isfinite(a) || !std::isfinite(b) || !std::isfinite(dd) || b == 0.0)
    return fail(FH_E_ARG, "CollinsProp needs finite A, B, D with B != 0");
  if (!(dx_out > 0) || !(dy_out > 0) || !std::isfinite(dx_out) || !std::isfinite(dy_out))
    return fail(FH_E_ARG, "output sampling must be positive");
  fh_plan* p = new fh_plan();
  int rc = plan_common(ctx, d->nx, d->ny, d->batch, d->dx, d->dy, d->lambdas, d->nlambdas, p, nullptr, 0,
                       d->tilts_x, d->tilts_y, d->ntilts);
  if (rc) { plan_free(p); return rc; }
  const int nx = d->nx, ny = d->ny;
  p->kind = 2;
  p->lx = ilog2_ceil(2 * nx - 1); p->ly = ilog2_ceil(2 * ny - 1);
  if (nx < 2 || ny < 2 || p->lx < kMinLog || p->lx > kMaxLog || p->ly < kMinLog || p->ly > kMaxLog) {
    plan_free(p);
    return fail(FH_E_UNSUPP, "CollinsProp needs 2 <= n with 2n - 1 <= %d (got %d x %d)", 1 << kMaxLog, nx, ny);
  }
  if ((rc = set_device(ctx))) { plan_free(p); return rc; }
  const int nx2 = 1 << p->lx, ny2 = 1 << p->ly;
  const double ax = dx_out / d->dx, ay = dy_out / d->dy;  // alpha = ds'/ds (collins_integral.jl:72)
  p->alpha_x = ax; p->alpha_y = ay;
  p->track = (d->flags & FH_PROP_TRACK_TILTS) ? 1 : 0;
  // the convolution kernel is even in x and y: quadrant tables
  p->qx = nx2 / 2 + 1; p->qy = ny2 / 2 + 1;
  p->tabsize = (long)p->qx * p->qy;
  const long nt = (long)p->lam_tables.size();
  if (hipMalloc(&p->tab, p->tabsize * nt * sizeof(float2)) != hipSuccess ||
      hipMalloc(&p->chirp, nt * 2 * (nx + ny) * sizeof(float2)) != hipSuccess ||
      hipMalloc(&p->scratch, (size_t)nx2 * mid_rows(p) * d->batch * sizeof(float2)) != hipSuccess) {
    plan_free(p);
    return fail(FH_E_NOMEM, "CollinsProp tables and workspace: out of device memory");
  }
  double2* dbuf = nullptr;
  if (hipMalloc(&dbuf, (size_t)nx2 * ny2 * sizeof(double2)) != hipSuccess) {
    plan_free(p);
    return fail(FH_E_NOMEM, "CollinsProp FP64 kernel buffer: out of device memory");
  }
  const double scale = 1.0 / ((double)nx2 * (double)ny2);
  const double nrm_im = -d->dx * d->dy / b;  // dx*dy/(i*b)
  const bool tilted = plan_tilted(p);
  float2* cax = p->chirp;
  float2* cay = cax + nt * nx;
  float2* cdx = cay + nt * ny;
  float2* cdy = cdx + nt * nx;
  for (long k = 0; k < nt && rc == FH_OK; ++k) {
    const double lam = p->lam_tables[k];
    hipLaunchKernelGGL(k_collins_sample, dim3((nx2 + 255) / 256, ny2), dim3(256), 0, ctx->stream, dbuf, nx2, ny2,
                       d->dx, d->dy, lam, ax, ay, b, 0.0, nrm_im);
    if (hipGetLastError() != hipSuccess) { rc = fail(FH_E_HIP, "Collins sample kernel launch failed"); break; }
    if ((rc = fft2_d(ctx, dbuf, nx2, ny2, p->lx, p->ly))) break;
    hipLaunchKernelGGL(k_tab_extract<float2>, dim3((p->qx + 255) / 256, p->qy), dim3(256), 0, ctx->stream,
                       p->tab + k * p->tabsize, dbuf, nx2, ny2, p->qx, p->qy, 0, scale);
    // chirps (collins_integral.jl:5-36) on the input grid, in cycles: A = cis(pi*x^2*(a -
    // alpha)/(b*lam))/lam, D = cis(pi*x^2*alpha*(d*alpha - 1)/(b*lam)); a tilted field whose
    // tilts are not tracked adds cis(+-2*pi*x*sin(theta)/lam) to A (+) and D (-)
    const bool lin = tilted && !p->track;
    const double fx = lin ? sin(p->tab_tx[k]) / lam : 0.0, fy = lin ? sin(p->tab_ty[k]) / lam : 0.0;
    const double bl = 2.0 * b * lam;
    hipLaunchKernelGGL(k_chirp_line, dim3((nx + 255) / 256), dim3(256), 0, ctx->stream, cax + k * nx, nx, d->dx,
                       (a - ax) / bl, fx, 1.0 / lam);
    hipLaunchKernelGGL(k_chirp_line, dim3((ny + 255) / 256), dim3(256), 0, ctx->stream, cay + k * ny, ny, d->dy,
                       (a - ay) / bl, fy, 1.0);
    hipLaunchKernelGGL(k_chirp_line, dim3((nx + 255) / 256), dim3(256), 0, ctx->stream, cdx + k * nx, nx, d->dx,
                       ax * (dd * ax - 1.0) / bl, -fx, 1.0);
    hipLaunchKernelGGL(k_chirp_line, dim3((ny + 255) / 256), dim3(256), 0, ctx->stream, cdy + k * ny, ny, d->dy,
                       ay * (dd * ay - 1.0) / bl, -fy, 1.0);
    if (hipGetLastError() != hipSuccess) { rc = fail(FH_E_HIP, "Collins table kernel launch failed"); break; }
  }
  if (rc == FH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(FH_E_HIP, "Collins table generation failed");
  (void)hipFree(dbuf);
  if (rc == FH_OK) rc = upload_tabidx(p);
  if (rc) { plan_free(p); return rc; }
  *out = p;
  return FH_OK;
}

int fh_czt_create(fh_ctx* ctx, int nx, int ny, int batch, int dims, const double* a, const double* w,
                  int center_on_grid, fh_plan** out) {
  if (!out) return fail(FH_E_ARG, "null output pointer");
  *out = nullptr;
  if (dims < 1 || dims > 3) return fail(FH_E_ARG, "dims must be 1 (x), 2 (y) or 3 (both), got %d", dims);
  if (!a || !w) return fail(FH_E_ARG, "null a or w");
  CztDim dd[2];
  for (int k = 0; k < 2; ++k) {
    const double ra = a[2 * k], ta = a[2 * k + 1], rw = w[2 * k], tw = w[2 * k + 1];
    dd[k].on = (dims >> k) & 1;
    dd[k].center = center_on_grid ? 1 : 0;
    if (dd[k].on && (!(ra > 0) || !(rw > 0) || !std::isfinite(ra) || !std::isfinite(rw) || !std::isfinite(ta) ||
                     !std::isfinite(tw)))
      return fail(FH_E_ARG, "CZT needs finite a and w with positive magnitudes");
    dd[k].la = dd[k].on ? log(ra) : 0.0; dd[k].ta = dd[k].on ? ta : 0.0;
    dd[k].lw = dd[k].on ? log(rw) : 0.0; dd[k].tw = dd[k].on ? tw : 0.0;
  }
  const double one = 1.0;
  fh_plan* p = new fh_plan();
  int rc = plan_common(ctx, nx, ny, batch, 1.0, 1.0, &one, 1, p);
  if (rc) { plan_free(p); return rc; }
  p->kind = 2;  // the padded chirp / convolution / chirp plan of CollinsProp, with unit scale
  p->lx = ilog2_ceil(2 * nx - 1); p->ly = ilog2_ceil(2 * ny - 1);
  if (nx < 2 || ny < 2 || p->lx < kMinLog || p->lx > kMaxLog || p->ly < kMinLog || p->ly > kMaxLog) {
    plan_free(p);
    return fail(FH_E_UNSUPP, "CZT needs 2 <= n with 2n - 1 <= %d (got %d x %d)", 1 << kMaxLog, nx, ny);
  }
  if ((rc = set_device(ctx))) { plan_free(p); return rc; }
  const int nx2 = 1 << p->lx, ny2 = 1 << p->ly;
  p->qx = nx2 / 2 + 1; p->qy = ny2 / 2 + 1;  // h depends on f^2: even, quadrant tables
  p->tabsize = (long)p->qx * p->qy;
  if (hipMalloc(&p->tab, p->tabsize * sizeof(float2)) != hipSuccess ||
      hipMalloc(&p->chirp, 2 * (nx + ny) * sizeof(float2)) != hipSuccess ||
      hipMalloc(&p->scratch, (size_t)nx2 * mid_rows(p) * batch * sizeof(float2)) != hipSuccess) {
    plan_free(p);
    return fail(FH_E_NOMEM, "CZT tables and workspace: out of device memory");
  }
  double2* dbuf = nullptr;
  if (hipMalloc(&dbuf, (size_t)nx2 * ny2 * sizeof(double2)) != hipSuccess) {
    plan_free(p);
    return fail(FH_E_NOMEM, "CZT FP64 kernel buffer: out of device memory");
  }
  hipLaunchKernelGGL(k_czt_sample, dim3((nx2 + 255) / 256, ny2), dim3(256), 0, ctx->stream, dbuf, nx, ny, nx2, ny2,
                     dd[0], dd[1]);
  if (hipGetLastError() != hipSuccess) rc = fail(FH_E_HIP, "CZT sample kernel launch failed");
  if (rc == FH_OK) rc = fft2_d(ctx, dbuf, nx2, ny2, p->lx, p->ly);
  if (rc == FH_OK) {
    hipLaunchKernelGGL(k_tab_extract<float2>, dim3((p->qx + 255) / 256, p->qy), dim3(256), 0, ctx->stream, p->tab, dbuf,
                       nx2, ny2, p->qx, p->qy, 0, 1.0 / ((double)nx2 * (double)ny2));
    // chirp lines in the plan's A_x, A_y (input) and D_x, D_y (output) slots
    float2* cax = p->chirp;
    float2* cay = cax + nx;
    float2* cdx = cay + ny;
    float2* cdy = cdx + nx;
    hipLaunchKernelGGL(k_czt_chirp, dim3((nx + 255) / 256), dim3(256), 0, ctx->stream, cax, nx, dd[0], 1);
    hipLaunchKernelGGL(k_czt_chirp, dim3((ny + 255) / 256), dim3(256), 0, ctx->stream, cay, ny, dd[1], 1);
    hipLaunchKernelGGL(k_czt_chirp, dim3((nx + 255) / 256), dim3(256), 0, ctx->stream, cdx, nx, dd[0], 0);
    hipLaunchKernelGGL(k_czt_chirp, dim3((ny + 255) / 256), dim3(256), 0, ctx->stream, cdy, ny, dd[1], 0);
    if (hipGetLastError() != hipSuccess) rc = fail(FH_E_HIP, "CZT table kernel launch failed");
  }
  if (rc == FH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(FH_E_HIP, "CZT table generation failed");
  (void)hipFree(dbuf);
  if (rc == FH_OK) rc = upload_tabidx(p);
  if (rc) { plan_free(p); return rc; }
  *out = p;
  return FH_OK;
}

int fh_rsprop_create(fh_ctx* ctx, int nx, int ny, int batch, double dx, double dy,
                     const double* lambdas, int nlambdas, double z, fh_plan** out) {
  fh_prop_desc d{};
  d.nx = nx; d.ny = ny; d.batch = batch; d.dx = dx; d.dy = dy;
  d.lambdas = lambdas; d.nlambdas = nlambdas; d.z = &z; d.nz = 1;
  return fh_rsprop_create_ex(ctx, &d, out);
}

int fh_plan_destroy(fh_plan* plan) {
  if (!plan) return FH_OK;
  (void)hipSetDevice(plan->ctx->device);
  plan_free(plan);
  return FH_OK;
}

int fh_propagate(fh_plan* p, void* u, int direction) {
  if (!p || !u) return fail(FH_E_ARG, "null plan or field");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(p->ctx))) return rc;
  float2* f = reinterpret_cast<float2*>(u);
  if (p->oned) return run_line_prop(p, f, direction);
  if ((rc = run_prop_plan(p, f, direction))) return rc;
  return run_xifft(p, f, direction);
}

int fh_propagate_pass(fh_plan* p, void* u, int direction, int pass) {
  if (!p || !u) return fail(FH_E_ARG, "null plan or field");
  if (pass < 0 || pass > 2) return fail(FH_E_ARG, "pass must be 0, 1 or 2, got %d", pass);
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(p->ctx))) return rc;
  float2* f = reinterpret_cast<float2*>(u);
  if (p->oned) return pass == 1 ? run_line_prop(p, f, direction) : FH_OK;  // one pass does it all
  if (pass < 2) return run_prop_plan(p, f, direction, pass);
  return run_xifft(p, f, direction);
}

int fh_plan_variants(fh_plan* p, int pass, int* count, int* current) {
  if (!p || !count || !current) return fail(FH_E_ARG, "null plan or output");
  if (pass < 0 || pass > 2) return fail(FH_E_ARG, "pass must be 0, 1 or 2, got %d", pass);
  int n = kNumXVariants;
  if (pass == 1) fast_conv_variants(p->ly, &n);
  if (p->oned) n = 1;
  *count = n > 0 ? n : 1;
  *current = pass == 1 ? p->conv_var : (pass == 0 ? p->xf_var : p->xi_var);
  return FH_OK;
}

int fh_plan_set_variant(fh_plan* p, int pass, int variant) {
  if (!p) return fail(FH_E_ARG, "null plan");
  if (pass < 0 || pass > 2) return fail(FH_E_ARG, "pass must be 0, 1 or 2, got %d", pass);
  int n = kNumXVariants;
  if (pass == 1) fast_conv_variants(p->ly, &n);
  if (p->oned) n = 1;
  if (variant < -1 || variant >= std::max(n, 1)) return fail(FH_E_ARG, "variant %d out of range (%d)", variant, n);
  if (pass == 1) p->conv_var = variant;
  else if (pass == 0) p->xf_var = variant;
  else p->xi_var = variant;
  return FH_OK;
}

int fh_autotune(fh_plan* p, void* u, int reps, char* report, int report_len) {
  if (!p || !u) return fail(FH_E_ARG, "null plan or field");
  if (reps < 1) return fail(FH_E_ARG, "reps must be >= 1");
  int rc;
  if ((rc = set_device(p->ctx))) return rc;
  if (p->oned) {  // one launch shape: nothing to choose
    if (report && report_len > 0) {
      strncpy(report, "{}", report_len - 1);
      report[report_len - 1] = 0;
    }
    return FH_OK;
  }
  float2* f = reinterpret_cast<float2*>(u);
  hipEvent_t e0, e1;
  FH_HIP(hipEventCreate(&e0));
  FH_HIP(hipEventCreate(&e1));
  std::string js = "{";
  static const char* const kPassName[] = {"x_fft", "y_fft_mul_ifft", "x_ifft"};
  for (int pass = 0; pass < 3 && rc == FH_OK; ++pass) {
    int n = kNumXVariants;
    const FastConvVariant* vars = pass == 1 ? fast_conv_variants(p->ly, &n) : nullptr;
    int* sel = pass == 0 ? &p->xf_var : (pass == 1 ? &p->conv_var : &p->xi_var);
    auto run = [&]() { return pass == 2 ? run_xifft(p, f, 1) : run_prop_plan(p, f, 1, pass); };
    char buf[200];
    snprintf(buf, sizeof buf, "%s\"%s\": {\"variants\": [", pass ? ", " : "", kPassName[pass]);
    js += buf;
    int best = -1;
    float best_ms = 0.f;
    for (int v = 0; v < n && rc == FH_OK; ++v) {
      *sel = v;
      if ((rc = run())) break;  // warm-up
      if (hipEventRecord(e0, p->ctx->stream) != hipSuccess) { rc = fail(FH_E_HIP, "event record failed"); break; }
      for (int r = 0; r < reps && rc == FH_OK; ++r) rc = run();
      if (rc) break;
      float ms = 0.f;
      if (hipEventRecord(e1, p->ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = fail(FH_E_HIP, "event timing failed"); break; }
      ms /= reps;
      snprintf(buf, sizeof buf, "%s{\"index\": %d, \"name\": \"%s\", \"ms\": %.6f}", v ? ", " : "", v,
               pass == 1 ? vars[v].name : kXVariants[v], ms);
      js += buf;
      if (best < 0 || ms < best_ms) { best = v; best_ms = ms; }
    }
    *sel = rc ? -1 : best;
    snprintf(buf, sizeof buf, "], \"best\": %d}", best);
    js += buf;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  js += "}";
  if (report && report_len > 0) {
    strncpy(report, js.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  return FH_OK;
}

// The Phase pass's operator: cis(s*phi) into the plan's factor array (allocated at the first call;
// its size is fixed by the plan, so graphs recorded later keep a valid pointer), then the op.
static int phase_time_op(fh_plan* p, const float* phi, void* saved, int direction, PhaseTimeOp* op) {
  const long npix = (long)p->nx * p->ny;
  if (!p->cis && hipMalloc(&p->cis, npix * sizeof(float2)) != hipSuccess)
    return fail(FH_E_NOMEM, "Phase factors: out of device memory");
  hipLaunchKernelGGL(k_phase_cis, dim3((int)std::min<long>((npix + 255) / 256, 4096)), dim3(256), 0, p->ctx->stream, phi,
                     (float)direction, p->cis, npix);
  FH_HIP(hipGetLastError());
  *op = PhaseTimeOp{p->cis, reinterpret_cast<float2*>(saved), npix, p->nx};
  return FH_OK;
}

int fh_phase_propagate(fh_plan* p, void* u, const float* phi, void* saved, int direction) {
  if (!p || !u || !phi) return fail(FH_E_ARG, "null plan, field or mask");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(p->ctx))) return rc;
  float2* f = reinterpret_cast<float2*>(u);
  if (p->kind == 2) return fail(FH_E_UNSUPP, "Phase fuses with ASProp and RSProp plans only");
  if (p->oned) {  // 1-D: the Phase, then the one-pass propagation
    if ((rc = fh_phase_apply(p->ctx, u, phi, p->nx, 1, p->batch, direction, saved))) return rc;
    return run_line_prop(p, f, direction);
  }
  PhaseTimeOp op{};
  if ((rc = phase_time_op(p, phi, saved, direction, &op)) || (rc = run_prop(p, f, direction, op))) return rc;
  return run_xifft(p, f, direction);
}

int fh_window(fh_ctx* ctx, const void* src, int snx, int sny, void* dst, int dnx, int dny, int batch, int x0,
              int y0, double fill_re, double fill_im) {
  if (!ctx || !src || !dst) return fail(FH_E_ARG, "null context, source or destination");
  if (snx <= 0 || sny <= 0 || dnx <= 0 || dny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  const size_t sb = (size_t)snx * sny * batch * sizeof(float2), db = (size_t)dnx * dny * batch * sizeof(float2);
  const char *a = static_cast<const char*>(src), *d = static_cast<const char*>(dst);
  if (a < d + db && d < a + sb) return fail(FH_E_ARG, "source and destination must not overlap");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const long total = (long)dnx * dny * batch;
  const int blocks = (int)std::min<long>((total + 255) / 256, 8192);
  hipLaunchKernelGGL(k_window, dim3(blocks), dim3(256), 0, ctx->stream, reinterpret_cast<const float2*>(src), snx,
                     sny, reinterpret_cast<float2*>(dst), dnx, dny, batch, x0, y0,
                     make_float2((float)fill_re, (float)fill_im));
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_phase_apply(fh_ctx* ctx, void* u, const float* phi, int nx, int ny, int batch, int direction,
                   void* saved) {
  if (!ctx || !u || !phi) return fail(FH_E_ARG, "null context, field or mask");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(ctx))) return rc;
  const long npix = (long)nx * ny;
  const int blocks = (int)std::min<long>((npix + 255) / 256, 4096);
  hipLaunchKernelGGL(k_phase_apply, dim3(blocks), dim3(256), 0, ctx->stream,
                     reinterpret_cast<float2*>(u), phi, reinterpret_cast<float2*>(saved), npix, batch,
                     (float)direction);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_phase_backward(fh_ctx* ctx, void* dv, const float* phi, const void* saved, float* dphi, int nx,
                      int ny, int batch, int direction, int accumulate) {
  if (!ctx || !dv || !phi) return fail(FH_E_ARG, "null context, cotangent or mask");
  if (dphi && !saved) return fail(FH_E_ARG, "the mask gradient needs the saved mask input");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(ctx))) return rc;
  const long npix = (long)nx * ny;
  const int blocks = (int)std::min<long>((npix + 255) / 256, 4096);
  hipLaunchKernelGGL(k_phase_backward, dim3(blocks), dim3(256), 0, ctx->stream,
                     reinterpret_cast<float2*>(dv), phi, reinterpret_cast<const float2*>(saved), dphi,
                     npix, batch, (float)direction, accumulate);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_fft2(fh_ctx* ctx, void* u, int nx, int ny, int batch, int sign, double scale) {
  if (!ctx || !u) return fail(FH_E_ARG, "null context or field");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  if (sign != 1 && sign != -1) return fail(FH_E_ARG, "sign must be -1 (forward) or +1 (inverse)");
  if (!std::isfinite(scale)) return fail(FH_E_ARG, "scale must be finite");
  const int lx = ilog2_exact(nx), ly = ilog2_exact(ny);
  int rc;
  if (ny == 1) {  // 1-D fields: line FFTs of the batch (power-of-two nx)
    if (lx < kMinLog || lx > kMaxLog)
      return fail(FH_E_UNSUPP, "1-D FFT needs a power-of-two nx in [%d, %d] (got %d)", 1 << kMinLog, 1 << kMaxLog, nx);
    if ((rc = set_device(ctx))) return rc;
    const float2* tw = nullptr;
    if ((rc = get_twiddles(ctx, nx, &tw))) return rc;
    LineIO<float2> io{};
    io.in = io.out = reinterpret_cast<float2*>(u); io.in_f = io.out_f = 0;
    io.in_l = io.out_l = nx; io.in_e = io.out_e = 1; io.nlines = batch; io.in_valid = io.out_count = nx;
    const ScaleOp sc{(float)scale};
    if (sign < 0) {
      if (scale == 1.0) FH_HIP((launch_fft<float2, float2, false, false>(lx, io, 1, tw, NoOp{}, ctx->stream)));
      else FH_HIP((launch_fft<float2, float2, false, false>(lx, io, 1, tw, sc, ctx->stream)));
    } else {
      if (scale == 1.0) FH_HIP((launch_fft<float2, float2, false, true>(lx, io, 1, tw, NoOp{}, ctx->stream)));
      else FH_HIP((launch_fft<float2, float2, false, true>(lx, io, 1, tw, sc, ctx->stream)));
    }
    return FH_OK;
  }
  if (lx < kMinLog || lx > kMaxLog || ly < kMinLog || ly > kMaxLog) {
    // other sizes: the DFT as a chirp z-transform (a = 1, w = e^{-2 pi i/n}; its adjoint plan is
    // the unnormalised inverse), one cached plan per (nx, ny, batch)
    const std::vector<int> key{nx, ny, batch};
    fh_plan* dp = nullptr;
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      auto it = ctx->dft_plans.find(key);
      if (it != ctx->dft_plans.end()) dp = it->second;
    }
    if (!dp) {
      const double a[4] = {1.0, 0.0, 1.0, 0.0};
      const double w[4] = {1.0, -2.0 * M_PI / nx, 1.0, -2.0 * M_PI / ny};
      if ((rc = fh_czt_create(ctx, nx, ny, batch, 3, a, w, 0, &dp))) return rc;
      std::lock_guard<std::mutex> lk(ctx->mu);
      ctx->dft_plans[key] = dp;
    }
    if ((rc = fh_propagate(dp, u, sign < 0 ? 1 : -1))) return rc;
    if (scale != 1.0) {
      const long n = (long)nx * ny * batch;
      hipLaunchKernelGGL(k_scale_c64, dim3((int)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, ctx->stream,
                         reinterpret_cast<float2*>(u), n, (float)scale);
      FH_HIP(hipGetLastError());
    }
    return FH_OK;
  }
  if ((rc = set_device(ctx))) return rc;
  const float2 *twx = nullptr, *twy = nullptr;
  if ((rc = get_twiddles(ctx, nx, &twx)) || (rc = get_twiddles(ctx, ny, &twy))) return rc;
  float2* f = reinterpret_cast<float2*>(u);
  const long npix = (long)nx * ny;
  LineIO<float2> io{};
  io.in = f; io.out = f; io.in_f = io.out_f = npix;
  io.in_l = io.out_l = nx; io.in_e = io.out_e = 1; io.nlines = ny; io.in_valid = io.out_count = nx;
  LineIO<float2> iy = io;
  iy.in_l = iy.out_l = 1; iy.in_e = iy.out_e = nx; iy.nlines = nx; iy.in_valid = iy.out_count = ny;
  const ScaleOp sc{(float)scale};
  // the scale rides on the first pass of a forward transform (prologue) and on the last
  // pass of an inverse one (epilogue)
  if (sign < 0) {
    if (scale == 1.0) FH_HIP((launch_fft<float2, float2, false, false>(lx, io, batch, twx, NoOp{}, ctx->stream)));
    else FH_HIP((launch_fft<float2, float2, false, false>(lx, io, batch, twx, sc, ctx->stream)));
    FH_HIP((launch_fft<float2, float2, true, false>(ly, iy, batch, twy, NoOp{}, ctx->stream)));
  } else {
    FH_HIP((launch_fft<float2, float2, false, true>(lx, io, batch, twx, NoOp{}, ctx->stream)));
    if (scale == 1.0) FH_HIP((launch_fft<float2, float2, true, true>(ly, iy, batch, twy, NoOp{}, ctx->stream)));
    else FH_HIP((launch_fft<float2, float2, true, true>(ly, iy, batch, twy, sc, ctx->stream)));
  }
  return FH_OK;
}

static int metric_partials(fh_ctx* ctx, size_t bytes) {
  // Sized generously on first use.  A larger request allocates a new buffer and retires the
  // old one without freeing it: a graph recorded before the growth (a captured training step)
  // keeps the old address and must still find valid memory when it replays.  Work on the
  // stream that still reads the old buffer is unaffected too.
  const size_t need = std::max<size_t>((bytes + sizeof(double) - 1) / sizeof(double), 1 << 16);
  if (ctx->metric_part_n >= need) return FH_OK;
  double* fresh = nullptr;
  if (hipMalloc(&fresh, need * sizeof(double)) != hipSuccess)
    return fail(FH_E_NOMEM, "metric partials: out of device memory");
  if (ctx->metric_part) ctx->metric_part_retired.push_back(ctx->metric_part);
  ctx->metric_part = fresh;
  ctx->metric_part_n = need;
  return FH_OK;
}

int fh_dot_product(fh_ctx* ctx, const void* u, const void* v, int nx, int ny, int batch, double ds,
                   void* c, double* p) {
  if (!ctx || !u || !v || !c) return fail(FH_E_ARG, "null context, fields or output");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  if (!std::isfinite(ds)) return fail(FH_E_ARG, "ds must be finite");
  int rc;
  if ((rc = set_device(ctx)) || (rc = metric_partials(ctx, (size_t)batch * kMetricBlocks * sizeof(double2))))
    return rc;
  double2* part = reinterpret_cast<double2*>(ctx->metric_part);
  const long npix = (long)nx * ny;
  hipLaunchKernelGGL(k_dot_partial, dim3(kMetricBlocks, batch), dim3(kMetricThreads), 0, ctx->stream,
                     reinterpret_cast<const float2*>(u), reinterpret_cast<const float2*>(v), npix, part);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dot_final, dim3(batch), dim3(kMetricThreads), 0, ctx->stream, (const double2*)part,
                     kMetricBlocks, ds, reinterpret_cast<double2*>(c), p);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_dot_matrix(fh_ctx* ctx, const void* u, const void* v, int nx, int ny, int batch, double ds, void* c,
                  double* p) {
  if (!ctx || !u || !v || !c) return fail(FH_E_ARG, "null context, fields or output");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  if (batch > kDotMaxModes) return fail(FH_E_UNSUPP, "the overlap matrix takes up to %d modes (got %d)", kDotMaxModes, batch);
  if (!std::isfinite(ds)) return fail(FH_E_ARG, "ds must be finite");
  const int npair = batch * batch;
  int rc;
  if ((rc = set_device(ctx)) || (rc = metric_partials(ctx, (size_t)kMetricBlocks * npair * sizeof(double2)))) return rc;
  double2* part = reinterpret_cast<double2*>(ctx->metric_part);
  const long npix = (long)nx * ny;
  hipLaunchKernelGGL(k_dotm_partial, dim3(kMetricBlocks), dim3(256), 0, ctx->stream, reinterpret_cast<const float2*>(u),
                     reinterpret_cast<const float2*>(v), npix, batch, part);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dotm_final, dim3(npair), dim3(kMetricThreads), 0, ctx->stream, (const double2*)part,
                     kMetricBlocks, npair, ds, reinterpret_cast<double2*>(c), p);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_dot_matrix_pullback(fh_ctx* ctx, const void* v, const void* dc, const void* c, const double* dp, int nx, int ny,
                           int batch, void* du) {
  if (!ctx || !v || !du) return fail(FH_E_ARG, "null context, reference fields or output");
  if (!dc && !c) return fail(FH_E_ARG, "give the cotangent dc (DotProduct) or the overlaps c (PowerCoupling)");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  if (batch > kDotMaxModes) return fail(FH_E_UNSUPP, "the overlap matrix takes up to %d modes (got %d)", kDotMaxModes, batch);
  const long npix = (long)nx * ny;
  const size_t bytes = (size_t)npix * batch * sizeof(float2);
  const char *a = static_cast<const char*>(v), *d = static_cast<const char*>(du);
  if (a < d + bytes && d < a + bytes) return fail(FH_E_ARG, "du must not overlap the reference fields");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const int blocks = (int)std::min<long>((npix + 255) / 256, 4096);
  hipLaunchKernelGGL(k_dotm_pullback, dim3(blocks), dim3(256), 0, ctx->stream, reinterpret_cast<const float2*>(v),
                     reinterpret_cast<const double2*>(dc), reinterpret_cast<const double2*>(c), dp, npix, batch,
                     reinterpret_cast<float2*>(du));
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_dot_pullback_propagate(fh_plan* p, const void* v, const void* dc, const void* c, const double* dp, void* du,
                              int direction, const float* phi, const void* saved, float* dphi, int accumulate) {
  if (!p || !v || !du) return fail(FH_E_ARG, "null plan, reference fields or output");
  if (!dc && !c) return fail(FH_E_ARG, "give the cotangent dc (DotProduct) or the overlaps c (PowerCoupling)");
  if ((saved || dphi) && !phi) return fail(FH_E_ARG, "a saved input or mask gradient needs the mask");
  if (dphi && !saved) return fail(FH_E_ARG, "the mask gradient needs the saved mask input");
  if (p->kind == 2) return fail(FH_E_UNSUPP, "the fused pullback takes ASProp and RSProp plans");
  const size_t bytes = (size_t)p->nx * p->ny * p->batch * sizeof(float2);
  const char *a = static_cast<const char*>(v), *d = static_cast<const char*>(du);
  if (a < d + bytes && d < a + bytes) return fail(FH_E_ARG, "du must not overlap the reference fields");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(p->ctx))) return rc;
  float2* f = reinterpret_cast<float2*>(du);
  if (p->oned) {  // 1-D: the pullback, then the one-pass adjoint propagation (and the Phase adjoint)
    const long npix = p->nx;
    const int blocks = (int)std::min<long>((npix + 255) / 256, 1024);
    hipLaunchKernelGGL(k_dot_pullback, dim3(blocks, p->batch), dim3(256), 0, p->ctx->stream,
                       reinterpret_cast<const float2*>(v), reinterpret_cast<const double2*>(dc),
                       reinterpret_cast<const double2*>(c), dp, npix, f);
    FH_HIP(hipGetLastError());
    if ((rc = run_line_prop(p, f, -direction))) return rc;
    return phi ? fh_phase_backward(p->ctx, du, phi, saved, dphi, p->nx, 1, p->batch, direction, accumulate) : FH_OK;
  }
  const DotGainOp op{reinterpret_cast<const double2*>(dc), reinterpret_cast<const double2*>(c), dp};
  const bool split = phi && grad_split(p);
  if (split && (rc = grad_scratch(p))) return rc;
  if ((rc = run_prop(p, f, -direction, op, -1, reinterpret_cast<const float2*>(v), split))) return rc;
  if (phi) return phase_grad_pass(p, f, phi, saved, dphi, direction, accumulate);
  return run_xifft(p, f, -direction);
}

int fh_dot_product_pullback(fh_ctx* ctx, const void* v, const void* dc, const void* c, const double* dp,
                            int nx, int ny, int batch, void* du) {
  if (!ctx || !v || !du) return fail(FH_E_ARG, "null context, field or output");
  if (!dc && !c) return fail(FH_E_ARG, "need the cotangents dc, or the dot products c for a coupling pullback");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const long npix = (long)nx * ny;
  const int blocks = (int)std::min<long>((npix + 255) / 256, 1024);
  hipLaunchKernelGGL(k_dot_pullback, dim3(blocks, batch), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const float2*>(v), reinterpret_cast<const double2*>(dc),
                     reinterpret_cast<const double2*>(c), dp, npix, reinterpret_cast<float2*>(du));
  FH_HIP(hipGetLastError());
  return FH_OK;
}

static int elementwise_blocks(long npix) { return (int)std::min<long>((npix + 255) / 256, 4096); }

int fh_offset_tilts(fh_ctx* ctx, void* u, int nx, int ny, int batch, double dx, double dy, const double* lambdas,
                    int nlambdas, const double* from_x, const double* from_y, const double* to_x,
                    const double* to_y, int ntilts) {
  if (!ctx || !u) return fail(FH_E_ARG, "null context or field");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  if (!(dx > 0) || !(dy > 0)) return fail(FH_E_ARG, "sampling must be positive");
  if (!lambdas || (nlambdas != 1 && nlambdas != batch))
    return fail(FH_E_ARG, "need 1 or batch (=%d) wavelengths, got %d", batch, nlambdas);
  if (!from_x || !from_y || !to_x || !to_y || (ntilts != 1 && ntilts != batch))
    return fail(FH_E_ARG, "need 1 or batch (=%d) tilt pairs on both sides, got %d", batch, ntilts);
  for (int k = 0; k < nlambdas; ++k)
    if (!(lambdas[k] > 0)) return fail(FH_E_ARG, "wavelengths must be positive");
  for (int k = 0; k < ntilts; ++k)
    for (double a : {from_x[k], from_y[k], to_x[k], to_y[k]})
      if (!std::isfinite(a) || fabs(a) >= M_PI / 2) return fail(FH_E_ARG, "tilt angles must be finite radians in (-pi/2, pi/2)");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const long npix = (long)nx * ny;
  for (int b0 = 0; b0 < batch; b0 += kTiltChunk) {
    const int nb = std::min(kTiltChunk, batch - b0);
    TiltCoef c{};
    for (int j = 0; j < nb; ++j) {
      const int b = b0 + j, t = ntilts == 1 ? 0 : b;
      const double lam = lambdas[nlambdas == 1 ? 0 : b];
      c.cx[j] = (sin(from_x[t]) - sin(to_x[t])) / lam;
      c.cy[j] = (sin(from_y[t]) - sin(to_y[t])) / lam;
    }
    hipLaunchKernelGGL(k_offset_tilts, dim3(elementwise_blocks(npix)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<float2*>(u) + b0 * npix, nx, ny, dx, dy, c, nb);
    FH_HIP(hipGetLastError());
  }
  return FH_OK;
}

int fh_mask_apply(fh_ctx* ctx, void* u, const void* m, int nx, int ny, int batch, int direction, void* saved) {
  if (!ctx || !u || !m) return fail(FH_E_ARG, "null context, field or mask");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(ctx))) return rc;
  const long npix = (long)nx * ny;
  hipLaunchKernelGGL(k_mask_apply, dim3(elementwise_blocks(npix)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<float2*>(u), reinterpret_cast<const float2*>(m),
                     reinterpret_cast<float2*>(saved), npix, batch, direction < 0);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_mask_backward(fh_ctx* ctx, void* dv, const void* m, const void* saved, void* dm, int nx, int ny,
                     int batch, int direction, int accumulate) {
  if (!ctx || !dv || !m) return fail(FH_E_ARG, "null context, cotangent or mask");
  if (dm && !saved) return fail(FH_E_ARG, "the mask gradient needs the saved mask input");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(ctx))) return rc;
  const long npix = (long)nx * ny;
  hipLaunchKernelGGL(k_mask_backward, dim3(elementwise_blocks(npix)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<float2*>(dv), reinterpret_cast<const float2*>(m),
                     reinterpret_cast<const float2*>(saved), reinterpret_cast<float2*>(dm), npix, batch,
                     direction > 0, accumulate);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_tea_apply(fh_ctx* ctx, void* u, const float* h, int nx, int ny, int batch, const double* kl,
                 const double* r, int direction, void* saved) {
  if (!ctx || !u || !h || !kl) return fail(FH_E_ARG, "null context, field, height map or coefficients");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(ctx))) return rc;
  const long npix = (long)nx * ny;
  hipLaunchKernelGGL(k_tea_apply, dim3(elementwise_blocks(npix)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<float2*>(u), h, reinterpret_cast<float2*>(saved), npix, batch, kl,
                     reinterpret_cast<const double2*>(r), (float)direction);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_tea_backward(fh_ctx* ctx, void* dv, const float* h, const void* saved, float* dh, int nx, int ny,
                    int batch, const double* kl, const double* r, int direction, int accumulate) {
  if (!ctx || !dv || !h || !kl) return fail(FH_E_ARG, "null context, cotangent, height map or coefficients");
  if (dh && !saved) return fail(FH_E_ARG, "the height gradient needs the saved input");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(ctx))) return rc;
  const long npix = (long)nx * ny;
  hipLaunchKernelGGL(k_tea_backward, dim3(elementwise_blocks(npix)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<float2*>(dv), h, reinterpret_cast<const float2*>(saved), dh, npix, batch,
                     kl, reinterpret_cast<const double2*>(r), (float)direction, accumulate);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_optim_step(fh_ctx* ctx, int rule, float* x, const float* dx, float* state, long n, double eta,
                  double rho, double beta, int prox, double p0, double p1) {
  if (!ctx || !x || !dx) return fail(FH_E_ARG, "null context, parameter or gradient");
  if (rule < 0 || rule > 3) return fail(FH_E_ARG, "unknown rule %d", rule);
  if (rule > 0 && !state) return fail(FH_E_ARG, "rule %d needs a state buffer", rule);
  if (prox < 0 || prox > 3) return fail(FH_E_ARG, "unknown proximal operator %d", prox);
  if (n <= 0) return fail(FH_E_ARG, "n must be positive");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const int blocks = (int)std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_optim_step, dim3(blocks), dim3(256), 0, ctx->stream, x, dx, state, n, rule, (float)eta,
                     (float)rho, (float)beta, prox, (float)p0, (float)p1, (const double*)nullptr);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_fista_step(fh_ctx* ctx, float* x, const float* dx, float* state, long n, double eta, double* tb, int prox,
                  double p0, double p1) {
  if (!ctx || !x || !dx || !state || !tb) return fail(FH_E_ARG, "null context, parameter, gradient, state or coefficients");
  if (prox < 0 || prox > 3) return fail(FH_E_ARG, "unknown proximal operator %d", prox);
  if (n <= 0) return fail(FH_E_ARG, "n must be positive");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  hipLaunchKernelGGL(k_fista_coef, dim3(1), dim3(64), 0, ctx->stream, tb);
  FH_HIP(hipGetLastError());
  const int blocks = (int)std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_optim_step, dim3(blocks), dim3(256), 0, ctx->stream, x, dx, state, n, 3, (float)eta, 0.f,
                     0.f, prox, (float)p0, (float)p1, (const double*)(tb + 1));
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_adam_step(fh_ctx* ctx, float* x, const float* dx, float* state, long n, double eta, double beta1,
                 double beta2, double eps, double* bt, int prox, double p0, double p1) {
  if (!ctx || !x || !dx || !state || !bt) return fail(FH_E_ARG, "null context, parameter, gradient, state or powers");
  if (n <= 0) return fail(FH_E_ARG, "n must be positive");
  if (prox < 0 || prox > 3) return fail(FH_E_ARG, "unknown proximal operator %d", prox);
  if (!(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0))
    return fail(FH_E_ARG, "Adam needs 0 <= beta < 1");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const int blocks = (int)std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_adam_step, dim3(blocks), dim3(256), 0, ctx->stream, x, dx, state, n, eta, beta1, beta2, eps,
                     (const double*)bt, prox, (float)p0, (float)p1);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_adam_advance, dim3(1), dim3(64), 0, ctx->stream, bt, beta1, beta2);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_sq_intensity_diff(fh_ctx* ctx, const void* u, const float* target, int nx, int ny, int batch,
                         int target_batch, const double* dc, double* loss, void* du) {
  if (!ctx || !u || !target || !loss) return fail(FH_E_ARG, "null context, field, target or loss");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  if (target_batch != batch && target_batch != 1)
    return fail(FH_E_ARG, "the target holds %d intensities; expected %d (per field) or 1 (spatial only)",
                target_batch, batch);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  const int spatial = target_batch == 1 && batch > 1;
  const int nout = spatial ? 1 : batch;
  if ((rc = metric_partials(ctx, (size_t)nout * kMetricBlocks * sizeof(double)))) return rc;
  const long npix = (long)nx * ny;
  hipLaunchKernelGGL(k_sqint_partial, dim3(kMetricBlocks, nout), dim3(kMetricThreads), 0, ctx->stream,
                     reinterpret_cast<const float2*>(u), target, npix, batch, spatial, dc,
                     reinterpret_cast<float2*>(du), ctx->metric_part);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(nout), dim3(kMetricThreads), 0, ctx->stream,
                     (const double*)ctx->metric_part, kMetricBlocks, loss);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_sq_intensity_diff_noalloc(fh_ctx* ctx, const void* u, const float* target, int nx, int ny, int batch,
                                 int target_batch, const double* dc, double* loss, void* du) {
  if (!ctx || !u || !target || !loss) return fail(FH_E_ARG, "null context, field, target or loss");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  if (target_batch != batch && target_batch != 1)
    return fail(FH_E_ARG, "the target holds %d intensities; expected %d (per field) or 1 (spatial only)",
                target_batch, batch);
  const int spatial = target_batch == 1 && batch > 1;
  const int nout = spatial ? 1 : batch;
  if (nout > kNoAllocSlots)
    return fail(FH_E_ARG, "%d loss values exceed the no-allocation maximum of %d per call", nout, kNoAllocSlots);
  double* part = scratch_addr(ctx->device);
  if (!part) return fail(FH_E_HIP, "no scratch address for device %d (context not created by fh_ctx_create)", ctx->device);
  int rc;
  if ((rc = set_device(ctx))) return rc;
  // every partial of the nout slots is written before k_sum_partials reads it: nothing to clear
  const long npix = (long)nx * ny;
  hipLaunchKernelGGL(k_sqint_partial, dim3(kMetricBlocks, nout), dim3(kMetricThreads), 0, ctx->stream,
                     reinterpret_cast<const float2*>(u), target, npix, batch, spatial, dc,
                     reinterpret_cast<float2*>(du), part);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(nout), dim3(kMetricThreads), 0, ctx->stream,
                     (const double*)part, kMetricBlocks, loss);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_as_z_gradient(fh_ctx* ctx, const void* V, const void* DV, int nx, int ny, int batch, double dx, double dy,
                     const double* lambdas, int nlambdas, double n0, double z, double* out) {
  if (!ctx || !V || !DV || !lambdas || !out) return fail(FH_E_ARG, "null context, spectra, wavelengths or output");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  if (nlambdas != 1 && nlambdas != batch) return fail(FH_E_ARG, "give 1 or batch wavelengths");
  if (!(dx > 0) || !(dy > 0) || !(n0 > 0)) return fail(FH_E_ARG, "sampling and n0 must be positive");
  for (int k = 0; k < nlambdas; ++k)
    if (!(lambdas[k] > 0)) return fail(FH_E_ARG, "wavelengths must be positive");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = metric_partials(ctx, (size_t)batch * kMetricBlocks * sizeof(double)))) return rc;
  const long npix = (long)nx * ny;
  for (int b = 0; b < batch; ++b) {
    const double k = n0 / lambdas[nlambdas == 1 ? 0 : b];
    hipLaunchKernelGGL(k_as_zgrad_partial, dim3(kMetricBlocks), dim3(kMetricThreads), 0, ctx->stream,
                       reinterpret_cast<const float2*>(V) + b * npix, reinterpret_cast<const float2*>(DV) + b * npix,
                       nx, ny, dx, dy, k * k, z < 0 ? 1 : 0, ctx->metric_part + (long)b * kMetricBlocks);
    FH_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_sum_partials, dim3(batch), dim3(kMetricThreads), 0, ctx->stream,
                     (const double*)ctx->metric_part, kMetricBlocks, out);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_sq_field_diff(fh_ctx* ctx, const void* u, const void* v, int nx, int ny, int batch, const double* dc,
                     double* loss, void* du) {
  if (!ctx || !u || !v || !loss) return fail(FH_E_ARG, "null context, field, target field or loss");
  if (nx <= 0 || ny <= 0 || batch <= 0) return fail(FH_E_ARG, "sizes must be positive");
  int rc;
  if ((rc = set_device(ctx))) return rc;
  if ((rc = metric_partials(ctx, (size_t)batch * kMetricBlocks * sizeof(double)))) return rc;
  hipLaunchKernelGGL(k_sqfield_partial, dim3(kMetricBlocks, batch), dim3(kMetricThreads), 0, ctx->stream,
                     reinterpret_cast<const float2*>(u), reinterpret_cast<const float2*>(v), (long)nx * ny, dc,
                     reinterpret_cast<float2*>(du), ctx->metric_part);
  FH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(batch), dim3(kMetricThreads), 0, ctx->stream,
                     (const double*)ctx->metric_part, kMetricBlocks, loss);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

int fh_propagate_phase_backward(fh_plan* p, void* dv, const float* phi, const void* saved, float* dphi,
                                int direction, int accumulate) {
  if (!p || !dv || !phi) return fail(FH_E_ARG, "null plan, cotangent or mask");
  if (dphi && !saved) return fail(FH_E_ARG, "the mask gradient needs the saved mask input");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(p->ctx))) return rc;
  float2* f = reinterpret_cast<float2*>(dv);
  if (p->kind == 2) return fail(FH_E_UNSUPP, "Phase fuses with ASProp and RSProp plans only");
  if (p->oned) {  // 1-D: the one-pass adjoint propagation, then the Phase adjoint and gradient
    if ((rc = run_line_prop(p, f, -direction))) return rc;
    return fh_phase_backward(p->ctx, dv, phi, saved, dphi, p->nx, 1, p->batch, direction, accumulate);
  }
  // adjoint propagation = the propagator in the reverse direction
  if (grad_split(p) && (rc = grad_scratch(p))) return rc;
  if ((rc = run_prop(p, f, -direction, NoOp{}, -1, nullptr, grad_split(p)))) return rc;
  return phase_grad_pass(p, f, phi, saved, dphi, direction, accumulate);
}

// The fused metric block; src (nullable) is the block's input when it is not u itself: the first
// pass reads it there and the Phase gradient takes it as the saved input (nothing is saved).
static int sqint_block(fh_plan* p, const float2* src, float2* f, const float* target, int target_batch,
                       const double* dc, double* loss, int direction, const float* phi, void* saved, float* dphi,
                       int accumulate, int cotangent = 1) {
  if (p->kind == 2 || p->oned) return fail(FH_E_UNSUPP, "the fused metric block takes 2-D ASProp and RSProp plans");
  if (target_batch != p->batch && !(target_batch == 1 && p->batch == 1))
    return fail(FH_E_UNSUPP, "the fused metric block needs one target per field (target batch %d, batch %d)",
                target_batch, p->batch);
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(p->ctx))) return rc;
  fh_ctx* c = p->ctx;
  // forward: [Phase in the first pass] x FFT, y pass
  if (phi) {
    PhaseTimeOp op{};
    if ((rc = phase_time_op(p, phi, saved, direction, &op)) || (rc = run_prop(p, f, direction, op, -1, src)))
      return rc;
  } else if ((rc = run_prop(p, f, direction, NoOp{}, -1, src))) {
    return rc;
  }
  // fused x IFFT + metric + pullback + x FFT of the adjoint, in place on the spectrum
  const int nxl = 1 << p->lx;
  const bool rs = p->kind != 0;
  float2* mid = rs ? p->scratch : f;
  FastIO io{};
  io.in = mid; io.out = mid; io.in_f = io.out_f = (long)nxl * mid_rows(p);
  io.in_l = io.out_l = nxl; io.in_e = io.out_e = 1; io.in_blk = io.out_blk = 1;
  io.nlines = p->ny; io.in_bytes = io.out_bytes = io.in_f * 8;
  const long tstride = target_batch == 1 ? 0 : (long)p->nx * p->ny;
  long nparts = 0;
  if ((rc = fast_sqint(c, p->lx, rs, io, p->batch, target, tstride, p->nx, dc, nullptr, &nparts))) return rc;
  if ((rc = metric_partials(c, (size_t)p->batch * nparts * sizeof(double)))) return rc;
  if ((rc = fast_sqint(c, p->lx, rs, io, p->batch, target, tstride, p->nx, dc, c->metric_part, &nparts))) return rc;
  hipLaunchKernelGGL(k_sum_partials, dim3(p->batch), dim3(kMetricThreads), 0, c->stream,
                     (const double*)c->metric_part, (int)nparts, loss);
  FH_HIP(hipGetLastError());
  // adjoint: y pass with the conjugate kernel, then x IFFT [un-applying the Phase, its gradient];
  // none of it when neither the cotangent nor a mask gradient is asked for
  if (!cotangent && !(phi && dphi)) return FH_OK;
  const bool split = phi && grad_split(p);
  if (split && (rc = grad_scratch(p))) return rc;
  if ((rc = run_prop(p, f, -direction, NoOp{}, 1, nullptr, split))) return rc;
  if (!phi) return run_xifft(p, f, -direction);
  return phase_grad_pass(p, f, phi, src ? (const void*)src : saved, dphi, direction, accumulate,
                         cotangent || !dphi);
}

int fh_sqint_block(fh_plan* p, void* u, const float* target, int target_batch, const double* dc, double* loss,
                   int direction, const float* phi, void* saved, float* dphi, int accumulate) {
  if (!p || !u || !target || !loss) return fail(FH_E_ARG, "null plan, field, target or loss");
  if ((saved || dphi) && !phi) return fail(FH_E_ARG, "a saved input or mask gradient needs the mask");
  if (dphi && !saved) return fail(FH_E_ARG, "the mask gradient needs the saved mask input");
  return sqint_block(p, nullptr, reinterpret_cast<float2*>(u), target, target_batch, dc, loss, direction, phi,
                     saved, dphi, accumulate);
}

int fh_sqint_block_src(fh_plan* p, const void* src, void* u, const float* target, int target_batch,
                       const double* dc, double* loss, int direction, const float* phi, float* dphi,
                       int accumulate, int cotangent) {
  if (!p || !src || !u || !target || !loss) return fail(FH_E_ARG, "null plan, source, field, target or loss");
  if (dphi && !phi) return fail(FH_E_ARG, "a mask gradient needs the mask");
  const size_t bytes = (size_t)p->nx * p->ny * p->batch * sizeof(float2);
  const char *a = static_cast<const char*>(src), *b = static_cast<const char*>(u);
  if (a < b + bytes && b < a + bytes) return fail(FH_E_ARG, "the source and the field must not overlap");
  return sqint_block(p, reinterpret_cast<const float2*>(src), reinterpret_cast<float2*>(u), target, target_batch,
                     dc, loss, direction, phi, nullptr, dphi, accumulate, cotangent);
}

int fh_propagate_plain_fp64(fh_plan* p, void* u_c128, int direction) {
  if (!p || !u_c128) return fail(FH_E_ARG, "null plan or field");
  if (p->kind != 0 || p->mode != 0 || p->filtered || p->oned)
    return fail(FH_E_UNSUPP, "the plain FP64 comparator covers on-axis 2-D ASProp without a filter only");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(p->ctx))) return rc;
  fh_ctx* c = p->ctx;
  const int nx = p->nx, ny = p->ny;
  const long npix = (long)nx * ny;
  if (!p->plain_tab) {
    if (hipMalloc(&p->plain_tab, npix * p->lam_tables.size() * sizeof(double2)) != hipSuccess)
      return fail(FH_E_NOMEM, "plain FP64 table: out of device memory");
    for (size_t k = 0; k < p->lam_tables.size(); ++k) {
      hipLaunchKernelGGL(k_as_table_full_d, dim3((nx + 255) / 256, ny), dim3(256), 0, c->stream,
                         p->plain_tab + k * npix, nx, ny, p->dx, p->dy, p->lam_tables[k], p->n0, p->tab_z[k],
                         1.0 / ((double)nx * ny));
      FH_HIP(hipGetLastError());
    }
  }
  const double2 *twx = nullptr, *twy = nullptr;
  if ((rc = get_twiddles(c, nx, &twx)) || (rc = get_twiddles(c, ny, &twy))) return rc;
  double2* u = reinterpret_cast<double2*>(u_c128);
  for (int b = 0; b < p->batch; ++b) {
    double2* f = u + b * npix;
    LineIO<double2> io{};
    io.in = f; io.out = f; io.in_f = io.out_f = npix;
    io.in_l = io.out_l = nx; io.in_e = io.out_e = 1; io.nlines = ny; io.in_valid = io.out_count = nx;
    LineIO<double2> iy = io;
    iy.in_l = iy.out_l = 1; iy.in_e = iy.out_e = nx; iy.nlines = nx; iy.in_valid = iy.out_count = ny;
    FH_HIP((launch_fft<double2, double2, false, false>(p->lx, io, 1, twx, NoOp{}, c->stream)));
    FH_HIP((launch_fft<double2, double2, true, false>(p->ly, iy, 1, twy, NoOp{}, c->stream)));
    const int blocks = (int)std::min<long>((npix + 255) / 256, 8192);
    hipLaunchKernelGGL(k_mul_full_d, dim3(blocks), dim3(256), 0, c->stream, f,
                       p->plain_tab + p->field_table[b] * npix, npix, direction < 0);
    FH_HIP(hipGetLastError());
    FH_HIP((launch_fft<double2, double2, true, true>(p->ly, iy, 1, twy, NoOp{}, c->stream)));
    FH_HIP((launch_fft<double2, double2, false, true>(p->lx, io, 1, twx, NoOp{}, c->stream)));
  }
  return FH_OK;
}

}  // extern "C"

// ================================================================ AS_BPM

namespace fh {

// cis (nullable): this slice's phases from k_bpm_cis ([B][ny][nx]), read instead of evaluated
template <typename RD, int L>
static hipError_t launch_bpm_step_l(const LineIO<float2>& io, int batch, const double2* tw,
                                    const DnPhaseOp& ph, float2* save, float2* ckspec, float* ddn,
                                    const float2* saved, const double2* cis, hipStream_t st) {
  if constexpr (L > 1) {
    // the gradient kernel loops over the batch inside a workgroup
    const long wgs = (long)((io.nlines + L - 1) / L) * (ddn ? 1 : batch);
    if (wgs < kFillWGs) return launch_bpm_step_l<RD, L / 2>(io, batch, tw, ph, save, ckspec, ddn, saved, cis, st);
  }
  const size_t lds = (RD::NST > 1) ? (size_t)L * LdsLine<RD::n, L, sizeof(double2)>::size * sizeof(double2) : 0;
  const unsigned threads = L * RD::P;
  const dim3 grid((io.nlines + L - 1) / L, ddn ? 1 : batch);
  hipError_t e;
  if (ddn) {
    auto kern = cis ? k_bpm_xstep_grad<double2, RD, L, true> : k_bpm_xstep_grad<double2, RD, L, false>;
    if ((e = ensure_lds(kern, lds)) != hipSuccess) return e;
    hipLaunchKernelGGL(kern, grid, dim3(threads), lds, st, io, batch, tw, ph, saved, ddn, cis);
  } else if (cis) {
    auto kern = k_bpm_xstep<double2, RD, L, TabPhaseOp>;
    if ((e = ensure_lds(kern, lds)) != hipSuccess) return e;
    const TabPhaseOp tp{cis, io.in_f, ph.nx};
    hipLaunchKernelGGL(kern, grid, dim3(threads), lds, st, io, tw, tp, save, ckspec);
  } else {
    auto kern = k_bpm_xstep<double2, RD, L, DnPhaseOp>;
    if ((e = ensure_lds(kern, lds)) != hipSuccess) return e;
    hipLaunchKernelGGL(kern, grid, dim3(threads), lds, st, io, tw, ph, save, ckspec);
  }
  return hipGetLastError();
}

template <int LOGN>
static hipError_t launch_bpm_step_t(const LineIO<float2>& io, int batch, const double2* tw,
                                    const DnPhaseOp& ph, float2* save, float2* ckspec, float* ddn,
                                    const float2* saved, const double2* cis, hipStream_t st) {
  using CF = Cfg<double, LOGN, false>;
  return launch_bpm_step_l<typename CF::RD, CF::L>(io, batch, tw, ph, save, ckspec, ddn, saved, cis, st);
}

static hipError_t launch_bpm_step(int logn, const LineIO<float2>& io, int batch, const double2* tw,
                                  const DnPhaseOp& ph, float2* save, float2* ckspec, float* ddn,
                                  const float2* saved, const double2* cis, hipStream_t st) {
  FH_LOGN_SWITCH(logn, (launch_bpm_step_t<LG>(io, batch, tw, ph, save, ckspec, ddn, saved, cis, st)))
}

}  // namespace fh

// Bound on the two segments' FP64 slice-phase tables of a BPM (512 x 512, 20 slices a segment:
// 2 x 84 MB)
constexpr size_t kBpmPhaseTabBytes = (size_t)512 << 20;

struct fh_bpm {
  fh_ctx* ctx = nullptr;
  int nx = 0, ny = 0, batch = 0, nslices = 0, every = 1, nck = 0;
  int lx = 0, ly = 0;
  double thickness = 0, n0 = 1, dz = 0;
  std::vector<double> lam_tables, tab_tx, tab_ty;
  std::vector<int> field_table;
  int mode = 0, track = 0, full = 0;  // slice propagator: as k_as_table's mode; full layout if tilted
  double2* tfull = nullptr;  // FP64 tables of P(dz), one per (wavelength, tilt) key
  double2* thalf = nullptr;  // ... of P(dz/2)
  int* d_tabidx = nullptr;
  long tabsize = 0;
  int qx = 0, qy = 0;
  double* kl = nullptr;      // [B] kdz / lambda_b (vacuum wavelength)
  float2* ckpt = nullptr;    // [nck][B][ny][nx] x-spectrum entering slice k, k % every == 0
  float2* seg = nullptr;     // [every][B][ny][nx] slice inputs of one segment (recompute)
  float2* work = nullptr;    // [B][ny][nx] recompute work field
  int saved_dir = 0;         // direction of the last saving forward (0 = none)
  size_t bytes = 0;
  // grids that are not powers of two: the slice propagators are general plans (padded form)
  // and the slice phase is an elementwise kernel; checkpoints then hold fields, not x-spectra
  int generic = 0;
  fh_plan* pfull = nullptr;
  fh_plan* phalf = nullptr;
  float* gpart = nullptr;    // [B][ny][nx] per-field dn gradient terms (batches of several fields)
  // the dn adjoint recomputes segment s-1 on a second stream while segment s's gradient steps
  // run on the caller's (double-buffered segments, event-ordered)
  float2* seg2 = nullptr;
  // 1-D fields: one workgroup per field line runs every slice (k_bpm_line); seg keeps every slice
  // input of a saving forward, [nslices][B][nx], and gpart the gradient terms [nslices][B][nx]
  int oned = 0;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_used[2] = {nullptr, nullptr};
  // slice phases of one segment, cis(s kl_b dn) in FP64 ([every][B][ny][nx], two buffers), filled on
  // the side stream ahead of the slices that read them (null: the slice steps evaluate the phase)
  double2* cis[2] = {nullptr, nullptr};
  hipEvent_t ev_tab[2] = {nullptr, nullptr}, ev_tused[2] = {nullptr, nullptr};
};

namespace fh {

static LineIO<float2> xrows(fh_bpm* m, float2* f) {
  LineIO<float2> io{};
  io.in = f; io.out = f; io.in_f = io.out_f = (long)m->nx * m->ny;
  io.in_l = io.out_l = m->nx; io.in_e = io.out_e = 1;
  io.nlines = m->ny; io.in_valid = io.out_count = m->nx;
  return io;
}

static void bpm_free(fh_bpm* m) {
  if (!m) return;
  for (void* q : {(void*)m->tfull, (void*)m->thalf, (void*)m->d_tabidx, (void*)m->kl, (void*)m->ckpt,
                  (void*)m->seg, (void*)m->work})
    if (q) (void)hipFree(q);
  plan_free(m->pfull);
  plan_free(m->phalf);
  if (m->gpart) (void)hipFree(m->gpart);
  if (m->seg2) (void)hipFree(m->seg2);
  for (double2* q : m->cis)
    if (q) (void)hipFree(q);
  for (hipEvent_t e : {m->ev_fork, m->ev_rec[0], m->ev_rec[1], m->ev_used[0], m->ev_used[1], m->ev_tab[0],
                       m->ev_tab[1], m->ev_tused[0], m->ev_tused[1]})
    if (e) (void)hipEventDestroy(e);
  if (m->side) (void)hipStreamDestroy(m->side);
  delete m;
}

// Slice phase of the generic BPM (apply_dn_slice!, bpm.jl:235-242): u_b *= cis(s kl_b dn),
// the phase formed in FP64; save (nullable) receives the input first.
__global__ void k_bpm_phase(float2* __restrict__ u, const float* __restrict__ dn, const double* __restrict__ kl,
                            double s, float2* __restrict__ save, long npix, int batch) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    const double d = dn[i];
    for (int b = 0; b < batch; ++b) {
      const long k = b * npix + i;
      const float2 v = u[k];
      if (save) save[k] = v;
      double sn, cs;
      sincos(s * kl[b] * d, &sn, &cs);
      u[k] = make_float2((float)(v.x * cs - v.y * sn), (float)(v.x * sn + v.y * cs));
    }
  }
}

// Its adjoint (bpm.jl:275-312), as the epilogue of the generic BPM's adjoint x IFFT: du_b *=
// cis(-s kl_b dn), and with saved set, part[b] = s kl_b Im(du_b conj(saved_b)), the field's term of
// the slice's dn gradient.  Both come from the FP64 transform values: the product cancels where
// du and the slice input are nearly in phase, so it is formed before du is rounded to c64.
struct DnPhaseGradOp {
  const float* dn;      // [ny][nx] slice
  const double* kl;     // [B]
  const float2* saved;  // [B][ny][nx] slice inputs (nullable: no gradient)
  float* part;          // [B][ny][nx] gradient terms (the slice's gradient itself for one field)
  long npix;
  double s;
  int nx;
  template <typename C>
  __device__ __forceinline__ C operator()(C v, int b, int y, int x) const {
    if (x >= nx) return v;
    const long i = (long)y * nx + x;
    double sn, cs;
    sincos(-s * kl[b] * (double)dn[i], &sn, &cs);
    const double wx = v.x * cs - v.y * sn, wy = v.x * sn + v.y * cs;
    if (saved) {
      const float2 sv = saved[b * npix + i];
      part[b * npix + i] = (float)(s * kl[b] * (wy * sv.x - wx * sv.y));
    }
    return cmake<C>(wx, wy);
  }
};

// out[i] = sum over fields of part[b][i], in field order (the batch sum of the dn gradient)
__global__ void k_sum_fields(const float* __restrict__ part, float* __restrict__ out, long npix, int batch) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npix; i += (long)gridDim.x * blockDim.x) {
    double g = 0.0;
    for (int b = 0; b < batch; ++b) g += part[b * npix + i];
    out[i] = (float)g;
  }
}

// x FFT of rows, optionally with slice k's phase as prologue
static int bpm_xfft(fh_bpm* m, float2* f, const float* dn_k, double sgn) {
  const double2* tw = nullptr;
  int rc;
  if ((rc = get_twiddles(m->ctx, m->nx, &tw))) return rc;
  LineIO<float2> io = xrows(m, f);
  if (dn_k) {
    DnPhaseOp ph{dn_k, m->kl, sgn, m->nx};
    FH_HIP((launch_fft<float2, double2, false, false>(m->lx, io, m->batch, tw, ph, m->ctx->stream)));
  } else {
    FH_HIP((launch_fft<float2, double2, false, false>(m->lx, io, m->batch, tw, NoOp{}, m->ctx->stream)));
  }
  return FH_OK;
}

static int bpm_xifft(fh_bpm* m, float2* f) {
  const double2* tw = nullptr;
  int rc;
  if ((rc = get_twiddles(m->ctx, m->nx, &tw))) return rc;
  FH_HIP((launch_fft<float2, double2, false, true>(m->lx, xrows(m, f), m->batch, tw, NoOp{}, m->ctx->stream)));
  return FH_OK;
}

static int bpm_ypass(fh_bpm* m, float2* f, const double2* tab, int direction) {
  const double2* tw = nullptr;
  int rc;
  if ((rc = get_twiddles(m->ctx, m->ny, &tw))) return rc;
  LineIO<float2> io{};
  io.in = f; io.out = f; io.in_f = io.out_f = (long)m->nx * m->ny;
  io.in_l = io.out_l = 1; io.in_e = io.out_e = m->nx;
  io.nlines = m->nx; io.in_valid = io.out_count = m->ny;
  QuadMulT<double2> q{tab, m->d_tabidx, m->tabsize, m->nx, m->ny, m->qy, direction < 0, m->lam_tables.size() == 1,
                      m->full};
  FH_HIP((launch_conv<float2, double2>(m->ly, io, m->batch, tw, q, m->ctx->stream)));
  return FH_OK;
}

// cis_k (nullable): slice k's phases from k_bpm_cis, [B][ny][nx]
static int bpm_step(fh_bpm* m, float2* f, const float* dn_k, double sgn, float2* save, float2* ckspec,
                    const double2* cis_k = nullptr) {
  const double2* tw = nullptr;
  int rc;
  if ((rc = get_twiddles(m->ctx, m->nx, &tw))) return rc;
  DnPhaseOp ph{dn_k, m->kl, sgn, m->nx};
  FH_HIP(launch_bpm_step(m->lx, xrows(m, f), m->batch, tw, ph, save, ckspec, nullptr, nullptr, cis_k,
                         m->ctx->stream));
  return FH_OK;
}

static int bpm_step_grad(fh_bpm* m, float2* f, const float* dn_k, double sgn, const float2* saved,
                         float* ddn_k, const double2* cis_k = nullptr) {
  const double2* tw = nullptr;
  int rc;
  if ((rc = get_twiddles(m->ctx, m->nx, &tw))) return rc;
  DnPhaseOp ph{dn_k, m->kl, sgn, m->nx};
  FH_HIP(launch_bpm_step(m->lx, xrows(m, f), m->batch, tw, ph, nullptr, nullptr, ddn_k, saved, cis_k,
                         m->ctx->stream));
  return FH_OK;
}

// The phases of slices k0 .. k1-1 (sign sgn) into out, on the context's current stream.
static int bpm_cis(fh_bpm* m, const float* dn, int k0, int k1, double sgn, double2* out) {
  const long npix = (long)m->nx * m->ny;
  const long total = npix * m->batch * (k1 - k0);
  const int blocks = (int)std::min<long>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(k_bpm_cis, dim3(blocks), dim3(256), 0, m->ctx->stream, dn + (long)k0 * npix, m->kl, sgn, out,
                     npix, m->batch, k1 - k0);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

// Slice inputs k0 .. k1-1 into seg[0 ..], replayed from the x-spectrum checkpoint of slice
// k0 with the forward's own kernels, so they equal the forward's values bit for bit.
static int bpm_recompute(fh_bpm* m, const float* dn, int k0, int k1, double sgn, int direction, float2* seg,
                         const double2* cis = nullptr) {
  const long npix = (long)m->nx * m->ny;
  const long fsz = npix * m->batch;
  FH_HIP(hipMemcpyAsync(m->work, m->ckpt + (long)(k0 / m->every) * fsz, fsz * sizeof(float2),
                        hipMemcpyDeviceToDevice, m->ctx->stream));
  int rc;
  for (int k = k0; k < k1; ++k) {
    if ((rc = bpm_step(m, m->work, dn + (long)k * npix, sgn, seg + (long)(k - k0) * fsz, nullptr,
                       cis ? cis + (long)(k - k0) * fsz : nullptr)))
      return rc;
    if (k + 1 < k1 && (rc = bpm_ypass(m, m->work, m->tfull, direction))) return rc;
  }
  return FH_OK;
}

static int bpm_phase(fh_bpm* m, float2* f, const float* dn_k, double sgn, float2* save) {
  const long npix = (long)m->nx * m->ny;
  hipLaunchKernelGGL(k_bpm_phase, dim3(elementwise_blocks(npix)), dim3(256), 0, m->ctx->stream, f, dn_k,
                     (const double*)m->kl, sgn, save, npix, m->batch);
  FH_HIP(hipGetLastError());
  return FH_OK;
}

// Slice propagator of the generic BPM: the padded convolution of a general ASProp plan (zero
// padding to the plan's power-of-two lengths, FFT, transfer function, IFFT, crop) computed in FP64
// with the plan's unrounded tables.  Hundreds of near-identity steps would repeat one FP32
// rounding of the butterflies and tables coherently; the values stored between the three
// launches are c64, as in the power-of-two slice steps.  p->scratch holds the x-spectrum rows,
// [B][ny][nx2].  gop (nullable) is the epilogue of the last launch: a slice phase un-applied in
// the adjoint, with its gradient terms.
static int bpm_generic_prop(fh_bpm* m, float2* f, const fh_plan* p, int direction,
                            const DnPhaseGradOp* gop = nullptr) {
  const int nx2 = 1 << p->lx, ny2 = 1 << p->ly;
  const double2 *twx = nullptr, *twy = nullptr;
  int rc;
  if ((rc = get_twiddles(m->ctx, nx2, &twx)) || (rc = get_twiddles(m->ctx, ny2, &twy))) return rc;
  LineIO<float2> xf{};
  xf.in = f; xf.out = p->scratch;
  xf.in_f = (long)m->nx * m->ny; xf.out_f = (long)nx2 * m->ny;
  xf.in_l = m->nx; xf.out_l = nx2; xf.in_e = xf.out_e = 1;
  xf.nlines = m->ny; xf.in_valid = m->nx; xf.out_count = nx2;
  FH_HIP((launch_fft<float2, double2, false, false>(p->lx, xf, m->batch, twx, NoOp{}, m->ctx->stream)));
  LineIO<float2> yc{};
  yc.in = p->scratch; yc.out = p->scratch;
  yc.in_f = yc.out_f = (long)nx2 * m->ny;
  yc.in_l = yc.out_l = 1; yc.in_e = yc.out_e = nx2;
  yc.nlines = nx2; yc.in_valid = yc.out_count = m->ny;
  QuadMulT<double2> q{p->tab_d, p->d_tabidx, p->tabsize, nx2, ny2, p->qy, direction < 0,
                      p->lam_tables.size() == 1, p->full};
  FH_HIP((launch_conv<float2, double2>(p->ly, yc, m->batch, twy, q, m->ctx->stream)));
  LineIO<float2> xi{};
  xi.in = p->scratch; xi.out = f;
  xi.in_f = (long)nx2 * m->ny; xi.out_f = (long)m->nx * m->ny;
  xi.in_l = nx2; xi.out_l = m->nx; xi.in_e = xi.out_e = 1;
  xi.nlines = m->ny; xi.in_valid = nx2; xi.out_count = m->nx;
  if (gop) FH_HIP((launch_fft<float2, double2, false, true>(p->lx, xi, m->batch, twx, *gop, m->ctx->stream)));
  else FH_HIP((launch_fft<float2, double2, false, true>(p->lx, xi, m->batch, twx, NoOp{}, m->ctx->stream)));
  return FH_OK;
}

// Generic forward (bpm.jl:244-260): P(dz/2); per slice [checkpoint], phase, P(dz) (P(dz/2) after
// the last slice).  Checkpoint c holds the field entering slice c * every.
static int bpm_generic_forward(fh_bpm* m, float2* f, const float* dn, int direction, int save) {
  const long npix = (long)m->nx * m->ny;
  const long fsz = npix * m->batch;
  int rc;
  if ((rc = bpm_generic_prop(m, f, m->phalf, direction))) return rc;
  for (int k = 0; k < m->nslices; ++k) {
    if (save && k % m->every == 0)
      FH_HIP(hipMemcpyAsync(m->ckpt + (long)(k / m->every) * fsz, f, fsz * sizeof(float2), hipMemcpyDeviceToDevice,
                            m->ctx->stream));
    if ((rc = bpm_phase(m, f, dn + (long)k * npix, direction, nullptr))) return rc;
    if ((rc = bpm_generic_prop(m, f, k + 1 < m->nslices ? m->pfull : m->phalf, direction))) return rc;
  }
  m->saved_dir = save ? direction : 0;
  return FH_OK;
}

// Generic adjoint (bpm.jl:289-312) with the dn gradient: the adjoint of the propagator before
// slice k (P(dz/2) for the last slice, P(dz) otherwise) un-applies slice k's phase in its last
// launch and writes the slice's gradient terms; a final P(dz/2) adjoint follows.  Slice inputs are
// recomputed one segment at a time from the checkpoints with the forward's own launches, before
// the propagation that needs them (one stream: the plans' scratch rows are reused in order).
static int bpm_generic_backward(fh_bpm* m, float2* f, const float* dn, float* ddn, int direction) {
  const long npix = (long)m->nx * m->ny;
  const long fsz = npix * m->batch;
  int rc;
  for (int k = m->nslices - 1; k >= 0; --k) {
    const float2* saved = nullptr;
    if (ddn) {
      const int k0 = (k / m->every) * m->every;
      if (k == m->nslices - 1 || k % m->every == m->every - 1) {
        const int k1 = std::min(k0 + m->every, m->nslices);
        FH_HIP(hipMemcpyAsync(m->work, m->ckpt + (long)(k0 / m->every) * fsz, fsz * sizeof(float2),
                              hipMemcpyDeviceToDevice, m->ctx->stream));
        for (int j = k0; j < k1; ++j) {
          if ((rc = bpm_phase(m, m->work, dn + (long)j * npix, direction, m->seg + (long)(j - k0) * fsz))) return rc;
          if (j + 1 < k1 && (rc = bpm_generic_prop(m, m->work, m->pfull, direction))) return rc;
        }
      }
      saved = m->seg + (long)(k - k0) * fsz;
    }
    float* ddn_k = ddn ? ddn + (long)k * npix : nullptr;
    const DnPhaseGradOp gop{dn + (long)k * npix, m->kl, saved, m->batch > 1 ? m->gpart : ddn_k, npix,
                            (double)direction, m->nx};
    if ((rc = bpm_generic_prop(m, f, k == m->nslices - 1 ? m->phalf : m->pfull, -direction, &gop))) return rc;
    if (ddn && m->batch > 1) {
      hipLaunchKernelGGL(k_sum_fields, dim3(elementwise_blocks(npix)), dim3(256), 0, m->ctx->stream,
                         (const float*)m->gpart, ddn_k, npix, m->batch);
      FH_HIP(hipGetLastError());
    }
  }
  return bpm_generic_prop(m, f, m->phalf, -direction);
}

template <int LOGN, bool ADJ>
static hipError_t bpm_line_launch_t(const BpmLine& a, const double2* tw, hipStream_t st) {
  using RD = typename Cfg<double, LOGN, false>::RD;
  auto kern = k_bpm_line<RD, ADJ>;
  const size_t lds = (size_t)LdsLine<RD::n, 1, sizeof(double2)>::size * sizeof(double2);
  hipError_t e = ensure_lds(kern, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(a.batch), dim3(RD::P), lds, st, a, tw);
  return hipGetLastError();
}

template <bool ADJ>
static hipError_t bpm_line_launch(int logn, const BpmLine& a, const double2* tw, hipStream_t st) {
  FH_LOGN_SWITCH(logn, (bpm_line_launch_t<LG, ADJ>(a, tw, st)))
}

// 1-D AS_BPM / Shift_BPM: one launch for the forward, one (plus the batch sum of the gradient
// terms) for the adjoint.
static int bpm_line_run(fh_bpm* m, float2* f, const float* dn, int direction, bool adj, bool with_save, float* ddn) {
  const double2* tw = nullptr;
  int rc;
  if ((rc = get_twiddles(m->ctx, 1 << m->lx, &tw))) return rc;
  BpmLine a{};
  a.u = f; a.dn = dn; a.kl = m->kl; a.tfull = m->tfull; a.thalf = m->thalf; a.tabidx = m->d_tabidx;
  a.n = m->nx; a.batch = m->batch; a.ns = m->nslices; a.dir = direction;
  if (!adj && with_save) a.save = m->seg;
  if (adj && ddn) {
    a.saved = m->seg;
    a.gpart = m->batch > 1 ? m->gpart : ddn;
  }
  FH_HIP(adj ? bpm_line_launch<true>(m->lx, a, tw, m->ctx->stream) : bpm_line_launch<false>(m->lx, a, tw, m->ctx->stream));
  if (adj && ddn && m->batch > 1) {
    const long total = (long)m->nslices * m->nx;
    hipLaunchKernelGGL(k_sum_line_grads, dim3((int)std::min<long>((total + 255) / 256, 4096)), dim3(256), 0,
                       m->ctx->stream, (const float*)m->gpart, ddn, m->nx, m->batch, m->nslices);
    FH_HIP(hipGetLastError());
  }
  return FH_OK;
}

// 1-D set-up: FP64 line tables of P(dz) and P(dz/2) per (wavelength, tilt), kdz per field, the
// slice-input stack and the gradient terms.
static int bpm_line_create(fh_bpm* m, const fh_prop_desc* d, bool shift) {
  const int nx = d->nx;
  for (double ty : m->tab_ty)
    if (ty != 0.0) return fail(FH_E_ARG, "a 1-D field has no y tilt");
  const bool pow2 = ilog2_exact(nx) >= 0;
  m->lx = pow2 ? ilog2_exact(nx) : ilog2_ceil(2 * nx - 1);
  m->ly = 0;
  m->oned = 1;
  if (nx < 2 || m->lx < kMinLog || m->lx > kMaxLog)
    return fail(FH_E_UNSUPP, "1-D BPM needs power-of-two nx in [%d, %d], or another nx with 2 nx - 1 <= %d (got %d)",
                1 << kMinLog, 1 << kMaxLog, 1 << kMaxLog, nx);
  bool tilted = false;
  for (double tx : m->tab_tx) tilted = tilted || tx != 0.0;
  m->mode = shift ? 3 : ((d->flags & FH_PROP_PARAXIAL) ? 1 : (tilted ? 2 : 0));
  m->track = (d->flags & FH_PROP_TRACK_TILTS) ? 1 : 0;
  const int N = 1 << m->lx;
  m->tabsize = N;
  const size_t ntab = m->lam_tables.size();
  const size_t stack = (size_t)m->nslices * m->batch * nx;
  if (hipMalloc(&m->tfull, ntab * N * sizeof(double2)) != hipSuccess ||
      hipMalloc(&m->thalf, ntab * N * sizeof(double2)) != hipSuccess ||
      hipMalloc(&m->d_tabidx, m->batch * sizeof(int)) != hipSuccess ||
      hipMalloc(&m->kl, m->batch * sizeof(double)) != hipSuccess ||
      hipMalloc(&m->seg, stack * sizeof(float2)) != hipSuccess ||
      (m->batch > 1 && hipMalloc(&m->gpart, stack * sizeof(float)) != hipSuccess))
    return fail(FH_E_NOMEM, "1-D BPM: out of device memory");
  std::vector<LineKey> kf, kh;
  for (size_t k = 0; k < ntab; ++k) {
    kf.push_back({m->lam_tables[k], m->dz, m->tab_tx[k]});
    kh.push_back({m->lam_tables[k], m->dz / 2, m->tab_tx[k]});
  }
  int rc;
  if ((rc = line_tables(m->ctx, nx, m->lx, d->dx, m->n0, m->mode, m->track, kf, nullptr, 0, nullptr, m->tfull)) ||
      (rc = line_tables(m->ctx, nx, m->lx, d->dx, m->n0, m->mode, m->track, kh, nullptr, 0, nullptr, m->thalf)))
    return rc;
  std::vector<double> kl(m->batch);
  for (int b = 0; b < m->batch; ++b) {
    const int t = m->field_table[b];
    kl[b] = 2.0 * M_PI * m->dz / cos(m->tab_tx[t]) / m->lam_tables[t];
  }
  if (hipMemcpy(m->kl, kl.data(), m->batch * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_tabidx, m->field_table.data(), m->batch * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return fail(FH_E_HIP, "1-D BPM: set-up failed");
  m->bytes = 2 * ntab * N * sizeof(double2) + stack * (sizeof(float2) + (m->batch > 1 ? sizeof(float) : 0));
  return FH_OK;
}

}  // namespace fh

extern "C" {

int fh_bpm_create_ex(fh_ctx* ctx, const fh_prop_desc* d, double thickness, int nslices, int checkpoint_every,
                     fh_bpm** out) {
  if (!out || !d) return fail(FH_E_ARG, "null descriptor or output pointer");
  *out = nullptr;
  if (!ctx) return fail(FH_E_ARG, "null context");
  if (d->flags & ~(FH_PROP_PARAXIAL | FH_PROP_TRACK_TILTS | FH_PROP_SHIFT)) return fail(FH_E_ARG, "unknown flags 0x%x", d->flags);
  if ((d->flags & FH_PROP_SHIFT) && (d->flags & (FH_PROP_PARAXIAL | FH_PROP_TRACK_TILTS)))
    return fail(FH_E_ARG, "Shift_BPM takes neither FH_PROP_PARAXIAL nor FH_PROP_TRACK_TILTS");
  if (d->z || d->nz) return fail(FH_E_ARG, "BPM slice distances follow from thickness and nslices (z must be null)");
  if (nslices < 2) return fail(FH_E_ARG, "BPM needs at least 2 slices, got %d", nslices);
  if (checkpoint_every < 1) return fail(FH_E_ARG, "checkpoint_every must be >= 1");
  if (!(thickness > 0) || !std::isfinite(thickness)) return fail(FH_E_ARG, "thickness must be positive");
  const bool shift = (d->flags & FH_PROP_SHIFT) != 0;
  const double n0 = shift ? 1.0 : d->n0;
  if (!(n0 > 0)) return fail(FH_E_ARG, "n0 must be positive");
  const int nx = d->nx, ny = d->ny, batch = d->batch;
  fh_plan tmp;  // wavelength/tilt bookkeeping and argument checks shared with the propagators
  int rc = plan_common(ctx, nx, ny, batch, d->dx, d->dy, d->lambdas, d->nlambdas, &tmp, nullptr, 0, d->tilts_x,
                       d->tilts_y, d->ntilts);
  if (rc) return rc;
  const int lx = ilog2_exact(nx), ly = ilog2_exact(ny);
  const bool pow2 = !(lx < kMinLog || lx > kMaxLog || ly < kMinLog || ly > kMaxLog);
  if ((rc = set_device(ctx))) return rc;
  fh_bpm* m = new fh_bpm();
  m->ctx = ctx; m->nx = nx; m->ny = ny; m->batch = batch; m->nslices = nslices; m->lx = lx; m->ly = ly;
  m->every = std::min(checkpoint_every, nslices);
  m->thickness = thickness; m->n0 = n0; m->dz = thickness / nslices;
  m->lam_tables = tmp.lam_tables; m->tab_tx = tmp.tab_tx; m->tab_ty = tmp.tab_ty;
  m->field_table = tmp.field_table;
  if (ny == 1) {  // 1-D fields
    if ((rc = bpm_line_create(m, d, shift))) { bpm_free(m); return rc; }
    *out = m;
    return FH_OK;
  }
  if (!pow2) {
    // other grids: general plans for P(dz) and P(dz/2) (the padded convolution form, run in
    // FP64 from their unrounded tables by bpm_generic_prop) and an elementwise slice phase
    m->generic = 1;
    fh_prop_desc pd = *d;
    double zf = m->dz, zh = m->dz / 2;
    pd.n0 = n0; pd.nz = 1;
    pd.z = &zf;
    if ((rc = asprop_create(ctx, &pd, &m->pfull, true))) { bpm_free(m); return rc; }
    pd.z = &zh;
    if ((rc = asprop_create(ctx, &pd, &m->phalf, true))) { bpm_free(m); return rc; }
    m->nck = (nslices + m->every - 1) / m->every;
    const size_t fsz = (size_t)nx * ny * batch * sizeof(float2);
    const size_t gbytes = batch > 1 ? (size_t)nx * ny * batch * sizeof(float) : 0;
    if (hipMalloc(&m->kl, batch * sizeof(double)) != hipSuccess || hipMalloc(&m->ckpt, m->nck * fsz) != hipSuccess ||
        hipMalloc(&m->seg, m->every * fsz) != hipSuccess || hipMalloc(&m->work, fsz) != hipSuccess ||
        (gbytes && hipMalloc(&m->gpart, gbytes) != hipSuccess)) {
      bpm_free(m);
      return fail(FH_E_NOMEM, "BPM: out of device memory");
    }
    std::vector<double> kl(batch);
    for (int b = 0; b < batch; ++b) {
      const int t = m->field_table[b];
      kl[b] = 2.0 * M_PI * m->dz / (cos(m->tab_tx[t]) * cos(m->tab_ty[t])) / m->lam_tables[t];
    }
    if (hipMemcpy(m->kl, kl.data(), batch * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      bpm_free(m);
      return fail(FH_E_HIP, "BPM: set-up failed");
    }
    // checkpoints, segment, work field, and per plan its x-spectrum rows and c64 + FP64 tables
    m->bytes = (m->nck + m->every + 1) * fsz + gbytes;
    for (const fh_plan* q : {m->pfull, m->phalf})
      m->bytes += ((size_t)1 << q->lx) * mid_rows(q) * batch * sizeof(float2) +
                  q->tabsize * q->lam_tables.size() * (sizeof(float2) + sizeof(double2));
    *out = m;
    return FH_OK;
  }
  // P is ASProp(dz; n0, paraxial) or ShiftProp(dz) (bpm.jl:54-73,140-142,207-217); the
  // kernel choice of angular_spectrum.jl:126-139
  const bool tilted = plan_tilted(&tmp);
  m->mode = shift ? 3 : ((d->flags & FH_PROP_PARAXIAL) ? 1 : (tilted ? 2 : 0));
  m->track = (d->flags & FH_PROP_TRACK_TILTS) ? 1 : 0;
  m->full = tilted ? 1 : 0;
  m->qx = m->full ? nx : nx / 2 + 1;
  m->qy = m->full ? ny + 1 : ny / 2 + 1;
  m->tabsize = (long)m->qx * m->qy;
  m->nck = (nslices + m->every - 1) / m->every;
  const size_t ntab = m->lam_tables.size();
  const size_t fsz = (size_t)nx * ny * batch * sizeof(float2);
  const size_t tbytes = m->tabsize * ntab * sizeof(double2);
  if (hipMalloc(&m->tfull, tbytes) != hipSuccess || hipMalloc(&m->thalf, tbytes) != hipSuccess ||
      hipMalloc(&m->d_tabidx, batch * sizeof(int)) != hipSuccess ||
      hipMalloc(&m->kl, batch * sizeof(double)) != hipSuccess ||
      hipMalloc(&m->ckpt, m->nck * fsz) != hipSuccess || hipMalloc(&m->seg, m->every * fsz) != hipSuccess ||
      hipMalloc(&m->seg2, m->every * fsz) != hipSuccess || hipMalloc(&m->work, fsz) != hipSuccess) {
    bpm_free(m);
    return fail(FH_E_NOMEM, "BPM: out of device memory");
  }
  if (hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_rec[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_rec[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_used[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_used[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_tab[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_tab[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_tused[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev_tused[1], hipEventDisableTiming) != hipSuccess) {
    bpm_free(m);
    return fail(FH_E_HIP, "BPM: stream or event creation failed");
  }
  // FP64 phase tables of two segments, when they fit kBpmPhaseTabBytes (otherwise the slice steps
  // evaluate their phases)
  const size_t cbytes = (size_t)m->every * nx * ny * batch * sizeof(double2);
  if (2 * cbytes <= kBpmPhaseTabBytes) {
    if (hipMalloc(&m->cis[0], cbytes) != hipSuccess || hipMalloc(&m->cis[1], cbytes) != hipSuccess) {
      bpm_free(m);
      return fail(FH_E_NOMEM, "BPM: out of device memory");
    }
  }
  const double scale = 1.0 / ((double)nx * (double)ny);
  for (size_t k = 0; k < ntab; ++k) {
    dim3 grid((m->qx + 255) / 256, m->qy);
    hipLaunchKernelGGL(k_as_table<double2>, grid, dim3(256), 0, ctx->stream, m->tfull + k * m->tabsize, nx, ny,
                       m->qx, m->qy, m->full, d->dx, d->dy, m->lam_tables[k], n0, m->dz, m->tab_tx[k],
                       m->tab_ty[k], m->mode, m->track, scale);
    hipLaunchKernelGGL(k_as_table<double2>, grid, dim3(256), 0, ctx->stream, m->thalf + k * m->tabsize, nx, ny,
                       m->qx, m->qy, m->full, d->dx, d->dy, m->lam_tables[k], n0, m->dz / 2, m->tab_tx[k],
                       m->tab_ty[k], m->mode, m->track, scale);
  }
  // kdz = 2*pi*dz / (cos(theta_x) cos(theta_y)) per field (bpm.jl:3-6,70), over the vacuum
  // wavelength (bpm.jl:235-242)
  std::vector<double> kl(batch);
  for (int b = 0; b < batch; ++b) {
    const int t = m->field_table[b];
    kl[b] = 2.0 * M_PI * m->dz / (cos(m->tab_tx[t]) * cos(m->tab_ty[t])) / m->lam_tables[t];
  }
  if (hipGetLastError() != hipSuccess ||
      hipMemcpy(m->kl, kl.data(), batch * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_tabidx, m->field_table.data(), batch * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    bpm_free(m);
    return fail(FH_E_HIP, "BPM: table set-up failed");
  }
  m->bytes = (m->nck + 2 * m->every + 1) * fsz + 2 * tbytes + (m->cis[0] ? 2 * cbytes : 0);
  *out = m;
  return FH_OK;
}

int fh_bpm_create(fh_ctx* ctx, int nx, int ny, int batch, double dx, double dy, const double* lambdas,
                  int nlambdas, double thickness, double n0, int nslices, int checkpoint_every,
                  fh_bpm** out) {
  fh_prop_desc d{};
  d.nx = nx; d.ny = ny; d.batch = batch; d.dx = dx; d.dy = dy;
  d.lambdas = lambdas; d.nlambdas = nlambdas; d.n0 = n0;
  return fh_bpm_create_ex(ctx, &d, thickness, nslices, checkpoint_every, out);
}

int fh_bpm_destroy(fh_bpm* m) {
  if (!m) return FH_OK;
  (void)hipSetDevice(m->ctx->device);
  bpm_free(m);
  return FH_OK;
}

int fh_bpm_workspace_bytes(fh_bpm* m, size_t* bytes) {
  if (!m || !bytes) return fail(FH_E_ARG, "null BPM or output");
  *bytes = m->bytes;
  return FH_OK;
}

int fh_bpm_forward(fh_bpm* m, void* u, const float* dn, int direction, int save) {
  if (!m || !u || !dn) return fail(FH_E_ARG, "null BPM, field or index profile");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(m->ctx))) return rc;
  float2* f = reinterpret_cast<float2*>(u);
  if (m->oned) {
    if ((rc = bpm_line_run(m, f, dn, direction, false, save != 0, nullptr))) return rc;
    m->saved_dir = save ? direction : 0;
    return FH_OK;
  }
  if (m->generic) return bpm_generic_forward(m, f, dn, direction, save);
  const long npix = (long)m->nx * m->ny;
  const long fsz = npix * m->batch;
  const double sgn = direction;
  // P(dz/2); per slice [IFFT_x, phase, FFT_x] + y pass (P(dz), or P(dz/2) after the
  // last slice); IFFT_x (bpm.jl:244-260)
  if ((rc = bpm_xfft(m, f, nullptr, sgn))) return rc;
  if ((rc = bpm_ypass(m, f, m->thalf, direction))) return rc;
  // segment by segment (checkpoint spacing); with phase tables the side stream fills segment
  // sg + 1's table (buffer (sg + 1) % 2, once segment sg - 1 has finished reading it) while the
  // caller's stream runs segment sg, which waits for its own table.  Every side-stream operation is
  // waited on by the caller's stream, so the work also records into one graph.
  const bool tab = m->cis[0] != nullptr;
  hipStream_t main = m->ctx->stream;
  const int nseg = m->nck;
  auto tables = [&](int sg) -> int {
    const int k0 = sg * m->every, k1 = std::min(k0 + m->every, m->nslices);
    if (sg >= 2) FH_HIP(hipStreamWaitEvent(m->side, m->ev_tused[sg % 2], 0));
    m->ctx->stream = m->side;
    const int r = bpm_cis(m, dn, k0, k1, sgn, m->cis[sg % 2]);
    m->ctx->stream = main;
    if (r) return r;
    FH_HIP(hipEventRecord(m->ev_tab[sg % 2], m->side));
    return FH_OK;
  };
  if (tab) {
    FH_HIP(hipEventRecord(m->ev_fork, main));
    FH_HIP(hipStreamWaitEvent(m->side, m->ev_fork, 0));
    if ((rc = tables(0))) return rc;
  }
  for (int sg = 0; sg < nseg; ++sg) {
    const int k0 = sg * m->every, k1 = std::min(k0 + m->every, m->nslices);
    if (tab) {
      if (sg + 1 < nseg && (rc = tables(sg + 1))) return rc;
      FH_HIP(hipStreamWaitEvent(main, m->ev_tab[sg % 2], 0));
    }
    for (int k = k0; k < k1; ++k) {
      float2* ck = (save && k == k0) ? m->ckpt + (long)sg * fsz : nullptr;
      const double2* ci = tab ? m->cis[sg % 2] + (long)(k - k0) * fsz : nullptr;
      if ((rc = bpm_step(m, f, dn + (long)k * npix, sgn, nullptr, ck, ci))) return rc;
      if ((rc = bpm_ypass(m, f, k + 1 < m->nslices ? m->tfull : m->thalf, direction))) return rc;
    }
    if (tab) FH_HIP(hipEventRecord(m->ev_tused[sg % 2], main));
  }
  if ((rc = bpm_xifft(m, f))) return rc;
  m->saved_dir = save ? direction : 0;
  return FH_OK;
}

int fh_bpm_backward(fh_bpm* m, void* du, const float* dn, float* ddn, int direction) {
  if (!m || !du || !dn) return fail(FH_E_ARG, "null BPM, cotangent or index profile");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(m->ctx))) return rc;
  if (ddn && m->saved_dir != direction)
    return fail(FH_E_ARG, "the dn gradient needs a saving fh_bpm_forward in the same direction first");
  float2* f = reinterpret_cast<float2*>(du);
  if (m->oned) return bpm_line_run(m, f, dn, direction, true, false, ddn);
  if (m->generic) return bpm_generic_backward(m, f, dn, ddn, direction);
  const long npix = (long)m->nx * m->ny;
  const long fsz = npix * m->batch;
  const double sgn = direction;
  // reverse loop of bpm.jl:289-312: adjoint of the trailing P(dz/2), then per slice
  // [IFFT_x, un-apply phase (+ gradient), FFT_x] + adjoint y pass, then IFFT_x
  if (!ddn) {
    if ((rc = bpm_xfft(m, f, nullptr, sgn))) return rc;
    if ((rc = bpm_ypass(m, f, m->thalf, -direction))) return rc;
    for (int k = m->nslices - 1; k >= 0; --k) {
      if ((rc = bpm_step(m, f, dn + (long)k * npix, -sgn, nullptr, nullptr))) return rc;
      if ((rc = bpm_ypass(m, f, k > 0 ? m->tfull : m->thalf, -direction))) return rc;
    }
    return bpm_xifft(m, f);
  }
  // with the gradient: segment s's slice inputs are recomputed on the second stream into
  // buffer s % 2 while the caller's stream runs segment s+1's gradient steps from the other
  // buffer; events order each recompute after the forward's checkpoints (fork) and after the
  // last use of its buffer, and each segment's steps after its recompute.  Every side-stream
  // operation is waited on by the caller's stream, so the work also records into one graph.
  hipStream_t main = m->ctx->stream;
  float2* bufs[2] = {m->seg, m->seg2};
  const int nseg = m->nck;
  FH_HIP(hipEventRecord(m->ev_fork, main));
  FH_HIP(hipStreamWaitEvent(m->side, m->ev_fork, 0));
  // with phase tables, the side stream fills segment sg's table (buffer sg % 2, ordered like the
  // segment buffer) before recomputing the segment from it; its gradient steps read the same table
  auto recompute = [&](int sg) -> int {
    const int k0 = sg * m->every, k1 = std::min(k0 + m->every, m->nslices);
    m->ctx->stream = m->side;
    const double2* ci = m->cis[sg % 2];
    int r = ci ? bpm_cis(m, dn, k0, k1, sgn, m->cis[sg % 2]) : FH_OK;
    if (!r) r = bpm_recompute(m, dn, k0, k1, sgn, direction, bufs[sg % 2], ci);
    m->ctx->stream = main;
    if (r) return r;
    FH_HIP(hipEventRecord(m->ev_rec[sg % 2], m->side));
    return FH_OK;
  };
  if ((rc = recompute(nseg - 1))) return rc;
  if ((rc = bpm_xfft(m, f, nullptr, sgn))) return rc;
  if ((rc = bpm_ypass(m, f, m->thalf, -direction))) return rc;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    if (sg >= 1) {
      if (sg + 1 < nseg) FH_HIP(hipStreamWaitEvent(m->side, m->ev_used[(sg + 1) % 2], 0));
      if ((rc = recompute(sg - 1))) return rc;
    }
    FH_HIP(hipStreamWaitEvent(main, m->ev_rec[sg % 2], 0));
    const int k0 = sg * m->every, k1 = std::min(k0 + m->every, m->nslices);
    for (int k = k1 - 1; k >= k0; --k) {
      const double2* ci = m->cis[sg % 2] ? m->cis[sg % 2] + (long)(k - k0) * fsz : nullptr;
      if ((rc = bpm_step_grad(m, f, dn + (long)k * npix, sgn, bufs[sg % 2] + (long)(k - k0) * fsz,
                              ddn + (long)k * npix, ci)))
        return rc;
      if ((rc = bpm_ypass(m, f, k > 0 ? m->tfull : m->thalf, -direction))) return rc;
    }
    FH_HIP(hipEventRecord(m->ev_used[sg % 2], main));
  }
  return bpm_xifft(m, f);
}

}  // extern "C"

// ================================================================ propagation groups
// A group runs several plans (fields of different sizes, wavelengths and distances) as one
// hipGraph: the plans' launch chains are captured on parallel streams (a fork/join), so
// small fields fill the GPU next to large ones and the whole batch costs one graph launch.
// Graphs are cached per (direction, field pointers).

struct fh_group {
  fh_ctx* ctx = nullptr;
  std::vector<fh_plan*> plans;
  hipStream_t cap = nullptr;
  std::vector<hipStream_t> side;
  std::vector<int> lane;       // side stream of plans[i]: (index of the plan's first listing) mod nside
  std::vector<hipEvent_t> ev;  // ev[0]: fork; ev[1 + s]: join of side stream s
  struct Graph {
    int direction;
    std::vector<void*> fields;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
};

namespace fh {

static size_t plan_field_bytes(const fh_plan* p) { return (size_t)p->nx * p->ny * p->batch * sizeof(float2); }

static void group_free(fh_group* g) {
  if (!g) return;
  for (auto& gr : g->graphs) {
    (void)hipGraphExecDestroy(gr.exec);
    (void)hipGraphDestroy(gr.graph);
  }
  for (auto e : g->ev) (void)hipEventDestroy(e);
  for (auto s : g->side) (void)hipStreamDestroy(s);
  if (g->cap) (void)hipStreamDestroy(g->cap);
  delete g;
}

// Capture every plan's propagation of fields[i] on side stream g->lane[i].  A plan listed more
// than once (one propagator for several same-shaped fields) always gets the same side stream:
// its launches share the plan's workspace and must run one after the other.
static int group_capture(fh_group* g, void* const* fields, int direction, fh_group::Graph& out) {
  fh_ctx* c = g->ctx;
  hipStream_t user = c->stream;
  int rc = FH_OK;
  FH_HIP(hipStreamBeginCapture(g->cap, hipStreamCaptureModeRelaxed));
  hipError_t e = hipEventRecord(g->ev[0], g->cap);
  for (size_t s = 0; e == hipSuccess && s < g->side.size(); ++s) e = hipStreamWaitEvent(g->side[s], g->ev[0], 0);
  for (size_t i = 0; e == hipSuccess && rc == FH_OK && i < g->plans.size(); ++i) {
    c->stream = g->side[g->lane[i]];
    float2* f = reinterpret_cast<float2*>(fields[i]);
    rc = run_prop_plan(g->plans[i], f, direction);
    if (rc == FH_OK) rc = run_xifft(g->plans[i], f, direction);
  }
  c->stream = user;
  for (size_t s = 0; e == hipSuccess && s < g->side.size(); ++s) {
    e = hipEventRecord(g->ev[1 + s], g->side[s]);
    if (e == hipSuccess) e = hipStreamWaitEvent(g->cap, g->ev[1 + s], 0);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(g->cap, &graph);
  if (rc != FH_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess || ee != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(FH_E_HIP, "graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : ee));
  }
  hipGraphExec_t exec = nullptr;
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail(FH_E_HIP, "graph instantiation failed");
  }
  out.direction = direction;
  out.fields.assign(fields, fields + g->plans.size());
  out.graph = graph;
  out.exec = exec;
  return FH_OK;
}

}  // namespace fh

extern "C" {

int fh_group_create(fh_ctx* ctx, int nplans, fh_plan* const* plans, int autotune_reps, fh_group** out) {
  if (!out) return fail(FH_E_ARG, "null output pointer");
  *out = nullptr;
  if (!ctx || !plans || nplans < 1) return fail(FH_E_ARG, "need a context and at least one plan");
  if (autotune_reps < 0) return fail(FH_E_ARG, "autotune_reps must be >= 0");
  for (int i = 0; i < nplans; ++i) {
    if (!plans[i]) return fail(FH_E_ARG, "plan %d is null", i);
    if (plans[i]->ctx != ctx) return fail(FH_E_ARG, "plan %d belongs to another context", i);
    if (plans[i]->oned) return fail(FH_E_UNSUPP, "plan %d is a 1-D plan: one launch already covers its batch", i);
  }
  int rc;
  if ((rc = set_device(ctx))) return rc;
  fh_group* g = new fh_group();
  g->ctx = ctx;
  g->plans.assign(plans, plans + nplans);
  std::vector<int> first(nplans);  // distinct plans in order of first listing
  int ndistinct = 0;
  for (int i = 0; i < nplans; ++i) {
    int j = 0;
    while (j < i && plans[j] != plans[i]) ++j;
    first[i] = j == i ? ndistinct++ : first[j];
  }
  const int nside = ndistinct < 4 ? ndistinct : 4;
  for (int i = 0; i < nplans; ++i) g->lane.push_back(first[i] % nside);
  // streams and events at the default priority
  if (hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking) != hipSuccess) { group_free(g); return fail(FH_E_HIP, "stream creation failed"); }
  for (int s = 0; s < nside; ++s) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) { group_free(g); return fail(FH_E_HIP, "stream creation failed"); }
    g->side.push_back(st);
  }
  for (int k = 0; k < 1 + nside; ++k) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { group_free(g); return fail(FH_E_HIP, "event creation failed"); }
    g->ev.push_back(e);
  }
  // One eager propagation per plan on a scratch field fills every lazily built table
  // (twiddles, stage tables, LDS attributes) so that capture records kernels only; with
  // autotune_reps > 0 each plan's y pass is also tuned on that scratch field.
  size_t big = 0;
  for (auto* p : g->plans) big = std::max(big, plan_field_bytes(p));
  void* tmp = nullptr;
  if (hipMalloc(&tmp, big) != hipSuccess) { group_free(g); return fail(FH_E_NOMEM, "group warm-up field: out of device memory"); }
  rc = FH_OK;
  if (hipMemsetAsync(tmp, 0, big, ctx->stream) != hipSuccess) rc = fail(FH_E_HIP, "warm-up clear failed");
  for (auto* p : g->plans) {
    if (rc == FH_OK && autotune_reps > 0) rc = fh_autotune(p, tmp, autotune_reps, nullptr, 0);
    if (rc == FH_OK) rc = fh_propagate(p, tmp, 1);
    if (rc == FH_OK) rc = fh_propagate(p, tmp, -1);
  }
  if (rc == FH_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(FH_E_HIP, "group warm-up failed");
  (void)hipFree(tmp);
  if (rc) { group_free(g); return rc; }
  *out = g;
  return FH_OK;
}

int fh_group_propagate(fh_group* g, void* const* fields, int direction) {
  if (!g || !fields) return fail(FH_E_ARG, "null group or field list");
  int rc;
  if ((rc = check_dir(direction)) || (rc = set_device(g->ctx))) return rc;
  for (size_t i = 0; i < g->plans.size(); ++i) {
    if (!fields[i]) return fail(FH_E_ARG, "field %zu is null", i);
    for (size_t j = 0; j < i; ++j)  // side streams would propagate one buffer twice at once
      if (fields[j] == fields[i]) return fail(FH_E_ARG, "fields %zu and %zu are the same buffer", j, i);
  }
  fh_group::Graph* gr = nullptr;
  for (auto& x : g->graphs)
    if (x.direction == direction && std::equal(x.fields.begin(), x.fields.end(), fields)) { gr = &x; break; }
  if (!gr) {
    if (g->graphs.size() >= 8) {  // bounded cache: drop the oldest graph once it has finished
      FH_HIP(hipStreamSynchronize(g->ctx->stream));
      (void)hipGraphExecDestroy(g->graphs.front().exec);
      (void)hipGraphDestroy(g->graphs.front().graph);
      g->graphs.erase(g->graphs.begin());
    }
    fh_group::Graph ng{};
    if ((rc = group_capture(g, fields, direction, ng))) return rc;
    g->graphs.push_back(ng);
    gr = &g->graphs.back();
  }
  FH_HIP(hipGraphLaunch(gr->exec, g->ctx->stream));
  return FH_OK;
}

int fh_group_graph_count(fh_group* g, int* count) {
  if (!g || !count) return fail(FH_E_ARG, "null group or output");
  *count = (int)g->graphs.size();
  return FH_OK;
}

int fh_group_destroy(fh_group* g) {
  if (!g) return FH_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  group_free(g);
  return FH_OK;
}

}  // extern "C"
