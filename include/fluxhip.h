/* libfluxhip: MI355X (gfx950) propagation path of FluxOptics.jl behind a plain C ABI.
 *
 * Fields are complex FP32 (c64, interleaved re/im) device arrays laid out like the
 * reference's column-major ScalarField electric[nx, ny, batch]: element (i, j, b) at
 * offset i + nx*(j + ny*b) (src/Fields.jl:118-149).  Masks are float[ny][nx] (the
 * reference's phi[nx, ny]), shared by every field of the batch.
 *
 * Conventions:
 *  - every function returns 0 on success and a negative FH_E* code on failure;
 *    fh_last_error() gives the message of the calling thread's last failure;
 *  - direction is +1 (Forward) or -1 (Backward), as sign(::Type{<:Direction})
 *    (src/optical_components/OpticalComponents.jl:34-99);
 *  - wavelengths, distances and sampling are doubles; a ComplexF32 ScalarField stores
 *    lambda as Float32 (src/Fields.jl:28-32), so a caller reproducing the reference
 *    passes (double)(float)lambda;
 *  - all work is asynchronous on the context's stream; the caller owns field, mask and
 *    gradient buffers; the library owns plans, tables and workspace;
 *  - 1-D fields (a ScalarField with Nd = 1, electric[nx, batch], src/Fields.jl:135-136) are
 *    passed with ny = 1 (their masks and dn slices are rows of nx values).  The reference's 1-D
 *    kernels are its 2-D ones at fy = 0 and theta_y = 0 (angular_spectrum.jl:14-25,37-43,64-76,
 *    shift_kernel.jl:8-12), so ASProp, ShiftProp, AS_BPM and Shift_BPM accept ny = 1 for any
 *    power-of-two nx in [16, 4096] or any other nx with 2 nx - 1 <= 4096 (a propagation is one
 *    launch: FFT, transfer function and IFFT per line); RSProp, CollinsProp and the CZT have no
 *    1-D form in the reference and refuse them; fh_fft2 takes 1-D fields of power-of-two nx.
 */
#ifndef FLUXHIP_H
#define FLUXHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FH_OK 0
#define FH_E_ARG -1      /* invalid argument (sizes, directions, null pointers) */
#define FH_E_HIP -2      /* a HIP runtime call failed */
#define FH_E_NOMEM -3    /* device allocation failed */
#define FH_E_UNSUPP -4   /* valid but unsupported (e.g. a non power-of-two grid) */

typedef struct fh_ctx fh_ctx;
typedef struct fh_plan fh_plan;
typedef struct fh_bpm fh_bpm;
typedef struct fh_group fh_group;

/* Library version (major*10000 + minor*100 + patch). */
int fh_version(void);

/* Message of the calling thread's last failure ("" if none). */
const char* fh_last_error(void);

/* Largest supported transform length per axis (ASProp: n itself, a power of two; RS pads
 * to the power of two >= 2n - 1, so RS fields are limited to half of it). */
int fh_max_length(void);

/* Context: one HIP device and one stream (NULL = the device's default stream).
 * Replaces the CUDA extension hook (src/cuda/CUDAExt.jl:10-37). */
int fh_ctx_create(int device, void* stream, fh_ctx** out);
int fh_ctx_set_stream(fh_ctx* ctx, void* stream);
int fh_ctx_synchronize(fh_ctx* ctx);
int fh_ctx_destroy(fh_ctx* ctx);

/* ASProp: angular-spectrum propagator for fields of shape [nx, ny, batch].
 * lambdas has nlambdas entries: 1 (shared) or batch (one per field).  Power-of-two grids run
 * the three-pass FFT form; any other grid (2n - 1 <= fh_max_length()) runs as the equivalent
 * zero-padded linear convolution with h = IDFT_n(H) extended periodically (the n-point
 * circular convolution the reference's FFT computes), with tables built in FP64.
 * Replaces ASProp(u, ds, z; n0) with use_cache=true, double_precision_kernel=true
 * (src/optical_components/freespace_propagators/angular_spectrum.jl:193-229,1-12):
 * the transfer function is computed once in FP64 and rounded to c64. */
int fh_asprop_create(fh_ctx* ctx, int nx, int ny, int batch, double dx, double dy,
                     const double* lambdas, int nlambdas, double z, double n0,
                     fh_plan** out);

/* RSProp: Rayleigh-Sommerfeld propagator (rayleigh_sommerfeld.jl:128-164), any nx, ny with
 * 2n - 1 <= fh_max_length().  The reference pads to (2n-1)^2; this library pads each axis to
 * the power of two >= 2n - 1 ((2n)^2 for power-of-two n), which yields the same cropped linear
 * convolution.  The kernel's FFT is computed in FP64 on the device and rounded once. */
int fh_rsprop_create(fh_ctx* ctx, int nx, int ny, int batch, double dx, double dy,
                     const double* lambdas, int nlambdas, double z, fh_plan** out);

/* Plan descriptor for fields with per-field distances and tilts, and paraxial kernels.
 * lambdas: 1 or batch values; z: 1 or batch values (one transfer function per distinct
 * (lambda, z, tilt) in the batch); tilts_x/tilts_y: 0, 1 or batch angles in radians, the
 * field's Fourier offset f0 = sin(theta)/lambda' (Fields.jl:34-37,168-180; a ComplexF32
 * field stores them as Float32, pass Float64(Float32(theta)) for parity).  n0 is used by
 * ASProp only. */
#define FH_PROP_PARAXIAL 1    /* ASProp(...; paraxial=true): as_paraxial_kernel */
#define FH_PROP_TRACK_TILTS 2 /* track_tilts=true */
#define FH_PROP_SHIFT 4       /* ShiftProp (shift_kernel.jl): H = cis(-2 pi z (tan(tx) fx + tan(ty) fy)) */
typedef struct fh_prop_desc {
  int nx, ny, batch;
  double dx, dy;
  const double* lambdas;
  int nlambdas;
  const double* z;
  int nz;
  const double* tilts_x;
  const double* tilts_y;
  int ntilts;
  double n0;
  int flags;
} fh_prop_desc;

/* ASProp with the reference's kernel choice (angular_spectrum.jl:126-139): paraxial
 * (as_paraxial_kernel, :27-35) when FH_PROP_PARAXIAL is set, else on-axis (as_kernel) when
 * every tilt is zero, else tilted (tilted_as_kernel, :45-62), with per-field z.  With
 * FH_PROP_SHIFT the plan is a ShiftProp (shift_kernel.jl:1-117): the geometric shift by
 * z tan(theta) of each field, as a Fourier-domain linear phase. */
int fh_asprop_create_ex(fh_ctx* ctx, const fh_prop_desc* desc, fh_plan** out);

/* ASProp with a frequency filter: ASProp(u, ds, z; filter) (angular_spectrum.jl:1-76,193-229),
 * whose kernel multiplies the transfer function by Complex(filter(...)).  The filter is a Julia
 * function, so the caller samples it on the host, in FP64, on the fftfreq grid of the plan:
 * filter holds nfilters grids of ny*nx complex doubles (re, im interleaved), grid f element
 * (i, j) at 2*(f*nx*ny + j*nx + i), sampled at the arguments the reference's kernel passes for
 * frequency indices (i, j) in FFT order (fx_i = fftfreq(nx, 1/dx)[i], likewise fy_j):
 * filter(fx_i, fy_j) for on-axis and paraxial kernels, filter(fx_i + f0x, fy_j + f0y) with
 * f0 = sin(theta)/lambda' per field for the tilted kernel.  nfilters is 1 (shared) or batch (one
 * grid per field, e.g. tilted fields).  The samples multiply the FP64 transfer function before
 * it is rounded; as in the reference, the on-axis and tilted kernels conjugate them with the rest
 * of the kernel when z <= 0, and Backward conjugates the whole kernel.  Not for FH_PROP_SHIFT. */
int fh_asprop_create_filtered(fh_ctx* ctx, const fh_prop_desc* desc, const double* filter, int nfilters,
                              fh_plan** out);

/* RSProp with per-field z and tilted fields (rs_tilted_kernel, rayleigh_sommerfeld.jl:14-27,
 * selected at :90-96 when a tilt is nonzero).  flags: FH_PROP_TRACK_TILTS only. */
int fh_rsprop_create_ex(fh_ctx* ctx, const fh_prop_desc* desc, fh_plan** out);

/* CollinsProp(u, ds, ds', (A, B, D)) (collins_integral.jl:200-301; FourierLens is (0, f, 0),
 * fourier_lens.jl:38-48; ParaxialProp with resampling is (1, z, 1)): the A chirp, the
 * zero-padded convolution with cis(pi*(ax x^2 + ay y^2)/(B lambda)) normalised by
 * prod(ds ./ sqrt(i B)), and the D chirp, alpha = ds'/ds, chirps on the input grid.  In the
 * plan the chirps ride on the first and last x passes, so it costs what an RSProp does; any n
 * with 2n - 1 <= fh_max_length().
 * Backward conjugates every factor and normalises by prod(ds' ./ sqrt(-i B)), as the
 * reference does (collins_integral.jl:189).  desc->z must be NULL; flags:
 * FH_PROP_TRACK_TILTS only (untracked tilts add cis(+-2 pi x sin(theta)/lambda) to A, D).
 * The caller relabels the field's sampling (ds' after Forward, ds after Backward). */
int fh_collins_create(fh_ctx* ctx, const fh_prop_desc* desc, double dx_out, double dy_out, double a,
                      double b, double d, fh_plan** out);

/* plan_czt(x, dims, a, w; center_on_grid) (FFTutils.jl:30-177): the chirp z-transform
 * X[m] = sum_k x[k] a^-k w^(m k) along the dimensions in the bitmask dims (1 = x, 2 = y),
 * with k, m shifted by (n-1)/2 when center_on_grid, by Bluestein's identity on the padded
 * chirp / convolution / chirp plan (any n with 2n - 1 <= fh_max_length()).  a and w are host
 * arrays in polar form
 * {rho_x, theta_x, rho_y, theta_y} (convert_czt_arg); the reference's defaults are
 * a = (1, 0) and w = (1, -2 pi / n) per dimension, for which the CZT is the DFT.
 * fh_propagate(plan, x, +1) applies czt!, fh_propagate(plan, x, -1) the adjoint plan's czt!
 * (FFTutils.jl:143-155). */
int fh_czt_create(fh_ctx* ctx, int nx, int ny, int batch, int dims, const double* a, const double* w,
                  int center_on_grid, fh_plan** out);

int fh_plan_destroy(fh_plan* plan);

/* propagate!(u, p, direction): in place on a c64 [nx, ny, batch] device array
 * (freespace.jl:29-63; Backward multiplies by the conjugate kernel, which is the
 * exact adjoint of Forward). */
int fh_propagate(fh_plan* plan, void* u, int direction);

/* One pass of fh_propagate: pass 0 = x-line FFT, 1 = y-line FFT * kernel * IFFT,
 * 2 = x-line IFFT.  Running 0, 1, 2 in order equals fh_propagate.  Consecutive
 * propagators may skip a pass 2 followed by a pass 0 on the same field (the IFFT/FFT
 * cancellation of merge_rules.jl:104-108); used too for per-kernel timing. */
int fh_propagate_pass(fh_plan* plan, void* u, int direction, int pass);

/* Launch variants of a pass (pass 1: radix split and lines per workgroup; passes 0 and 2:
 * lane mapping).  count receives the number of variants, current the selected one
 * (-1 = built-in default). */
int fh_plan_variants(fh_plan* plan, int pass, int* count, int* current);
int fh_plan_set_variant(fh_plan* plan, int pass, int variant);

/* Autotuner: times every variant of each of the plan's three passes on the device array u
 * (its contents are overwritten), keeps the fastest and writes a JSON summary into report. */
int fh_autotune(fh_plan* plan, void* u, int reps, char* report, int report_len);

/* Phase then propagation fused: u <- P(u * cis(s*phi)); if saved != NULL the mask
 * input u is stored there first (propagate_and_save!, phasemask.jl:100-113). */
int fh_phase_propagate(fh_plan* plan, void* u, const float* phi, void* saved, int direction);

/* Phase on its own: u <- u * cis(s*phi), saving the input if saved != NULL
 * (phasemask.jl:90-113). */
int fh_phase_apply(fh_ctx* ctx, void* u, const float* phi, int nx, int ny, int batch,
                   int direction, void* saved);

/* Phase adjoint with the mask gradient (backpropagate_with_gradient!,
 * phasemask.jl:142-150 and compute_phase_gradient!, :115-140):
 *   du <- dv * cis(-s*phi)   (in place in dv)
 *   dphi <- s * sum_b Im(du * conj(saved))   (added to dphi when accumulate != 0). */
int fh_phase_backward(fh_ctx* ctx, void* dv, const float* phi, const void* saved,
                      float* dphi, int nx, int ny, int batch, int direction, int accumulate);

/* Adjoint propagation then Phase adjoint with gradient, fused:
 *   du <- P^H(dv) * cis(-s*phi), dphi (+)= s * sum_b Im(du * conj(saved)). */
int fh_propagate_phase_backward(fh_plan* plan, void* dv, const float* phi, const void* saved,
                                float* dphi, int direction, int accumulate);

/* The plain GPU form used as a comparator: the same propagation in complex FP64
 * (c128 [nx, ny, batch]), per field, unfused (five launches per field: x FFT, y FFT,
 * multiply by a full FP64 transfer-function table, y IFFT, x IFFT). */
int fh_propagate_plain_fp64(fh_plan* plan, void* u_c128, int direction);

/* FourierOperator (fourier_operator.jl:65-79, FFTutils.jl:13-28): in-place 2-D DFT of a c64
 * [nx, ny, batch] device array, u <- scale * sum u e^{sign 2 pi i jk/n}.  The reference's FFT
 * is sign = -1, scale = 1; its IFFT is sign = +1, scale = 1/(nx*ny); the adjoint of
 * (sign, scale) is (-sign, scale).  Power-of-two sizes run the line FFTs; other sizes (2n - 1 <=
 * fh_max_length()) the DFT as a chirp z-transform (a plan cached in the context). */
int fh_fft2(fh_ctx* ctx, void* u, int nx, int ny, int batch, int sign, double scale);

/* SquaredIntensityDifference (squared_intensity_difference.jl:78-90, compute_metric and
 * backpropagate_metric).  target holds target_batch FP32 intensity planes [nx, ny]:
 *   target_batch == batch ("full dimensions"): loss[b] = sum (|u_b|^2 - I_b)^2, b < batch;
 *   target_batch == 1 < batch ("spatial only"): loss[0] = sum (sum_b |u_b|^2 - I)^2.
 * loss is a device array of doubles (batch or 1 values), summed in FP64 in a fixed order.
 * When du is not NULL the pullback du_b = 4 dc[slot] d u_b is written there (du may equal u);
 * dc is a device array with one cotangent per loss value, or NULL for ones. */
int fh_sq_intensity_diff(fh_ctx* ctx, const void* u, const float* target, int nx, int ny, int batch,
                         int target_batch, const double* dc, double* loss, void* du);

/* fh_sq_intensity_diff without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The block partials live in a fixed per-GPU array inside the library
 * (2 MiB, held until the process ends; its address is looked up when the first context on the
 * GPU is created), so a call writes at most 1024 loss values (batch fields with per-field
 * targets, or one spatial-only value): more returns FH_E_ARG.  Results equal
 * fh_sq_intensity_diff's bit for bit.  Every call on one GPU shares that array: calls that
 * overlap in time on one GPU (other streams, other contexts, other host threads) race; calls in
 * order on one stream do not.  No state carries over between calls (each call writes every
 * partial it reads). */
int fh_sq_intensity_diff_noalloc(fh_ctx* ctx, const void* u, const float* target, int nx, int ny, int batch,
                                 int target_batch, const double* dc, double* loss, void* du);

/* ASPropZ (angular_spectrum.jl:236-300), an on-axis ASProp whose distance z is trainable: dL/dz per
 * field (out: batch doubles, device) from V = FFT(v) of the output and DV = FFT(dv) of its cotangent
 * (unnormalised 2-D FFTs, e.g. fh_fft2 with sign -1): Re (1/N) sum conj(DV) 2 pi i w~ V with
 * w = sqrt((n0/lambda)^2 - fx^2 - fy^2) (principal root), w~ = w for z >= 0 and conj(w) for z < 0
 * (the kernel's H(z) = conj(H(-z)), angular_spectrum.jl:1-12).  lambdas: 1 or batch values (host). */
int fh_as_z_gradient(fh_ctx* ctx, const void* V, const void* DV, int nx, int ny, int batch, double dx, double dy,
                     const double* lambdas, int nlambdas, double n0, double z, double* out);

/* SquaredFieldDifference (squared_field_difference.jl, compute_metric and backpropagate_metric):
 * loss[b] = sum |u_b - v_b|^2 over the pixels of field b (summed in FP64 in a fixed order); when du
 * is not NULL the pullback du_b = 2 dc[b] (u_b - v_b) is written there (du may equal u); dc as in
 * fh_sq_intensity_diff. */
int fh_sq_field_diff(fh_ctx* ctx, const void* u, const void* v, int nx, int ny, int batch, const double* dc,
                     double* loss, void* du);

/* The tail of a differentiated system ending in [Phase ->] propagator -> SquaredIntensityDifference,
 * fused: propagate u (the Phase, when phi is given, rides on the first pass and saves its input to
 * saved), then one pass does the last x IFFT, the metric values (loss[b], one per field; the target
 * has the field's full dims, or both are single), the pullback 4 dc (|u|^2 - I) u and the first x FFT
 * of the adjoint propagation; the adjoint then finishes (un-applying the Phase and accumulating its
 * gradient into dphi when given, as fh_propagate_phase_backward).  u ends as the cotangent at the
 * block's input.  Equals fh_phase_propagate / fh_propagate + fh_sq_intensity_diff +
 * fh_propagate_phase_backward / fh_propagate(-direction) with two HBM passes fewer.  ASProp and
 * RSProp plans. */
int fh_sqint_block(fh_plan* plan, void* u, const float* target, int target_batch, const double* dc,
                   double* loss, int direction, const float* phi, void* saved, float* dphi, int accumulate);

/* fh_sqint_block on a block whose input is a system's source field: ScalarSource's propagate copies
 * u0 into its output field before the first component runs (scalar_source.jl:68-71); here the first
 * pass reads src itself and writes u (which need not hold anything), and the Phase gradient takes src
 * as its saved input, so neither the copy nor the save is made.  src is left unchanged and must not
 * overlap u.  cotangent != 0: u ends as the cotangent at the block's input, as in fh_sqint_block (the
 * gradient of a trainable source, scalar_source.jl:76-79); cotangent == 0: u ends unspecified, and
 * the last pass writes only the mask gradient (no adjoint at all without a mask gradient). */
int fh_sqint_block_src(fh_plan* plan, const void* src, void* u, const float* target, int target_batch,
                       const double* dc, double* loss, int direction, const float* phi, float* dphi,
                       int accumulate, int cotangent);

/* AS_BPM: split-step beam propagation through nslices slices of index perturbation dn
 * (AS_BPM(u, thickness, n0, dn), bpm.jl:130-142 and :8-74): dz = thickness/nslices,
 * P(dz) and P(dz/2) are ASProps in index n0, the slice phase is
 * cis(s*kdz/lambda*dn_k) with kdz = 2*pi*dz and the vacuum wavelength.
 * The reference keeps every slice input for the gradient (bpm.jl:49,227-231); this
 * library keeps one checkpoint every checkpoint_every slices and recomputes a segment
 * at a time in the adjoint, bitwise equal to a forward that keeps them all. */
int fh_bpm_create(fh_ctx* ctx, int nx, int ny, int batch, double dx, double dy,
                  const double* lambdas, int nlambdas, double thickness, double n0,
                  int nslices, int checkpoint_every, fh_bpm** out);

/* 1-D fields (ny = 1) run every slice of one field in one workgroup, in FP64 registers across
 * the slices, and a saving forward keeps every slice input (nslices * batch * nx c64, as the
 * reference's saved stack), so checkpoint_every does not apply to them. */
/* The general BPM constructor (bpm.jl:54-73): the field template's tilts (per field or
 * shared) and the slice propagator from desc->flags -- ASProp(dz; n0) with the reference's
 * kernel choice and FH_PROP_PARAXIAL / FH_PROP_TRACK_TILTS (AS_BPM, bpm.jl:130-142), or
 * ShiftProp(dz) with FH_PROP_SHIFT (Shift_BPM, bpm.jl:207-217; desc->n0 unused).  kdz is
 * 2*pi*dz / (cos(theta_x) cos(theta_y)) per field (bpm.jl:3-6,70).  desc->z must be NULL:
 * the slice distances follow from thickness and nslices. */
int fh_bpm_create_ex(fh_ctx* ctx, const fh_prop_desc* desc, double thickness, int nslices,
                     int checkpoint_every, fh_bpm** out);
int fh_bpm_destroy(fh_bpm* bpm);

/* Device memory owned by the BPM object (checkpoints, two segment buffers, transfer-function
 * tables and, when two segments' FP64 slice phases fit 512 MiB, the two phase tables), in bytes. */
int fh_bpm_workspace_bytes(fh_bpm* bpm, size_t* bytes);

/* propagate!(u, bpm, direction) (bpm.jl:244-260) in place; dn is float
 * [nslices][ny][nx] (the reference's dn[nx, ny, nslices]).  save != 0 keeps the
 * checkpoints that fh_bpm_backward needs for the dn gradient
 * (propagate_and_save!, bpm.jl:262-273).  Work the BPM runs on its own second stream (the
 * next segment's slice phases; in fh_bpm_backward the recomputed segments) is ordered after
 * the context's stream and joined back into it before the call returns its last launch. */
int fh_bpm_forward(fh_bpm* bpm, void* u, const float* dn, int direction, int save);

/* backpropagate!(du, bpm, direction; u_saved, dp) (bpm.jl:289-312): du <- adjoint(du)
 * in place; if ddn != NULL it receives the dn gradient [nslices][ny][nx], summed over
 * the batch (requires a saving forward in the same direction with the same dn). */
int fh_bpm_backward(fh_bpm* bpm, void* du, const float* dn, float* ddn, int direction);

/* Propagation groups (no counterpart in the reference, which launches per field): many
 * plans (fields of uneven sizes, wavelengths, distances) run as one hipGraph whose plan
 * chains are captured on parallel streams.  fh_group_create warms every plan up on a
 * scratch field (and, with autotune_reps > 0, tunes each plan's y pass there); all plans
 * must share ctx.  fh_group_propagate(g, fields, direction) propagates fields[i] with
 * plans[i] in place, launching the graph on the context's stream; graphs are captured on
 * first use and cached per (direction, field pointers), at most 8. */
int fh_group_create(fh_ctx* ctx, int nplans, fh_plan* const* plans, int autotune_reps, fh_group** out);
int fh_group_propagate(fh_group* group, void* const* fields, int direction);
int fh_group_graph_count(fh_group* group, int* count);
int fh_group_destroy(fh_group* group);

/* offset_tilts!(u, tilts) (Fields.jl:280-293), the operation of TiltAnchor (tilt_anchor.jl:
 * 56-77): u_b <- u_b * cis(2*pi/lambda_b * ((sin(fx_b) - sin(tx_b)) x + (sin(fy_b) - sin(ty_b)) y))
 * on the centred grid (GridUtils.jl:56-60), from the field's tilts (from_x, from_y) to new ones
 * (to_x, to_y); host arrays of 1 or batch entries (ntilts), lambdas of 1 or batch. */
int fh_offset_tilts(fh_ctx* ctx, void* u, int nx, int ny, int batch, double dx, double dy,
                    const double* lambdas, int nlambdas, const double* from_x, const double* from_y,
                    const double* to_x, const double* to_y, int ntilts);

/* Mask (modulators/mask.jl:95-141): u <- u * m (Forward) or u * conj(m) (Backward) with m a
 * c64 [nx, ny] device transmission shared by the batch; saved (nullable) receives the input.
 * fh_mask_backward: dm (+)= sum_b conj_direction(dv * conj(saved)) (complex gradient, from the
 * cotangent at the mask output), then dv <- dv * conj_direction(m, reverse(direction)). */
int fh_mask_apply(fh_ctx* ctx, void* u, const void* m, int nx, int ny, int batch, int direction, void* saved);
int fh_mask_backward(fh_ctx* ctx, void* dv, const void* m, const void* saved, void* dm, int nx, int ny,
                     int batch, int direction, int accumulate);

/* TeaDOE / TeaReflector (modulators/tea_doe.jl:146-199): per field b, with the device arrays
 * kl[b] = 2 pi dn(lambda_b) / lambda_b and r (2*batch doubles: re, im; NULL = 1), u_b <- u_b *
 * r_b * cis(kl_b h) (Forward) or conj(r_b) * cis(-kl_b h) (Backward); h is an FP32 [nx, ny]
 * height map.  fh_tea_backward: dv <- dv * conj(t), then dh (+)= s * sum_b kl_b Im(dv conj(saved)). */
int fh_tea_apply(fh_ctx* ctx, void* u, const float* h, int nx, int ny, int batch, const double* kl,
                 const double* r, int direction, void* saved);
int fh_tea_backward(fh_ctx* ctx, void* dv, const float* h, const void* saved, float* dh, int nx, int ny,
                    int batch, const double* kl, const double* r, int direction, int accumulate);

/* One optimiser step fused with its proximal operator over n FP32 parameter values (complex
 * parameters as interleaved re, im).  rule: 0 Descent, 1 Momentum(eta, rho), 2 Nesterov(eta,
 * rho) (Optimisers.jl 0.4.6, the reference's dependency), 3 Fista (OptimisersExt.jl:12-40;
 * pass eta = the squared rate the reference stores and beta = (t_k - 1)/t_{k+1}); state holds
 * the rule's buffer (momentum, velocity or the previous iterate; NULL for Descent).  prox
 * (proximal_operators/): 0 none, 1 ClampProx(p0, p1), 2 IstaProx(p0, p1), 3 PositiveProx. */
int fh_optim_step(fh_ctx* ctx, int rule, float* x, const float* dx, float* state, long n, double eta,
                  double rho, double beta, int prox, double p0, double p1);

/* pad! / crop! (pad_crop_operators.jl:1-16,61-79) on c64 fields of batch fields: dst (dnx x dny per
 * field) receives dst(x, y) = src(x - x0, y - y0) where that lies inside src (snx x sny per field),
 * else fill.  pad!(u_pad, u; offset, pad_val) is x0, y0 = offset >= 0, fill = pad_val;
 * crop!(u_crop, u; offset) is x0, y0 = -offset.  src and dst must not overlap. */
int fh_window(fh_ctx* ctx, const void* src, int snx, int sny, void* dst, int dnx, int dny, int batch, int x0,
              int y0, double fill_re, double fill_im);

/* Fista (OptimisersExt.jl:12-40) with its coefficients in device memory, so that a recorded step
 * replays correctly: tb holds {t_k, beta} (start with t_1 = 1); each call sets beta = (t_k - 1) /
 * t_{k+1}, t_k <- t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2, then runs fh_optim_step's Fista update with
 * that beta (eta = the squared rate the reference stores; state = the previous iterate). */
int fh_fista_step(fh_ctx* ctx, float* x, const float* dx, float* state, long n, double eta, double* tb, int prox,
                  double p0, double p1);

/* Adam (Optimisers.jl 0.4.6, the reference's dependency; the rule of its published
 * beam-splitter example): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, x -= eta m/(1 - bt1) /
 * (sqrt(v/(1 - bt2)) + eps), then the proximal operator as in fh_optim_step.  state holds 2n
 * FP32 values (m, then v), zero before the first step; bt is a device array {b1^t, b2^t} for
 * step t ({b1, b2} before the first step) that the call advances on the stream, so a recorded
 * training step replays correctly. */
int fh_adam_step(fh_ctx* ctx, float* x, const float* dx, float* state, long n, double eta, double beta1,
                 double beta2, double eps, double* bt, int prox, double p0, double p1);

/* DotProduct and PowerCoupling, mode-selective (metrics/dot_product.jl:60-106,
 * power_coupling.jl:55-69): c[b] = ds * sum u_b conj(v_b) into the device complex128 array c
 * (FP64, fixed summation order), and p[b] = |c[b]|^2 when p is not NULL.  The pullback writes
 * du_b = v_b * g_b with g_b = dc[b] (DotProduct; dc complex128) or, when dc is NULL,
 * g_b = 2 c[b] dp[b] (PowerCoupling; dp doubles, NULL = ones).  As in the reference, the
 * pullback carries no ds factor. */
int fh_dot_product(fh_ctx* ctx, const void* u, const void* v, int nx, int ny, int batch, double ds,
                   void* c, double* p);
int fh_dot_product_pullback(fh_ctx* ctx, const void* v, const void* dc, const void* c, const double* dp,
                            int nx, int ny, int batch, void* du);

/* The mode-selective DotProduct / PowerCoupling pullback fused with the adjoint propagation before
 * it in a differentiated system ([Phase ->] propagator -> metric, ChainRulesCoreExt.jl:21-111):
 * du <- P^H(v_b g_b) with g_b = dc[b] (DotProduct) or 2 c[b] dp[b] (PowerCoupling; dp NULL = ones),
 * the first pass of the adjoint reading the reference fields v itself, so the cotangent v g is never
 * stored; with phi the Phase adjoint and its gradient follow in the last pass, as in
 * fh_propagate_phase_backward.  direction is the forward propagation's.  Equals
 * fh_dot_product_pullback then fh_propagate(plan, du, -direction) [then fh_phase_backward] with one
 * pass fewer; du must not overlap v.  ASProp and RSProp plans. */
int fh_dot_pullback_propagate(fh_plan* plan, const void* v, const void* dc, const void* c, const double* dp, void* du,
                              int direction, const float* phi, const void* saved, float* dphi, int accumulate);

/* DotProduct and PowerCoupling with mode_selective=false (metrics/dot_product.jl:53-66,85-97,
 * power_coupling.jl:55-69): the overlap matrix of the batch's n <= 32 modes,
 * c[i + n*j] = ds * sum_x conj(v_i(x)) u_j(x) (the reference's column-major c[i, j] = V'U, complex128,
 * FP64 sums in a fixed order), and p[i + n*j] = |c[i + n*j]|^2 when p is not NULL.  The pullback
 * writes du_j = sum_i v_i G[i + n*j] with G = dc (DotProduct; complex128, same layout) or, when dc is
 * NULL, G = 2 c o dp (PowerCoupling; dp doubles, NULL = ones); no ds factor, as in the reference.
 * du must not overlap v.  More than 32 modes return FH_E_UNSUPP. */
int fh_dot_matrix(fh_ctx* ctx, const void* u, const void* v, int nx, int ny, int batch, double ds, void* c,
                  double* p);
int fh_dot_matrix_pullback(fh_ctx* ctx, const void* v, const void* dc, const void* c, const double* dp, int nx, int ny,
                           int batch, void* du);

#ifdef __cplusplus
}
#endif

#endif /* FLUXHIP_H */
