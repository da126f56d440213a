/*
 * fr_oracle.h -- CPU oracle for the fastrecon gridding / image hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and nothing else.  The product (fastrecon_amd) never links or calls it.
 *
 * It restates the reference CPU path of pengwg/fastrecon in plain C:
 *   - FP32 restatement, operation-for-operation faithful to the reference
 *     (same evaluation order, same float/double promotions, no contraction),
 *   - an FP64 variant that takes the same FP32 tap decisions (same LUT index,
 *     same accept test) but accumulates in double, isolating FP32 rounding,
 *   - an FP64 mixed-radix FFT, fftshift, crop and sum-of-squares.
 * Each function cites the reference file:line it follows.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   or_kb_lut        -- scipy.special.i0 fixture + closed forms     (pinned)
 *   or_i0            -- scipy.special.i0 at many points             (pinned)
 *   or_normalize     -- closed-form affine map, min->0, max->N-1    (pinned)
 *   or_grid_f32/f64  -- single-sample known answers, whole-grid brute
 *                       force (numpy), linearity, 3D-slab == 2D     (pinned)
 *   or_fft_c2c       -- numpy.fft (independent library), Parseval   (pinned)
 *   or_fftshift      -- numpy.fft.fftshift, involution              (pinned)
 *   or_sos           -- closed form in numpy                        (pinned)
 *   or_lowfilter     -- float64 exp(-floor(r/att*100)/100) window,
 *                       centre unchanged, zero beyond the cut-off    (pinned)
 *   or_normalize     -- unit coil norm per voxel, phase ratios kept (pinned)
 *   or_sos_map       -- map == data gives SOS^2 (real), conj-linear
 *                       in the map                                 (pinned)
 *   or_degrid_f64    -- adjoint identity against or_grid_f64        (pinned)
 *   or_degrid_f32    -- FP64 degrid to FP32 rounding, one-tap exact  (pinned)
 */
#ifndef FR_ORACLE_H
#define FR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Modified Bessel function of the first kind, order 0 (FP64 power series). */
double or_i0(double x);

/* KB LUT, ConvKernel::ConvKernel (ReconData/ConvKernel.cpp:7-30).
 * Writes L floats; returns beta (float) through *beta_out if non-null. */
void or_kb_lut(float kWidth, float og, int L, float *lut, float *beta_out);

/* Component bounds (ReconData.cpp:56-81 storeTrajComponent: minmax_element)
 * and ReconData::normalizeTraj / transformLocalTraj (ReconData.cpp:201-225).
 * traj is AoS with `stride` floats per sample; dim components are
 * transformed in place.  bounds[2*d] = min, bounds[2*d+1] = max of comp d
 * (before normalisation); they are updated like m_bounds. */
void or_traj_bounds(const float *traj, int64_t M, int stride, int dim, float *bounds);
void or_normalize_traj(float *traj, int64_t M, int stride, int dim, float *bounds, unsigned size);

/* GridLut<float>::griddingChannel (ReconData/GridLut.cpp:80-147), FP32,
 * for C coils.  traj: AoS, `stride` floats per sample (normalised grid
 * coordinates).  data: C x M complex64 (re,im interleaved, coil-major).
 * out: C x N^dim complex64, zeroed by the caller or not (we zero it).
 * nthreads: OpenMP threads over coils (GridLut.cpp:56-76); <=0 means 1. */
void or_grid_f32(int dim, unsigned N, int64_t M, int C,
                 const float *traj, int stride, const float *dcf,
                 const float *data, float kWidth, const float *lut, int L,
                 float *out, int nthreads);

/* Same tap decisions as or_grid_f32, accumulation in double. */
void or_grid_f64(int dim, unsigned N, int64_t M, int C,
                 const float *traj, int stride, const float *dcf,
                 const float *data, float kWidth, const float *lut, int L,
                 double *out, int nthreads);

/* FP32 degridding in tap order (the device kernels' arithmetic). */
void or_degrid_f32(int dim, unsigned N, int64_t M, int C,
                   const float *traj, int stride,
                   const float *grid, float kWidth, const float *lut, int L,
                   float *out);

/* Adjoint of or_grid_f64 (degridding): out[c][s] = sum_taps w * grid[c][cell]
 * (no dcf).  Same taps as gridding. */
void or_degrid_f64(int dim, unsigned N, int64_t M, int C,
                   const float *traj, int stride,
                   const double *grid, float kWidth, const float *lut, int L,
                   double *out);

/* In-place unnormalised C2C FFT of `batch` arrays of shape (nz,ny,nx)
 * (x fastest), sign = +1 (FFTW_BACKWARD, ImageFilter/FFT.cpp:44-47) or -1.
 * dim=2 uses (ny,nx).  FP64. */
void or_fft_c2c(int dim, unsigned N, int batch, double *data, int sign);

/* fftShift2/fftShift3 (ImageFilter/ImageFilter.cpp:231-280), in place,
 * `batch` images of N^dim complex (float or double variants). */
void or_fftshift_f32(int dim, unsigned N, int batch, float *data);
void or_fftshift_f64(int dim, unsigned N, int batch, double *data);

/* ImageFilter::SOS(ImageSize) (ImageFilter.cpp:159-209) followed by the
 * magnitude written by Main.cpp:175-181: out[rc] = |sqrt(sum_c v conj v)|.
 * FP32 semantics, coils summed in order. in: C x N^dim complex64. */
void or_sos_f32(int dim, unsigned N, int C, const float *in,
                unsigned rcx, unsigned rcy, unsigned rcz, float *out);
void or_sos_f64(int dim, unsigned N, int C, const double *in,
                unsigned rcx, unsigned rcy, unsigned rcz, double *out);

/* ImageFilter::lowFilter (ImageFilter.cpp:56-93): Gaussian-like k-space
 * window exp(-idx/100), idx = (int)((float)r / att * 100.0),
 * att = (float)(2.0*res*res/4.0), r = |(x,y,z) - size/2|^2 (integer),
 * zero where idx >= 2000.  In place on C x N^dim complex64. */
void or_lowfilter_f32(int dim, unsigned N, int C, float *data, int res);
/* ImageFilter::normalize (ImageFilter.cpp:95-118): per voxel
 * mag = sum_c std::norm(v_c) (libstdc++: fl(cabsf(v)^2)), v_c /= sqrtf(mag). */
void or_normalize_f32(int dim, unsigned N, int C, float *data);
/* ImageFilter::SOS(map, size) (ImageFilter.cpp:170-209), the PILS combine:
 * out[rc] = sum_c v_c * conj(map_c) over the centred crop, FP32 complex
 * arithmetic in coil order.  out: rc complex64. */
void or_sos_map_f32(int dim, unsigned N, int C, const float *in, const float *map,
                    unsigned rcx, unsigned rcy, unsigned rcz, float *out);

/* Pipe-Menon iteration w <- w / G^H G w on the oracle's own FP64 gridding
 * (single channel, real weights).  Not in the reference (net-new). */
void or_pipe_menon_f64(int dim, unsigned N, int64_t M, const float *traj, int stride,
                       float kWidth, const float *lut, int L, int iters, double *w);

#ifdef __cplusplus
}
#endif
#endif
