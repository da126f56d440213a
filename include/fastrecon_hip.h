/*
 * fastrecon_hip.h -- C ABI of libfastrecon_hip.so, the MI355X (gfx950)
 * backend behind the fastrecon C++ API (ReconData / GridLut / ImageFilter /
 * ImageData).
 *
 * The reference (pengwg/fastrecon) has no C ABI: its GPU backend is a set of
 * C++ subclasses (cuReconData, cuGridLut, cuImageData, cuImageFilter) that
 * the factories return.  Each entry point below replaces one piece of that
 * backend; the reference interface it stands in for is cited on each.
 *
 * Conventions
 *   - Every call returns fr_status (FR_OK == 0) and never exits the process;
 *     fr_last_error() gives a message for the last failure on a context.
 *   - Handles are opaque and own their device memory.
 *   - All device work of a context runs in order on the context's one HIP
 *     stream (default priority).  Calls that read or write host memory
 *     synchronise that stream before returning, so host pointers are only
 *     borrowed for the duration of the call.  Calls taking device pointers
 *     are asynchronous on the context stream.
 *   - One context per host thread.
 *   - Results are deterministic: the gridding kernels use no global atomics.
 *     With exact_order every grid cell accumulates in ascending sample order,
 *     the order of the reference CPU loop (ReconData/GridLut.cpp:96-145), and
 *     the grid is bit-identical to it; otherwise sums are grouped (sample-
 *     range slices, four-product matrix-core steps) in a fixed order.
 */
#ifndef FASTRECON_HIP_H
#define FASTRECON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FR_OK = 0,
    FR_ERR_INVALID = 1,      /* bad argument / size mismatch                */
    FR_ERR_HIP = 2,          /* HIP runtime error                           */
    FR_ERR_NOMEM = 3,        /* device allocation failed                    */
    FR_ERR_STATE = 4,        /* call out of order (e.g. execute before plan)*/
    FR_ERR_UNSUPPORTED = 5,  /* configuration this build does not handle   */
    FR_ERR_LIBRARY = 6       /* rocFFT or another library failed            */
} fr_status;

typedef enum { FR_SAMPLES_C64 = 0, FR_SAMPLES_C16 = 1 } fr_sample_type;
typedef enum { FR_LAYOUT_COIL_MAJOR = 0, FR_LAYOUT_SAMPLE_MAJOR = 1 } fr_layout;

typedef struct fr_ctx fr_ctx;
typedef struct fr_samples fr_samples;
typedef struct fr_plan fr_plan;
typedef struct fr_grids fr_grids;
typedef struct fr_fft fr_fft;

/* Tuning knobs for a gridding plan; zero fields mean "choose automatically".
 * The LUT length is never tuned in parity mode: it is the L passed in. */
typedef struct {
    int tile_x, tile_y, tile_z;  /* base work-unit box (one wave owns a unit) */
    int slabs;                   /* waves per workgroup                        */
    int accumulate;              /* where a unit's cells accumulate:
                                    0: automatic (matrix cores; LDS read-add-write
                                       for exact_order plans),
                                    1: LDS float atomic add, 2: LDS read-add-write,
                                    3: VGPRs -- refused (FR_ERR_UNSUPPORTED),
                                    4: gather -- per-batch cell-sorted taps, the
                                       batch's sample rows in LDS, one VGPR per
                                       cell (unit box of at most 32 cells),
                                    5: matrix cores -- 16-cell blocks of a unit
                                       accumulate four records per
                                       v_mfma_f32_16x16x4_f32, weights from the
                                       plan's taps (reference tap set and LUT
                                       index); sums group four products, so
                                       exact_order is refused,
                                    6: matrix cores, weights computed from the
                                       sample positions with the reference tap
                                       test and LUT index (no taps stored; unit
                                       box of at most 16 cells, LUT of at most
                                       2048 entries); the automatic choice when
                                       these hold; exact_order is refused */
    int split_taps;              /* heavier units are split (0: automatic); for
                                    accumulate 6 the unit is records, not taps:
                                    a unit's records are cut into slices of at
                                    most this many (rounded up to 16)          */
    int exact_order;             /* 1: split units by space only, so every cell
                                    sums in ascending sample order (bit-identical
                                    to the CPU path); 0: heavy units are also split
                                    by sample range into slices combined in a fixed
                                    order (deterministic, not bitwise CPU order) */
    int items_per_wave;          /* accumulate 6: light items one wave grids in
                                    turn (0: 1)                                 */
    int xcd_run;                 /* accumulate 6: consecutive items each of the 8
                                    XCDs takes in launch order (0: 256)         */
    int stores;                  /* accumulate 6: grid stores 0: automatic
                                    (cached), 1: streaming (nontemporal), 2: cached */
    int waves_per_simd;          /* accumulate 6: 0: automatic (4), 3 or 4 (the
                                    kernel variant compiled for that occupancy) */
    int tile_writes;             /* accumulate 6: 0: automatic (pairs), 1: four
                                    x-adjacent units per workgroup written out
                                    together in whole rows, 2: off (each unit
                                    written by its wave), 3: pairs (a wave
                                    grids two x-adjacent units and writes them
                                    out together in rows of two units)     */
} fr_tune;

typedef struct {
    int64_t samples, taps, records, streams;
    int tile_x, tile_y, tile_z, slabs, tiles, lds_bytes;
    int64_t plan_bytes;
    double build_ms;
    int64_t items, heavy_units, max_unit_taps;
    int exact_order;
    int accumulate;              /* the plan's resolved accumulation mode (1, 2, 4, 5 or 6) */
    int64_t scratch_bytes;       /* heavy-unit slice partials of one pass, per 32 coils */
} fr_plan_info;

/* ------------------------------------------------------------ host helpers */
/* ConvKernel::ConvKernel (ReconData/ConvKernel.cpp:7-30): the L-entry radial
 * Kaiser-Bessel LUT, normalised so lut[0] == 1; beta (may be NULL) gets the
 * Beatty et al. shape parameter.  Host-only, no device needed. */
fr_status fr_kb_lut(float kWidth, float og, int L, float *lut, float *beta);
/* ReconData::storeTrajComponent bounds (ReconData.cpp:56-81) followed by
 * normalizeTraj (ReconData.cpp:211-225), in place on a host AoS trajectory:
 * every component uses component 0's bounds.  bounds (may be NULL) receives
 * the normalised (min, max) per component, like m_bounds.  Host-only. */
fr_status fr_traj_normalize(float *traj, int64_t M, int stride, int dim, unsigned N, float *bounds);
/* transformLocalTraj (ReconData.cpp:201-209): x <- (x + tr) * scale. */
fr_status fr_traj_transform(float *traj, int64_t M, int stride, int dim, float tr, float scale);

/* ---------------------------------------------------------------- context */
/* No reference counterpart (the reference uses the CUDA default stream and
 * checks no errors, cuGridLut.cu:303). */
fr_status fr_ctx_create(int device, fr_ctx **out);
fr_status fr_ctx_destroy(fr_ctx *ctx);
const char *fr_last_error(const fr_ctx *ctx);
fr_status fr_ctx_synchronize(fr_ctx *ctx);
void *fr_ctx_stream(fr_ctx *ctx);              /* the hipStream_t */

/* ---------------------------------------------------------------- samples */
/* Replaces cuReconData (ReconData/cuReconData.h:6-47): device copies of the
 * trajectory, DCF and all coils (the reference keeps one coil resident,
 * cuReconData.cu:12-31).  Coordinates must already be normalised to grid
 * units by ReconData::normalizeTraj (ReconData.cpp:211-225), which the C++
 * layer does on the host with the reference's exact FP32 arithmetic. */
fr_status fr_samples_create(fr_ctx *ctx, int dim, int64_t M, int coils,
                            int sample_type, fr_samples **out);
fr_status fr_samples_destroy(fr_samples *s);
/* traj: M samples, `stride` floats apart, `dim` coordinates each (the
 * reference's Point<float> AoS is stride 4, ReconData.h:12-15). */
fr_status fr_samples_set_traj(fr_samples *s, const float *traj, int stride);
/* ReconData::setDcf (ReconData.cpp:45-54) / cuGetDcf (cuReconData.cu:39-50). */
fr_status fr_samples_set_dcf(fr_samples *s, const float *dcf);
/* ReconData::addChannelData (ReconData.cpp:33-43) / cuGetChannelData
 * (cuReconData.cu:12-31): M complex64 values of coil c (converted to FP16
 * on the device for FR_SAMPLES_C16). */
fr_status fr_samples_set_coil(fr_samples *s, int coil, const void *data);
/* ReconData::updateSingleAcquisition (ReconData.cpp:179-189,
 * cuReconData.cu:52-62): overwrite n samples of coil c starting at first. */
fr_status fr_samples_update(fr_samples *s, int coil, int64_t first, int64_t n, const void *data);
/* Device-pointer variants (asynchronous): data in complex64, either coil-
 * major [C][M] or sample-major [M][C]; dcf and traj in device memory. */
fr_status fr_samples_set_data_device(fr_samples *s, const void *data, int layout);
fr_status fr_samples_set_dcf_device(fr_samples *s, const float *dcf);
fr_status fr_samples_set_traj_device(fr_samples *s, const float *traj, int stride);
fr_status fr_samples_get_dcf(const fr_samples *s, float *host_dcf);

/* ------------------------------------------------------------------- plan */
/* Replaces cuGridLut::plan (ReconData/cuGridLut.cu:151-226), the sorted-bin
 * gridding LUT.  N = unsigned(reconSize * og) (GridLut.cpp:45); lut is the
 * ConvKernel table of length L (ConvKernel.cpp:7-30); kWidth the full
 * kernel width.  The plan depends on the trajectory only. */
fr_status fr_plan_create(fr_ctx *ctx, const fr_samples *s, unsigned N, float kWidth,
                         const float *lut, int L, const fr_tune *tune, fr_plan **out);
fr_status fr_plan_destroy(fr_plan *p);
fr_status fr_plan_get_info(const fr_plan *p, fr_plan_info *info);

/* ------------------------------------------------------------------ grids */
/* Replaces cuImageData (ImageData/cuImageData.h:7-36): C device-resident
 * grids of N^dim complex64, index x + N*y + N*N*z (GridLut.cpp:111). */
fr_status fr_grids_create(fr_ctx *ctx, int dim, unsigned N, int coils, fr_grids **out);
/* Batched 2D frames (net-new: the reference reconstructs a series as
 * separate ReconData/GridLut instances that share one plan, GridLut.h:24-26
 * setIndex, cuGridLut.h:50-56).  Frame f owns frame_samples[f] consecutive
 * samples; fr_samples_set_traj takes the 2D points of all frames (each
 * frame normalised with its own bounds, as its own ReconData would be).  A
 * plan built on such samples grids every frame in one launch into grids from
 * fr_grids_create_frames, laid out [coil][frame][N^2]; FFT plans for them use
 * batch = coils * frames; fr_sos writes one rcx x rcy image per frame. */
fr_status fr_samples_create_frames(fr_ctx *ctx, int frames, const int64_t *frame_samples, int coils,
                                   int sample_type, fr_samples **out);
fr_status fr_grids_create_frames(fr_ctx *ctx, unsigned N, int coils, int frames, fr_grids **out);
fr_status fr_grids_destroy(fr_grids *g);
fr_status fr_grids_download(const fr_grids *g, int coil, void *host);
fr_status fr_grids_upload(fr_grids *g, int coil, const void *host);
void *fr_grids_device_ptr(fr_grids *g);

/* Replaces cuGridLut::execute / griddingChannel / gridding_kernel
 * (cuGridLut.cu:28-48, 231-309) and GridLut::execute (GridLut.cpp:53-78):
 * grids every coil of s into out (all coils in one launch). */
fr_status fr_grid_execute(fr_plan *p, const fr_samples *s, fr_grids *out);

/* fr_grid_execute without any device allocation: the heavy-unit slices'
 * partial sums, the only device memory a pass needs beyond its inputs and
 * outputs (which fr_grid_execute allocates in the plan on first use), go to
 * a static 32 MiB scratch array that each GPU holds for the life of the
 * process.  No counterpart in the reference; for callers that cannot
 * allocate (HIP graph capture, serving loops).  Same kernels and results as
 * fr_grid_execute.  Fails with FR_ERR_UNSUPPORTED when the plan needs more
 * scratch than the array holds (fr_plan_info.scratch_bytes per 32 coils).
 * Calls that overlap on one GPU share that array and race. */
fr_status fr_grid_execute_noalloc(fr_plan *p, const fr_samples *s, fr_grids *out);

/* Degridding (grid -> samples), the adjoint of fr_grid_execute with the
 * same taps, no DCF.  Net-new (the reference has none).  s supplies the
 * trajectory the plan was built on; out: device complex64, sample-major
 * [M][g->C]. */
fr_status fr_degrid_execute(fr_plan *p, const fr_samples *s, const fr_grids *g, void *out_dev);

/* -------------------------------------------------------------------- FFT */
/* Replaces cuFFT (ImageFilter/cuFFT.cpp:15-37) and FFT (FFT.cpp:30-91):
 * one batched in-place C2C rocFFT plan over all coils, unnormalised.
 * sign +1 is FFTW_BACKWARD / CUFFT_INVERSE, the reference default. */
fr_status fr_fft_create(fr_ctx *ctx, int dim, unsigned N, int batch, int sign, fr_fft **out);
fr_status fr_fft_destroy(fr_fft *f);
fr_status fr_fft_execute(fr_fft *f, fr_grids *g);

/* FFT + fftShift + crop + sum-of-squares in one call: out_dev receives what
 * fr_fft_execute, then fr_sos(..., apply_shift, deapod_dev, out_dev) would
 * give (FFT.cpp:82-84, ImageFilter.cpp:40-53,159-209, Main.cpp:175-181), to
 * FP32 rounding.  For N = 256, 384 or 512 and rcx a multiple of 16 it runs
 * cropped transforms that keep only the rows and columns reaching the crop
 * (the grids are left untransformed); otherwise it runs exactly those two
 * calls (the grids are left transformed).  No counterpart in the reference. */
fr_status fr_fft_sos(fr_fft *f, fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz, int apply_shift,
                     const float *deapod_dev, float *out_dev);

/* ---------------------------------------------------------- image stage */
/* ImageFilter::fftShift (ImageFilter.cpp:40-53, 231-280), in place, even N. */
fr_status fr_fftshift(fr_ctx *ctx, fr_grids *g);
/* Fused crop + sum-of-squares + magnitude (ImageFilter.cpp:159-209 and
 * Main.cpp:175-181): out_dev[rcz][rcy][rcx] = sqrt(sum_c |v_c|^2) / deapod,
 * coils summed in order in FP32.  apply_shift != 0 folds fftShift into the
 * read (the grids are left unshifted).  deapod_dev may be NULL. */
fr_status fr_sos(fr_ctx *ctx, const fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz,
                 int apply_shift, const float *deapod_dev, float *out_dev);
/* ImageFilter::lowFilter (ImageFilter.cpp:55-93), in place on all coils;
 * the 2000-entry expf table is evaluated on the host exactly as the
 * reference does. */
fr_status fr_lowfilter(fr_ctx *ctx, fr_grids *g, int res);
/* ImageFilter::normalize (ImageFilter.cpp:95-118): v /= sqrtf(sum_c |v_c|^2). */
fr_status fr_normalize(fr_ctx *ctx, fr_grids *g);
/* ImageFilter::crop (ImageFilter.cpp:120-157): out_dev = C crops of
 * rcx*rcy*rcz complex64 (coil-major). */
fr_status fr_crop(fr_ctx *ctx, const fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz, void *out_dev);
/* ImageFilter::SOS(map, size) (ImageFilter.cpp:170-209), the PILS combine:
 * out_dev[rc] = sum_c v_c * conj(map_c) (complex64, no square root);
 * apply_shift folds fftShift into the reads of both inputs. */
fr_status fr_sos_map(fr_ctx *ctx, const fr_grids *g, const fr_grids *map, unsigned rcx, unsigned rcy, unsigned rcz,
                     int apply_shift, void *out_dev);
/* Device copy of all coils between grids of the same shape (ImageData copy,
 * ImageData.cpp:83-95). */
fr_status fr_grids_copy(fr_ctx *ctx, const fr_grids *src, fr_grids *dst);
/* Deapodisation map (net-new): |FFT| of one unit sample gridded at the
 * normalised position `centre` with the same LUT, shifted and cropped.
 * out_dev: rcx*rcy*rcz floats. */
fr_status fr_deapod_map(fr_ctx *ctx, int dim, unsigned N, float kWidth, const float *lut, int L,
                        const float *centre, unsigned rcx, unsigned rcy, unsigned rcz,
                        float *out_dev);

/* --------------------------------------------------- density compensation */
/* Pipe-Menon iterative DCF (net-new): w <- w / G^H G w, `iters` times,
 * starting from the samples' current DCF (or 1 if init_ones), written back
 * to the samples' DCF.  Uses this plan's gridding and degridding kernels.
 * The first call (of this or fr_dcf_identity_scale) gives the plan a
 * single-channel workspace (the trajectory, unit data, one grid, M complex
 * densities) that it keeps until fr_plan_destroy. */
fr_status fr_dcf_pipe_menon(fr_plan *p, fr_samples *s, int iters, int init_ones);

/* Global scale of the samples' DCF (net-new) from the Pipe-Menon fixed-point
 * identity sum(w) = |G w|^2 (G^H G w = 1 at convergence): one single-channel
 * gridding of w on this plan, then w <- alpha w with alpha = sum(w)/|G w|^2,
 * exact when w is the fixed point times a constant.  tile_stride 1: the whole
 * grid; k > 1 on a position-weight plan with unit pairs: |G w|^2 estimated
 * from a sample of the grid in two strata, each sampled sum scaled by its
 * units over its sampled units -- the light units of one run of four
 * consecutive unit pairs in every k (a regular sub-lattice in Morton order),
 * and the heavy (densest) units whose rank by record count is a multiple of
 * k -- about 1/k of the gridding work (other plans take the whole grid).
 * alpha_out (host, may be NULL) receives alpha.  Not for batched frames. */
fr_status fr_dcf_identity_scale(fr_plan *p, fr_samples *s, int tile_stride, float *alpha_out);

/* ------------------------------------------------- graphs and autotuning
 * Net-new (the reference launches per coil and partition with a device
 * synchronisation after each, cuGridLut.cu:261-309).  fr_graph_create runs
 * one eager pass (gridding; the FFT if f and no out_dev; fr_fft_sos into
 * out_dev if out_dev, which leaves the grids' contents unspecified), then
 * captures the same pass on the context's stream into a HIP graph;
 * fr_graph_launch replays it on the current sample data. */
typedef struct fr_graph fr_graph;
fr_status fr_graph_create(fr_plan *p, const fr_samples *s, fr_grids *g, fr_fft *f, unsigned rcx, unsigned rcy,
                          unsigned rcz, int apply_shift, const float *deapod_dev, float *out_dev, fr_graph **out);
fr_status fr_graph_launch(fr_graph *gr);
void fr_graph_destroy(fr_graph *gr);
/* Times `reps` gridding passes for each candidate fr_tune (NULL/0: the
 * eight structural candidates of fr_autotune_search -- 16-cell position-
 * weight boxes in four aspect ratios, 16-cell matrix-core tap boxes and a
 * 32-cell LDS read-add-write box) on these samples and returns the
 * fastest in *best (and its time); cand_ms (optional, ncand or 8 entries)
 * gets each candidate's time, -1 for candidates the plan refuses.  The LUT length is
 * not tuned: tap weights are resolved when the plan is built, so L does not
 * change the pass time, and L = 512 keeps reference parity. */
fr_status fr_autotune(fr_ctx *ctx, const fr_samples *s, unsigned N, float kWidth, const float *lut, int L,
                      const fr_tune *candidates, int ncand, int reps, fr_tune *best, float *best_ms,
                      float *cand_ms);

/* Key of the autotune cache: the device and the workload's shape (not the
 * trajectory itself: workloads of one shape share their tune). */
typedef struct {
    char arch[32];      /* device architecture, e.g. "gfx950"             */
    int cus;            /* compute units                                   */
    int dim;
    unsigned N;
    int64_t M;
    int C;
    int type;           /* FR_SAMPLES_*                                    */
    int frames;         /* batched 2D frames (0: one grid)                 */
    float kwidth;
    int L;
} fr_tune_key;

fr_status fr_tune_key_of(fr_ctx *ctx, const fr_samples *s, unsigned N, float kWidth, int L, fr_tune_key *key);

/* The autotune cache file (text, one line per key; host only, no device
 * work).  lookup: *found = 1 and the stored tune and times on a hit, 0 on a
 * miss or a missing file.  store: adds or replaces the key's line (written
 * to path.tmp, then renamed over path). */
fr_status fr_tune_cache_lookup(const char *path, const fr_tune_key *key, fr_tune *tune, float *best_ms,
                               float *default_ms, int *found);
fr_status fr_tune_cache_store(const char *path, const fr_tune_key *key, const fr_tune *tune, float best_ms,
                              float default_ms);

/* Staged autotuning search on these samples (net-new): the default plan and
 * the structural candidates; then, around the fastest, the heavy-unit slice
 * size (split_taps 256 ... 65536) and, for a position-weight plan, the
 * write-out grouping (tile_writes 1 and 2; pairs are the default); then that
 * plan's launch-only knobs (items_per_wave 1/2/4, xcd_run
 * 64 ... 4096, waves_per_simd 3/4, stores 1/2), one at a time; finally the
 * default and the winner in turn, three rounds each, and the winner is kept
 * only if it beats the default by 1%.  *best gets the winner as an explicit
 * tune (the default's settings spelled out when it wins), *best_ms its time,
 * *default_ms the default plan's, *timed the number of timed passes.  With a
 * cache_path, a cached key returns the stored tune and times with *timed =
 * 0 and nothing built or timed; a miss searches and stores the result. */
fr_status fr_autotune_search(fr_ctx *ctx, const fr_samples *s, unsigned N, float kWidth, const float *lut, int L,
                             int reps, const char *cache_path, fr_tune *best, float *best_ms, float *default_ms,
                             int *timed);

/* --------------------------------------------- learned density compensation
 * Net-new (the reference has no DCF estimation).  Features per sample, 8
 * floats per row: log1p(rho)/5, azimuth/pi, polar/pi, log1p(n_1)/8,
 * log1p(n_2)/8, log1p(n_4)/8, min(rho_max - rho, 4)/4, log1p(c_b/V_b)/8,
 * with u = (t - (N-1)/2)/(N/2) for the samples' normalised trajectory t,
 * rho = |u| N/2 (rho_max its largest value over the samples), n_r the
 * number of other samples whose nearest cell lies within Chebyshev distance
 * r cells, and c_b / V_b the samples per cell of the sample's radius shell
 * b = floor(2 rho) (c_b samples with that b, V_b the cells of the shell
 * b/2 <= rho < (b+1)/2: (4 pi/3)(hi^3 - lo^3) in 3D, pi (hi^2 - lo^2) in
 * 2D).  feat_dev: M*8 floats. */
fr_status fr_ldcf_features(fr_ctx *ctx, const fr_samples *s, unsigned N, float *feat_dev);

/* MLP 8 -> 64 -> 64 -> 1 (SiLU), BF16 weight matrices, FP32 biases and
 * accumulation.  Parameters are a flat FP32 vector of fr_mlp_param_count()
 * values: W1[8][64], b1[64], W2[64][64] ([in][out]), b2[64], w3[64], b3.  The
 * predicted DCF is exp(mu + sigma * y), divided by 1 + n_k when the network
 * has a base feature k (fr_mlp_set_base): exp(mu + sigma * y - 8 f_k). */
typedef struct fr_mlp fr_mlp;
int fr_mlp_param_count(void);
fr_status fr_mlp_create(fr_ctx *ctx, fr_mlp **out);
void fr_mlp_destroy(fr_mlp *m);
fr_status fr_mlp_set_params(fr_mlp *m, const float *params, float mu, float sigma);
fr_status fr_mlp_get_params(fr_mlp *m, float *params, float *mu, float *sigma);
/* Base feature of the DCF head: a box-count feature k = 3, 4 or 5 (the
 * network then predicts the DCF relative to 1 / (1 + n_k), the inverse local
 * density), or -1 for none (the default). */
fr_status fr_mlp_set_base(fr_mlp *m, int feature);
int fr_mlp_get_base(const fr_mlp *m);
/* y_dev[M] = network output for feature rows feat_dev[M*8]. */
fr_status fr_mlp_forward(fr_mlp *m, const float *feat_dev, int64_t M, float *y_dev);
/* The samples' DCF <- the DCF head (feature rows of the same samples). */
fr_status fr_mlp_dcf(fr_mlp *m, const float *feat_dev, fr_samples *s);
/* fr_ldcf_features + fr_mlp_dcf in one pass from the samples' trajectory on
 * an N-point grid, without feature rows in memory (bitwise the same DCF).
 * The cell-count workspace (20 bytes per grid cell, plus 6 per sample on
 * grids of at most 8192 tiles, i.e. 3D up to 320^3: about 415 MB at 256^3
 * with 13.1M samples) is kept in the network object and grows to the
 * largest call; no host sync. */
fr_status fr_ldcf_dcf(fr_mlp *m, fr_samples *s, unsigned N);
/* Mean squared error of y against targ_dev over the batch rows idx_dev[B]
 * (or rows 0..B-1 if idx_dev is NULL), its gradient (host, param order) and
 * the loss; deterministic. */
fr_status fr_mlp_grad(fr_mlp *m, const float *feat_dev, const float *targ_dev, const int64_t *idx_dev, int64_t B,
                      float *grad_host, float *loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on that loss; *loss (optional)
 * is the batch loss before the step. */
fr_status fr_mlp_train_step(fr_mlp *m, const float *feat_dev, const float *targ_dev, const int64_t *idx_dev,
                            int64_t B, float lr, float *loss);

/* --------------------------------------------------------- device memory */
fr_status fr_malloc(fr_ctx *ctx, size_t bytes, void **dev);
fr_status fr_free(fr_ctx *ctx, void *dev);
fr_status fr_memcpy_d2h(fr_ctx *ctx, void *host, const void *dev, size_t bytes);
fr_status fr_memcpy_h2d(fr_ctx *ctx, void *dev, const void *host, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
