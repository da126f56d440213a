// Internal definitions shared by the libfastrecon_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <vector>

#include "fastrecon_hip.h"

struct fr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    int cu_count = 0;
    std::string arch;          // gcnArchName up to the first ':' (e.g. gfx950)
    void *scratch = nullptr;   // this device's copy of the static scratch (fr_noalloc.hip)
};

struct fr_samples {
    fr_ctx *ctx = nullptr;
    int dim = 0;
    int64_t M = 0;
    int C = 0;
    int type = FR_SAMPLES_C64;
    float *traj = nullptr;     // SoA: x[M], y[M], z[M] (z only for dim 3)
    float *dcf = nullptr;      // M
    void *data = nullptr;      // sample-major [M][C] complex (float2 or __half2)
    void *staging = nullptr;   // M complex64 scratch for host uploads
    bool has_traj = false, has_dcf = false;
    std::vector<char> coil_set;
    // batched frames: samples [frame_off[f], frame_off[f+1]) belong to 2D
    // frame f, and traj holds a third row with the frame index (as float)
    int frames = 0;
    std::vector<int64_t> frame_off;
    // identity of this handle (never reused) and a counter bumped on every
    // change of dcf[]: a position-weight plan keeps a copy of the DCF with
    // its positions and refreshes it only when these differ from its copy's
    uint64_t uid = 0;
    uint64_t dcf_version = 0;
};

uint64_t fr_next_uid();
inline void fr_dcf_changed(fr_samples *s) { ++s->dcf_version; }

struct fr_grids {
    fr_ctx *ctx = nullptr;
    int dim = 0;
    unsigned N = 0;
    int C = 0;
    size_t cells = 0;          // N^dim, or frames * N^2 for batched frames
    float2 *data = nullptr;    // [C][cells] (frames: [C][frame][N^2])
    int frames = 0;            // batched 2D frames (0: one grid)
};

// Accumulator row of one grid cell in LDS: 64 components (32 coils, re/im)
// plus one pad word so rows start in different banks.
#define FR_ACC_STRIDE 65
// Units hold at most 64 cells; the row after a plan's largest unit
// (row cells_max) absorbs the padding taps that fill records up to a
// multiple of FR_TAP_GROUP.
#define FR_TAP_GROUP 4
// accumulation modes of a plan (fr_tune.accumulate after resolution)
#define FR_ACC_LDS_ATOMIC 1
#define FR_ACC_LDS_RMW 2
#define FR_ACC_REGS 3       // refused: see fr_plan_create
#define FR_ACC_GATHER 4     // cell-sorted batches, sample rows in LDS, VGPR accumulators
#define FR_GATHER_BATCH 32  // records per gather batch
#define FR_GATHER_CELLS 32  // unit cells held in VGPRs by the gather kernel
#define FR_ACC_MFMA 5       // matrix-core products of 16-cell blocks (weights from the plan's taps)
#define FR_ACC_POS 6        // matrix-core products of 16-cell units, weights computed from sample positions
#define FR_POS_MAX_L 2048   // longest LUT of a position-weight plan (its selection table is staged in LDS)
#define FR_POS_WAVES 4      // waves per workgroup of the position-weight kernel
#define FR_POS_BATCH 16     // records per batch of the position-weight kernel; units are padded to it
#define FR_MFMA_MAX_L 4096  // LUT lengths for which the plan builds exact index thresholds (degridding)
#define FR_CT_MAX_L 1023    // longest LUT for compact 2-byte taps (10-bit index, index L = zero weight)
#ifndef FR_MFMA_XCD_RUN
#define FR_MFMA_XCD_RUN 128       // items per XCD run in the matrix-core kernel's launch order
#endif

// One batch of a gather-mode item: records [r0, r0 + nrec) and their taps
// T[toff ..) sorted by cell; cnt byte k = tap count of cell k (a multiple
// of 4; padding taps read the batch's zero row with weight 0).
struct fr_batch {
    uint64_t toff;
    uint32_t r0, nrec;
    uint32_t cnt[FR_GATHER_CELLS / 4];
};

// A work unit is a box of grid cells owned by one wave.  Its records (one
// per sample that has taps in the box, ascending sample order) are H[hdr ..
// hdr + nrec) and its taps T[tap .. tap + ntap).
struct fr_unit {
    int32_t ox, oy, oz;          // origin (global cell coordinates)
    int32_t ext;                 // ux | uy << 8 | uz << 16
    uint64_t hdr;                // first record
    uint64_t tap;                // first tap
    uint32_t nrec, ntap;
};

// Unit geometry shared by plan and kernels.
struct fr_geom {
    int dim;
    int N;
    int bx, by, bz;              // base unit box (before splitting)
    int cells_max;               // bx * by * bz
    int waves;                   // waves per workgroup
};

// A work item is one wave's share of a unit: records [r0, r1) starting at
// tap t0.  Light units have one item that writes the grid; heavy units are
// cut by sample range into several items that write partial boxes to
// scratch slot `slot`, summed afterwards in item order.
struct fr_item {
    uint32_t unit;
    int32_t slot;                // -1: write the grid directly
    uint32_t r0, r1, t0;         // records [r0, r1), first tap t0
    uint32_t pad;
};

// Position-weight mode: a tile is tile_units (4 or 2) x-adjacent units
// written out together: four gridded by one workgroup (one wave each) and
// stored in whole rows, or a pair gridded by one wave and stored in rows of
// two units; item[w] is the item of unit w, or -1 when that unit is written
// elsewhere (front units) -- its cells are then left alone.  The tile's
// box: box0 = ox | oy << 16, box1 = oz.
struct fr_tile {
    int32_t item[4];
    uint32_t box0, box1;
};

// A heavy unit: its items are [item0, item0 + nitems), partial slots
// [slot0, slot0 + nitems).
struct fr_heavy {
    uint32_t unit;
    uint32_t nitems;
    uint32_t slot0;
    uint32_t pad;
};

struct fr_plan {
    fr_ctx *ctx = nullptr;
    // single-channel workspace of Pipe-Menon and the identity scale (unit
    // data on the plan's trajectory, a one-channel grid, degridded
    // densities): created on first use, kept until the plan is destroyed
    fr_samples *dcf_u = nullptr;
    fr_grids *dcf_g = nullptr;
    float2 *dcf_den = nullptr;
    void *dcf_scr = nullptr;        // the sampled identity's sub-slice sums (grown on demand)
    size_t dcf_scr_bytes = 0;
    int dim = 0;
    unsigned N = 0;
    int64_t M = 0;
    float kWidth = 0.f, kHW = 0.f;
    int L = 0;
    float *lut = nullptr;              // device copy of the LUT
    std::vector<float> host_lut;
    fr_geom geom{};
    int accumulate = 0;
    fr_unit *units = nullptr;          // device, heaviest first
    int64_t nunits = 0;
    uint2 *hdrs = nullptr;             // records: (sample, ntaps)
    uint32_t *tstart = nullptr;        // records: index of the first tap (padding entries: the tap count)
    uint2 *taps = nullptr;             // taps: (LDS byte offset of the cell's row, LUT weight bits), or
                                       // uint16_t (cell | LUT index << 6) when compact_taps
    bool compact_taps = false;
    int64_t nrecords = 0, ntaps = 0;
    int64_t max_unit_taps = 0;
    int exact_order = 0;               // 1: no sample slicing (bitwise CPU order)
    int frames = 0;                    // batched 2D frames (0: one grid)
    fr_item *items = nullptr;          // device, launch order
    int64_t nitems = 0;
    int64_t nitems_front = 0;          // items launched first, one per workgroup
    fr_heavy *heavy = nullptr;         // device
    int64_t nheavy = 0, nslots = 0;
    float *partial = nullptr;          // scratch for heavy-unit slices
    size_t partial_bytes = 0;
    fr_batch *batches = nullptr;       // gather mode: item i owns batches [items[i].t0, ..)
    // matrix-core mode: exact LUT-index thresholds on the squared distance,
    // pairs (B[m], B[m+1]) for m = 0 .. L, B[m] = least d2 whose reference
    // LUT index is >= m (fr_plan.hip, mfma_thresholds)
    float2 *mfma_pairs = nullptr;
    float mfma_dacc = 0.f;             // least d2 with sqrtf(d2) >= hw (rejected from here on)
    float mfma_cidx = 0.f;             // approximately (L - 1) / hw
    int64_t nbatches = 0;
    // position-weight mode (FR_ACC_POS): no taps.  hdrs is unused; recs holds
    // each unit's sample indices in ascending order, padded with the index M
    // to a multiple of FR_POS_BATCH (every unit starts at such a multiple),
    // then 64 more entries M; pos[M + 1] the samples' grid positions and
    // DCF (x, y, z, w; w copied from the samples before every pass), entry
    // M far outside the grid with w = 0; tsel[L] the LUT-index selection
    // table (B[m+1], LUT[m], LUT[m+1], 0) for the estimate m
    // (fr_plan_pos.hip).  Items carry their unit's box: t0 = ox | oy << 16,
    // pad = oz | (ux | uy << 5 | uz << 10) << 16.
    uint32_t *recs = nullptr;
    float4 *pos = nullptr;
    uint64_t pos_dcf_uid = 0, pos_dcf_version = 0;   // whose DCF pos[].w holds (uid 0: none)
    float4 *tsel = nullptr;
    int pos_ipw = 1;                   // items per wave (light items)
    int pos_stores = 2;                // grid stores: 1 streaming, 2 cached
    int pos_waves = 4;                 // kernel variant: 3 or 4 waves per SIMD
    int xcd_run = FR_MFMA_XCD_RUN;     // items per XCD run in launch order
    int64_t nrecs_padded = 0;
    fr_tile *tiles = nullptr;          // position-weight mode: light units by tile, launch order
    int64_t ntiles = 0;                // 0: light items are written unit by unit
    int tile_units = 0;                // units per tile: 4 (workgroup tiles) or 2 (wave pairs)
    double build_ms = 0.0;
};

struct fr_fft {
    fr_ctx *ctx = nullptr;
    int dim = 0;
    unsigned N = 0;
    int batch = 0;
    int sign = 1;
    void *plan = nullptr;              // rocfft_plan
    void *info = nullptr;              // rocfft_execution_info
    void *work = nullptr;
    size_t work_bytes = 0;
    // cropped FFT + SOS path (fr_fftcrop.hip): twiddles and pass buffers
    float2 *tw = nullptr;
    void *crop_ws = nullptr;
    size_t crop_ws_bytes = 0;
};

// ----------------------------------------------------------------- errors
fr_status fr_fail(fr_ctx *ctx, fr_status st, const char *fmt, ...);

#define FR_HIP(ctx, expr)                                                        \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess)                                                    \
            return fr_fail((ctx), FR_ERR_HIP, "%s failed: %s (%s:%d)", #expr,     \
                           hipGetErrorString(_e), __FILE__, __LINE__);           \
    } while (0)

#define FR_CHECK(ctx, cond, st, ...)                                             \
    do {                                                                         \
        if (!(cond)) return fr_fail((ctx), (st), __VA_ARGS__);                   \
    } while (0)

fr_status fr_alloc(fr_ctx *ctx, void **p, size_t bytes);

static inline size_t fr_cells(int dim, unsigned N)
{
    return dim == 3 ? (size_t)N * N * N : (size_t)N * N;
}

// cells of one coil of a plan's grid (all frames of a batched plan)
static inline size_t fr_plan_cells(int dim, unsigned N, int frames)
{
    return frames > 0 ? (size_t)frames * N * N : fr_cells(dim, N);
}

// Launch helpers implemented in the kernel translation units.
// force_mode: 0 = the plan's accumulation mode, else that mode (any plan's
// records serve FR_ACC_MFMA when its LUT has at most FR_MFMA_MAX_L entries)
// scratch: the heavy-unit slice partials go there (of scratch_bytes; no
// allocation), else into the plan's own buffer, grown on demand
// dcf: position-weight mode only (see fr_launch_grid_pos)
fr_status fr_launch_grid(fr_plan *p, const fr_samples *s, float2 *out, int force_mode, float2 *scratch = nullptr,
                         size_t scratch_bytes = 0, const fr_samples *dcf = nullptr);
fr_status fr_launch_grid_mfma(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial);
// A sample of the unit pairs with stride k: runs of FR_POS_WAVES consecutive
// pairs (one workgroup's; neighbouring units share samples, which then come
// from cache), one run in every k.  Sampled pair j is tile
// fr_sampled_tile(j, k); fr_sampled_tiles(n, k) of n tiles are sampled (all
// of them for k = 1).
__host__ __device__ inline int64_t fr_sampled_tile(int64_t j, int64_t k)
{
    return (j / FR_POS_WAVES) * (FR_POS_WAVES * k) + j % FR_POS_WAVES;
}
__host__ __device__ inline int64_t fr_sampled_tiles(int64_t n, int64_t k)
{
    const int64_t run = (int64_t)FR_POS_WAVES * k, rest = n % run;
    return (n / run) * FR_POS_WAVES + (rest < FR_POS_WAVES ? rest : FR_POS_WAVES);
}
// tile_stride > 1 (plans with unit pairs, single-channel samples): a sample
// of the grid for estimates, not a gridding pass -- the sampled pairs of
// light units (fr_sampled_tile) into `out` (other units' cells are left
// alone), and the
// front (heavy) units whose rank is a multiple of tile_stride, each front
// item cut into front_sub sub-slices: wave w = item * front_sub + sub writes
// its sums to virtual slot w of `partial` (cells q < the unit's cell count
// at partial[w * 32 * cells_max + q]; nitems_front * front_sub slots), with
// no slice combine.
// dcf: the samples handle whose DCF weights the pass (same M; default s)
fr_status fr_launch_grid_pos(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial, int64_t tile_stride = 1,
                             const fr_samples *dcf = nullptr, int front_sub = 1);
// bytes of heavy-unit slice partials a pass of C coils needs
size_t fr_grid_scratch_bytes(const fr_plan *p, int C);
fr_status fr_grid_check(const fr_plan *p, const fr_samples *s, const fr_grids *out);
// the no-allocation entry point's static scratch (fr_noalloc.hip): its
// address on the context's device, looked up once per device
fr_status fr_scratch_init(fr_ctx *ctx);
// cropped FFT + SOS helpers (fr_fftcrop.hip)
fr_status fr_fft_workspace(fr_fft *f, size_t bytes);
// FR_ERR_INVALID unless [p, p + bytes) lies inside the device allocation of p
fr_status fr_check_device_extent(fr_ctx *ctx, const void *p, size_t bytes, const char *what);
#define FR_EXTENT(ctx, p, bytes, what)                                          \
    do {                                                                        \
        fr_status fr_extent_st_ = fr_check_device_extent((ctx), (p), (bytes), (what)); \
        if (fr_extent_st_) return fr_extent_st_;                                \
    } while (0)
fr_status fr_fft_tw(fr_fft *f);
bool fr_fft_sos_fused(const fr_fft *f, const fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz);
fr_status fr_plan_build(fr_plan *p, const fr_samples *s, const fr_tune *tune);
fr_status fr_plan_build_pos(fr_plan *p, const fr_samples *s, const fr_tune *tune);
// exact LUT-index thresholds of a plan (mfma_pairs, mfma_dacc, mfma_cidx) and,
// when sel is set, the position-weight selection table (tsel)
std::vector<fr_tune> fr_tune_structural_candidates(int dim);
fr_status fr_plan_thresholds(fr_plan *p, bool sel);
