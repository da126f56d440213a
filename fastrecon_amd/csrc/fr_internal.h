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
};

struct fr_grids {
    fr_ctx *ctx = nullptr;
    int dim = 0;
    unsigned N = 0;
    int C = 0;
    size_t cells = 0;          // N^dim
    float2 *data = nullptr;    // [C][cells]
};

// Geometry of the tile decomposition shared by plan and kernels.
struct fr_tiling {
    int dim;
    int N;
    int tx, ty, tz;            // tile extent (tz = 1 in 2D)
    int ntx, nty, ntz;         // tiles per axis
    int slabs;                 // waves per workgroup; slab = slow-axis band
    int slab_depth;            // slow-axis cells per slab
    __host__ __device__ int tiles() const { return ntx * nty * ntz; }
    __host__ __device__ int tile_cells() const { return tx * ty * tz; }
};

struct fr_plan {
    fr_ctx *ctx = nullptr;
    int dim = 0;
    unsigned N = 0;
    int64_t M = 0;
    float kWidth = 0.f, kHW = 0.f;
    int L = 0;
    float *lut = nullptr;              // device copy of the LUT
    std::vector<float> host_lut;
    fr_tiling tl{};
    int accumulate = 0;
    uint32_t *words = nullptr;         // record streams
    ulonglong2 *ranges = nullptr;      // [tiles*slabs] (begin, end) word offsets
    int64_t nwords = 0, nrecords = 0, ntaps = 0;
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

// Launch helpers implemented in the kernel translation units.
fr_status fr_launch_grid(fr_plan *p, const fr_samples *s, float2 *out, int C_out_offset);
fr_status fr_plan_build(fr_plan *p, const fr_samples *s, const fr_tune *tune);
