// Image stage: batched rocFFT over all coils, in-place fftshift, and the
// fused shift + crop + sum-of-squares (+ deapodisation) epilogue.
#include "fr_internal.h"

#include <algorithm>

#include <rocfft/rocfft.h>

#include <mutex>

namespace {

std::once_flag g_rocfft_once;
rocfft_status g_rocfft_setup = rocfft_status_success;

// In-place fftshift for even N (ImageFilter.cpp:231-280): element (z,y,x)
// with z < N/2 swaps with (z+N/2, (y+N/2)%N, (x+N/2)%N); in 2D the row
// index plays the role of z.  Each pair is visited by exactly one thread.
__global__ void k_fftshift(float2 *__restrict__ g, int dim, int N, int C, size_t cells)
{
    const int h = N / 2;
    const size_t half = cells / 2;                       // elements with slow coord < N/2
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half * C) return;
    const int c = (int)(i / half);
    const size_t r = i - (size_t)c * half;
    size_t j;
    if (dim == 3) {
        const int x = (int)(r % N), y = (int)((r / N) % N), z = (int)(r / ((size_t)N * N));
        const int x1 = x < h ? x + h : x - h, y1 = y < h ? y + h : y - h, z1 = z + h;
        j = ((size_t)z1 * N + y1) * N + x1;
    } else {
        const int x = (int)(r % N), y = (int)(r / N);
        const int x1 = x < h ? x + h : x - h, y1 = y + h;
        j = (size_t)y1 * N + x1;
    }
    float2 *base = g + (size_t)c * cells;
    const float2 a = base[r], b = base[j];
    base[r] = b;
    base[j] = a;
}

// out[z][y][x] over the rc crop at offset (N - rc)/2 (ImageFilter.cpp:
// 173-176); source index optionally fftshifted; sum over coils in order of
// fl(re*re) + fl(im*im) (the real part of v * conj(v)), then sqrt
// (ImageFilter.cpp:162-167) and |.| (Main.cpp:177), divided by deapod.
// Batched frames (frames > 0): z is the frame, taken whole and unshifted,
// and the deapodisation map is one rcx x rcy image shared by the frames.
__global__ void k_sos(const float2 *__restrict__ g, int dim, int N, int C, size_t cells, int rcx, int rcy, int rcz,
                      int shift, const float *__restrict__ deapod, float *__restrict__ out, int frames)
{
    const size_t n = (size_t)rcx * rcy * rcz;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % rcx), y = (int)((i / rcx) % rcy), z = (int)(i / ((size_t)rcx * rcy));
    const int nz = frames > 0 ? frames : (dim == 3 ? N : 1);
    int sx = (N - rcx) / 2 + x, sy = (N - rcy) / 2 + y, sz = (nz - rcz) / 2 + z;
    if (shift) {
        const int h = N / 2;
        sx = sx < h ? sx + h : sx - h;
        sy = sy < h ? sy + h : sy - h;
        if (dim == 3) sz = sz < h ? sz + h : sz - h;
    }
    const size_t src = ((size_t)sz * N + sy) * N + sx;
    float acc = 0.f;
    for (int c = 0; c < C; ++c) {
        const float2 v = g[(size_t)c * cells + src];
        acc = __fadd_rn(acc, __fadd_rn(__fmul_rn(v.x, v.x), __fmul_rn(v.y, v.y)));
    }
    float m = sqrtf(acc);
    if (deapod) m = __fdiv_rn(m, deapod[frames > 0 ? i % ((size_t)rcx * rcy) : i]);
    out[i] = m;
}

inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" fr_status fr_fft_create(fr_ctx *ctx, int dim, unsigned N, int batch, int sign, fr_fft **out)
{
    if (!ctx || !out) return FR_ERR_INVALID;
    *out = nullptr;
    FR_CHECK(ctx, dim == 2 || dim == 3, FR_ERR_INVALID, "dim must be 2 or 3");
    FR_CHECK(ctx, batch >= 1, FR_ERR_INVALID, "batch must be >= 1");
    FR_CHECK(ctx, sign == 1 || sign == -1, FR_ERR_INVALID, "sign must be +1 or -1");
    std::call_once(g_rocfft_once, [] { g_rocfft_setup = rocfft_setup(); });
    FR_CHECK(ctx, g_rocfft_setup == rocfft_status_success, FR_ERR_LIBRARY, "rocfft_setup failed");
    fr_fft *f = new (std::nothrow) fr_fft;
    if (!f) return FR_ERR_NOMEM;
    f->ctx = ctx; f->dim = dim; f->N = N; f->batch = batch; f->sign = sign;
    size_t lengths[3] = {N, N, N};
    rocfft_plan plan = nullptr;
    rocfft_status rs = rocfft_plan_create(&plan, rocfft_placement_inplace,
                                          sign > 0 ? rocfft_transform_type_complex_inverse
                                                   : rocfft_transform_type_complex_forward,
                                          rocfft_precision_single, (size_t)dim, lengths, (size_t)batch, nullptr);
    if (rs != rocfft_status_success) { delete f; return fr_fail(ctx, FR_ERR_LIBRARY, "rocfft_plan_create: %d", (int)rs); }
    f->plan = plan;
    rocfft_plan_get_work_buffer_size(plan, &f->work_bytes);
    rocfft_execution_info info = nullptr;
    rocfft_execution_info_create(&info);
    f->info = info;
    rocfft_execution_info_set_stream(info, ctx->stream);
    if (f->work_bytes) {
        fr_status st = fr_alloc(ctx, &f->work, f->work_bytes);
        if (st) { fr_fft_destroy(f); return st; }
        rocfft_execution_info_set_work_buffer(info, f->work, f->work_bytes);
    }
    *out = f;
    return FR_OK;
}

extern "C" fr_status fr_fft_destroy(fr_fft *f)
{
    if (!f) return FR_OK;
    hipStreamSynchronize(f->ctx->stream);
    if (f->info) rocfft_execution_info_destroy((rocfft_execution_info)f->info);
    if (f->plan) rocfft_plan_destroy((rocfft_plan)f->plan);
    hipFree(f->work);
    hipFree(f->tw);
    hipFree(f->crop_ws);
    delete f;
    return FR_OK;
}

extern "C" fr_status fr_fft_execute(fr_fft *f, fr_grids *g)
{
    if (!f || !g) return FR_ERR_INVALID;
    fr_ctx *ctx = f->ctx;
    FR_CHECK(ctx, g->dim == f->dim && g->N == f->N && (int64_t)g->C * std::max(g->frames, 1) == f->batch,
             FR_ERR_INVALID,
             "grids (dim %d, N %u, C %d) do not match the FFT plan (dim %d, N %u, batch %d)", g->dim, g->N, g->C,
             f->dim, f->N, f->batch);
    void *bufs[1] = {g->data};
    rocfft_status rs = rocfft_execute((rocfft_plan)f->plan, bufs, nullptr, (rocfft_execution_info)f->info);
    FR_CHECK(ctx, rs == rocfft_status_success, FR_ERR_LIBRARY, "rocfft_execute: %d", (int)rs);
    return FR_OK;
}

extern "C" fr_status fr_fftshift(fr_ctx *ctx, fr_grids *g)
{
    if (!ctx || !g) return FR_ERR_INVALID;
    FR_CHECK(ctx, g->N % 2 == 0, FR_ERR_UNSUPPORTED, "fftshift needs an even grid size (got %u)", g->N);
    // batched frames: every (coil, frame) is its own contiguous 2D grid
    const int batch = g->C * std::max(g->frames, 1);
    const size_t cells = g->frames > 0 ? (size_t)g->N * g->N : g->cells;
    const size_t n = cells / 2 * batch;
    hipLaunchKernelGGL(k_fftshift, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, g->data, g->dim, (int)g->N, batch,
                       cells);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_sos(fr_ctx *ctx, const fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz, int apply_shift,
                            const float *deapod_dev, float *out_dev)
{
    if (!ctx || !g || !out_dev) return FR_ERR_INVALID;
    if (g->dim == 2) rcz = g->frames > 0 ? (unsigned)g->frames : 1u;    // frames: one image per frame
    FR_CHECK(ctx, rcx >= 1 && rcy >= 1 && rcz >= 1 && rcx <= g->N && rcy <= g->N && (g->dim == 2 || rcz <= g->N),
             FR_ERR_INVALID, "crop %ux%ux%u outside grid %u", rcx, rcy, rcz, g->N);
    FR_CHECK(ctx, !apply_shift || g->N % 2 == 0, FR_ERR_UNSUPPORTED, "fftshift needs an even grid size");
    const size_t n = (size_t)rcx * rcy * rcz;
    FR_EXTENT(ctx, out_dev, sizeof(float) * n, "fr_sos image");
    FR_EXTENT(ctx, deapod_dev, sizeof(float) * rcx * rcy * (g->dim == 3 ? rcz : 1u), "fr_sos deapodisation map");
    hipLaunchKernelGGL(k_sos, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, g->data, g->dim, (int)g->N, g->C,
                       g->cells, (int)rcx, (int)rcy, (int)rcz, apply_shift, deapod_dev, out_dev, g->frames);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_deapod_map(fr_ctx *ctx, int dim, unsigned N, float kWidth, const float *lut, int L,
                                   const float *centre, unsigned rcx, unsigned rcy, unsigned rcz, float *out_dev)
{
    if (!ctx || !lut || !centre || !out_dev) return FR_ERR_INVALID;
    fr_samples *s = nullptr;
    fr_plan *p = nullptr;
    fr_grids *g = nullptr;
    fr_fft *f = nullptr;
    fr_status st = fr_samples_create(ctx, dim, 1, 1, FR_SAMPLES_C64, &s);
    const float one[2] = {1.f, 0.f}, w = 1.f;
    if (!st) st = fr_samples_set_traj(s, centre, dim);
    if (!st) st = fr_samples_set_dcf(s, &w);
    if (!st) st = fr_samples_set_coil(s, 0, one);
    if (!st) st = fr_plan_create(ctx, s, N, kWidth, lut, L, nullptr, &p);
    if (!st) st = fr_grids_create(ctx, dim, N, 1, &g);
    if (!st) st = fr_grid_execute(p, s, g);
    if (!st) st = fr_fft_create(ctx, dim, N, 1, +1, &f);
    if (!st) st = fr_fft_execute(f, g);
    if (!st) st = fr_sos(ctx, g, rcx, rcy, rcz, 1, nullptr, out_dev);
    if (!st) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = fr_fail(ctx, FR_ERR_HIP, "deapod: %s", hipGetErrorString(e));
    }
    fr_fft_destroy(f);
    fr_grids_destroy(g);
    fr_plan_destroy(p);
    fr_samples_destroy(s);
    return st;
}

// ------------------------------------------------------- PILS / crop ops
namespace {

// ImageFilter::lowFilter (ImageFilter.cpp:55-93): x0 = size/2 per axis,
// r = squared integer distance, idx = int(float(r) / att * 100.0) in
// double, zero past the 2000-entry table, else scale by coeff[idx].
__global__ void k_lowfilter(float2 *__restrict__ g, int dim, int N, int C, size_t cells, float att,
                            const float *__restrict__ coeff)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells * C) return;
    const size_t e = i % cells;
    const int x = (int)(e % N), y = (int)((e / N) % N), z = dim == 3 ? (int)(e / ((size_t)N * N)) : 0;
    const int h = N / 2, hz = dim == 3 ? N / 2 : 0;   // z0 = size.z / 2 with size.z = 1 in 2D
    const int r = (z - hz) * (z - hz) + (y - h) * (y - h) + (x - h) * (x - h);
    const int idx = (int)((double)__fdiv_rn((float)r, att) * 100.0);
    float2 v = g[i];
    if (idx >= 2000) {
        v = make_float2(0.f, 0.f);
    } else {
        const float c = coeff[idx];
        v = make_float2(__fmul_rn(v.x, c), __fmul_rn(v.y, c));
    }
    g[i] = v;
}

// ImageFilter::normalize (ImageFilter.cpp:95-118): v /= sqrtf(sum_c norm(v_c)).
// libstdc++'s std::norm(complex<float>) is abs(v)^2 with abs = cabsf, which
// glibc evaluates as the float rounding of the double sqrt(re^2 + im^2); the
// same is done here, then squared in float.
__global__ void k_normalize(float2 *__restrict__ g, int C, size_t cells)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    float m = 0.f;
    for (int c = 0; c < C; ++c) {
        const float2 v = g[(size_t)c * cells + i];
        const double re = v.x, im = v.y;
        const float a = (float)__dsqrt_rn(__dadd_rn(__dmul_rn(re, re), __dmul_rn(im, im)));
        m = __fadd_rn(m, __fmul_rn(a, a));
    }
    const float s = sqrtf(m);
    for (int c = 0; c < C; ++c) {
        float2 v = g[(size_t)c * cells + i];
        g[(size_t)c * cells + i] = make_float2(__fdiv_rn(v.x, s), __fdiv_rn(v.y, s));
    }
}

// ImageFilter::crop (ImageFilter.cpp:120-157): channel-wise centre crop
__global__ void k_crop(const float2 *__restrict__ in, int dim, int N, int C, size_t cells, int rcx, int rcy, int rcz,
                       float2 *__restrict__ out)
{
    const size_t n = (size_t)rcx * rcy * rcz;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * C) return;
    const int c = (int)(i / n);
    const size_t o = i - (size_t)c * n;
    const int x = (int)(o % rcx), y = (int)((o / rcx) % rcy), z = (int)(o / ((size_t)rcx * rcy));
    const int nz = dim == 3 ? N : 1;
    const size_t src = ((size_t)((nz - rcz) / 2 + z) * N + (N - rcy) / 2 + y) * N + (N - rcx) / 2 + x;
    out[i] = in[(size_t)c * cells + src];
}

// ImageFilter::SOS(map, size) (ImageFilter.cpp:170-209): out = sum over
// coils in order of v * conj(m) on the crop, complex FP32 (no sqrt).
__global__ void k_sos_map(const float2 *__restrict__ g, const float2 *__restrict__ m, int dim, int N, int C,
                          size_t cells, int rcx, int rcy, int rcz, int shift, float2 *__restrict__ out)
{
    const size_t n = (size_t)rcx * rcy * rcz;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % rcx), y = (int)((i / rcx) % rcy), z = (int)(i / ((size_t)rcx * rcy));
    const int nz = dim == 3 ? N : 1;
    int sx = (N - rcx) / 2 + x, sy = (N - rcy) / 2 + y, sz = (nz - rcz) / 2 + z;
    if (shift) {
        const int h = N / 2;
        sx = sx < h ? sx + h : sx - h;
        sy = sy < h ? sy + h : sy - h;
        if (dim == 3) sz = sz < h ? sz + h : sz - h;
    }
    const size_t src = ((size_t)sz * N + sy) * N + sx;
    float ar = 0.f, ai = 0.f;
    for (int c = 0; c < C; ++c) {
        const float2 v = g[(size_t)c * cells + src];
        const float2 w = m[(size_t)c * cells + src];
        // (vr + i vi) * (wr - i wi): real vr*wr - vi*(-wi), imag vr*(-wi) + vi*wr
        const float re = __fsub_rn(__fmul_rn(v.x, w.x), __fmul_rn(v.y, -w.y));
        const float im = __fadd_rn(__fmul_rn(v.x, -w.y), __fmul_rn(v.y, w.x));
        ar = __fadd_rn(ar, re);
        ai = __fadd_rn(ai, im);
    }
    out[i] = make_float2(ar, ai);
}

}  // namespace

extern "C" fr_status fr_lowfilter(fr_ctx *ctx, fr_grids *g, int res)
{
    if (!ctx || !g) return FR_ERR_INVALID;
    FR_CHECK(ctx, g->frames == 0, FR_ERR_UNSUPPORTED, "not available for batched frames");
    FR_CHECK(ctx, res > 0, FR_ERR_INVALID, "lowFilter resolution must be positive");
    // coefficients exactly as the reference's host loop: expf(-r / 100.0)
    float coeff[2000];
    for (int r = 0; r < 2000; ++r) coeff[r] = expf((float)(-r / 100.0));
    const float att = (float)(2.0 * res * res / 4.0);
    float *dc = nullptr;
    fr_status st = fr_alloc(ctx, (void **)&dc, sizeof coeff);
    if (st) return st;
    hipError_t e = hipMemcpyAsync(dc, coeff, sizeof coeff, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        const size_t n = g->cells * g->C;
        hipLaunchKernelGGL(k_lowfilter, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, g->data, g->dim, (int)g->N,
                           g->C, g->cells, att, dc);
        e = hipStreamSynchronize(ctx->stream);
    }
    hipFree(dc);
    FR_CHECK(ctx, e == hipSuccess, FR_ERR_HIP, "lowFilter: %s", hipGetErrorString(e));
    return FR_OK;
}

extern "C" fr_status fr_normalize(fr_ctx *ctx, fr_grids *g)
{
    if (!ctx || !g) return FR_ERR_INVALID;
    FR_CHECK(ctx, g->frames == 0, FR_ERR_UNSUPPORTED, "not available for batched frames");
    hipLaunchKernelGGL(k_normalize, dim3(nblk(g->cells, 256)), dim3(256), 0, ctx->stream, g->data, g->C, g->cells);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_crop(fr_ctx *ctx, const fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz, void *out_dev)
{
    if (!ctx || !g || !out_dev) return FR_ERR_INVALID;
    FR_CHECK(ctx, g->frames == 0, FR_ERR_UNSUPPORTED, "not available for batched frames");
    if (g->dim == 2) rcz = 1;
    FR_CHECK(ctx, rcx >= 1 && rcy >= 1 && rcz >= 1 && rcx <= g->N && rcy <= g->N && (g->dim == 2 || rcz <= g->N),
             FR_ERR_INVALID, "crop %ux%ux%u outside grid %u", rcx, rcy, rcz, g->N);
    const size_t n = (size_t)rcx * rcy * rcz * g->C;
    FR_EXTENT(ctx, out_dev, sizeof(float2) * n, "fr_crop output");
    hipLaunchKernelGGL(k_crop, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, g->data, g->dim, (int)g->N, g->C,
                       g->cells, (int)rcx, (int)rcy, (int)rcz, (float2 *)out_dev);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_sos_map(fr_ctx *ctx, const fr_grids *g, const fr_grids *map, unsigned rcx, unsigned rcy,
                                unsigned rcz, int apply_shift, void *out_dev)
{
    if (!ctx || !g || !map || !out_dev) return FR_ERR_INVALID;
    FR_CHECK(ctx, g->frames == 0, FR_ERR_UNSUPPORTED, "not available for batched frames");
    if (g->dim == 2) rcz = 1;
    FR_CHECK(ctx, map->dim == g->dim && map->N == g->N && map->C >= g->C, FR_ERR_INVALID,
             "map grids do not match the image grids");
    FR_CHECK(ctx, rcx >= 1 && rcy >= 1 && rcz >= 1 && rcx <= g->N && rcy <= g->N && (g->dim == 2 || rcz <= g->N),
             FR_ERR_INVALID, "crop %ux%ux%u outside grid %u", rcx, rcy, rcz, g->N);
    FR_CHECK(ctx, !apply_shift || g->N % 2 == 0, FR_ERR_UNSUPPORTED, "fftshift needs an even grid size");
    const size_t n = (size_t)rcx * rcy * rcz;
    FR_EXTENT(ctx, out_dev, sizeof(float2) * n, "fr_sos_map output");
    hipLaunchKernelGGL(k_sos_map, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, g->data, map->data, g->dim,
                       (int)g->N, g->C, g->cells, (int)rcx, (int)rcy, (int)rcz, apply_shift, (float2 *)out_dev);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_grids_copy(fr_ctx *ctx, const fr_grids *src, fr_grids *dst)
{
    if (!ctx || !src || !dst) return FR_ERR_INVALID;
    FR_CHECK(ctx, src->dim == dst->dim && src->N == dst->N && src->C == dst->C, FR_ERR_INVALID,
             "grid copy between different shapes");
    FR_HIP(ctx, hipMemcpyAsync(dst->data, src->data, sizeof(float2) * src->cells * src->C, hipMemcpyDeviceToDevice,
                               ctx->stream));
    return FR_OK;
}
