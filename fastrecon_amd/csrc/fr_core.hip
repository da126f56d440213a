// Context, sample containers, grid containers and copies.
#include "fr_internal.h"

#include <atomic>
#include <cstring>
#include <new>

uint64_t fr_next_uid()
{
    static std::atomic<uint64_t> next{0};
    return ++next;
}

fr_status fr_fail(fr_ctx *ctx, fr_status st, const char *fmt, ...)
{
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return st;
}

fr_status fr_alloc(fr_ctx *ctx, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return FR_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fr_fail(ctx, FR_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return FR_OK;
}

// ------------------------------------------------------------------ context
extern "C" fr_status fr_ctx_create(int device, fr_ctx **out)
{
    if (!out) return FR_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return FR_ERR_HIP;
    if (device < 0 || device >= n) return FR_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return FR_ERR_HIP;
    fr_ctx *c = new (std::nothrow) fr_ctx;
    if (!c) return FR_ERR_NOMEM;
    c->device = device;
    if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return FR_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->cu_count = prop.multiProcessorCount;
        c->arch = prop.gcnArchName;
        c->arch = c->arch.substr(0, c->arch.find(':'));
    }
    // this device's static scratch for fr_grid_execute_noalloc
    if (fr_scratch_init(c) != FR_OK) {
        hipStreamDestroy(c->stream);
        delete c;
        return FR_ERR_HIP;
    }
    *out = c;
    return FR_OK;
}

extern "C" fr_status fr_ctx_destroy(fr_ctx *ctx)
{
    if (!ctx) return FR_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return FR_OK;
}

extern "C" const char *fr_last_error(const fr_ctx *ctx)
{
    return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" fr_status fr_ctx_synchronize(fr_ctx *ctx)
{
    if (!ctx) return FR_ERR_INVALID;
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FR_OK;
}

extern "C" void *fr_ctx_stream(fr_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

// ------------------------------------------------------------------ samples
extern "C" fr_status fr_samples_create(fr_ctx *ctx, int dim, int64_t M, int coils, int type,
                                       fr_samples **out)
{
    if (!ctx || !out) return FR_ERR_INVALID;
    *out = nullptr;
    FR_CHECK(ctx, dim == 2 || dim == 3, FR_ERR_INVALID, "dim must be 2 or 3 (got %d)", dim);
    FR_CHECK(ctx, M > 0 && M < (int64_t(1) << 27), FR_ERR_UNSUPPORTED,
             "sample count %lld outside [1, 2^27)", (long long)M);
    FR_CHECK(ctx, coils > 0 && coils <= 1024, FR_ERR_INVALID, "coil count %d outside [1,1024]", coils);
    FR_CHECK(ctx, type == FR_SAMPLES_C64 || type == FR_SAMPLES_C16, FR_ERR_INVALID, "bad sample type");
    fr_samples *s = new (std::nothrow) fr_samples;
    if (!s) return FR_ERR_NOMEM;
    s->ctx = ctx; s->dim = dim; s->M = M; s->C = coils; s->type = type;
    s->coil_set.assign(coils, 0);
    size_t esz = type == FR_SAMPLES_C64 ? 8 : 4;
    fr_status st;
    // data and DCF hold one more sample, M, kept at zero: the null sample
    // that pads the position-weight plans' record lists
    if ((st = fr_alloc(ctx, (void **)&s->traj, sizeof(float) * dim * M)) ||
        (st = fr_alloc(ctx, (void **)&s->dcf, sizeof(float) * (M + 1))) ||
        (st = fr_alloc(ctx, &s->data, esz * (M + 1) * coils)) ||
        (st = fr_alloc(ctx, &s->staging, 8 * M + 16 * M))) {
        fr_samples_destroy(s);
        return st;
    }
    hipMemsetAsync(s->data, 0, esz * (M + 1) * coils, ctx->stream);
    hipMemsetAsync(s->dcf, 0, sizeof(float) * (M + 1), ctx->stream);
    s->uid = fr_next_uid();
    *out = s;
    return FR_OK;
}

__global__ void k_frame_rows(float *__restrict__ row, const int64_t *__restrict__ off, int frames, int64_t M)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    int lo = 0, hi = frames - 1;               // last frame whose offset <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    row[i] = (float)lo;
}

extern "C" fr_status fr_samples_create_frames(fr_ctx *ctx, int frames, const int64_t *frame_samples, int coils,
                                              int type, fr_samples **out)
{
    if (!ctx || !out || !frame_samples) return FR_ERR_INVALID;
    *out = nullptr;
    FR_CHECK(ctx, frames >= 1 && frames <= (1 << 16), FR_ERR_INVALID, "frame count %d outside [1, 65536]", frames);
    std::vector<int64_t> off(frames + 1, 0);
    for (int f = 0; f < frames; ++f) {
        FR_CHECK(ctx, frame_samples[f] >= 1, FR_ERR_INVALID, "frame %d has no samples", f);
        off[f + 1] = off[f] + frame_samples[f];
    }
    const int64_t M = off[frames];
    FR_CHECK(ctx, M < (int64_t(1) << 27), FR_ERR_UNSUPPORTED, "sample count %lld outside [1, 2^27)", (long long)M);
    FR_CHECK(ctx, coils > 0 && coils <= 1024, FR_ERR_INVALID, "coil count %d outside [1,1024]", coils);
    FR_CHECK(ctx, type == FR_SAMPLES_C64 || type == FR_SAMPLES_C16, FR_ERR_INVALID, "bad sample type");
    fr_samples *s = new (std::nothrow) fr_samples;
    if (!s) return FR_ERR_NOMEM;
    s->ctx = ctx; s->dim = 2; s->M = M; s->C = coils; s->type = type;
    s->frames = frames;
    s->frame_off = off;
    s->coil_set.assign(coils, 0);
    const size_t esz = type == FR_SAMPLES_C64 ? 8 : 4;
    int64_t *off_d = nullptr;
    fr_status st;
    if ((st = fr_alloc(ctx, (void **)&s->traj, sizeof(float) * 3 * M)) ||
        (st = fr_alloc(ctx, (void **)&s->dcf, sizeof(float) * (M + 1))) ||
        (st = fr_alloc(ctx, &s->data, esz * (M + 1) * coils)) ||
        (st = fr_alloc(ctx, &s->staging, 8 * M + 16 * M)) ||
        (st = fr_alloc(ctx, (void **)&off_d, sizeof(int64_t) * (frames + 1)))) {
        hipFree(off_d);
        fr_samples_destroy(s);
        return st;
    }
    hipError_t e = hipMemcpyAsync(off_d, off.data(), sizeof(int64_t) * (frames + 1), hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_frame_rows, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, ctx->stream,
                           s->traj + 2 * M, off_d, frames, M);
        e = hipMemsetAsync(s->data, 0, esz * (M + 1) * coils, ctx->stream);
    }
    if (e == hipSuccess) e = hipMemsetAsync(s->dcf, 0, sizeof(float) * (M + 1), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(off_d);
    if (e != hipSuccess) {
        fr_samples_destroy(s);
        return fr_fail(ctx, FR_ERR_HIP, "frame samples: %s", hipGetErrorString(e));
    }
    s->uid = fr_next_uid();
    *out = s;
    return FR_OK;
}

extern "C" fr_status fr_samples_destroy(fr_samples *s)
{
    if (!s) return FR_OK;
    hipStreamSynchronize(s->ctx->stream);
    hipFree(s->traj); hipFree(s->dcf); hipFree(s->data); hipFree(s->staging);
    delete s;
    return FR_OK;
}

__global__ void k_deinterleave_traj(const float *__restrict__ in, int stride, int dim, int64_t M,
                                    float *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    for (int d = 0; d < dim; ++d) out[d * M + i] = in[i * stride + d];
}

// one coil (contiguous complex64) -> column c of the sample-major array
__global__ void k_scatter_coil_c64(const float2 *__restrict__ in, int64_t n, int C, int c,
                                   float2 *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i * C + c] = in[i];
}

__global__ void k_scatter_coil_c16(const float2 *__restrict__ in, int64_t n, int C, int c,
                                   __half2 *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float2 v = in[i];
        out[i * C + c] = __halves2half2(__float2half_rn(v.x), __float2half_rn(v.y));
    }
}

// coil-major [C][M] -> sample-major [M][C], tiled through LDS
template <typename TO>
__global__ void k_transpose_cm_to_sm(const float2 *__restrict__ in, int64_t M, int C, TO *__restrict__ out)
{
    __shared__ float2 tile[32][33];
    int64_t s0 = (int64_t)blockIdx.x * 32;
    int c0 = blockIdx.y * 32;
    int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: ty 0..7
    for (int r = ty; r < 32; r += 8) {
        int c = c0 + r;
        int64_t s = s0 + tx;
        if (c < C && s < M) tile[r][tx] = in[(int64_t)c * M + s];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int64_t s = s0 + r;
        int c = c0 + tx;
        if (c < C && s < M) {
            float2 v = tile[tx][r];
            if constexpr (sizeof(TO) == 8) {
                out[s * C + c] = v;
            } else {
                out[s * C + c] = __halves2half2(__float2half_rn(v.x), __float2half_rn(v.y));
            }
        }
    }
}

__global__ void k_convert_sm_c16(const float2 *__restrict__ in, int64_t n, __half2 *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float2 v = in[i];
        out[i] = __halves2half2(__float2half_rn(v.x), __float2half_rn(v.y));
    }
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

extern "C" fr_status fr_samples_set_traj(fr_samples *s, const float *traj, int stride)
{
    if (!s || !traj) return FR_ERR_INVALID;
    fr_ctx *ctx = s->ctx;
    FR_CHECK(ctx, stride >= s->dim, FR_ERR_INVALID, "traj stride %d < dim %d", stride, s->dim);
    size_t bytes = sizeof(float) * (size_t)stride * s->M;
    void *tmp = nullptr;
    fr_status st = fr_alloc(ctx, &tmp, bytes);
    if (st) return st;
    hipError_t e = hipMemcpyAsync(tmp, traj, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_deinterleave_traj, dim3(nblk(s->M, 256)), dim3(256), 0, ctx->stream,
                           (const float *)tmp, stride, s->dim, s->M, s->traj);
        e = hipStreamSynchronize(ctx->stream);
    }
    hipFree(tmp);
    FR_CHECK(ctx, e == hipSuccess, FR_ERR_HIP, "set_traj: %s", hipGetErrorString(e));
    s->has_traj = true;
    return FR_OK;
}

// A caller's device buffer must hold what the call reads or writes: [p, p +
// bytes) has to lie inside the allocation p points into.  Pointers the
// runtime cannot place (e.g. imported or host memory) are not checked, nor
// are calls made while the context's stream is being captured (the eager
// pass before a capture checks them).
fr_status fr_check_device_extent(fr_ctx *ctx, const void *p, size_t bytes, const char *what)
{
    if (!p || bytes == 0) return FR_OK;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cap) != hipSuccess) {
        (void)hipGetLastError();
        return FR_OK;
    }
    if (cap != hipStreamCaptureStatusNone) return FR_OK;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess || base == nullptr) {
        (void)hipGetLastError();
        return FR_OK;
    }
    const size_t avail = size - (size_t)((const char *)p - (const char *)base);
    FR_CHECK(ctx, bytes <= avail, FR_ERR_INVALID, "%s: the call accesses %zu bytes but the device allocation holds "
             "%zu from this pointer", what, bytes, avail);
    return FR_OK;
}

extern "C" fr_status fr_samples_set_traj_device(fr_samples *s, const float *traj, int stride)
{
    if (!s || !traj) return FR_ERR_INVALID;
    fr_ctx *ctx = s->ctx;
    FR_CHECK(ctx, stride >= s->dim, FR_ERR_INVALID, "traj stride %d < dim %d", stride, s->dim);
    if (s->M > 0) {
        fr_status e = fr_check_device_extent(ctx, traj, sizeof(float) * ((size_t)(s->M - 1) * stride + s->dim),
                                          "fr_samples_set_traj_device");
        if (e) return e;
    }
    hipLaunchKernelGGL(k_deinterleave_traj, dim3(nblk(s->M, 256)), dim3(256), 0, ctx->stream,
                       traj, stride, s->dim, s->M, s->traj);
    FR_HIP(ctx, hipGetLastError());
    s->has_traj = true;
    return FR_OK;
}

extern "C" fr_status fr_samples_set_dcf(fr_samples *s, const float *dcf)
{
    if (!s || !dcf) return FR_ERR_INVALID;
    fr_ctx *ctx = s->ctx;
    FR_HIP(ctx, hipMemcpyAsync(s->dcf, dcf, sizeof(float) * s->M, hipMemcpyHostToDevice, ctx->stream));
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->has_dcf = true;
    fr_dcf_changed(s);
    return FR_OK;
}

extern "C" fr_status fr_samples_set_dcf_device(fr_samples *s, const float *dcf)
{
    if (!s || !dcf) return FR_ERR_INVALID;
    fr_ctx *ctx = s->ctx;
    fr_status e = fr_check_device_extent(ctx, dcf, sizeof(float) * s->M, "fr_samples_set_dcf_device");
    if (e) return e;
    FR_HIP(ctx, hipMemcpyAsync(s->dcf, dcf, sizeof(float) * s->M, hipMemcpyDeviceToDevice, ctx->stream));
    s->has_dcf = true;
    fr_dcf_changed(s);
    return FR_OK;
}

extern "C" fr_status fr_samples_get_dcf(const fr_samples *s, float *host)
{
    if (!s || !host) return FR_ERR_INVALID;
    fr_ctx *ctx = s->ctx;
    FR_HIP(ctx, hipMemcpyAsync(host, s->dcf, sizeof(float) * s->M, hipMemcpyDeviceToHost, ctx->stream));
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FR_OK;
}

static fr_status scatter_coil(fr_samples *s, int coil, int64_t first, int64_t n)
{
    fr_ctx *ctx = s->ctx;
    if (s->type == FR_SAMPLES_C64) {
        hipLaunchKernelGGL(k_scatter_coil_c64, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                           (const float2 *)s->staging, n, s->C, coil, (float2 *)s->data + first * s->C);
    } else {
        hipLaunchKernelGGL(k_scatter_coil_c16, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                           (const float2 *)s->staging, n, s->C, coil, (__half2 *)s->data + first * s->C);
    }
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_samples_set_coil(fr_samples *s, int coil, const void *data)
{
    if (!s || !data) return FR_ERR_INVALID;
    fr_ctx *ctx = s->ctx;
    FR_CHECK(ctx, coil >= 0 && coil < s->C, FR_ERR_INVALID, "coil %d outside [0,%d)", coil, s->C);
    FR_HIP(ctx, hipMemcpyAsync(s->staging, data, 8 * (size_t)s->M, hipMemcpyHostToDevice, ctx->stream));
    fr_status st = scatter_coil(s, coil, 0, s->M);
    if (st) return st;
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->coil_set[coil] = 1;
    return FR_OK;
}

extern "C" fr_status fr_samples_update(fr_samples *s, int coil, int64_t first, int64_t n, const void *data)
{
    if (!s || !data) return FR_ERR_INVALID;
    fr_ctx *ctx = s->ctx;
    FR_CHECK(ctx, coil >= 0 && coil < s->C, FR_ERR_INVALID, "coil %d outside [0,%d)", coil, s->C);
    FR_CHECK(ctx, first >= 0 && n >= 0 && first + n <= s->M, FR_ERR_INVALID,
             "update range [%lld,%lld) outside [0,%lld)", (long long)first, (long long)(first + n),
             (long long)s->M);
    if (n == 0) return FR_OK;
    FR_HIP(ctx, hipMemcpyAsync(s->staging, data, 8 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    fr_status st = scatter_coil(s, coil, first, n);
    if (st) return st;
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FR_OK;
}

extern "C" fr_status fr_samples_set_data_device(fr_samples *s, const void *data, int layout)
{
    if (!s || !data) return FR_ERR_INVALID;
    fr_ctx *ctx = s->ctx;
    FR_CHECK(ctx, layout == FR_LAYOUT_COIL_MAJOR || layout == FR_LAYOUT_SAMPLE_MAJOR, FR_ERR_INVALID,
             "bad layout %d", layout);
    int64_t n = s->M * s->C;
    // complex64 input whatever the sample type (FP16 samples are converted here)
    fr_status e = fr_check_device_extent(ctx, data, sizeof(float2) * (size_t)n, "fr_samples_set_data_device");
    if (e) return e;
    if (layout == FR_LAYOUT_SAMPLE_MAJOR) {
        if (s->type == FR_SAMPLES_C64) {
            FR_HIP(ctx, hipMemcpyAsync(s->data, data, 8 * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            hipLaunchKernelGGL(k_convert_sm_c16, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream,
                               (const float2 *)data, n, (__half2 *)s->data);
        }
    } else {
        dim3 grid(nblk(s->M, 32), nblk(s->C, 32));
        if (s->type == FR_SAMPLES_C64)
            hipLaunchKernelGGL(k_transpose_cm_to_sm<float2>, grid, dim3(256), 0, ctx->stream,
                               (const float2 *)data, s->M, s->C, (float2 *)s->data);
        else
            hipLaunchKernelGGL(k_transpose_cm_to_sm<__half2>, grid, dim3(256), 0, ctx->stream,
                               (const float2 *)data, s->M, s->C, (__half2 *)s->data);
    }
    FR_HIP(ctx, hipGetLastError());
    for (auto &c : s->coil_set) c = 1;
    return FR_OK;
}

// -------------------------------------------------------------------- grids
extern "C" fr_status fr_grids_create(fr_ctx *ctx, int dim, unsigned N, int coils, fr_grids **out)
{
    if (!ctx || !out) return FR_ERR_INVALID;
    *out = nullptr;
    FR_CHECK(ctx, dim == 2 || dim == 3, FR_ERR_INVALID, "dim must be 2 or 3");
    FR_CHECK(ctx, N >= 2 && N <= 4096, FR_ERR_INVALID, "grid size %u outside [2,4096]", N);
    FR_CHECK(ctx, coils >= 1, FR_ERR_INVALID, "coils must be >= 1");
    fr_grids *g = new (std::nothrow) fr_grids;
    if (!g) return FR_ERR_NOMEM;
    g->ctx = ctx; g->dim = dim; g->N = N; g->C = coils; g->cells = fr_cells(dim, N);
    fr_status st = fr_alloc(ctx, (void **)&g->data, sizeof(float2) * g->cells * coils);
    if (st) { delete g; return st; }
    *out = g;
    return FR_OK;
}

extern "C" fr_status fr_grids_create_frames(fr_ctx *ctx, unsigned N, int coils, int frames, fr_grids **out)
{
    if (!ctx || !out) return FR_ERR_INVALID;
    *out = nullptr;
    FR_CHECK(ctx, N >= 2 && N <= 4096, FR_ERR_INVALID, "grid size %u outside [2,4096]", N);
    FR_CHECK(ctx, coils >= 1 && frames >= 1, FR_ERR_INVALID, "coils and frames must be >= 1");
    fr_grids *g = new (std::nothrow) fr_grids;
    if (!g) return FR_ERR_NOMEM;
    g->ctx = ctx; g->dim = 2; g->N = N; g->C = coils; g->frames = frames;
    g->cells = (size_t)frames * N * N;
    fr_status st = fr_alloc(ctx, (void **)&g->data, sizeof(float2) * g->cells * coils);
    if (st) { delete g; return st; }
    *out = g;
    return FR_OK;
}

extern "C" fr_status fr_grids_destroy(fr_grids *g)
{
    if (!g) return FR_OK;
    hipStreamSynchronize(g->ctx->stream);
    hipFree(g->data);
    delete g;
    return FR_OK;
}

extern "C" fr_status fr_grids_download(const fr_grids *g, int coil, void *host)
{
    if (!g || !host) return FR_ERR_INVALID;
    fr_ctx *ctx = g->ctx;
    FR_CHECK(ctx, coil >= 0 && coil < g->C, FR_ERR_INVALID, "coil %d outside [0,%d)", coil, g->C);
    FR_HIP(ctx, hipMemcpyAsync(host, g->data + g->cells * coil, sizeof(float2) * g->cells,
                               hipMemcpyDeviceToHost, ctx->stream));
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FR_OK;
}

extern "C" fr_status fr_grids_upload(fr_grids *g, int coil, const void *host)
{
    if (!g || !host) return FR_ERR_INVALID;
    fr_ctx *ctx = g->ctx;
    FR_CHECK(ctx, coil >= 0 && coil < g->C, FR_ERR_INVALID, "coil %d outside [0,%d)", coil, g->C);
    FR_HIP(ctx, hipMemcpyAsync(g->data + g->cells * coil, host, sizeof(float2) * g->cells,
                               hipMemcpyHostToDevice, ctx->stream));
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FR_OK;
}

extern "C" void *fr_grids_device_ptr(fr_grids *g) { return g ? (void *)g->data : nullptr; }

// ------------------------------------------------------------ device memory
extern "C" fr_status fr_malloc(fr_ctx *ctx, size_t bytes, void **dev)
{
    if (!ctx || !dev) return FR_ERR_INVALID;
    return fr_alloc(ctx, dev, bytes);
}

extern "C" fr_status fr_free(fr_ctx *ctx, void *dev)
{
    if (!ctx) return FR_ERR_INVALID;
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FR_HIP(ctx, hipFree(dev));
    return FR_OK;
}

extern "C" fr_status fr_memcpy_d2h(fr_ctx *ctx, void *host, const void *dev, size_t bytes)
{
    if (!ctx) return FR_ERR_INVALID;
    FR_HIP(ctx, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FR_OK;
}

extern "C" fr_status fr_memcpy_h2d(fr_ctx *ctx, void *dev, const void *host, size_t bytes)
{
    if (!ctx) return FR_ERR_INVALID;
    FR_HIP(ctx, hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FR_OK;
}
