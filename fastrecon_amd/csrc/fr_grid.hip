// Gridding execute: one workgroup per (grid tile, chunk of <= 32 coils).
//
// Replaces cuGridLut::griddingChannel / gridding_kernel
// (ReconData/cuGridLut.cu:231-309) and the CPU loop GridLut::griddingChannel
// (GridLut.cpp:80-147).
//
// Layout and mapping (gfx950, wave64):
//   * samples are sample-major [M][C] complex, so the 32 coils of a sample
//     are one contiguous 256-byte read;
//   * the tile's accumulator lives in LDS as [cell][2*coils (+1 pad)] FP32;
//   * wave w of the workgroup owns slab w of the tile and walks that slab's
//     record stream (scalar loads: the stream is wave-uniform); lane p holds
//     component p (coil p/2, re or im) of the current sample, already scaled
//     by its DCF exactly like the reference's `dcf * sample`;
//   * every tap is one multiply w * d and one LDS add into [cell][p]; lanes
//     address consecutive words, so the add is bank-conflict free.
// A cell is only ever touched by the wave owning its slab, and that wave
// walks samples in ascending order, so each cell sums its contributions in
// the reference CPU order with the same FP32 operations: the result is
// deterministic, uses no global atomics, and matches the CPU path bit for
// bit.  The finished tile is written once, coalesced along x.
#include "fr_internal.h"

namespace {

template <typename TS>
__device__ __forceinline__ float load_comp(const TS *__restrict__ data, uint32_t s, int C, int coil, int comp);

template <>
__device__ __forceinline__ float load_comp<float2>(const float2 *__restrict__ data, uint32_t s, int C, int coil,
                                                   int comp)
{
    const float *f = reinterpret_cast<const float *>(data);
    return f[((size_t)s * C + coil) * 2 + comp];
}

template <>
__device__ __forceinline__ float load_comp<__half2>(const __half2 *__restrict__ data, uint32_t s, int C, int coil,
                                                    int comp)
{
    const __half *h = reinterpret_cast<const __half *>(data);
    return __half2float(h[((size_t)s * C + coil) * 2 + comp]);
}

template <int ACC, typename TS>
__global__ void __launch_bounds__(1024) k_grid_tiles(const uint32_t *__restrict__ words,
                                                     const ulonglong2 *__restrict__ ranges,
                                                     const TS *__restrict__ data, const float *__restrict__ dcf,
                                                     int C, int stride, fr_tiling tl, float2 *__restrict__ out,
                                                     size_t cells)
{
    extern __shared__ float acc[];
    const int ncell = tl.tx * tl.ty * tl.tz;
    for (int i = threadIdx.x; i < ncell * stride; i += blockDim.x) acc[i] = 0.f;
    __syncthreads();

    const int tile = blockIdx.x;
    const int c0 = blockIdx.y * 32;
    const int nc = min(32, C - c0);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const bool active = lane < 2 * nc;
    const int coil = c0 + (lane >> 1);
    const int comp = lane & 1;

    const ulonglong2 r = ranges[(size_t)tile * tl.slabs + wave];
    uint64_t b = r.x;
    const uint64_t e = r.y;
    while (b < e) {
        const uint32_t s = words[b];
        const uint32_t n = words[b + 1];
        float d = 0.f;
        if (active) d = __fmul_rn(dcf[s], load_comp<TS>(data, s, C, coil, comp));
        const uint32_t *t = words + b + 2;
        for (uint32_t k = 0; k < n; ++k) {
            const uint32_t cell = t[2 * k];
            const float w = __uint_as_float(t[2 * k + 1]);
            const float v = __fmul_rn(w, d);
            if (active) {
                float *a = acc + cell * stride + lane;
                if constexpr (ACC == 0) {
                    atomicAdd(a, v);           // ds_add_f32; the slab's wave is the only writer
                } else {
                    *a = __fadd_rn(*a, v);
                }
            }
        }
        b += 2 + 2 * (uint64_t)n;
    }
    __syncthreads();

    // write the tile: coil-major grids [C][N^dim], x fastest
    const int ti = tile % tl.ntx;
    const int tj = (tile / tl.ntx) % tl.nty;
    const int tk = tile / (tl.ntx * tl.nty);
    const int N = tl.N;
    for (int i = threadIdx.x; i < nc * ncell; i += blockDim.x) {
        const int c = i / ncell;
        const int cell = i - c * ncell;
        const int lx = cell % tl.tx;
        const int ly = (cell / tl.tx) % tl.ty;
        const int lz = cell / (tl.tx * tl.ty);
        const int gx = ti * tl.tx + lx, gy = tj * tl.ty + ly, gz = tk * tl.tz + lz;
        if (gx < N && gy < N && gz < (tl.dim == 3 ? N : 1)) {
            const size_t g = ((size_t)gz * N + gy) * N + gx;
            out[(size_t)(c0 + c) * cells + g] = make_float2(acc[cell * stride + 2 * c], acc[cell * stride + 2 * c + 1]);
        }
    }
}

template <int ACC, typename TS>
fr_status launch(fr_plan *p, const fr_samples *s, float2 *out)
{
    fr_ctx *ctx = p->ctx;
    const fr_tiling &tl = p->tl;
    const int stride = 2 * std::min(s->C, 32) + 1;
    const size_t lds = (size_t)tl.tile_cells() * stride * sizeof(float);
    auto kern = k_grid_tiles<ACC, TS>;
    FR_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    dim3 grid(tl.tiles(), (s->C + 31) / 32);
    hipLaunchKernelGGL(kern, grid, dim3(64 * tl.slabs), lds, ctx->stream, p->words, p->ranges,
                       (const TS *)s->data, s->dcf, s->C, stride, tl, out, fr_cells(p->dim, p->N));
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

}  // namespace

fr_status fr_launch_grid(fr_plan *p, const fr_samples *s, float2 *out, int)
{
    if (s->type == FR_SAMPLES_C64)
        return p->accumulate ? launch<1, float2>(p, s, out) : launch<0, float2>(p, s, out);
    return p->accumulate ? launch<1, __half2>(p, s, out) : launch<0, __half2>(p, s, out);
}

extern "C" fr_status fr_grid_execute(fr_plan *p, const fr_samples *s, fr_grids *out)
{
    if (!p || !s || !out) return FR_ERR_INVALID;
    fr_ctx *ctx = p->ctx;
    FR_CHECK(ctx, s->M == p->M && s->dim == p->dim, FR_ERR_INVALID,
             "samples (dim %d, M %lld) do not match the plan (dim %d, M %lld)", s->dim, (long long)s->M, p->dim,
             (long long)p->M);
    FR_CHECK(ctx, out->dim == p->dim && out->N == p->N && out->C == s->C, FR_ERR_INVALID,
             "grids (dim %d, N %u, C %d) do not match plan/samples (dim %d, N %u, C %d)", out->dim, out->N, out->C,
             p->dim, p->N, s->C);
    FR_CHECK(ctx, s->has_dcf, FR_ERR_STATE, "DCF not set");
    return fr_launch_grid(p, s, out->data, 0);
}
