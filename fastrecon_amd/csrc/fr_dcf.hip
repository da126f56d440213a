// Degridding (grid -> samples, the adjoint of gridding) and the Pipe-Menon
// iterative density compensation.  Neither exists in the reference; both
// reuse the reference's exact tap decisions (fr_taps.h) so that degridding
// is the exact adjoint of fr_grid_execute.
#include "fr_internal.h"
#include "fr_taps.h"

#include <algorithm>
#include <new>
#include <vector>

namespace {

// One thread per (sample, coil): out[s][c] = sum over the sample's taps of
// LUT weight * grid[c][cell], in the reference's z, y, x tap order.
__global__ void k_degrid(const float *__restrict__ traj, int64_t M, int dim, int N, float hw, int L,
                         const float *__restrict__ lut, const float2 *__restrict__ grid, size_t cells, int C,
                         float2 *__restrict__ out, int frames)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * C) return;
    const int64_t s = i / C;
    const int c = (int)(i - s * C);
    fr_box b;
    fr_sample_box(dim, N, traj, M, s, hw, b, frames);
    const float2 *g = grid + (size_t)c * cells;
    float re = 0.f, im = 0.f;
    fr_box_taps(b, b.s, b.e, hw, L, [&](int x, int y, int z, int ki) {
        const float w = lut[ki];
        const float2 v = g[((size_t)z * N + y) * N + x];
        re = __fadd_rn(re, __fmul_rn(w, v.x));
        im = __fadd_rn(im, __fmul_rn(w, v.y));
    });
    out[s * C + c] = make_float2(re, im);
}

// Same sums as k_degrid, with the LUT index taken from a fast square root
// corrected by the plan's exact thresholds (B[m] <= d2 < B[m+1]; the tap test
// d2 < dacc is sqrtf(d2) < hw): identical taps, indices and summation order,
// so the result is bit-identical to k_degrid, at a fraction of its
// arithmetic.  The thresholds and the LUT are staged in LDS.
__global__ void __launch_bounds__(256) k_degrid_thr(const float *__restrict__ traj, int64_t M, int dim, int N,
                                                    float hw, int L, const float *__restrict__ lut,
                                                    const float2 *__restrict__ gpairs, float dacc, float cidx,
                                                    const float2 *__restrict__ grid, size_t cells, int C,
                                                    float2 *__restrict__ out, int frames)
{
    extern __shared__ float2 sm[];
    float2 *pairs = sm;
    float *slut = reinterpret_cast<float *>(sm + L + 1);
    for (int q = threadIdx.x; q <= L; q += blockDim.x) pairs[q] = gpairs[q];
    for (int q = threadIdx.x; q < L; q += blockDim.x) slut[q] = lut[q];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * C) return;
    const int64_t s = i / C;
    const int c = (int)(i - s * C);
    fr_box b;
    fr_sample_box(dim, N, traj, M, s, hw, b, frames);
    const float2 *g = grid + (size_t)c * cells;
    float re = 0.f, im = 0.f;
    for (int z = b.s[2]; z <= b.e[2]; ++z) {
        const float dz = __fsub_rn((float)z, b.c[2]);
        const float dz2 = __fmul_rn(dz, dz);
        for (int y = b.s[1]; y <= b.e[1]; ++y) {
            const float dy = __fsub_rn((float)y, b.c[1]);
            const float dy2 = __fmul_rn(dy, dy);
            for (int x = b.s[0]; x <= b.e[0]; ++x) {
                const float dx = __fsub_rn((float)x, b.c[0]);
                const float d2 = fr_tap_dk2(dz2, dy2, dx);
                if (d2 < dacc) {
                    const int ka = min((int)__builtin_fmaf(__builtin_amdgcn_sqrtf(d2), cidx, 0.5f), L);
                    const float2 bb = pairs[ka];
                    const int ki = min(max(ka - (d2 < bb.x ? 1 : 0) + (d2 >= bb.y ? 1 : 0), 0), L - 1);
                    const float w = slut[ki];
                    const float2 v = g[((size_t)z * N + y) * N + x];
                    re = __fadd_rn(re, __fmul_rn(w, v.x));
                    im = __fadd_rn(im, __fmul_rn(w, v.y));
                }
            }
        }
    }
    out[s * C + c] = make_float2(re, im);
}

__global__ void k_fill_unit(float2 *__restrict__ d, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = make_float2(1.f, 0.f);
}

__global__ void k_fill_f32(float *__restrict__ d, int64_t n, float v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = v;
}

// w <- w / Re(G^H G w); samples whose denominator is not positive get 0.
__global__ void k_pm_update(float *__restrict__ w, const float2 *__restrict__ den, int64_t M)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const float d = den[i].x;
    w[i] = d > 0.f ? __fdiv_rn(w[i], d) : 0.f;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Per-block FP64 partial sums of |g|^2 over n grid cells and of w over M
// samples: part[2b] and part[2b+1].  Fixed grid-stride order and a fixed
// reduction tree, so the result does not vary between runs.
constexpr int SUM_BLOCKS = 1024;
__global__ void __launch_bounds__(256) k_identity_partials(const float2 *__restrict__ g, int64_t n,
                                                           const float *__restrict__ w, int64_t M,
                                                           double *__restrict__ part)
{
    double a = 0.0, b = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float2 v = g[i];
        a += (double)v.x * v.x + (double)v.y * v.y;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += stride) b += (double)w[i];
    for (int m = 32; m >= 1; m >>= 1) {
        a += __shfl_xor(a, m);
        b += __shfl_xor(b, m);
    }
    __shared__ double sa[4], sb[4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sa[wv] = a;
        sb[wv] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (sa[0] + sa[1]) + (sa[2] + sa[3]);
        part[2 * blockIdx.x + 1] = (sb[0] + sb[1]) + (sb[2] + sb[3]);
    }
}

// alpha = sum(w) / sum |G w|^2 from the partials (one thread, in order),
// then w <- alpha w by every thread of the grid
__global__ void k_identity_alpha(const double *__restrict__ part, int nb, float *__restrict__ alpha)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nb; ++i) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    alpha[0] = a > 0.0 ? (float)(b / a) : 0.f;
}

__global__ void k_scale_by(float *__restrict__ w, int64_t M, const float *__restrict__ alpha)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) w[i] = __fmul_rn(w[i], alpha[0]);
}

}  // namespace

fr_status fr_degrid_raw(fr_plan *p, const fr_samples *s, const float2 *grid, int C, float2 *out)
{
    fr_ctx *ctx = p->ctx;
    if (p->mfma_pairs) {
        const size_t lds = sizeof(float2) * (p->L + 1) + sizeof(float) * p->L;
        hipLaunchKernelGGL(k_degrid_thr, dim3(nblk(p->M * C, 256)), dim3(256), lds, ctx->stream, s->traj, p->M,
                           p->dim, (int)p->N, p->kHW, p->L, p->lut, p->mfma_pairs, p->mfma_dacc, p->mfma_cidx, grid,
                           fr_plan_cells(p->dim, p->N, p->frames), C, out, p->frames);
        FR_HIP(ctx, hipGetLastError());
        return FR_OK;
    }
    hipLaunchKernelGGL(k_degrid, dim3(nblk(p->M * C, 256)), dim3(256), 0, ctx->stream, s->traj, p->M, p->dim,
                       (int)p->N, p->kHW, p->L, p->lut, grid, fr_plan_cells(p->dim, p->N, p->frames), C, out,
                       p->frames);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_degrid_execute(fr_plan *p, const fr_samples *s, const fr_grids *g, void *out_dev)
{
    if (!p || !s || !g || !out_dev) return FR_ERR_INVALID;
    fr_ctx *ctx = p->ctx;
    FR_CHECK(ctx, s->M == p->M && s->dim == p->dim, FR_ERR_INVALID, "samples do not match the plan");
    FR_CHECK(ctx, g->dim == p->dim && g->N == p->N && g->frames == p->frames && s->frames == p->frames,
             FR_ERR_INVALID, "grids do not match the plan");
    FR_EXTENT(ctx, out_dev, sizeof(float2) * (size_t)s->M * g->C, "fr_degrid_execute output");
    return fr_degrid_raw(p, s, g->data, g->C, (float2 *)out_dev);
}

// The plan's single-channel workspace, created on first use: unit data
// sharing the samples' trajectory (copied on every call, since the caller's
// samples may be another handle on the same trajectory), a one-channel grid
// and the degridded densities.  Its DCF is the caller's to fill.
static fr_status dcf_workspace(fr_plan *p, const fr_samples *s)
{
    fr_ctx *ctx = p->ctx;
    const int64_t M = s->M;
    fr_status st = FR_OK;
    if (!p->dcf_u) {
        if (s->frames > 0) {
            std::vector<int64_t> cnt(s->frames);
            for (int f = 0; f < s->frames; ++f) cnt[f] = s->frame_off[f + 1] - s->frame_off[f];
            st = fr_samples_create_frames(ctx, s->frames, cnt.data(), 1, FR_SAMPLES_C64, &p->dcf_u);
            if (!st) st = fr_grids_create_frames(ctx, p->N, 1, s->frames, &p->dcf_g);
        } else {
            st = fr_samples_create(ctx, s->dim, M, 1, FR_SAMPLES_C64, &p->dcf_u);
            if (!st) st = fr_grids_create(ctx, s->dim, p->N, 1, &p->dcf_g);
        }
        // (the identity scale keeps its partial sums in the same buffer)
        const size_t den_bytes = std::max(sizeof(float2) * (size_t)M, sizeof(double) * (2 * SUM_BLOCKS + 1));
        if (!st) st = fr_alloc(ctx, (void **)&p->dcf_den, den_bytes);
        if (st) {
            hipFree(p->dcf_den);
            fr_grids_destroy(p->dcf_g);
            fr_samples_destroy(p->dcf_u);
            p->dcf_den = nullptr;
            p->dcf_g = nullptr;
            p->dcf_u = nullptr;
            return st;
        }
        hipLaunchKernelGGL(k_fill_unit, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, (float2 *)p->dcf_u->data, M);
        FR_HIP(ctx, hipGetLastError());
    }
    const int rows = s->frames > 0 ? 3 : s->dim;
    FR_HIP(ctx, hipMemcpyAsync(p->dcf_u->traj, s->traj, sizeof(float) * rows * M, hipMemcpyDeviceToDevice,
                               ctx->stream));
    p->dcf_u->has_traj = p->dcf_u->has_dcf = true;
    return FR_OK;
}

// the single-channel gridding mode of a plan: the matrix-core kernels
// (the records and taps of plans in modes 1, 2 and 5 serve it; several
// times faster here than the LDS kernels, whose lanes are the coil
// components), a position-weight plan's own mode, an exact-order plan's
// sequential mode (CPU summation order)
static int density_mode(const fr_plan *p)
{
    return p->accumulate == FR_ACC_POS ? FR_ACC_POS
                                       : (p->accumulate != FR_ACC_GATHER && !p->exact_order ? FR_ACC_MFMA : 0);
}

extern "C" fr_status fr_dcf_pipe_menon(fr_plan *p, fr_samples *s, int iters, int init_ones)
{
    if (!p || !s) return FR_ERR_INVALID;
    fr_ctx *ctx = p->ctx;
    FR_CHECK(ctx, s->M == p->M && s->dim == p->dim, FR_ERR_INVALID, "samples do not match the plan");
    FR_CHECK(ctx, iters >= 0, FR_ERR_INVALID, "iters must be >= 0");
    FR_CHECK(ctx, s->frames == p->frames, FR_ERR_INVALID, "samples and plan differ in frames");
    const int64_t M = s->M;
    fr_status st = dcf_workspace(p, s);
    if (st) return st;
    fr_samples *u = p->dcf_u;
    if (init_ones || !s->has_dcf)
        hipLaunchKernelGGL(k_fill_f32, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, u->dcf, M, 1.f);
    else
        hipMemcpyAsync(u->dcf, s->dcf, sizeof(float) * M, hipMemcpyDeviceToDevice, ctx->stream);
    fr_dcf_changed(u);
    const int mode = density_mode(p);
    for (int it = 0; it < iters && !st; ++it) {
        st = fr_launch_grid(p, u, p->dcf_g->data, mode);
        if (!st) st = fr_degrid_raw(p, u, p->dcf_g->data, 1, p->dcf_den);
        if (!st) {
            hipLaunchKernelGGL(k_pm_update, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, u->dcf, p->dcf_den, M);
            fr_dcf_changed(u);
        }
    }
    if (st) return st;
    hipMemcpyAsync(s->dcf, u->dcf, sizeof(float) * M, hipMemcpyDeviceToDevice, ctx->stream);
    s->has_dcf = true;
    fr_dcf_changed(s);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    FR_CHECK(ctx, e == hipSuccess, FR_ERR_HIP, "pipe-menon: %s", hipGetErrorString(e));
    return FR_OK;
}

// Global scale of a DCF from the Pipe-Menon fixed-point identity: at
// convergence G^H G w = 1 on every sample, so sum(w) = <w, G^H G w> =
// |G w|^2.  alpha = sum(w) / |G w|^2 restores that for any w, and is exact
// for a DCF that differs from the fixed point by a constant factor.  One
// single-channel gridding pass of w on this plan, two reductions, no host
// round trip; w is scaled in place.
extern "C" fr_status fr_dcf_identity_scale(fr_plan *p, fr_samples *s, float *alpha_out)
{
    if (!p || !s) return FR_ERR_INVALID;
    fr_ctx *ctx = p->ctx;
    FR_CHECK(ctx, s->M == p->M && s->dim == p->dim, FR_ERR_INVALID, "samples do not match the plan");
    FR_CHECK(ctx, s->frames == 0 && p->frames == 0, FR_ERR_UNSUPPORTED,
             "identity scale of batched frames: scale each frame on its own plan");
    FR_CHECK(ctx, s->has_traj && s->has_dcf, FR_ERR_STATE, "samples need a trajectory and a DCF");
    const int64_t M = s->M;
    fr_status st = dcf_workspace(p, s);
    if (st) return st;
    fr_samples *u = p->dcf_u;
    hipMemcpyAsync(u->dcf, s->dcf, sizeof(float) * M, hipMemcpyDeviceToDevice, ctx->stream);
    fr_dcf_changed(u);
    st = fr_launch_grid(p, u, p->dcf_g->data, density_mode(p));
    if (st) return st;
    // the partial sums and alpha live in the degridded-density buffer (sized
    // for them)
    double *part = reinterpret_cast<double *>(p->dcf_den);
    float *alpha = reinterpret_cast<float *>(part + 2 * SUM_BLOCKS);
    hipLaunchKernelGGL(k_identity_partials, dim3(SUM_BLOCKS), dim3(256), 0, ctx->stream, p->dcf_g->data,
                       (int64_t)p->dcf_g->cells, s->dcf, M, part);
    hipLaunchKernelGGL(k_identity_alpha, dim3(1), dim3(64), 0, ctx->stream, part, SUM_BLOCKS, alpha);
    hipLaunchKernelGGL(k_scale_by, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, s->dcf, M, alpha);
    fr_dcf_changed(s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && alpha_out)
        e = hipMemcpyAsync(alpha_out, alpha, sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    FR_CHECK(ctx, e == hipSuccess, FR_ERR_HIP, "identity scale: %s", hipGetErrorString(e));
    return FR_OK;
}
