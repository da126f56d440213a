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

// Sums of K values per partial over one 256-thread block, in a fixed order
// (thread t: partials t, t + 256, ...; then the wave tree and the four waves
// in order); thread 0 gets the totals.
template <int K>
__device__ inline void block_sums(const double *__restrict__ part, int stride, int nb, double (&out)[K])
{
    double a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256)
#pragma unroll
        for (int k = 0; k < K; ++k) a[k] += part[(size_t)i * stride + k];
#pragma unroll
    for (int k = 0; k < K; ++k)
        for (int m = 32; m >= 1; m >>= 1) a[k] += __shfl_xor(a[k], m);
    __shared__ double s[K][4];
    __syncthreads();                 // a previous call's readers are done
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) s[k][threadIdx.x >> 6] = a[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = (s[k][0] + s[k][1]) + (s[k][2] + s[k][3]);
}

// alpha = sum(w) / sum |G w|^2 from the partials (one 256-thread block),
// then w <- alpha w by every thread of the grid
__global__ void __launch_bounds__(256) k_identity_alpha(const double *__restrict__ part, int nb,
                                                        float *__restrict__ alpha)
{
    double t[2];
    block_sums<2>(part, 2, nb, t);
    if (threadIdx.x == 0) alpha[0] = t[0] > 0.0 ? (float)(t[1] / t[0]) : 0.f;
}

// sum of |g|^2 over the cells of an item's unit box
__device__ inline double unit_energy(const fr_item &I, const float2 *__restrict__ g, int N)
{
    const int ox = (int)(I.t0 & 0xffffu), oy = (int)(I.t0 >> 16), oz = (int)(I.pad & 0xffffu);
    const int ux = (int)((I.pad >> 16) & 31u), uy = (int)((I.pad >> 21) & 31u), uz = (int)((I.pad >> 26) & 31u);
    double a = 0.0;
    for (int z = 0; z < uz; ++z)
        for (int y = 0; y < uy; ++y)
            for (int x = 0; x < ux; ++x) {
                const float2 v = g[((size_t)(oz + z) * N + (oy + y)) * N + (ox + x)];
                a += (double)v.x * v.x + (double)v.y * v.y;
            }
    return a;
}

// Sampled identity, light stratum: per-block FP64 sums (fixed order) of
// |g|^2 over the light units of the sampled pairs (fr_sampled_tile) and
// their count (part[3b], part[3b+1]), and the count of heavy (front) units
// (part[3b+2]).
// Threads [0, nfront): front items, one unit each at its first item (a
// unit's items are consecutive); then one thread per (sampled tile, unit).
constexpr int TILE_SUMS = 3;
constexpr int IDENTITY_SUB = 8;      // sub-slices per front item in the sampled identity
__global__ void __launch_bounds__(256) k_identity_tiles(const fr_tile *__restrict__ tiles, int64_t ntiles,
                                                        int64_t stride, const fr_item *__restrict__ items,
                                                        int64_t nfront, const float2 *__restrict__ g, int N,
                                                        double *__restrict__ part)
{
    double v[TILE_SUMS] = {0.0, 0.0, 0.0};
    const int64_t nsamp = fr_sampled_tiles(ntiles, stride);
    const int64_t total = nfront + 2 * nsamp;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        if (t < nfront) {
            if (t == 0 || items[t - 1].unit != items[t].unit) v[2] += 1.0;
            continue;
        }
        const int64_t q = t - nfront;
        const int32_t it = tiles[fr_sampled_tile(q >> 1, stride)].item[q & 1];
        if (it < 0) continue;
        v[0] += unit_energy(items[it], g, N);
        v[1] += 1.0;
    }
#pragma unroll
    for (int k = 0; k < TILE_SUMS; ++k)
        for (int m = 32; m >= 1; m >>= 1) v[k] += __shfl_xor(v[k], m);
    __shared__ double s[TILE_SUMS][4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < TILE_SUMS; ++k) s[k][wv] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < TILE_SUMS; ++k)
            part[TILE_SUMS * blockIdx.x + k] = (s[k][0] + s[k][1]) + (s[k][2] + s[k][3]);
}

// Sampled identity, heavy stratum: block b (front item b) sums, when item b
// is the first of a unit whose rank is a multiple of the stride, the unit's
// sub-slice sums (virtual slots [b fsub, e fsub) for its items [b, e)) per
// cell in FP64 (16 lanes per cell, then the lanes in order) and writes the
// unit's sum of |g|^2 and 1 to hp[2b], hp[2b+1]; other blocks write zeros.
__global__ void __launch_bounds__(256) k_identity_heavy(const fr_item *__restrict__ items, int64_t nfront,
                                                        int64_t stride, int fsub, const float2 *__restrict__ vs,
                                                        int cells_max, double *__restrict__ hp)
{
    const int64_t b = blockIdx.x;
    const fr_item I = items[b];
    if ((b > 0 && items[b - 1].unit == I.unit) || I.unit % (uint64_t)stride != 0) {
        if (threadIdx.x == 0) {
            hp[2 * b] = 0.0;
            hp[2 * b + 1] = 0.0;
        }
        return;
    }
    int64_t e = b + 1;
    while (e < nfront && items[e].unit == I.unit) ++e;
    const int ncell = (int)(((I.pad >> 16) & 31u) * ((I.pad >> 21) & 31u) * ((I.pad >> 26) & 31u));
    const int q = threadIdx.x & 15, part = threadIdx.x >> 4;
    double re = 0.0, im = 0.0;
    if (q < ncell)
        for (int64_t v = b * fsub + part; v < e * fsub; v += 16) {
            const float2 x = vs[(size_t)v * 32 * cells_max + q];
            re += x.x;
            im += x.y;
        }
    __shared__ double sr[16][16], si[16][16];
    sr[part][q] = re;
    si[part][q] = im;
    __syncthreads();
    if (threadIdx.x < 16) {
        double R = 0.0, J = 0.0;
        for (int k = 0; k < 16; ++k) {
            R += sr[k][threadIdx.x];
            J += si[k][threadIdx.x];
        }
        sr[0][threadIdx.x] = R * R + J * J;     // column threadIdx.x: this thread's alone
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int k = 0; k < 16; ++k) a += sr[0][k];
        hp[2 * b] = a;
        hp[2 * b + 1] = 1.0;
    }
}

// alpha = sum(w) / |G w|^2, |G w|^2 estimated per stratum: the sampled
// units' sum times the stratum's units over its sampled units
__global__ void __launch_bounds__(256) k_identity_alpha_sampled(const double *__restrict__ pt,
                                                                const double *__restrict__ pw, int nb,
                                                                const double *__restrict__ hp, int nh,
                                                                double light_units, float *__restrict__ alpha)
{
    double t[TILE_SUMS], u[2], h[2];
    block_sums<TILE_SUMS>(pt, TILE_SUMS, nb, t);
    block_sums<2>(pw, 2, nb, u);
    block_sums<2>(hp, 2, nh, h);
    if (threadIdx.x != 0) return;
    const double light = t[1] > 0.0 ? t[0] * light_units / t[1] : 0.0;
    const double heavy = h[1] > 0.0 ? h[0] * t[2] / h[1] : 0.0;
    const double est = light + heavy;
    alpha[0] = est > 0.0 ? (float)(u[1] / est) : 0.f;
}

__global__ void k_scale_by(float *__restrict__ w, int64_t M, const float *__restrict__ alpha)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) w[i] = __fmul_rn(w[i], alpha[0]);
}

// w <- alpha w, and the same value into the positions' copy of w
__global__ void k_scale_by_pos(float *__restrict__ w, float4 *__restrict__ pos, int64_t M,
                               const float *__restrict__ alpha)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const float v = __fmul_rn(w[i], alpha[0]);
    w[i] = v;
    pos[i].w = v;
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
// with_traj: copy the samples' trajectory into the workspace's handle (the
// passes that read it: degridding, and gridding in the modes that take
// positions from the trajectory)
static fr_status dcf_workspace(fr_plan *p, const fr_samples *s, bool with_traj = true)
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
        const size_t den_bytes = std::max(sizeof(float2) * (size_t)M, sizeof(double) * ((TILE_SUMS + 2) * SUM_BLOCKS + 1));
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
    if (with_traj) {
        const int rows = s->frames > 0 ? 3 : s->dim;
        FR_HIP(ctx, hipMemcpyAsync(p->dcf_u->traj, s->traj, sizeof(float) * rows * M, hipMemcpyDeviceToDevice,
                                   ctx->stream));
        p->dcf_u->has_traj = true;
    }
    p->dcf_u->has_dcf = true;
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
// round trip; w is scaled in place.  With tile_stride > 1 on a plan with
// unit pairs, |G w|^2 is estimated from a sample of the grid in two strata,
// each sampled sum scaled by the stratum's units over its sampled units: the
// light units of one run of four consecutive pairs in every tile_stride (a
// regular sub-lattice of the Morton-ordered pairs), and the heavy (front)
// units whose rank by record count is a multiple of tile_stride.  A heavy unit's slices are cut into
// IDENTITY_SUB sub-slices on as many waves (its slices would otherwise run
// one per wave and set the pass's duration) and summed per cell here.
extern "C" fr_status fr_dcf_identity_scale(fr_plan *p, fr_samples *s, int tile_stride, float *alpha_out)
{
    if (!p || !s) return FR_ERR_INVALID;
    fr_ctx *ctx = p->ctx;
    FR_CHECK(ctx, s->M == p->M && s->dim == p->dim, FR_ERR_INVALID, "samples do not match the plan");
    FR_CHECK(ctx, s->frames == 0 && p->frames == 0, FR_ERR_UNSUPPORTED,
             "identity scale of batched frames: scale each frame on its own plan");
    FR_CHECK(ctx, s->has_traj && s->has_dcf, FR_ERR_STATE, "samples need a trajectory and a DCF");
    FR_CHECK(ctx, tile_stride >= 1, FR_ERR_INVALID, "tile_stride must be >= 1");
    const int64_t M = s->M;
    // a position-weight pass grids the workspace's unit data with s's DCF
    // through the positions (no copies); the other modes read the DCF and
    // trajectory of the handle they grid
    const bool pos = p->accumulate == FR_ACC_POS;
    fr_status st = dcf_workspace(p, s, !pos);
    if (st) return st;
    fr_samples *u = p->dcf_u;
    if (!pos) {
        hipMemcpyAsync(u->dcf, s->dcf, sizeof(float) * M, hipMemcpyDeviceToDevice, ctx->stream);
        fr_dcf_changed(u);
    }
    const bool sampled = tile_stride > 1 && pos && p->ntiles > 0 && p->tile_units == 2;
    // the partial sums and alpha live in the degridded-density buffer (sized
    // for them)
    double *part = reinterpret_cast<double *>(p->dcf_den);
    double *part2 = part + TILE_SUMS * SUM_BLOCKS;
    float *alpha = reinterpret_cast<float *>(part2 + 2 * SUM_BLOCKS);
    if (sampled) {
        // the front items' sub-slice sums, then the front units' partials
        const int64_t nf = p->nitems_front;
        FR_CHECK(ctx, nf < INT32_MAX, FR_ERR_UNSUPPORTED, "identity scale: %lld front items", (long long)nf);
        const size_t vbytes = sizeof(float2) * (size_t)nf * IDENTITY_SUB * 32 * p->geom.cells_max;
        const size_t need = vbytes + sizeof(double) * 2 * (size_t)std::max<int64_t>(nf, 1);
        if (need > p->dcf_scr_bytes) {
            FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
            hipFree(p->dcf_scr);
            p->dcf_scr = nullptr;
            p->dcf_scr_bytes = 0;
            st = fr_alloc(ctx, &p->dcf_scr, need);
            if (st) return st;
            p->dcf_scr_bytes = need;
        }
        float2 *vs = static_cast<float2 *>(p->dcf_scr);
        double *hp = reinterpret_cast<double *>(static_cast<char *>(p->dcf_scr) + vbytes);
        st = fr_launch_grid_pos(p, u, p->dcf_g->data, vs, tile_stride, s, IDENTITY_SUB);
        if (st) return st;
        hipLaunchKernelGGL(k_identity_tiles, dim3(SUM_BLOCKS), dim3(256), 0, ctx->stream, p->tiles, p->ntiles,
                           (int64_t)tile_stride, p->items, nf, p->dcf_g->data, (int)p->N, part);
        if (nf > 0)
            hipLaunchKernelGGL(k_identity_heavy, dim3((unsigned)nf), dim3(256), 0, ctx->stream, p->items, nf,
                               (int64_t)tile_stride, IDENTITY_SUB, vs, p->geom.cells_max, hp);
        hipLaunchKernelGGL(k_identity_partials, dim3(SUM_BLOCKS), dim3(256), 0, ctx->stream, p->dcf_g->data,
                           (int64_t)0, s->dcf, M, part2);
        hipLaunchKernelGGL(k_identity_alpha_sampled, dim3(1), dim3(256), 0, ctx->stream, part, part2, SUM_BLOCKS, hp,
                           (int)nf, (double)(p->nitems - nf), alpha);
    } else {
        st = fr_launch_grid(p, u, p->dcf_g->data, density_mode(p), nullptr, 0, pos ? s : nullptr);
        if (st) return st;
        hipLaunchKernelGGL(k_identity_partials, dim3(SUM_BLOCKS), dim3(256), 0, ctx->stream, p->dcf_g->data,
                           (int64_t)p->dcf_g->cells, s->dcf, M, part);
        hipLaunchKernelGGL(k_identity_alpha, dim3(1), dim3(256), 0, ctx->stream, part, SUM_BLOCKS, alpha);
    }
    // the positions hold s's current DCF after the pass: scale both, and
    // the positions stay current
    const bool pos_current = pos && p->pos_dcf_uid == s->uid && p->pos_dcf_version == s->dcf_version;
    if (pos_current)
        hipLaunchKernelGGL(k_scale_by_pos, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, s->dcf, p->pos, M, alpha);
    else
        hipLaunchKernelGGL(k_scale_by, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, s->dcf, M, alpha);
    fr_dcf_changed(s);
    if (pos_current) p->pos_dcf_version = s->dcf_version;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && alpha_out)
        e = hipMemcpyAsync(alpha_out, alpha, sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    FR_CHECK(ctx, e == hipSuccess, FR_ERR_HIP, "identity scale: %s", hipGetErrorString(e));
    return FR_OK;
}
