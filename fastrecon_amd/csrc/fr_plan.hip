// Gridding plan: per-(tile, slab) record streams of taps.
//
// Replaces cuGridLut::plan (ReconData/cuGridLut.cu:151-226).  The reference
// materialises a (cell, {sample, distance}) tuple for every box cell of
// every sample, sorts by cell and streams 25 partitions over PCIe per coil.
// Here the grid is cut into tiles; each tile is owned by one workgroup whose
// waves each own a slab (a band of the slowest tile axis).  For every
// (tile, slab) the plan holds one stream of records, in ascending sample
// order:
//     [sample index][n] then n x [cell index inside the tile][LUT weight]
// The stream order is produced by a stable radix sort of (key, sample)
// pairs generated in sample order, so every grid cell later receives its
// contributions in the reference CPU order.  The plan lives in HBM and is
// reused for every coil, frame and DCF.
#include "fr_internal.h"
#include "fr_taps.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>

namespace {

struct key_ctx {
    fr_tiling tl;
    float hw;
    int L;
    int dim;
    int N;
};

// Enumerate the (tile, slab) keys a sample's box touches and the taps in
// each: f(key, lo, hi) is called per key with the sub-box of the tile/slab.
template <typename F>
__device__ __forceinline__ void for_each_key(const key_ctx &k, const fr_box &b, F &&f)
{
    const fr_tiling &tl = k.tl;
    if (b.s[0] > b.e[0] || b.s[1] > b.e[1] || b.s[2] > b.e[2]) return;
    const int slow = k.dim == 3 ? 2 : 1;
    const int txs = b.s[0] / tl.tx, txe = b.e[0] / tl.tx;
    const int tys = b.s[1] / tl.ty, tye = b.e[1] / tl.ty;
    const int tzs = b.s[2] / tl.tz, tze = b.e[2] / tl.tz;
    const int text[3] = {tl.tx, tl.ty, tl.tz};
    for (int tz = tzs; tz <= tze; ++tz)
        for (int ty = tys; ty <= tye; ++ty)
            for (int tx = txs; tx <= txe; ++tx) {
                const int t[3] = {tx, ty, tz};
                int lo[3], hi[3];
                for (int a = 0; a < 3; ++a) {
                    lo[a] = max(b.s[a], t[a] * text[a]);
                    hi[a] = min(b.e[a], t[a] * text[a] + text[a] - 1);
                }
                const int tile = (tz * tl.nty + ty) * tl.ntx + tx;
                // split the slow axis into slabs
                const int base = t[slow] * text[slow];
                const int sl0 = (lo[slow] - base) / tl.slab_depth;
                const int sl1 = (hi[slow] - base) / tl.slab_depth;
                for (int sl = sl0; sl <= sl1; ++sl) {
                    int l2[3] = {lo[0], lo[1], lo[2]}, h2[3] = {hi[0], hi[1], hi[2]};
                    l2[slow] = max(lo[slow], base + sl * tl.slab_depth);
                    h2[slow] = min(hi[slow], base + sl * tl.slab_depth + tl.slab_depth - 1);
                    f(tile * tl.slabs + sl, t, l2, h2);
                }
            }
}

__global__ void k_count_records(const float *__restrict__ traj, int64_t M, key_ctx k,
                                uint32_t *__restrict__ nrec, uint32_t *__restrict__ ntap)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    fr_box b;
    fr_sample_box(k.dim, k.N, traj, M, s, k.hw, b);
    uint32_t r = 0, t = 0;
    for_each_key(k, b, [&](int key, const int *tcoord, const int *lo, const int *hi) {
        uint32_t n = 0;
        fr_box_taps(b, lo, hi, k.hw, k.L, [&](int, int, int, int) { ++n; });
        if (n) { ++r; t += n; }
    });
    nrec[s] = r;
    ntap[s] = t;
}

__global__ void k_emit_records(const float *__restrict__ traj, int64_t M, key_ctx k,
                               const uint64_t *__restrict__ pos, uint32_t *__restrict__ keys,
                               uint64_t *__restrict__ vals)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    fr_box b;
    fr_sample_box(k.dim, k.N, traj, M, s, k.hw, b);
    uint64_t p = pos[s];
    for_each_key(k, b, [&](int key, const int *tcoord, const int *lo, const int *hi) {
        uint32_t n = 0;
        fr_box_taps(b, lo, hi, k.hw, k.L, [&](int, int, int, int) { ++n; });
        if (n) {
            keys[p] = (uint32_t)key;
            vals[p] = ((uint64_t)s << 32) | n;
            ++p;
        }
    });
}

__global__ void k_record_words(const uint64_t *__restrict__ vals, int64_t R, uint64_t *__restrict__ words)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) words[r] = 2 + 2 * (vals[r] & 0xffffffffu);
}

__global__ void k_key_ranges(const uint32_t *__restrict__ keys, const uint64_t *__restrict__ off,
                             const uint64_t *__restrict__ vals, int64_t R, ulonglong2 *__restrict__ ranges)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    uint32_t k = keys[r];
    if (r == 0 || keys[r - 1] != k) ranges[k].x = off[r];
    if (r == R - 1 || keys[r + 1] != k) ranges[k].y = off[r] + 2 + 2 * (vals[r] & 0xffffffffu);
}

__global__ void k_fill_records(const float *__restrict__ traj, int64_t M, key_ctx k,
                               const uint32_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                               const uint64_t *__restrict__ off, int64_t R, const float *__restrict__ lut,
                               uint32_t *__restrict__ words)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const uint32_t key = keys[r];
    const int64_t s = (int64_t)(vals[r] >> 32);
    const uint32_t n = (uint32_t)(vals[r] & 0xffffffffu);
    fr_box b;
    fr_sample_box(k.dim, k.N, traj, M, s, k.hw, b);
    uint32_t *w = words + off[r];
    w[0] = (uint32_t)s;
    w[1] = n;
    uint32_t i = 0;
    const fr_tiling tl = k.tl;
    for_each_key(k, b, [&](int kk, const int *t, const int *lo, const int *hi) {
        if ((uint32_t)kk != key) return;
        const int ox = t[0] * tl.tx, oy = t[1] * tl.ty, oz = t[2] * tl.tz;
        fr_box_taps(b, lo, hi, k.hw, k.L, [&](int x, int y, int z, int ki) {
            uint32_t cell = (uint32_t)(((z - oz) * tl.ty + (y - oy)) * tl.tx + (x - ox));
            w[2 + 2 * i] = cell;
            w[3 + 2 * i] = __float_as_uint(lut[ki]);
            ++i;
        });
    });
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
struct dbuf {
    T *p = nullptr;
    ~dbuf() { if (p) (void)hipFree(p); }
};

}  // namespace

// Choose the tile decomposition.  The LDS accumulator holds, per tile cell,
// 2 * min(C, 32) floats plus one pad float (conflict-free write-out).
static fr_tiling choose_tiling(int dim, unsigned N, int coils, const fr_tune *tune)
{
    fr_tiling t{};
    t.dim = dim;
    t.N = (int)N;
    const int P = 2 * std::min(coils, 32);
    if (dim == 3) {
        t.tx = t.ty = t.tz = P <= 4 ? 16 : 8;
        t.slabs = t.tz;
    } else {
        t.tx = P <= 4 ? 64 : 32;
        t.ty = P <= 4 ? 64 : 32;
        t.tz = 1;
        t.slabs = 16;
    }
    if (tune) {
        if (tune->tile_x > 0) t.tx = tune->tile_x;
        if (tune->tile_y > 0) t.ty = tune->tile_y;
        if (dim == 3 && tune->tile_z > 0) t.tz = tune->tile_z;
        if (tune->slabs > 0) t.slabs = tune->slabs;
    }
    const int slow = dim == 3 ? t.tz : t.ty;
    if (t.slabs > slow) t.slabs = slow;
    t.slab_depth = slow / t.slabs;
    t.ntx = ((int)N + t.tx - 1) / t.tx;
    t.nty = ((int)N + t.ty - 1) / t.ty;
    t.ntz = dim == 3 ? ((int)N + t.tz - 1) / t.tz : 1;
    return t;
}

fr_status fr_plan_build(fr_plan *p, const fr_samples *s, const fr_tune *tune)
{
    fr_ctx *ctx = p->ctx;
    hipStream_t st = ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    fr_tiling tl = choose_tiling(p->dim, p->N, s->C, tune);
    const int slow = p->dim == 3 ? tl.tz : tl.ty;
    FR_CHECK(ctx, tl.tx >= 1 && tl.ty >= 1 && tl.tz >= 1 && tl.slabs >= 1 && slow % tl.slabs == 0,
             FR_ERR_INVALID, "tile %dx%dx%d with %d slabs is not valid", tl.tx, tl.ty, tl.tz, tl.slabs);
    FR_CHECK(ctx, tl.slabs * 64 <= 1024, FR_ERR_INVALID, "at most 16 slabs per tile");
    const int P = 2 * std::min(s->C, 32);
    const int64_t lds = (int64_t)tl.tile_cells() * (P + 1) * 4;
    FR_CHECK(ctx, lds <= 160 * 1024, FR_ERR_INVALID,
             "tile %dx%dx%d needs %lld B of LDS for %d coils (max 163840)", tl.tx, tl.ty, tl.tz,
             (long long)lds, std::min(s->C, 32));
    const int64_t nkeys = (int64_t)tl.tiles() * tl.slabs;
    FR_CHECK(ctx, nkeys < (int64_t(1) << 31), FR_ERR_UNSUPPORTED, "too many tiles");
    p->tl = tl;
    p->accumulate = tune ? tune->accumulate : 0;

    key_ctx k{tl, p->kHW, p->L, p->dim, (int)p->N};
    const int64_t M = s->M;
    dbuf<uint32_t> nrec, ntap;
    dbuf<uint64_t> pos;
    fr_status e;
    if ((e = fr_alloc(ctx, (void **)&nrec.p, 4 * M)) || (e = fr_alloc(ctx, (void **)&ntap.p, 4 * M)) ||
        (e = fr_alloc(ctx, (void **)&pos.p, 8 * (M + 1))))
        return e;
    hipLaunchKernelGGL(k_count_records, dim3(nblk(M, 256)), dim3(256), 0, st, s->traj, M, k, nrec.p, ntap.p);
    FR_HIP(ctx, hipGetLastError());

    // exclusive scan of record counts -> positions (64-bit)
    size_t tmp_bytes = 0;
    auto rec64 = hipcub::TransformInputIterator<uint64_t, hipcub::CastOp<uint64_t>, const uint32_t *>(
        nrec.p, hipcub::CastOp<uint64_t>());
    FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, rec64, pos.p + 1, (int)M, st));
    dbuf<char> tmp;
    if ((e = fr_alloc(ctx, (void **)&tmp.p, tmp_bytes + 16))) return e;
    FR_HIP(ctx, hipMemsetAsync(pos.p, 0, 8, st));
    FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(tmp.p, tmp_bytes, rec64, pos.p + 1, (int)M, st));
    uint64_t R = 0;
    FR_HIP(ctx, hipMemcpyAsync(&R, pos.p + M, 8, hipMemcpyDeviceToHost, st));
    FR_HIP(ctx, hipStreamSynchronize(st));
    FR_CHECK(ctx, R < (uint64_t(1) << 31), FR_ERR_UNSUPPORTED, "plan has %llu records (max 2^31)",
             (unsigned long long)R);

    dbuf<uint32_t> keys_in, keys_out;
    dbuf<uint64_t> vals_in, vals_out;
    if ((e = fr_alloc(ctx, (void **)&keys_in.p, 4 * R + 4)) || (e = fr_alloc(ctx, (void **)&keys_out.p, 4 * R + 4)) ||
        (e = fr_alloc(ctx, (void **)&vals_in.p, 8 * R + 8)) || (e = fr_alloc(ctx, (void **)&vals_out.p, 8 * R + 8)))
        return e;
    hipLaunchKernelGGL(k_emit_records, dim3(nblk(M, 256)), dim3(256), 0, st, s->traj, M, k, pos.p, keys_in.p,
                       vals_in.p);
    FR_HIP(ctx, hipGetLastError());

    int end_bit = 1;
    while (end_bit < 32 && (int64_t(1) << end_bit) < nkeys) ++end_bit;
    size_t sort_bytes = 0;
    FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys_in.p, keys_out.p, vals_in.p,
                                                   vals_out.p, (int)R, 0, end_bit, st));
    dbuf<char> sort_tmp;
    if ((e = fr_alloc(ctx, (void **)&sort_tmp.p, sort_bytes + 16))) return e;
    FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(sort_tmp.p, sort_bytes, keys_in.p, keys_out.p, vals_in.p,
                                                   vals_out.p, (int)R, 0, end_bit, st));

    // word offsets of each record (exclusive scan of record sizes)
    dbuf<uint64_t> rw, off;
    if ((e = fr_alloc(ctx, (void **)&rw.p, 8 * R + 8)) || (e = fr_alloc(ctx, (void **)&off.p, 8 * R + 16)))
        return e;
    hipLaunchKernelGGL(k_record_words, dim3(nblk(R, 256)), dim3(256), 0, st, vals_out.p, (int64_t)R, rw.p);
    size_t scan_bytes = 0;
    FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, rw.p, off.p + 1, (int)R, st));
    dbuf<char> scan_tmp;
    if ((e = fr_alloc(ctx, (void **)&scan_tmp.p, scan_bytes + 16))) return e;
    FR_HIP(ctx, hipMemsetAsync(off.p, 0, 8, st));
    FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(scan_tmp.p, scan_bytes, rw.p, off.p + 1, (int)R, st));
    uint64_t W = 0;
    FR_HIP(ctx, hipMemcpyAsync(&W, off.p + R, 8, hipMemcpyDeviceToHost, st));
    FR_HIP(ctx, hipStreamSynchronize(st));

    if ((e = fr_alloc(ctx, (void **)&p->words, 4 * W + 16))) return e;
    if ((e = fr_alloc(ctx, (void **)&p->ranges, sizeof(ulonglong2) * nkeys))) return e;
    FR_HIP(ctx, hipMemsetAsync(p->ranges, 0, sizeof(ulonglong2) * nkeys, st));
    if (R > 0) {
        hipLaunchKernelGGL(k_key_ranges, dim3(nblk(R, 256)), dim3(256), 0, st, keys_out.p, off.p, vals_out.p,
                           (int64_t)R, (ulonglong2 *)p->ranges);
        hipLaunchKernelGGL(k_fill_records, dim3(nblk(R, 128)), dim3(128), 0, st, s->traj, M, k, keys_out.p,
                           vals_out.p, off.p, (int64_t)R, p->lut, p->words);
    }
    FR_HIP(ctx, hipGetLastError());
    FR_HIP(ctx, hipStreamSynchronize(st));
    p->nwords = (int64_t)W;
    p->nrecords = (int64_t)R;
    p->ntaps = ((int64_t)W - 2 * (int64_t)R) / 2;
    p->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FR_OK;
}

extern "C" fr_status fr_plan_create(fr_ctx *ctx, const fr_samples *s, unsigned N, float kWidth,
                                    const float *lut, int L, const fr_tune *tune, fr_plan **out)
{
    if (!ctx || !s || !lut || !out) return FR_ERR_INVALID;
    *out = nullptr;
    FR_CHECK(ctx, s->has_traj, FR_ERR_STATE, "trajectory not set");
    FR_CHECK(ctx, N >= 2 && N <= 4096, FR_ERR_INVALID, "grid size %u outside [2,4096]", N);
    FR_CHECK(ctx, s->dim == 2 || (uint64_t)N * N * N < (uint64_t(1) << 31), FR_ERR_UNSUPPORTED,
             "3D grid %u^3 exceeds 2^31 cells", N);
    FR_CHECK(ctx, L >= 2 && L <= (1 << 20), FR_ERR_INVALID, "LUT length %d invalid", L);
    FR_CHECK(ctx, kWidth > 0.f && kWidth <= 16.f, FR_ERR_INVALID, "kernel width %g invalid", (double)kWidth);
    fr_plan *p = new (std::nothrow) fr_plan;
    if (!p) return FR_ERR_NOMEM;
    p->ctx = ctx; p->dim = s->dim; p->N = N; p->M = s->M;
    p->kWidth = kWidth; p->kHW = kWidth / 2.0f; p->L = L;     // GridLut.cpp:86
    p->host_lut.assign(lut, lut + L);
    fr_status e = fr_alloc(ctx, (void **)&p->lut, sizeof(float) * L);
    if (!e) {
        hipError_t he = hipMemcpyAsync(p->lut, lut, sizeof(float) * L, hipMemcpyHostToDevice, ctx->stream);
        if (he != hipSuccess) e = fr_fail(ctx, FR_ERR_HIP, "lut upload: %s", hipGetErrorString(he));
    }
    if (!e) e = fr_plan_build(p, s, tune);
    if (e) { fr_plan_destroy(p); return e; }
    *out = p;
    return FR_OK;
}

extern "C" fr_status fr_plan_destroy(fr_plan *p)
{
    if (!p) return FR_OK;
    hipStreamSynchronize(p->ctx->stream);
    hipFree(p->lut); hipFree(p->words); hipFree(p->ranges);
    delete p;
    return FR_OK;
}

extern "C" fr_status fr_plan_get_info(const fr_plan *p, fr_plan_info *info)
{
    if (!p || !info) return FR_ERR_INVALID;
    info->samples = p->M;
    info->taps = p->ntaps;
    info->records = p->nrecords;
    info->streams = (int64_t)p->tl.tiles() * p->tl.slabs;
    info->tile_x = p->tl.tx; info->tile_y = p->tl.ty; info->tile_z = p->tl.tz;
    info->slabs = p->tl.slabs;
    info->tiles = p->tl.tiles();
    info->lds_bytes = 0;
    info->plan_bytes = 4 * p->nwords + (int64_t)sizeof(ulonglong2) * info->streams;
    info->build_ms = p->build_ms;
    return FR_OK;
}
