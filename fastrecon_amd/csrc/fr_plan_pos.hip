// Gridding plan of the position-weight matrix-core mode (FR_ACC_POS).
//
// Replaces cuGridLut::plan (ReconData/cuGridLut.cu:151-226) for this mode.
// The grid is tiled by fixed boxes of at most 16 cells (4 x 2 x 2 by
// default; 8 x 2 in 2D and for batched frames, one frame deep).  For every
// box the plan lists the samples with at least one accepted tap in it
// (tap test of GridLut.cpp:99-136), in ascending sample order, padded to a
// multiple of FR_POS_BATCH (16) with the null sample M.  No taps are stored: the gridding
// kernel recomputes each weight from the sample position.
//
// Everything runs on the device; the host reads three small totals.
//   1. per-sample record counts and per-box counts (one pass over samples);
//   2. box order: boxes with more than a quarter slice of records first,
//      heaviest first, then Morton order (a 64-bit key radix sort);
//   3. per-box padded offsets, item and slot counts (scans);
//   4. (box rank, sample) pairs in sample order, stable-sorted by rank, then
//      placed at the padded offsets;
//   5. items: each box's records cut into slices of at most S records
//      (a multiple of 16); boxes with several slices write partial sums to
//      scratch slots, combined afterwards in slice order (k_combine);
//   6. tiles: the light boxes grouped by TW x-adjacent boxes, in Morton
//      order of the tiles: pairs (TW = 2, the default), which a wave grids
//      one after the other and writes out in rows of two units (64 bytes for
//      4-cell-wide units), or fours (TW = 4), which a workgroup grids and
//      writes out in whole rows.
#include "fr_internal.h"
#include "fr_taps.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

namespace {

struct pgeom {
    float hw;
    int L, dim, N, frames;
    int bx, by, bz;
    int nbx, nby, nbz;
};

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
struct dbuf {
    T *p = nullptr;
    ~dbuf() { if (p) (void)hipFree(p); }
};

// Grid positions as float4 (x, y, z, 0); 2D samples get z = 0 and batched
// frames their frame index (the box of fr_sample_box); entry M is the null
// sample, far outside the grid.
__global__ void k_positions(const float *__restrict__ traj, int64_t M, int dim, int frames,
                            float4 *__restrict__ pos)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > M) return;
    if (i == M) {
        pos[i] = make_float4(3e30f, 3e30f, 3e30f, 0.f);
        return;
    }
    float z = 0.f;
    if (frames > 0)
        z = (float)min(max((int)traj[2 * M + i], 0), frames - 1);
    else if (dim == 3)
        z = traj[2 * M + i];
    pos[i] = make_float4(traj[i], traj[M + i], z, 0.f);     // .w: the DCF, set before each pass
}

// Does the sample of box b have an accepted tap in [lo, hi] (inclusive)?
// Same operations as fr_box_taps (the reference tap test).
__device__ __forceinline__ bool any_tap(const fr_box &b, const int lo[3], const int hi[3], float hw)
{
    for (int z = lo[2]; z <= hi[2]; ++z) {
        const float dz = __fsub_rn((float)z, b.c[2]);
        const float dz2 = __fmul_rn(dz, dz);
        for (int y = lo[1]; y <= hi[1]; ++y) {
            const float dy = __fsub_rn((float)y, b.c[1]);
            const float dy2 = __fmul_rn(dy, dy);
            for (int x = lo[0]; x <= hi[0]; ++x) {
                const float dx = __fsub_rn((float)x, b.c[0]);
                if (sqrtf(fr_tap_dk2(dz2, dy2, dx)) < hw) return true;
            }
        }
    }
    return false;
}

// f(box id) for every box in which sample s has an accepted tap, in box
// order (z, y, x); returns the count.
template <typename F>
__device__ __forceinline__ uint32_t sample_boxes(const float *__restrict__ traj, int64_t M, int64_t s,
                                                 const pgeom &g, F &&f)
{
    fr_box b;
    fr_sample_box(g.dim, g.N, traj, M, s, g.hw, b, g.frames);
    if (b.s[0] > b.e[0] || b.s[1] > b.e[1] || b.s[2] > b.e[2]) return 0;
    uint32_t n = 0;
    for (int Z = b.s[2] / g.bz; Z <= b.e[2] / g.bz; ++Z)
        for (int Y = b.s[1] / g.by; Y <= b.e[1] / g.by; ++Y)
            for (int X = b.s[0] / g.bx; X <= b.e[0] / g.bx; ++X) {
                const int lo[3] = {max(b.s[0], X * g.bx), max(b.s[1], Y * g.by), max(b.s[2], Z * g.bz)};
                const int hi[3] = {min(b.e[0], X * g.bx + g.bx - 1), min(b.e[1], Y * g.by + g.by - 1),
                                   min(b.e[2], Z * g.bz + g.bz - 1)};
                if (any_tap(b, lo, hi, g.hw)) {
                    f(((uint32_t)Z * g.nby + Y) * g.nbx + X, n);
                    ++n;
                }
            }
    return n;
}

// Per-box record counts, first added up per workgroup in an LDS hash table
// keyed by box (consecutive samples of a spoke share most of their boxes: on
// the 256^3 kooshball a workgroup's 256 samples make about 2000 hits on
// about 500 distinct boxes), then flushed with one global atomic per
// distinct box.  A hit that finds no free or matching slot within
// BOX_PROBES probes goes to the global counter directly, so every hit is
// counted once whatever the table holds.
constexpr int BOX_HASH_BITS = 12;
constexpr uint32_t BOX_HASH = 1u << BOX_HASH_BITS, BOX_PROBES = 16, BOX_EMPTY = 0xffffffffu;
__global__ void __launch_bounds__(256) k_box_count(const float *__restrict__ traj, int64_t M, pgeom g,
                                                   uint32_t *__restrict__ cnt, uint32_t *__restrict__ nrec)
{
    __shared__ uint32_t hkey[BOX_HASH], hcnt[BOX_HASH];
    for (uint32_t i = threadIdx.x; i < BOX_HASH; i += blockDim.x) {
        hkey[i] = BOX_EMPTY;
        hcnt[i] = 0u;
    }
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < M)
        nrec[s] = sample_boxes(traj, M, s, g, [&](uint32_t box, uint32_t) {
            uint32_t h = (box * 2654435761u) >> (32 - BOX_HASH_BITS);
            for (uint32_t probe = 0; probe < BOX_PROBES; ++probe) {
                const uint32_t k = atomicCAS(&hkey[h], BOX_EMPTY, box);
                if (k == BOX_EMPTY || k == box) {
                    atomicAdd(&hcnt[h], 1u);
                    return;
                }
                h = (h + 1u) & (BOX_HASH - 1u);
            }
            atomicAdd(&cnt[box], 1u);
        });
    __syncthreads();
    // a slot with a count holds its box (the thread that claims a slot adds
    // to it before the barrier)
    for (uint32_t i = threadIdx.x; i < BOX_HASH; i += blockDim.x)
        if (hcnt[i]) atomicAdd(&cnt[hkey[i]], hcnt[i]);
}

__device__ __forceinline__ uint64_t spread3(uint32_t v)
{
    uint64_t m = 0;
    for (int bit = 0; bit < 21; ++bit) m |= (uint64_t)((v >> bit) & 1u) << (3 * bit);
    return m;
}

__device__ __forceinline__ uint64_t spread2(uint32_t v)
{
    uint64_t m = 0;
    for (int bit = 0; bit < 21; ++bit) m |= (uint64_t)((v >> bit) & 1u) << (2 * bit);
    return m;
}

// Sort key of a box: boxes above the heavy cut first, most records first;
// the rest in Morton order (frame-major for batched frames, which share no
// samples).
__global__ void k_box_keys(const uint32_t *__restrict__ cnt, int64_t nbox, pgeom g, uint32_t heavy_cut,
                           uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const uint32_t X = (uint32_t)(b % g.nbx), Y = (uint32_t)((b / g.nbx) % g.nby), Z = (uint32_t)(b / ((int64_t)g.nbx * g.nby));
    const uint32_t c = cnt[b];
    uint64_t k;
    if (c > heavy_cut)
        k = (uint64_t)(0xffffffffu - c);
    else if (g.frames > 0)
        k = (uint64_t(1) << 63) | ((uint64_t)Z << 42) | spread2(X) | (spread2(Y) << 1);
    else
        k = (uint64_t(1) << 63) | spread3(X) | (spread3(Y) << 1) | (spread3(Z) << 2);
    key[b] = k;
    val[b] = (uint32_t)b;
}

// Per unit (rank order): record count, padded count, item count, heavy-slot
// count and flags.
__global__ void k_unit_stats(const uint32_t *__restrict__ order, int64_t nu, const uint32_t *__restrict__ cnt,
                             uint32_t S, uint32_t heavy_cut, uint32_t *__restrict__ rank, uint32_t *__restrict__ c_r,
                             uint32_t *__restrict__ c4_r, uint32_t *__restrict__ nit_r,
                             uint32_t *__restrict__ hs_r, uint32_t *__restrict__ hf_r, uint32_t *__restrict__ fr_r)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint32_t b = order[u];
    rank[b] = (uint32_t)u;
    const uint32_t c = cnt[b], c4 = (c + (FR_POS_BATCH - 1)) & ~(uint32_t)(FR_POS_BATCH - 1);
    const uint32_t nit = c4 == 0 ? 1u : (c4 + S - 1) / S;
    c_r[u] = c;
    c4_r[u] = c4;
    nit_r[u] = nit;
    hs_r[u] = nit > 1 ? nit : 0u;
    hf_r[u] = nit > 1;
    fr_r[u] = c > heavy_cut;
}

struct totals {
    uint64_t R, R4, nitems, nslots, nheavy, nfront_units, nfront_items;
};

__global__ void k_totals(int64_t nu, const uint64_t *__restrict__ Q, const uint64_t *__restrict__ P,
                         const uint64_t *__restrict__ I, const uint64_t *__restrict__ HS,
                         const uint64_t *__restrict__ HI, const uint64_t *__restrict__ FR,
                         const uint32_t *__restrict__ c_r, const uint32_t *__restrict__ c4_r,
                         const uint32_t *__restrict__ nit_r, const uint32_t *__restrict__ hs_r,
                         const uint32_t *__restrict__ hf_r, const uint32_t *__restrict__ fr_r, totals *__restrict__ t)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int64_t l = nu - 1;
    t->R = Q[l] + c_r[l];
    t->R4 = P[l] + c4_r[l];
    t->nitems = I[l] + nit_r[l];
    t->nslots = HS[l] + hs_r[l];
    t->nheavy = HI[l] + hf_r[l];
    t->nfront_units = FR[l] + fr_r[l];
    t->nfront_items = t->nfront_units < (uint64_t)nu ? I[t->nfront_units] : t->nitems;
}

__global__ void k_units(const uint32_t *__restrict__ order, int64_t nu, pgeom g, const uint64_t *__restrict__ P,
                        const uint32_t *__restrict__ c_r, const uint32_t *__restrict__ c4_r, fr_unit *__restrict__ units)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const int64_t b = order[u];
    const int X = (int)(b % g.nbx), Y = (int)((b / g.nbx) % g.nby), Z = (int)(b / ((int64_t)g.nbx * g.nby));
    const int ox = X * g.bx, oy = Y * g.by, oz = Z * g.bz;
    const int ux = min(g.bx, g.N - ox), uy = min(g.by, g.N - oy);
    const int uz = g.dim == 3 && g.frames == 0 ? min(g.bz, g.N - oz) : 1;
    units[u] = fr_unit{ox, oy, oz, ux | (uy << 8) | (uz << 16), P[u], 0, c4_r[u], c_r[u]};
}

__global__ void k_emit(const float *__restrict__ traj, int64_t M, pgeom g, const uint64_t *__restrict__ spos,
                       const uint32_t *__restrict__ rank, uint32_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    const uint64_t p = spos[s];
    sample_boxes(traj, M, s, g, [&](uint32_t box, uint32_t n) {
        keys[p + n] = rank[box];
        vals[p + n] = (uint32_t)s;
    });
}

__global__ void k_place(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals, int64_t R,
                        const uint64_t *__restrict__ P, const uint64_t *__restrict__ Q, uint32_t *__restrict__ recs)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const uint32_t u = keys[r];
    recs[P[u] + ((uint64_t)r - Q[u])] = vals[r];
}

// padding records (null sample M) after each unit's records, and 64 past the end
__global__ void k_pad(int64_t nu, const uint64_t *__restrict__ P, const uint32_t *__restrict__ c_r,
                      const uint32_t *__restrict__ c4_r, uint64_t R4, uint32_t null_sample, uint32_t *__restrict__ recs)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nu)
        for (uint32_t i = c_r[u]; i < c4_r[u]; ++i) recs[P[u] + i] = null_sample;
    if (u < 64) recs[R4 + u] = null_sample;
}

__global__ void k_items(int64_t nu, uint32_t S, const uint64_t *__restrict__ P, const uint64_t *__restrict__ I,
                        const uint64_t *__restrict__ HS, const uint64_t *__restrict__ HI,
                        const uint32_t *__restrict__ c4_r, const uint32_t *__restrict__ nit_r,
                        const fr_unit *__restrict__ units, fr_item *__restrict__ items, fr_heavy *__restrict__ heavy)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint32_t nit = nit_r[u];
    const uint64_t end = P[u] + c4_r[u];
    // the unit's box travels with its items (the kernel reads no unit record)
    const fr_unit U = units[u];
    const uint32_t ux = U.ext & 255, uy = (U.ext >> 8) & 255, uz = U.ext >> 16;
    const uint32_t box0 = (uint32_t)U.ox | ((uint32_t)U.oy << 16);
    const uint32_t box1 = (uint32_t)U.oz | ((ux | (uy << 5) | (uz << 10)) << 16);
    for (uint32_t q = 0; q < nit; ++q) {
        const uint64_t r0 = P[u] + (uint64_t)q * S;
        const uint64_t r1 = min(r0 + S, end);
        items[I[u] + q] = fr_item{(uint32_t)u, nit > 1 ? (int32_t)(HS[u] + q) : -1, (uint32_t)r0, (uint32_t)r1, box0,
                                  box1};
    }
    if (nit > 1) heavy[HI[u]] = fr_heavy{(uint32_t)u, nit, (uint32_t)HS[u], 0u};
}

// Tiles: TW x-adjacent units.  Sort key: Morton order of the tile
// coordinates (frame-major for batched frames).
__global__ void k_tile_keys(int64_t ntiles, int ntx, int nty, int frames, uint64_t *__restrict__ key,
                            uint32_t *__restrict__ val)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const uint32_t X = (uint32_t)(t % ntx), Y = (uint32_t)((t / ntx) % nty), Z = (uint32_t)(t / ((int64_t)ntx * nty));
    key[t] = frames > 0 ? (((uint64_t)Z << 42) | spread2(X) | (spread2(Y) << 1))
                        : (spread3(X) | (spread3(Y) << 1) | (spread3(Z) << 2));
    val[t] = (uint32_t)t;
}

// tiles in launch order: box and no items yet; tpos = inverse permutation
__global__ void k_tile_init(const uint32_t *__restrict__ torder, int64_t ntiles, int ntx, int nty, pgeom g, int tw,
                            fr_tile *__restrict__ tiles, uint32_t *__restrict__ tpos)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ntiles) return;
    const uint32_t t = torder[q];
    tpos[t] = (uint32_t)q;
    const uint32_t X = t % ntx, Y = (t / ntx) % nty, Z = t / (ntx * nty);
    tiles[q] = fr_tile{{-1, -1, -1, -1}, (X * tw * g.bx) | ((Y * g.by) << 16), Z * g.bz};
}

// every light unit's (single) item into its tile slot
__global__ void k_tile_fill(const uint32_t *__restrict__ order, int64_t nu, int64_t nfront_units, pgeom g, int ntx, int tw,
                            const uint32_t *__restrict__ tpos, const uint64_t *__restrict__ I,
                            fr_tile *__restrict__ tiles)
{
    const int64_t u = nfront_units + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const int64_t b = order[u];
    const uint32_t X = (uint32_t)(b % g.nbx), Y = (uint32_t)((b / g.nbx) % g.nby),
                   Z = (uint32_t)(b / ((int64_t)g.nbx * g.nby));
    const uint32_t t = (Z * g.nby + Y) * ntx + X / tw;
    tiles[tpos[t]].item[X % tw] = (int32_t)I[u];
}

template <typename In>
fr_status excl_scan64(fr_ctx *ctx, In in, uint64_t *out, int64_t n, dbuf<char> &tmp, size_t &tmp_bytes)
{
    size_t need = 0;
    auto in64 = hipcub::TransformInputIterator<uint64_t, hipcub::CastOp<uint64_t>, In>(in, hipcub::CastOp<uint64_t>());
    FR_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, need, in64, out, (int)n, ctx->stream));
    if (need > tmp_bytes) {
        FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (tmp.p) (void)hipFree(tmp.p);
        tmp.p = nullptr;
        tmp_bytes = 0;
        fr_status e = fr_alloc(ctx, (void **)&tmp.p, need);
        if (e) return e;
        tmp_bytes = need;
    }
    FR_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp.p, need, in64, out, (int)n, ctx->stream));
    return FR_OK;
}

}  // namespace

fr_status fr_plan_build_pos(fr_plan *p, const fr_samples *s, const fr_tune *tune)
{
    fr_ctx *ctx = p->ctx;
    hipStream_t st = ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    const int N = (int)p->N;
    const int64_t M = s->M;
    FR_CHECK(ctx, p->L <= FR_POS_MAX_L, FR_ERR_UNSUPPORTED,
             "position-weight gridding (accumulate 6) takes LUTs of at most %d entries (have %d)", FR_POS_MAX_L, p->L);
    FR_CHECK(ctx, !(tune && tune->exact_order), FR_ERR_INVALID,
             "exact_order needs a sequential accumulation mode (1, 2 or 4); the matrix cores sum four products at "
             "a time");

    // unit box: 4 x 2 x 2 in 3D; 8 x 2 in 2D and per frame (the 64 uneven
    // spiral frames on 384^2: 1.01 ms against 1.12 with 4 x 4; the 402-spoke
    // radial case on 512^2 is the same with both)
    fr_geom geom{};
    geom.dim = p->dim;
    geom.N = N;
    const bool flat = p->dim == 2;
    geom.bx = flat ? 8 : 4;
    geom.by = 2;
    geom.bz = flat ? 1 : 2;
    if (tune) {
        if (tune->tile_x > 0) geom.bx = tune->tile_x;
        if (tune->tile_y > 0) geom.by = tune->tile_y;
        if (!flat && tune->tile_z > 0) geom.bz = tune->tile_z;
    }
    geom.bx = std::min(geom.bx, N);
    geom.by = std::min(geom.by, N);
    geom.bz = flat ? 1 : std::min(geom.bz, N);
    geom.cells_max = geom.bx * geom.by * geom.bz;
    geom.waves = FR_POS_WAVES;
    FR_CHECK(ctx, geom.bx >= 1 && geom.by >= 1 && geom.bz >= 1 && geom.cells_max <= 16, FR_ERR_INVALID,
             "position-weight gridding (accumulate 6) takes unit boxes of at most 16 cells (have %dx%dx%d)",
             geom.bx, geom.by, geom.bz);
    p->geom = geom;
    p->accumulate = FR_ACC_POS;
    p->exact_order = 0;
    // defaults measured on the 32-coil 256^3 kooshball (4.70 ms with pairs;
    // 4.95 unit by unit, 5.41 with tiles of four; streaming stores slower)
    p->pos_ipw = 1;
    p->xcd_run = 256;
    p->pos_stores = 2;
    p->pos_waves = 4;
    if (tune && tune->items_per_wave > 0) p->pos_ipw = tune->items_per_wave;
    if (tune && tune->xcd_run > 0) p->xcd_run = tune->xcd_run;
    if (tune && tune->stores > 0) p->pos_stores = tune->stores;
    if (tune && tune->waves_per_simd > 0) p->pos_waves = tune->waves_per_simd;
    const int tile_writes = tune ? tune->tile_writes : 0;
    FR_CHECK(ctx, tile_writes >= 0 && tile_writes <= 3, FR_ERR_INVALID, "tile_writes %d out of range", tile_writes);
    FR_CHECK(ctx, (p->pos_stores == 1 || p->pos_stores == 2) && (p->pos_waves == 3 || p->pos_waves == 4),
             FR_ERR_INVALID, "stores %d / waves_per_simd %d out of range", p->pos_stores, p->pos_waves);
    FR_CHECK(ctx, p->pos_ipw <= 64 && p->xcd_run <= (1 << 20), FR_ERR_INVALID, "items_per_wave %d / xcd_run %d "
             "out of range", p->pos_ipw, p->xcd_run);

    fr_status e = fr_plan_thresholds(p, true);
    if (e) return e;

    pgeom g{p->kHW, p->L, p->dim, N, p->frames, geom.bx, geom.by, geom.bz, 0, 0, 0};
    g.nbx = (N + geom.bx - 1) / geom.bx;
    g.nby = (N + geom.by - 1) / geom.by;
    g.nbz = p->frames > 0 ? p->frames : (flat ? 1 : (N + geom.bz - 1) / geom.bz);
    const int64_t nbox = (int64_t)g.nbx * g.nby * g.nbz;
    FR_CHECK(ctx, nbox < (int64_t(1) << 31), FR_ERR_UNSUPPORTED, "%lld unit boxes", (long long)nbox);
    FR_CHECK(ctx, N <= 65535 && g.nbz * geom.bz <= 65535, FR_ERR_UNSUPPORTED, "grid too large for packed unit boxes");

    // positions (entry M: the null sample)
    if ((e = fr_alloc(ctx, (void **)&p->pos, sizeof(float4) * (M + 1)))) return e;
    hipLaunchKernelGGL(k_positions, dim3(nblk(M + 1, 256)), dim3(256), 0, st, s->traj, M, p->dim, p->frames, p->pos);

    // 1. counts
    dbuf<uint32_t> cnt, nrec;
    dbuf<uint64_t> spos;
    if ((e = fr_alloc(ctx, (void **)&cnt.p, 4 * nbox)) || (e = fr_alloc(ctx, (void **)&nrec.p, 4 * M)) ||
        (e = fr_alloc(ctx, (void **)&spos.p, 8 * M)))
        return e;
    FR_HIP(ctx, hipMemsetAsync(cnt.p, 0, 4 * nbox, st));
    hipLaunchKernelGGL(k_box_count, dim3(nblk(M, 256)), dim3(256), 0, st, s->traj, M, g, cnt.p, nrec.p);
    FR_HIP(ctx, hipGetLastError());
    dbuf<char> tmp;
    size_t tmp_bytes = 0;
    if ((e = excl_scan64(ctx, nrec.p, spos.p, M, tmp, tmp_bytes))) return e;
    uint64_t last[2] = {0, 0};
    FR_HIP(ctx, hipMemcpyAsync(&last[0], spos.p + M - 1, 8, hipMemcpyDeviceToHost, st));
    uint32_t lastn = 0;
    FR_HIP(ctx, hipMemcpyAsync(&lastn, nrec.p + M - 1, 4, hipMemcpyDeviceToHost, st));
    FR_HIP(ctx, hipStreamSynchronize(st));
    const uint64_t R = last[0] + lastn;
    FR_CHECK(ctx, R + 4 * (uint64_t)nbox + 64 < (uint64_t(1) << 31), FR_ERR_UNSUPPORTED,
             "plan has %llu records (max about 2^31)", (unsigned long long)R);

    // slice size S (records, a multiple of 16): about 32 slices' worth of
    // work per CU, at least 64 and at most 16384 (the 384^3 FP16 kooshball's
    // 300M records would give 36,600: 16384 grids it 3% faster; the 256^3
    // kooshball's 12,600 stays)
    const int cus = std::max(1, ctx->cu_count);
    uint64_t S = std::min<uint64_t>(16384, std::max<uint64_t>(64, R / ((uint64_t)cus * 32)));
    if (tune && tune->split_taps > 0) S = (uint64_t)tune->split_taps;
    S = std::min<uint64_t>((S + 15) / 16 * 16, uint64_t(1) << 30);
    const uint32_t heavy_cut = (uint32_t)std::max<uint64_t>(4, S / 4);

    // 2. box order
    dbuf<uint64_t> key_in, key_out;
    dbuf<uint32_t> val_in, order;
    if ((e = fr_alloc(ctx, (void **)&key_in.p, 8 * nbox)) || (e = fr_alloc(ctx, (void **)&key_out.p, 8 * nbox)) ||
        (e = fr_alloc(ctx, (void **)&val_in.p, 4 * nbox)) || (e = fr_alloc(ctx, (void **)&order.p, 4 * nbox)))
        return e;
    hipLaunchKernelGGL(k_box_keys, dim3(nblk(nbox, 256)), dim3(256), 0, st, cnt.p, nbox, g, heavy_cut, key_in.p,
                       val_in.p);
    {
        size_t need = 0;
        FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, need, key_in.p, key_out.p, val_in.p, order.p,
                                                       (int)nbox, 0, 64, st));
        dbuf<char> stmp;
        if ((e = fr_alloc(ctx, (void **)&stmp.p, need + 16))) return e;
        FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(stmp.p, need, key_in.p, key_out.p, val_in.p, order.p,
                                                       (int)nbox, 0, 64, st));
        FR_HIP(ctx, hipStreamSynchronize(st));          // before stmp is freed
    }

    // 3. per-unit statistics and their scans
    const int64_t nu = nbox;
    dbuf<uint32_t> rank, c_r, c4_r, nit_r, hs_r, hf_r, fr_r;
    dbuf<uint64_t> Q, P, I, HS, HI, FRs;
    for (auto *b : {&rank, &c_r, &c4_r, &nit_r, &hs_r, &hf_r, &fr_r})
        if ((e = fr_alloc(ctx, (void **)&b->p, 4 * nu))) return e;
    for (auto *b : {&Q, &P, &I, &HS, &HI, &FRs})
        if ((e = fr_alloc(ctx, (void **)&b->p, 8 * nu))) return e;
    hipLaunchKernelGGL(k_unit_stats, dim3(nblk(nu, 256)), dim3(256), 0, st, order.p, nu, cnt.p, (uint32_t)S,
                       heavy_cut, rank.p, c_r.p, c4_r.p, nit_r.p, hs_r.p, hf_r.p, fr_r.p);
    FR_HIP(ctx, hipGetLastError());
    if ((e = excl_scan64(ctx, c_r.p, Q.p, nu, tmp, tmp_bytes)) || (e = excl_scan64(ctx, c4_r.p, P.p, nu, tmp, tmp_bytes)) ||
        (e = excl_scan64(ctx, nit_r.p, I.p, nu, tmp, tmp_bytes)) ||
        (e = excl_scan64(ctx, hs_r.p, HS.p, nu, tmp, tmp_bytes)) ||
        (e = excl_scan64(ctx, hf_r.p, HI.p, nu, tmp, tmp_bytes)) ||
        (e = excl_scan64(ctx, fr_r.p, FRs.p, nu, tmp, tmp_bytes)))
        return e;
    dbuf<totals> tot_d;
    if ((e = fr_alloc(ctx, (void **)&tot_d.p, sizeof(totals)))) return e;
    hipLaunchKernelGGL(k_totals, dim3(1), dim3(64), 0, st, nu, Q.p, P.p, I.p, HS.p, HI.p, FRs.p, c_r.p, c4_r.p,
                       nit_r.p, hs_r.p, hf_r.p, fr_r.p, tot_d.p);
    totals tot{};
    FR_HIP(ctx, hipMemcpyAsync(&tot, tot_d.p, sizeof(totals), hipMemcpyDeviceToHost, st));
    FR_HIP(ctx, hipStreamSynchronize(st));
    FR_CHECK(ctx, tot.R == R, FR_ERR_HIP, "plan: record counts disagree (%llu vs %llu)", (unsigned long long)tot.R,
             (unsigned long long)R);

    // 4. records
    p->nunits = nu;
    p->nrecords = (int64_t)R;
    p->nrecs_padded = (int64_t)tot.R4;
    p->nitems = (int64_t)tot.nitems;
    p->nitems_front = (int64_t)tot.nfront_items;
    p->nheavy = (int64_t)tot.nheavy;
    p->nslots = (int64_t)tot.nslots;
    if ((e = fr_alloc(ctx, (void **)&p->units, sizeof(fr_unit) * nu)) ||
        (e = fr_alloc(ctx, (void **)&p->recs, sizeof(uint32_t) * (tot.R4 + 64))) ||
        (e = fr_alloc(ctx, (void **)&p->items, sizeof(fr_item) * std::max<uint64_t>(1, tot.nitems))) ||
        (e = fr_alloc(ctx, (void **)&p->heavy, sizeof(fr_heavy) * std::max<uint64_t>(1, tot.nheavy))))
        return e;
    hipLaunchKernelGGL(k_units, dim3(nblk(nu, 256)), dim3(256), 0, st, order.p, nu, g, P.p, c_r.p, c4_r.p, p->units);
    if (R > 0) {
        dbuf<uint32_t> keys, vals, keys2, vals2;
        if ((e = fr_alloc(ctx, (void **)&keys.p, 4 * R)) || (e = fr_alloc(ctx, (void **)&vals.p, 4 * R)) ||
            (e = fr_alloc(ctx, (void **)&keys2.p, 4 * R)) || (e = fr_alloc(ctx, (void **)&vals2.p, 4 * R)))
            return e;
        hipLaunchKernelGGL(k_emit, dim3(nblk(M, 256)), dim3(256), 0, st, s->traj, M, g, spos.p, rank.p, keys.p,
                           vals.p);
        int end_bit = 1;
        while (end_bit < 32 && (int64_t(1) << end_bit) < nu) ++end_bit;
        size_t need = 0;
        FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys.p, keys2.p, vals.p, vals2.p, (int)R, 0,
                                                       end_bit, st));
        dbuf<char> stmp;
        if ((e = fr_alloc(ctx, (void **)&stmp.p, need + 16))) return e;
        FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(stmp.p, need, keys.p, keys2.p, vals.p, vals2.p, (int)R, 0,
                                                       end_bit, st));
        hipLaunchKernelGGL(k_place, dim3(nblk((int64_t)R, 256)), dim3(256), 0, st, keys2.p, vals2.p, (int64_t)R, P.p,
                           Q.p, p->recs);
        FR_HIP(ctx, hipGetLastError());
        FR_HIP(ctx, hipStreamSynchronize(st));          // before the sort buffers are freed
    }
    hipLaunchKernelGGL(k_pad, dim3(nblk(std::max<int64_t>(nu, 64), 256)), dim3(256), 0, st, nu, P.p, c_r.p, c4_r.p,
                       tot.R4, (uint32_t)M, p->recs);

    // 5. items and heavy units
    hipLaunchKernelGGL(k_items, dim3(nblk(nu, 256)), dim3(256), 0, st, nu, (uint32_t)S, P.p, I.p, HS.p, HI.p,
                       c4_r.p, nit_r.p, p->units, p->items, p->heavy);
    FR_HIP(ctx, hipGetLastError());

    // 6. tiles of TW x-adjacent whole units (unit box of 16 cells with
    // power-of-two sides; the grid divides into whole tiles): pairs by
    // default (tile_writes 0 or 3), rows of 2 bx cells; fours with
    // tile_writes 1, rows of 4 bx cells; none with tile_writes 2
    const bool pow2 = (geom.bx & (geom.bx - 1)) == 0 && (geom.by & (geom.by - 1)) == 0 &&
                      (geom.bz & (geom.bz - 1)) == 0;
    const int TW = tile_writes == 1 ? FR_POS_WAVES : (tile_writes == 2 ? 0 : 2);
    p->ntiles = 0;
    p->tile_units = 0;
    if (TW > 0 && pow2 && geom.cells_max == 16 && N % (TW * geom.bx) == 0 && N % geom.by == 0 &&
        (flat || N % geom.bz == 0) && nu - (int64_t)tot.nfront_units > 0) {
        const int ntx = g.nbx / TW;
        const int64_t ntiles = (int64_t)ntx * g.nby * g.nbz;
        dbuf<uint64_t> tk, tk2;
        dbuf<uint32_t> tv, torder, tpos;
        if ((e = fr_alloc(ctx, (void **)&tk.p, 8 * ntiles)) || (e = fr_alloc(ctx, (void **)&tk2.p, 8 * ntiles)) ||
            (e = fr_alloc(ctx, (void **)&tv.p, 4 * ntiles)) || (e = fr_alloc(ctx, (void **)&torder.p, 4 * ntiles)) ||
            (e = fr_alloc(ctx, (void **)&tpos.p, 4 * ntiles)) ||
            (e = fr_alloc(ctx, (void **)&p->tiles, sizeof(fr_tile) * ntiles)))
            return e;
        hipLaunchKernelGGL(k_tile_keys, dim3(nblk(ntiles, 256)), dim3(256), 0, st, ntiles, ntx, g.nby, p->frames,
                           tk.p, tv.p);
        size_t need = 0;
        FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, need, tk.p, tk2.p, tv.p, torder.p, (int)ntiles, 0, 64,
                                                       st));
        dbuf<char> stmp;
        if ((e = fr_alloc(ctx, (void **)&stmp.p, need + 16))) return e;
        FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(stmp.p, need, tk.p, tk2.p, tv.p, torder.p, (int)ntiles, 0, 64,
                                                       st));
        hipLaunchKernelGGL(k_tile_init, dim3(nblk(ntiles, 256)), dim3(256), 0, st, torder.p, ntiles, ntx, g.nby, g,
                           TW, p->tiles, tpos.p);
        const int64_t nlight = nu - (int64_t)tot.nfront_units;
        hipLaunchKernelGGL(k_tile_fill, dim3(nblk(nlight, 256)), dim3(256), 0, st, order.p, nu,
                           (int64_t)tot.nfront_units, g, ntx, TW, tpos.p, I.p, p->tiles);
        FR_HIP(ctx, hipGetLastError());
        FR_HIP(ctx, hipStreamSynchronize(st));
        p->ntiles = ntiles;
        p->tile_units = TW;
    }
    FR_HIP(ctx, hipStreamSynchronize(st));
    p->max_unit_taps = 0;
    p->ntaps = 0;
    p->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FR_OK;
}
