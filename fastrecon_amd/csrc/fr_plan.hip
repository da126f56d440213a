// Gridding plan: adaptive work units and their record streams.
//
// Replaces cuGridLut::plan (ReconData/cuGridLut.cu:151-226), which builds a
// (cell, {sample, distance}) tuple for every box cell of every sample, sorts
// it by cell and streams 25 partitions of it over PCIe for every coil.
//
// Here the grid is covered by "units", boxes of cells that one wave owns
// (its accumulator lives in that wave's slice of LDS).  Units start as base
// boxes sized to the LDS budget and are split in halves (a k-d split driven
// by a per-cell tap histogram) until no unit carries more than a threshold
// number of taps, so the 1/|k|^2 density peak of a radial trajectory is
// spread over many waves; units are ordered heaviest first.  For every unit
// the plan holds its records, in ascending sample order:
//     H[r] = (sample, n),  T[t .. t+n) = (LDS byte offset of the cell's
//                                          accumulator row, LUT weight)
// Records come from a stable radix sort of (unit, sample) pairs generated in
// sample order, so each cell later receives its contributions in the
// reference CPU order (GridLut.cpp:96-145).  Tap sets and LUT indices use the
// reference arithmetic (fr_taps.h).  The plan depends only on the
// trajectory, lives in HBM and serves every coil, frame and DCF.
#include "fr_internal.h"
#include "fr_taps.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <numeric>
#include <cmath>
#include <cstring>

namespace {

struct gctx {
    float hw;
    int L;
    int dim;
    int N;
    int frames;    // batched 2D frames along z (0: one grid)
};

__device__ __forceinline__ size_t cell_index(int x, int y, int z, int N)
{
    return ((size_t)z * N + y) * N + x;
}

// per-cell histogram of accepted taps
__global__ void k_count_cells(const float *__restrict__ traj, int64_t M, gctx g, uint32_t *__restrict__ cnt)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    fr_box b;
    fr_sample_box(g.dim, g.N, traj, M, s, g.hw, b, g.frames);
    fr_box_taps(b, b.s, b.e, g.hw, g.L, [&](int x, int y, int z, int) {
        atomicAdd(&cnt[cell_index(x, y, z, g.N)], 1u);
    });
}

// Visit the units a sample's box overlaps: a unit is met for the first time
// at the first cell of (box ∩ unit) in z, y, x order, i.e. where each
// coordinate equals max(box start, unit origin).  f(unit, lo, hi) gets the
// sub-box box ∩ unit.
template <typename F>
__device__ __forceinline__ void for_each_unit(const fr_box &b, int N, const int32_t *__restrict__ cell_unit,
                                              const fr_unit *__restrict__ units, F &&f)
{
    for (int z = b.s[2]; z <= b.e[2]; ++z)
        for (int y = b.s[1]; y <= b.e[1]; ++y)
            for (int x = b.s[0]; x <= b.e[0]; ++x) {
                const int u = cell_unit[cell_index(x, y, z, N)];
                const fr_unit U = units[u];
                const int ux = U.ext & 255, uy = (U.ext >> 8) & 255, uz = U.ext >> 16;
                if (x != max(b.s[0], U.ox) || y != max(b.s[1], U.oy) || z != max(b.s[2], U.oz)) continue;
                const int lo[3] = {x, y, z};
                const int hi[3] = {min(b.e[0], U.ox + ux - 1), min(b.e[1], U.oy + uy - 1), min(b.e[2], U.oz + uz - 1)};
                f(u, U, lo, hi);
            }
}

__global__ void k_count_records(const float *__restrict__ traj, int64_t M, gctx g,
                                const int32_t *__restrict__ cell_unit, const fr_unit *__restrict__ units,
                                uint32_t *__restrict__ nrec)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    fr_box b;
    fr_sample_box(g.dim, g.N, traj, M, s, g.hw, b, g.frames);
    uint32_t r = 0;
    if (b.s[0] <= b.e[0] && b.s[1] <= b.e[1] && b.s[2] <= b.e[2])
        for_each_unit(b, g.N, cell_unit, units, [&](int, const fr_unit &, const int *lo, const int *hi) {
            uint32_t n = 0;
            fr_box_taps(b, lo, hi, g.hw, g.L, [&](int, int, int, int) { ++n; });
            r += n ? 1u : 0u;
        });
    nrec[s] = r;
}

__global__ void k_emit_records(const float *__restrict__ traj, int64_t M, gctx g,
                               const int32_t *__restrict__ cell_unit, const fr_unit *__restrict__ units,
                               const uint64_t *__restrict__ pos, uint32_t *__restrict__ keys,
                               uint64_t *__restrict__ vals)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    fr_box b;
    fr_sample_box(g.dim, g.N, traj, M, s, g.hw, b, g.frames);
    uint64_t p = pos[s];
    if (b.s[0] <= b.e[0] && b.s[1] <= b.e[1] && b.s[2] <= b.e[2])
        for_each_unit(b, g.N, cell_unit, units, [&](int u, const fr_unit &, const int *lo, const int *hi) {
            uint32_t n = 0;
            fr_box_taps(b, lo, hi, g.hw, g.L, [&](int, int, int, int) { ++n; });
            if (n) {
                keys[p] = (uint32_t)u;
                vals[p] = ((uint64_t)s << 32) | n;
                ++p;
            }
        });
}

// Records are padded to a multiple of FR_TAP_GROUP taps with dummy taps
// (the spare accumulator row, weight 0), so the kernel applies taps in
// fixed-size groups without remainder handling.
__host__ __device__ __forceinline__ uint32_t padded_taps(uint32_t n)
{
    return (n + FR_TAP_GROUP - 1) / FR_TAP_GROUP * FR_TAP_GROUP;
}

__global__ void k_record_taps(const uint64_t *__restrict__ vals, int64_t R, uint64_t *__restrict__ n)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) n[r] = padded_taps((uint32_t)(vals[r] & 0xffffffffu));
}

__global__ void k_unit_ranges(const uint32_t *__restrict__ keys, const uint64_t *__restrict__ toff, int64_t R,
                              fr_unit *__restrict__ units)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const uint32_t k = keys[r];
    if (r == 0 || keys[r - 1] != k) {
        units[k].hdr = (uint64_t)r;
        units[k].tap = toff[r];
    }
    if (r == R - 1 || keys[r + 1] != k) {
        // the first record of this unit was written by another thread; the
        // counts are finished by k_unit_counts once all ranges are known
        units[k].nrec = (uint32_t)(r + 1);          // provisional: one past the last record
        units[k].ntap = (uint32_t)0;
    }
}

__global__ void k_unit_counts(fr_unit *__restrict__ units, int64_t nunits, const uint64_t *__restrict__ toff)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    fr_unit U = units[u];
    if (U.nrec == 0) return;                         // no records
    const uint64_t end = U.nrec;                     // one past the last record
    U.nrec = (uint32_t)(end - U.hdr);
    U.ntap = (uint32_t)(toff[end] - U.tap);
    units[u] = U;
}

__global__ void k_fill_records(const float *__restrict__ traj, int64_t M, gctx g,
                               const int32_t *__restrict__ cell_unit, const fr_unit *__restrict__ units,
                               const uint32_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                               const uint64_t *__restrict__ toff, int64_t R, const float *__restrict__ lut,
                               uint32_t tap_scale, uint32_t pad_x, uint2 *__restrict__ H, uint2 *__restrict__ T,
                               uint16_t *__restrict__ T16)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const uint32_t key = keys[r];
    const int64_t s = (int64_t)(vals[r] >> 32);
    const uint32_t n = (uint32_t)(vals[r] & 0xffffffffu);
    const uint32_t np = padded_taps(n);
    H[r] = make_uint2((uint32_t)s, np);
    if (!T && !T16) return;                     // records only
    fr_box b;
    fr_sample_box(g.dim, g.N, traj, M, s, g.hw, b, g.frames);
    uint32_t i = 0;
    if (T16) {
        // compact taps: cell | LUT index << 6; padding: cell pad_x, index L (a zero weight)
        uint16_t *w = T16 + toff[r];
        for_each_unit(b, g.N, cell_unit, units, [&](int u, const fr_unit &U, const int *lo, const int *hi) {
            if ((uint32_t)u != key) return;
            const int ux = U.ext & 255, uy = (U.ext >> 8) & 255;
            fr_box_taps(b, lo, hi, g.hw, g.L, [&](int x, int y, int z, int ki) {
                const uint32_t cell = (uint32_t)(((z - U.oz) * uy + (y - U.oy)) * ux + (x - U.ox));
                w[i++] = (uint16_t)(cell | ((uint32_t)ki << 6));
            });
        });
        for (; i < np; ++i) w[i] = (uint16_t)(pad_x | ((uint32_t)g.L << 6));
        return;
    }
    uint2 *w = T + toff[r];
    for_each_unit(b, g.N, cell_unit, units, [&](int u, const fr_unit &U, const int *lo, const int *hi) {
        if ((uint32_t)u != key) return;
        const int ux = U.ext & 255, uy = (U.ext >> 8) & 255;
        fr_box_taps(b, lo, hi, g.hw, g.L, [&](int x, int y, int z, int ki) {
            const uint32_t cell = (uint32_t)(((z - U.oz) * uy + (y - U.oy)) * ux + (x - U.ox));
            w[i++] = make_uint2(cell * tap_scale, __float_as_uint(lut[ki]));
        });
    });
    for (; i < np; ++i) w[i] = make_uint2(pad_x, 0u);
}

// the tap array's tail (past the last record's taps) holds padding taps
template <typename TT>
__global__ void k_pad_tail(TT *__restrict__ T, uint64_t from, int64_t n, TT pad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) T[from + i] = pad;
}

// first tap of each record as 32 bits, and the tap count past the last record
// (the headers' padding)
__global__ void k_tap_starts(const uint64_t *__restrict__ toff, int64_t R, int64_t n, uint32_t *__restrict__ tstart)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) tstart[r] = (uint32_t)toff[r < R ? r : R];
}

// A record opens a new work item when it is its unit's first record or when
// the unit's running tap count crosses a multiple of the slice size.
__global__ void k_item_flags(const uint32_t *__restrict__ keys, const uint64_t *__restrict__ toff, int64_t R,
                             const fr_unit *__restrict__ units, uint64_t slice, uint32_t *__restrict__ flag)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const fr_unit U = units[keys[r]];
    uint32_t f = (uint64_t)r == U.hdr;
    if (!f) f = (toff[r] - U.tap) / slice != (toff[r - 1] - U.tap) / slice;
    flag[r] = f;
}

__global__ void k_make_items(const uint32_t *__restrict__ keys, const uint64_t *__restrict__ toff,
                             const uint32_t *__restrict__ flag, const uint32_t *__restrict__ idx, int64_t R,
                             const fr_unit *__restrict__ units, fr_item *__restrict__ items,
                             uint32_t *__restrict__ first_item, uint32_t *__restrict__ last_item)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R || !flag[r]) return;
    const uint32_t u = keys[r];
    const fr_unit U = units[u];
    const uint32_t i = idx[r] - 1;                    // inclusive scan of flags
    // the item ends at the next flagged record of the same unit or the unit end
    uint64_t r1 = U.hdr + U.nrec;
    for (uint64_t q = (uint64_t)r + 1; q < U.hdr + U.nrec; ++q)
        if (flag[q]) { r1 = q; break; }
    items[i] = fr_item{u, -1, (uint32_t)r, (uint32_t)r1, (uint32_t)toff[r]};
    if ((uint64_t)r == U.hdr) first_item[u] = i;
    if (r1 == U.hdr + U.nrec) last_item[u] = i;
}

// empty units still own cells: give each one item that writes zeros
__global__ void k_empty_items(const fr_unit *__restrict__ units, int64_t nunits, const uint32_t *__restrict__ eidx,
                              uint32_t base, fr_item *__restrict__ items)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits || units[u].nrec != 0) return;
    items[base + eidx[u] - 1] = fr_item{(uint32_t)u, -1, 0u, 0u, 0u};
}

__global__ void k_empty_flags(const fr_unit *__restrict__ units, int64_t nunits, uint32_t *__restrict__ f)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nunits) f[u] = units[u].nrec == 0;
}

__global__ void k_heavy_flags(const fr_unit *__restrict__ units, int64_t nunits, const uint32_t *__restrict__ first,
                              const uint32_t *__restrict__ last, uint32_t *__restrict__ f, uint32_t *__restrict__ nsl)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    const bool h = units[u].nrec != 0 && last[u] != first[u];
    f[u] = h;
    nsl[u] = h ? last[u] - first[u] + 1 : 0;
}

__global__ void k_make_heavy(int64_t nunits, const uint32_t *__restrict__ hflag, const uint32_t *__restrict__ hidx,
                             const uint32_t *__restrict__ nsl, const uint32_t *__restrict__ slot_scan,
                             const uint32_t *__restrict__ first, fr_heavy *__restrict__ heavy,
                             fr_item *__restrict__ items)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits || !hflag[u]) return;
    const uint32_t h = hidx[u] - 1;
    const uint32_t slot0 = slot_scan[u] - nsl[u];       // exclusive
    heavy[h] = fr_heavy{(uint32_t)u, nsl[u], slot0, 0};
    for (uint32_t k = 0; k < nsl[u]; ++k) items[first[u] + k].slot = (int32_t)(slot0 + k);
}

// ---- gather mode: batches of FR_GATHER_BATCH records with cell-sorted taps
__global__ void k_item_nbatch(const fr_item *__restrict__ items, int64_t nitems, uint32_t *__restrict__ nb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    nb[i] = (items[i].r1 - items[i].r0 + FR_GATHER_BATCH - 1) / FR_GATHER_BATCH;
}

__global__ void k_make_batches(fr_item *__restrict__ items, int64_t nitems, const uint32_t *__restrict__ nb,
                               const uint32_t *__restrict__ scan, fr_batch *__restrict__ batches)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitems) return;
    const uint32_t b0 = scan[i] - nb[i];            // inclusive scan
    const fr_item it = items[i];
    for (uint32_t b = 0; b < nb[i]; ++b) {
        fr_batch B{};
        B.r0 = it.r0 + b * FR_GATHER_BATCH;
        B.nrec = min((uint32_t)FR_GATHER_BATCH, it.r1 - B.r0);
        batches[b0 + b] = B;
    }
    items[i].t0 = b0;
}

// One wave per batch; lane c < FR_GATHER_CELLS counts the batch's taps of
// cell c (at most one per record) and pads the count to a multiple of 4.
__global__ void k_batch_count(fr_batch *__restrict__ batches, int64_t nbat, const uint2 *__restrict__ H,
                              const uint64_t *__restrict__ toff, const uint2 *__restrict__ T,
                              uint64_t *__restrict__ bsize)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= nbat) return;
    const fr_batch B = batches[w];
    uint32_t n = 0;
    for (uint32_t j = 0; j < B.nrec; ++j) {
        const uint64_t r = B.r0 + j;
        const uint64_t t0 = toff[r];
        const uint32_t np = H[r].y;
        for (uint32_t t = 0; t < np; ++t) n += T[t0 + t].x == (uint32_t)lane;
    }
    const uint32_t p = lane < FR_GATHER_CELLS ? (n + 3u) & ~3u : 0u;
    if (lane < FR_GATHER_CELLS) reinterpret_cast<uint8_t *>(batches[w].cnt)[lane] = (uint8_t)p;
    uint32_t sum = p;
    for (int m = 32; m >= 1; m >>= 1) sum += __shfl_xor(sum, m);
    if (lane == 0) bsize[w] = sum;
}

__global__ void k_batch_write(fr_batch *__restrict__ batches, int64_t nbat, const uint2 *__restrict__ H,
                              const uint64_t *__restrict__ toff, const uint2 *__restrict__ T,
                              const uint64_t *__restrict__ bstart, uint2 *__restrict__ T2)
{
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= nbat) return;
    const fr_batch B = batches[w];
    const uint64_t base = bstart[w] - [&] {                    // inclusive scan -> exclusive start
        uint32_t s = 0;
        for (int k = 0; k < FR_GATHER_CELLS; ++k) s += reinterpret_cast<const uint8_t *>(B.cnt)[k];
        return (uint64_t)s;
    }();
    if (lane == 0) batches[w].toff = base;
    if (lane >= FR_GATHER_CELLS) return;
    uint64_t o = base;
    for (int k = 0; k < lane; ++k) o += reinterpret_cast<const uint8_t *>(B.cnt)[k];
    const uint64_t end = o + reinterpret_cast<const uint8_t *>(B.cnt)[lane];
    for (uint32_t j = 0; j < B.nrec; ++j) {
        const uint64_t r = B.r0 + j;
        const uint64_t t0 = toff[r];
        const uint32_t np = H[r].y;
        for (uint32_t t = 0; t < np; ++t) {
            const uint2 a = T[t0 + t];
            if (a.x == (uint32_t)lane) T2[o++] = make_uint2(j * 256u, a.y);
        }
    }
    for (; o < end; ++o) T2[o] = make_uint2(FR_GATHER_BATCH * 256u, 0u);
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename T>
struct dbuf {
    T *p = nullptr;
    ~dbuf() { if (p) (void)hipFree(p); }
};

// The reference's LUT index of a squared distance (GridLut.cpp:112-115):
// dk = sqrtf(d2), roundf(dk / hw * (L - 1)), in IEEE FP32 on the host.
int ref_lut_index(float d2, float hw, int L)
{
    const float dk = sqrtf(d2);
    const float q = dk / hw;
    const float r = q * (float)(L - 1);
    return (int)roundf(r);
}

// Least non-negative float x with pred(x), by bisection on the bit pattern
// (non-negative floats order like their bits); pred is monotone.
template <typename P>
float least_float(uint32_t hi_bits, P &&pred)
{
    uint32_t lo = 0, hi = hi_bits;               // pred(hi) holds
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        float x;
        std::memcpy(&x, &mid, 4);
        if (pred(x)) hi = mid; else lo = mid + 1;
    }
    float x;
    std::memcpy(&x, &lo, 4);
    return x;
}

int floor_pow2(int v)
{
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}

}  // namespace

// Base unit box: 32 cells (4 x 4 x 2 in 3D, 8 x 4 in 2D).  A unit never
// exceeds 64 cells, so a record (one sample's taps inside one unit) never
// carries more than 64 taps.  Smaller units cost more records but use less
// LDS per wave, so more waves per CU hide the LDS and memory latency; on
// MI355X 16-32 cells per unit measured best for the 32-coil kooshball.
//
// The matrix-core mode (FR_ACC_MFMA) pays per (record, 16-cell block) and
// writes rows of ux cells per coil: its base unit is one x-elongated block,
// 8 x 2 x 1 in 3D and 8 x 2 in 2D (of the 16-64-cell shapes measured on
// MI355X, 8 x 2 x 1 was fastest for the 32-coil 256^3 kooshball).
static fr_geom choose_geom(int dim, unsigned N, int coils, int accumulate, const fr_tune *tune)
{
    fr_geom g{};
    g.dim = dim;
    g.N = (int)N;
    if (accumulate == FR_ACC_MFMA) {
        // Default unit by coil count (matrix-core kernel, measured on the 3D
        // kooshball at 256^3 and 384^3): with few coils a record's matrix
        // work is small next to its fixed per-record cost, so bigger units
        // (fewer records per sample) win; at 32 coils the smallest x-elongated
        // unit does (1 coil: 4x4x2 2.6 ms vs 8x2x1 3.7; 8: 8x2x2 3.2 vs 4.1;
        // 12: 8x2x2 3.7 vs 4.3; 32: 8x2x1 5.9 vs 8x2x2 7.1).  In 2D, 8x4 up
        // to 16 coils (the 12-coil spiral series: 1.16 vs 1.36 ms).
        if (dim == 3) {
            g.bx = coils <= 2 ? 4 : 8;
            g.by = coils <= 2 ? 4 : 2;
            g.bz = coils <= 16 ? 2 : 1;
        } else {
            g.bx = 8;
            g.by = coils <= 16 ? 4 : 2;
            g.bz = 1;
        }
    } else {
        g.bx = dim == 3 ? 4 : 8;
        g.by = 4;
        g.bz = dim == 3 ? 2 : 1;
    }
    g.waves = 1;
    if (tune) {
        if (tune->tile_x > 0) g.bx = tune->tile_x;
        if (tune->tile_y > 0) g.by = tune->tile_y;
        if (dim == 3 && tune->tile_z > 0) g.bz = tune->tile_z;
    }
    g.bx = std::min(g.bx, (int)N);
    g.by = std::min(g.by, (int)N);
    g.bz = dim == 3 ? std::min(g.bz, (int)N) : 1;
    g.cells_max = g.bx * g.by * g.bz;
    return g;
}

namespace {

struct host_unit {
    int ox, oy, oz, ux, uy, uz;
    uint64_t taps;
};

struct unit_builder {
    const std::vector<uint32_t> &cnt;
    int N, dim;
    uint64_t thr;
    std::vector<host_unit> out;

    uint64_t sum(int ox, int oy, int oz, int ux, int uy, int uz) const
    {
        uint64_t t = 0;
        for (int z = oz; z < oz + uz; ++z)
            for (int y = oy; y < oy + uy; ++y) {
                const uint32_t *row = cnt.data() + ((size_t)z * N + y) * N + ox;
                for (int x = 0; x < ux; ++x) t += row[x];
            }
        return t;
    }

    void split(host_unit b)
    {
        if (b.taps <= thr || b.ux * b.uy * b.uz == 1) {
            out.push_back(b);
            return;
        }
        host_unit a = b, c = b;
        if (b.uz >= b.uy && b.uz >= b.ux && b.uz > 1) {
            a.uz = b.uz / 2; c.oz = b.oz + a.uz; c.uz = b.uz - a.uz;
        } else if (b.uy >= b.ux && b.uy > 1) {
            a.uy = b.uy / 2; c.oy = b.oy + a.uy; c.uy = b.uy - a.uy;
        } else {
            a.ux = b.ux / 2; c.ox = b.ox + a.ux; c.ux = b.ux - a.ux;
        }
        a.taps = sum(a.ox, a.oy, a.oz, a.ux, a.uy, a.uz);
        c.taps = b.taps - a.taps;
        split(a);
        split(c);
    }
};

}  // namespace

fr_status fr_plan_build(fr_plan *p, const fr_samples *s, const fr_tune *tune)
{
    fr_ctx *ctx = p->ctx;
    hipStream_t st = ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    // automatic accumulation: the matrix cores, or LDS read-add-write for an
    // exact-order plan (the only mode that sums in the CPU loop's order)
    int acc_mode = tune ? tune->accumulate : 0;
    const bool exact = tune && tune->exact_order;
    if (acc_mode == 0) acc_mode = exact ? FR_ACC_LDS_RMW : FR_ACC_MFMA;
    fr_geom geom = choose_geom(p->dim, p->N, s->C, acc_mode, tune);
    const int stride = 2 * std::min(s->C, 32) + 1;
    FR_CHECK(ctx, geom.bx >= 1 && geom.by >= 1 && geom.bz >= 1 && geom.bx <= 255 && geom.by <= 255 && geom.bz <= 255,
             FR_ERR_INVALID, "unit box %dx%dx%d out of range", geom.bx, geom.by, geom.bz);
    FR_CHECK(ctx, geom.cells_max <= 64, FR_ERR_INVALID, "unit box %dx%dx%d exceeds 64 cells", geom.bx, geom.by,
             geom.bz);
    (void)stride;
    p->geom = geom;
    p->accumulate = acc_mode;
    FR_CHECK(ctx, p->accumulate >= 0 && p->accumulate <= FR_ACC_MFMA, FR_ERR_INVALID, "unknown accumulate mode %d",
             p->accumulate);
    FR_CHECK(ctx, p->accumulate != FR_ACC_MFMA || !(tune && tune->exact_order), FR_ERR_INVALID,
             "exact_order needs a sequential accumulation mode (1, 2 or 4); the matrix cores sum four products at "
             "a time");
    FR_CHECK(ctx, p->accumulate != FR_ACC_GATHER || geom.cells_max <= FR_GATHER_CELLS, FR_ERR_INVALID,
             "gather accumulation needs a unit box of at most %d cells (have %dx%dx%d)", FR_GATHER_CELLS, geom.bx,
             geom.by, geom.bz);
    // VGPR accumulation (GPR-index addressed adds) gave wrong sums and a
    // memory fault on gfx950 whose cause is not established: it is refused.
    FR_CHECK(ctx, p->accumulate != FR_ACC_REGS, FR_ERR_UNSUPPORTED,
             "register accumulation (accumulate = 3) is disabled in this build");
    p->exact_order = exact ? 1 : 0;

    const int N = (int)p->N;
    const size_t cells = fr_plan_cells(p->dim, p->N, p->frames);
    const int64_t M = s->M;
    gctx g{p->kHW, p->L, p->dim, N, p->frames};
    fr_status e;

    if (p->L <= FR_MFMA_MAX_L && (e = fr_plan_thresholds(p, false))) return e;

    // 1. per-cell tap histogram
    dbuf<uint32_t> cnt_d;
    if ((e = fr_alloc(ctx, (void **)&cnt_d.p, 4 * cells))) return e;
    FR_HIP(ctx, hipMemsetAsync(cnt_d.p, 0, 4 * cells, st));
    hipLaunchKernelGGL(k_count_cells, dim3(nblk(M, 256)), dim3(256), 0, st, s->traj, M, g, cnt_d.p);
    FR_HIP(ctx, hipGetLastError());
    std::vector<uint32_t> cnt(cells);
    FR_HIP(ctx, hipMemcpyAsync(cnt.data(), cnt_d.p, 4 * cells, hipMemcpyDeviceToHost, st));
    FR_HIP(ctx, hipStreamSynchronize(st));

    // 2. units: base boxes split until each carries at most `thr` taps
    uint64_t total = 0;
    for (uint32_t c : cnt) total += c;
    const int cus = std::max(1, ctx->cu_count);
    uint64_t thr = std::max<uint64_t>(256, total / ((uint64_t)cus * 32));
    if (tune && tune->split_taps > 0) thr = (uint64_t)tune->split_taps;
    // exact order: heavy units are split in space (each cell keeps one wave
    // and the CPU summation order); otherwise they are split by sample range
    const uint64_t space_thr = p->exact_order ? thr : ~uint64_t(0);
    const uint64_t slice_thr = p->exact_order ? ~uint64_t(0) : thr;
    unit_builder ub{cnt, N, p->dim, space_thr, {}};
    const int nz = p->frames > 0 ? p->frames : (p->dim == 3 ? N : 1);
    for (int oz = 0; oz < nz; oz += geom.bz)
        for (int oy = 0; oy < N; oy += geom.by)
            for (int ox = 0; ox < N; ox += geom.bx) {
                host_unit b{ox, oy, oz, std::min(geom.bx, N - ox), std::min(geom.by, N - oy),
                            std::min(geom.bz, nz - oz), 0};
                b.taps = ub.sum(b.ox, b.oy, b.oz, b.ux, b.uy, b.uz);
                ub.split(b);
            }
    std::vector<host_unit> &hu = ub.out;
    // Launch order: units heavier than a quarter of the split threshold first
    // (longest first), then the rest along a Morton curve of their origins so
    // that items running at the same time share samples in cache.
    const uint64_t heavy_cut = thr / 4;
    // (batched frames share no samples: frame-major, Morton within a frame)
    auto morton = [&](const host_unit &u) {
        const uint32_t a = (uint32_t)u.ox / geom.bx, b = (uint32_t)u.oy / geom.by, c = (uint32_t)u.oz / geom.bz;
        uint64_t m = 0;
        if (p->frames > 0) {
            for (int bit = 0; bit < 21; ++bit)
                m |= ((uint64_t)((a >> bit) & 1) << (2 * bit)) | ((uint64_t)((b >> bit) & 1) << (2 * bit + 1));
            return ((uint64_t)c << 42) | m;
        }
        for (int bit = 0; bit < 21; ++bit)
            m |= ((uint64_t)((a >> bit) & 1) << (3 * bit)) | ((uint64_t)((b >> bit) & 1) << (3 * bit + 1)) |
                 ((uint64_t)((c >> bit) & 1) << (3 * bit + 2));
        return m;
    };
    std::stable_sort(hu.begin(), hu.end(), [&](const host_unit &a, const host_unit &b) {
        const bool ha = a.taps > heavy_cut, hb = b.taps > heavy_cut;
        if (ha != hb) return ha;
        if (ha) return a.taps > b.taps;
        return morton(a) < morton(b);
    });
    int64_t nheavy_units = 0;
    while (nheavy_units < (int64_t)hu.size() && hu[nheavy_units].taps > heavy_cut) ++nheavy_units;
    const int64_t nunits = (int64_t)hu.size();
    std::vector<fr_unit> units(nunits);
    std::vector<int32_t> cell_unit(cells);
    for (int64_t u = 0; u < nunits; ++u) {
        const host_unit &h = hu[u];
        units[u] = fr_unit{h.ox, h.oy, h.oz, h.ux | (h.uy << 8) | (h.uz << 16), 0, 0, 0, 0};
        for (int z = h.oz; z < h.oz + h.uz; ++z)
            for (int y = h.oy; y < h.oy + h.uy; ++y)
                for (int x = h.ox; x < h.ox + h.ux; ++x) cell_unit[((size_t)z * N + y) * N + x] = (int32_t)u;
    }
    p->max_unit_taps = nunits ? (int64_t)hu[0].taps : 0;
    dbuf<int32_t> cu_d;
    if ((e = fr_alloc(ctx, (void **)&cu_d.p, 4 * cells))) return e;
    if ((e = fr_alloc(ctx, (void **)&p->units, sizeof(fr_unit) * std::max<int64_t>(1, nunits)))) return e;
    FR_HIP(ctx, hipMemcpyAsync(cu_d.p, cell_unit.data(), 4 * cells, hipMemcpyHostToDevice, st));
    FR_HIP(ctx, hipMemcpyAsync(p->units, units.data(), sizeof(fr_unit) * nunits, hipMemcpyHostToDevice, st));
    p->nunits = nunits;

    // 3. records: (unit, sample) pairs in sample order, stable-sorted by unit
    dbuf<uint32_t> nrec;
    dbuf<uint64_t> pos;
    if ((e = fr_alloc(ctx, (void **)&nrec.p, 4 * M)) || (e = fr_alloc(ctx, (void **)&pos.p, 8 * (M + 1)))) return e;
    hipLaunchKernelGGL(k_count_records, dim3(nblk(M, 256)), dim3(256), 0, st, s->traj, M, g, cu_d.p, p->units,
                       nrec.p);
    FR_HIP(ctx, hipGetLastError());
    auto rec64 = hipcub::TransformInputIterator<uint64_t, hipcub::CastOp<uint64_t>, const uint32_t *>(
        nrec.p, hipcub::CastOp<uint64_t>());
    size_t tmp_bytes = 0;
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
    hipLaunchKernelGGL(k_emit_records, dim3(nblk(M, 256)), dim3(256), 0, st, s->traj, M, g, cu_d.p, p->units,
                       pos.p, keys_in.p, vals_in.p);
    FR_HIP(ctx, hipGetLastError());
    int end_bit = 1;
    while (end_bit < 32 && (int64_t(1) << end_bit) < nunits) ++end_bit;
    size_t sort_bytes = 0;
    FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys_in.p, keys_out.p, vals_in.p,
                                                   vals_out.p, (int)R, 0, end_bit, st));
    dbuf<char> sort_tmp;
    if ((e = fr_alloc(ctx, (void **)&sort_tmp.p, sort_bytes + 16))) return e;
    FR_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(sort_tmp.p, sort_bytes, keys_in.p, keys_out.p, vals_in.p,
                                                   vals_out.p, (int)R, 0, end_bit, st));

    // 4. tap offsets (exclusive scan of per-record tap counts)
    dbuf<uint64_t> rn, toff;
    if ((e = fr_alloc(ctx, (void **)&rn.p, 8 * R + 8)) || (e = fr_alloc(ctx, (void **)&toff.p, 8 * R + 16)))
        return e;
    if (R) hipLaunchKernelGGL(k_record_taps, dim3(nblk(R, 256)), dim3(256), 0, st, vals_out.p, (int64_t)R, rn.p);
    size_t scan_bytes = 0;
    FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, rn.p, toff.p + 1, (int)R, st));
    dbuf<char> scan_tmp;
    if ((e = fr_alloc(ctx, (void **)&scan_tmp.p, scan_bytes + 16))) return e;
    FR_HIP(ctx, hipMemsetAsync(toff.p, 0, 8, st));
    if (R) FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(scan_tmp.p, scan_bytes, rn.p, toff.p + 1, (int)R, st));
    uint64_t TT = 0;
    FR_HIP(ctx, hipMemcpyAsync(&TT, toff.p + R, 8, hipMemcpyDeviceToHost, st));
    FR_HIP(ctx, hipStreamSynchronize(st));
    FR_CHECK(ctx, TT + 1024 < (uint64_t(1) << 32), FR_ERR_UNSUPPORTED, "plan has %llu taps (max 2^32 - 1024)",
             (unsigned long long)TT);

    // 5. unit record/tap ranges, then the records themselves.  Matrix-core
    // plans store compact 2-byte taps (cell, LUT index) when they fit: cells
    // < 64 and L <= FR_CT_MAX_L (index L is the padding taps' zero weight).
    p->compact_taps = p->accumulate == FR_ACC_MFMA && geom.cells_max <= 63 && p->L <= FR_CT_MAX_L;
    const size_t tap_bytes = p->compact_taps ? sizeof(uint16_t) : sizeof(uint2);
    const uint64_t tap_alloc = TT + 576;
    if ((e = fr_alloc(ctx, (void **)&p->hdrs, sizeof(uint2) * (R + 64)))) return e;
    if ((e = fr_alloc(ctx, (void **)&p->tstart, sizeof(uint32_t) * (R + 64)))) return e;
    hipLaunchKernelGGL(k_tap_starts, dim3(nblk(R + 64, 256)), dim3(256), 0, st, toff.p, (int64_t)R, (int64_t)R + 64,
                       p->tstart);
    // padded so a wave may stage 512 taps from any 64-aligned chunk base
    if ((e = fr_alloc(ctx, (void **)&p->taps, tap_bytes * tap_alloc))) return e;
    FR_HIP(ctx, hipMemsetAsync(p->taps, 0, tap_bytes * tap_alloc, st));
    FR_HIP(ctx, hipMemsetAsync(p->hdrs, 0, sizeof(uint2) * (R + 64), st));
    // tap cell field: the byte offset of the cell's accumulator row; padding
    // taps go to the spare row cells_max with weight 0
    // (gather mode: the cell index, padding marked 0xffff; the batch build
    // below re-sorts them)
    const bool gather = p->accumulate == FR_ACC_GATHER;
    const uint32_t tap_scale = gather ? 1u : 4u * FR_ACC_STRIDE;
    const uint32_t pad_x = gather ? 0xffffu
                                  : (p->compact_taps ? (uint32_t)geom.cells_max
                                                     : (uint32_t)geom.cells_max * 4u * FR_ACC_STRIDE);
    // the matrix-core kernel reads entry TT for lanes without a tap
    if (p->compact_taps)
        hipLaunchKernelGGL(k_pad_tail<uint16_t>, dim3(nblk(tap_alloc - TT, 256)), dim3(256), 0, st,
                           reinterpret_cast<uint16_t *>(p->taps), TT, (int64_t)(tap_alloc - TT),
                           (uint16_t)(pad_x | ((uint32_t)p->L << 6)));
    else if (!gather)
        hipLaunchKernelGGL(k_pad_tail<uint2>, dim3(nblk(tap_alloc - TT, 256)), dim3(256), 0, st, p->taps, TT,
                           (int64_t)(tap_alloc - TT), make_uint2(pad_x, 0u));
    if (R) {
        hipLaunchKernelGGL(k_unit_ranges, dim3(nblk(R, 256)), dim3(256), 0, st, keys_out.p, toff.p, (int64_t)R,
                           p->units);
        hipLaunchKernelGGL(k_unit_counts, dim3(nblk(nunits, 256)), dim3(256), 0, st, p->units, nunits, toff.p);
        hipLaunchKernelGGL(k_fill_records, dim3(nblk(R, 128)), dim3(128), 0, st, s->traj, M, g, cu_d.p, p->units,
                           keys_out.p, vals_out.p, toff.p, (int64_t)R, p->lut, tap_scale, pad_x, p->hdrs,
                           p->compact_taps ? nullptr : p->taps,
                           p->compact_taps ? reinterpret_cast<uint16_t *>(p->taps) : nullptr);
    }
    FR_HIP(ctx, hipGetLastError());

    // 6. work items: one per unit, heavy units cut into slices of slice_thr taps
    {
        dbuf<uint32_t> flag, idx, first, last, eflag, eidx, hflag, hidx, nsl, sscan;
        const int64_t nu = std::max<int64_t>(1, nunits);
        if ((e = fr_alloc(ctx, (void **)&flag.p, 4 * R + 4)) || (e = fr_alloc(ctx, (void **)&idx.p, 4 * R + 4)) ||
            (e = fr_alloc(ctx, (void **)&first.p, 4 * nu)) || (e = fr_alloc(ctx, (void **)&last.p, 4 * nu)) ||
            (e = fr_alloc(ctx, (void **)&eflag.p, 4 * nu)) || (e = fr_alloc(ctx, (void **)&eidx.p, 4 * nu)) ||
            (e = fr_alloc(ctx, (void **)&hflag.p, 4 * nu)) || (e = fr_alloc(ctx, (void **)&hidx.p, 4 * nu)) ||
            (e = fr_alloc(ctx, (void **)&nsl.p, 4 * nu)) || (e = fr_alloc(ctx, (void **)&sscan.p, 4 * nu)))
            return e;
        size_t b1 = 0, b2 = 0;
        FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, b1, flag.p, idx.p, (int)std::max<uint64_t>(R, 1), st));
        FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, b2, eflag.p, eidx.p, (int)nu, st));
        dbuf<char> stmp;
        if ((e = fr_alloc(ctx, (void **)&stmp.p, std::max(b1, b2) + 16))) return e;
        uint32_t nrec_items = 0, nempty = 0, nheavy = 0, nslots = 0;
        if (R) {
            hipLaunchKernelGGL(k_item_flags, dim3(nblk(R, 256)), dim3(256), 0, st, keys_out.p, toff.p, (int64_t)R,
                               p->units, slice_thr, flag.p);
            FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(stmp.p, b1, flag.p, idx.p, (int)R, st));
            FR_HIP(ctx, hipMemcpyAsync(&nrec_items, idx.p + R - 1, 4, hipMemcpyDeviceToHost, st));
        }
        hipLaunchKernelGGL(k_empty_flags, dim3(nblk(nunits, 256)), dim3(256), 0, st, p->units, nunits, eflag.p);
        FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(stmp.p, b2, eflag.p, eidx.p, (int)nunits, st));
        FR_HIP(ctx, hipMemcpyAsync(&nempty, eidx.p + nunits - 1, 4, hipMemcpyDeviceToHost, st));
        FR_HIP(ctx, hipStreamSynchronize(st));
        p->nitems = (int64_t)nrec_items + nempty;

        if ((e = fr_alloc(ctx, (void **)&p->items, sizeof(fr_item) * std::max<int64_t>(1, p->nitems)))) return e;
        FR_HIP(ctx, hipMemsetAsync(first.p, 0, 4 * nu, st));
        FR_HIP(ctx, hipMemsetAsync(last.p, 0, 4 * nu, st));
        if (R)
            hipLaunchKernelGGL(k_make_items, dim3(nblk(R, 256)), dim3(256), 0, st, keys_out.p, toff.p, flag.p, idx.p,
                               (int64_t)R, p->units, p->items, first.p, last.p);
        hipLaunchKernelGGL(k_empty_items, dim3(nblk(nunits, 256)), dim3(256), 0, st, p->units, nunits, eidx.p,
                           nrec_items, p->items);
        hipLaunchKernelGGL(k_heavy_flags, dim3(nblk(nunits, 256)), dim3(256), 0, st, p->units, nunits, first.p,
                           last.p, hflag.p, nsl.p);
        FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(stmp.p, b2, hflag.p, hidx.p, (int)nunits, st));
        FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(stmp.p, b2, nsl.p, sscan.p, (int)nunits, st));
        FR_HIP(ctx, hipMemcpyAsync(&nheavy, hidx.p + nunits - 1, 4, hipMemcpyDeviceToHost, st));
        FR_HIP(ctx, hipMemcpyAsync(&nslots, sscan.p + nunits - 1, 4, hipMemcpyDeviceToHost, st));
        FR_HIP(ctx, hipStreamSynchronize(st));
        p->nheavy = nheavy;
        p->nslots = nslots;
        if ((e = fr_alloc(ctx, (void **)&p->heavy, sizeof(fr_heavy) * std::max<uint32_t>(1, nheavy)))) return e;
        hipLaunchKernelGGL(k_make_heavy, dim3(nblk(nunits, 256)), dim3(256), 0, st, nunits, hflag.p, hidx.p, nsl.p,
                           sscan.p, first.p, p->heavy, p->items);
        FR_HIP(ctx, hipGetLastError());
        FR_HIP(ctx, hipStreamSynchronize(st));
        // items of the units ordered first (the heavy ones) are launched in
        // order one per workgroup; the remaining items are spread over the
        // XCDs in contiguous ranges (see fr_grid.hip)
        p->nitems_front = 0;
        if (nheavy_units > 0 && nrec_items > 0) {
            std::vector<fr_item> hitems(nrec_items);
            FR_HIP(ctx, hipMemcpyAsync(hitems.data(), p->items, sizeof(fr_item) * nrec_items,
                                       hipMemcpyDeviceToHost, st));
            FR_HIP(ctx, hipStreamSynchronize(st));
            int64_t f = 0;
            while (f < (int64_t)nrec_items && (int64_t)hitems[f].unit < nheavy_units) ++f;
            p->nitems_front = f;
        }
    }
    p->nrecords = (int64_t)R;
    p->ntaps = (int64_t)TT;

    // 7. gather mode: split every item into batches of FR_GATHER_BATCH
    // records and re-sort each batch's taps by cell (stable: record order)
    if (gather && p->nitems > 0) {
        const int64_t ni = p->nitems;
        dbuf<uint32_t> nb, nbs;
        if ((e = fr_alloc(ctx, (void **)&nb.p, 4 * ni)) || (e = fr_alloc(ctx, (void **)&nbs.p, 4 * ni))) return e;
        hipLaunchKernelGGL(k_item_nbatch, dim3(nblk(ni, 256)), dim3(256), 0, st, p->items, ni, nb.p);
        size_t sb = 0;
        FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, sb, nb.p, nbs.p, (int)ni, st));
        dbuf<char> stmp2;
        if ((e = fr_alloc(ctx, (void **)&stmp2.p, sb + 16))) return e;
        FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(stmp2.p, sb, nb.p, nbs.p, (int)ni, st));
        uint32_t nbat = 0;
        FR_HIP(ctx, hipMemcpyAsync(&nbat, nbs.p + ni - 1, 4, hipMemcpyDeviceToHost, st));
        FR_HIP(ctx, hipStreamSynchronize(st));
        p->nbatches = nbat;
        if ((e = fr_alloc(ctx, (void **)&p->batches, sizeof(fr_batch) * std::max<uint32_t>(1, nbat)))) return e;
        hipLaunchKernelGGL(k_make_batches, dim3(nblk(ni, 256)), dim3(256), 0, st, p->items, ni, nb.p, nbs.p,
                           p->batches);
        dbuf<uint64_t> bsize, bscan;
        if ((e = fr_alloc(ctx, (void **)&bsize.p, 8 * (uint64_t)std::max<uint32_t>(1, nbat))) ||
            (e = fr_alloc(ctx, (void **)&bscan.p, 8 * (uint64_t)std::max<uint32_t>(1, nbat))))
            return e;
        uint64_t T2n = 0;
        if (nbat) {
            hipLaunchKernelGGL(k_batch_count, dim3(nblk((int64_t)nbat * 64, 256)), dim3(256), 0, st, p->batches,
                               (int64_t)nbat, p->hdrs, toff.p, p->taps, bsize.p);
            size_t sb2 = 0;
            FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(nullptr, sb2, bsize.p, bscan.p, (int)nbat, st));
            dbuf<char> stmp3;
            if ((e = fr_alloc(ctx, (void **)&stmp3.p, sb2 + 16))) return e;
            FR_HIP(ctx, hipcub::DeviceScan::InclusiveSum(stmp3.p, sb2, bsize.p, bscan.p, (int)nbat, st));
            FR_HIP(ctx, hipMemcpyAsync(&T2n, bscan.p + nbat - 1, 8, hipMemcpyDeviceToHost, st));
            FR_HIP(ctx, hipStreamSynchronize(st));
        }
        uint2 *T2 = nullptr;
        if ((e = fr_alloc(ctx, (void **)&T2, sizeof(uint2) * (T2n + 64)))) return e;
        FR_HIP(ctx, hipMemsetAsync(T2, 0, sizeof(uint2) * (T2n + 64), st));
        if (nbat)
            hipLaunchKernelGGL(k_batch_write, dim3(nblk((int64_t)nbat * 64, 256)), dim3(256), 0, st, p->batches,
                               (int64_t)nbat, p->hdrs, toff.p, p->taps, bscan.p, T2);
        FR_HIP(ctx, hipGetLastError());
        FR_HIP(ctx, hipStreamSynchronize(st));
        hipFree(p->taps);
        p->taps = T2;
        p->ntaps = (int64_t)T2n;
    }
    p->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FR_OK;
}


// Exact LUT-index thresholds for a fast (1 ulp) square root.  The reference
// index roundf(sqrtf(d2) / hw * (L - 1)) (GridLut.cpp:112-136) is monotone in
// the squared distance d2, so B[m] = least d2 whose index is >= m exists, and
// the index of d2 is the m with B[m] <= d2 < B[m + 1].  B is found by
// bisection over float bit patterns with the reference arithmetic on the
// host.  mfma_pairs[m] = (B[m], B[m + 1]) serves the degridding kernel
// (estimate within one of the index, corrected both ways); with sel,
// tsel[m] = (B[m + 1], LUT[m], LUT[m + 1], 0) serves the position-weight
// gridding kernel (estimate the index or one below it).
fr_status fr_plan_thresholds(fr_plan *p, bool sel)
{
    fr_ctx *ctx = p->ctx;
    hipStream_t st = ctx->stream;
    fr_status e;
    const float hw = p->kHW;
    const int L = p->L;
    const float top = 64.f * hw * hw + 64.f;        // far beyond every accepted tap
    uint32_t top_bits;
    std::memcpy(&top_bits, &top, 4);
    std::vector<float> B(L + 2, INFINITY);
    for (int m = 0; m < L; ++m) {
        if (ref_lut_index(top, hw, L) < m) break;           // not reached below the search bound
        B[m] = least_float(top_bits, [&](float d2) { return ref_lut_index(d2, hw, L) >= m; });
    }
    p->mfma_dacc = least_float(top_bits, [&](float d2) { return sqrtf(d2) >= hw; });
    p->mfma_cidx = (float)((double)(L - 1) / (double)hw);
    std::vector<float2> pairs(L + 1);
    for (int m = 0; m <= L; ++m) pairs[m] = make_float2(B[m], B[m + 1]);
    if (!p->mfma_pairs && (e = fr_alloc(ctx, (void **)&p->mfma_pairs, sizeof(float2) * (L + 1)))) return e;
    FR_HIP(ctx, hipMemcpyAsync(p->mfma_pairs, pairs.data(), sizeof(float2) * (L + 1), hipMemcpyHostToDevice, st));
    if (sel) {
        std::vector<float4> t(L);
        for (int m = 0; m < L; ++m)
            t[m] = make_float4(B[m + 1], p->host_lut[m], m + 1 < L ? p->host_lut[m + 1] : 0.f, 0.f);
        if (!p->tsel && (e = fr_alloc(ctx, (void **)&p->tsel, sizeof(float4) * L))) return e;
        FR_HIP(ctx, hipMemcpyAsync(p->tsel, t.data(), sizeof(float4) * L, hipMemcpyHostToDevice, st));
    }
    FR_HIP(ctx, hipStreamSynchronize(st));
    return FR_OK;
}

// The position-weight matrix-core mode (accumulate 6) is the default for
// plans that are not exact-order, unless the LUT is too long for its LDS
// table or the tuned unit box exceeds 16 cells (then accumulate 5).
static bool fr_plan_uses_pos(const fr_plan *p, const fr_samples *s, const fr_tune *tune)
{
    (void)s;
    const int acc = tune ? tune->accumulate : 0;
    if (acc == FR_ACC_POS) return true;
    if (acc != 0 || (tune && tune->exact_order) || p->L > FR_POS_MAX_L) return false;
    if (tune) {
        const int bx = tune->tile_x > 0 ? tune->tile_x : 4;
        const int by = tune->tile_y > 0 ? tune->tile_y : (p->dim == 2 ? 4 : 2);
        const int bz = p->dim == 3 && tune->tile_z > 0 ? tune->tile_z : (p->dim == 2 ? 1 : 2);
        if (bx * by * bz > 16) return false;
    }
    return true;
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
    FR_CHECK(ctx, s->frames == 0 || (uint64_t)s->frames * N * N < (uint64_t(1) << 31), FR_ERR_UNSUPPORTED,
             "%d frames of %u^2 exceed 2^31 cells", s->frames, N);
    FR_CHECK(ctx, L >= 2 && L <= (1 << 20), FR_ERR_INVALID, "LUT length %d invalid", L);
    FR_CHECK(ctx, kWidth > 0.f && kWidth <= 16.f, FR_ERR_INVALID, "kernel width %g invalid", (double)kWidth);
    fr_plan *p = new (std::nothrow) fr_plan;
    if (!p) return FR_ERR_NOMEM;
    p->ctx = ctx; p->dim = s->dim; p->N = N; p->M = s->M; p->frames = s->frames;
    p->kWidth = kWidth; p->kHW = kWidth / 2.0f; p->L = L;     // GridLut.cpp:86
    p->host_lut.assign(lut, lut + L);
    fr_status e = fr_alloc(ctx, (void **)&p->lut, sizeof(float) * L);
    if (!e) {
        hipError_t he = hipMemcpyAsync(p->lut, lut, sizeof(float) * L, hipMemcpyHostToDevice, ctx->stream);
        if (he != hipSuccess) e = fr_fail(ctx, FR_ERR_HIP, "lut upload: %s", hipGetErrorString(he));
    }
    if (!e) e = fr_plan_uses_pos(p, s, tune) ? fr_plan_build_pos(p, s, tune) : fr_plan_build(p, s, tune);
    if (e) { fr_plan_destroy(p); return e; }
    *out = p;
    return FR_OK;
}

extern "C" fr_status fr_plan_destroy(fr_plan *p)
{
    if (!p) return FR_OK;
    hipStreamSynchronize(p->ctx->stream);
    hipFree(p->lut); hipFree(p->units); hipFree(p->hdrs); hipFree(p->taps); hipFree(p->tstart);
    hipFree(p->items); hipFree(p->heavy); hipFree(p->partial); hipFree(p->batches); hipFree(p->mfma_pairs);
    hipFree(p->recs); hipFree(p->pos); hipFree(p->tsel); hipFree(p->tiles);
    hipFree(p->dcf_den);
    hipFree(p->dcf_scr);
    fr_grids_destroy(p->dcf_g);
    fr_samples_destroy(p->dcf_u);
    delete p;
    return FR_OK;
}

extern "C" fr_status fr_plan_get_info(const fr_plan *p, fr_plan_info *info)
{
    if (!p || !info) return FR_ERR_INVALID;
    info->samples = p->M;
    info->taps = p->ntaps;
    info->records = p->nrecords;
    info->streams = p->nunits;
    info->tile_x = p->geom.bx; info->tile_y = p->geom.by; info->tile_z = p->geom.bz;
    info->slabs = p->geom.waves;
    info->tiles = (int)std::min<int64_t>(p->nunits, INT32_MAX);
    info->lds_bytes = 0;
    if (p->accumulate == FR_ACC_POS)
        info->plan_bytes = (int64_t)sizeof(uint32_t) * (p->nrecs_padded + 64) + (int64_t)sizeof(float4) * (p->M + 1) +
                           (int64_t)sizeof(fr_unit) * p->nunits + (int64_t)sizeof(fr_item) * p->nitems;
    else
        info->plan_bytes = (int64_t)(sizeof(uint2) + sizeof(uint32_t)) * p->nrecords +
                           (int64_t)(p->compact_taps ? sizeof(uint16_t) : sizeof(uint2)) * p->ntaps +
                           (int64_t)sizeof(fr_unit) * p->nunits;
    info->build_ms = p->build_ms;
    info->items = p->nitems;
    info->heavy_units = p->nheavy;
    info->max_unit_taps = p->max_unit_taps;
    info->exact_order = p->exact_order;
    info->accumulate = p->accumulate;
    info->scratch_bytes = (int64_t)fr_grid_scratch_bytes(p, 32);
    return FR_OK;
}
