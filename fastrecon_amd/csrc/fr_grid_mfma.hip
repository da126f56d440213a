// Gridding on the matrix cores (accumulate mode FR_ACC_MFMA; the default for
// plans that are not exact-order).
//
// Same work items, records, taps and deterministic slice combine as the LDS
// kernel (fr_grid.hip), different arithmetic: a unit of at most 64 cells is
// cut into blocks of 16 cells, and its records are taken four at a time.  For
// each block the four records' contributions are one dense product
//
//     acc[16 cells][16 components] += A[16 cells][4 records] * B[4 records][16 components]
//
// on v_mfma_f32_16x16x4_f32, with
//   A[i][k] = LUT weight of record k's sample at cell i (0 where the sample
//             has no tap there): the plan's taps (reference tap set and LUT
//             index, GridLut.cpp:99-136) scattered into a per-wave LDS tile;
//   B[k][j] = dcf * sample component j of record k (GridLut.cpp:109).
// Every product w * (dcf * s) is the reference's; only the order of the sum
// differs (the matrix core adds four products at a time), so results agree
// with the CPU loop to FP32 rounding (relative L2 ~1e-7), and are
// deterministic.
//
// Lane layout of v_mfma_f32_16x16x4_f32 (wave64): lane l supplies A[l % 16][l
// / 16] and B[l / 16][l % 16]; accumulator register r of lane l holds
// D[4 (l / 16) + r][l % 16].  So lane l works on record l / 16 of a group and
// component slice l % 16.  Components: with NCB component blocks per 16
// lanes, lane j holds components [NCB j, NCB j + NCB) of a chunk of 8 NCB
// coils (NCB = 4: coils 2j and 2j+1; NCB = 2: coil j; NCB = 1: coil j/2, re
// or im).
//
// The accumulators leave through an LDS transpose, so that consecutive lanes
// store consecutive cells of one coil (rows of ux cells): x-elongated units
// (8 x 2 x 2 by default) give the longest contiguous runs.
//
// Non-finite sample values differ from the reference in one respect: a
// sample's row meets every cell of the blocks it touches (with weight 0), so
// a NaN or Inf spreads to those cells.
#include "fr_internal.h"
#include "fr_taps.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));


__device__ __forceinline__ float2 ld_coil(const float2 *__restrict__ row, int c) { return row[c]; }
__device__ __forceinline__ float2 ld_coil(const __half2 *__restrict__ row, int c) { return __half22float2(row[c]); }
__device__ __forceinline__ float4 ld_pair(const float2 *__restrict__ row, int c)
{
    return *reinterpret_cast<const float4 *>(row + c);
}
__device__ __forceinline__ float4 ld_pair(const __half2 *__restrict__ row, int c)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(row + c);
    const float2 a = __half22float2(*reinterpret_cast<const __half2 *>(&u.x));
    const float2 b = __half22float2(*reinterpret_cast<const __half2 *>(&u.y));
    return make_float4(a.x, a.y, b.x, b.y);
}

constexpr int RB = 16;                       // records per batch
constexpr int64_t XCD_RUN = FR_MFMA_XCD_RUN; // consecutive items one XCD takes in turn
// light items per workgroup, consecutive in an XCD's run: 4 for units of 32
// or more cells (one LUT staging for four items measured 5% faster with 8x2x2
// units), 1 for 16-cell units (4 per workgroup measured 6% slower there)
template <int NBLK> constexpr int items_per_wg() { return NBLK >= 2 ? 4 : 1; }
static_assert(XCD_RUN % 4 == 0, "XCD runs hold whole groups of items");

// The workgroup is one wave.  LDS data written by some lanes and read by
// others needs only the wave's in-order LDS queue and a compiler barrier, not
// a workgroup barrier (whose fence also waits for every outstanding global
// store, i.e. for the grid write-out).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
constexpr int TAP_ROW = 4 * FR_ACC_STRIDE;   // plan taps: (cell * TAP_ROW, LUT weight bits)
// the pipeline reads headers up to 3 RB - 1 records past a live record; the
// plan pads the headers and tap starts by 64 records
static_assert(3 * RB - 1 < 64, "record header padding");

// Every lane reads the headers of the records it works on (16 lanes share
// each one), so a batch needs no cross-lane exchange: tap pass p gives lane l
// tap u = l % TPR of record p * RPP + l / TPR (a record carries at most TPR =
// 16 NBLK taps), and row group q gives lane (j, k) the DCF and component
// slice of record 4 q + k (for NBLK = 1 the same record as pass q).
template <int NBLK>
struct hdr_st {
    uint2 hp[4 * NBLK];         // pass p: (sample, padded tap count) of the lane's record
    uint32_t t0[4 * NBLK];      // ... and the index of its first tap
    uint2 hg[NBLK == 1 ? 1 : 4];   // group q (NBLK > 1): (sample, count) of record 4 q + k
};

// One batch of RB records in flight: taps (padding taps where the record has
// no more), DCFs (0 past the item's end) and sample rows.
template <int NBLK, bool CT>
struct batch_st {
    typename std::conditional<CT, uint32_t, uint2>::type tp[4 * NBLK];   // compact or full taps
    float w[4];
    float2 a[4], b[4];
};

// tap -> (cell within the unit, LUT weight)
__device__ __forceinline__ int tap_cell(uint32_t t) { return (int)(t & 63u); }
__device__ __forceinline__ int tap_cell(uint2 t) { return (int)(t.x / TAP_ROW); }
__device__ __forceinline__ float tap_weight(uint32_t t, const float *slut) { return slut[t >> 6]; }
__device__ __forceinline__ float tap_weight(uint2 t, const float *) { return __uint_as_float(t.y); }

// element at a 32-bit byte offset from a uniform base: the load takes the base
// from SGPRs and only the offset from a VGPR (no 64-bit address arithmetic);
// the plan arrays and the DCF stay below 4 GiB (checked by the launcher)
template <typename E>
__device__ __forceinline__ E ld_off(const E *__restrict__ base, uint32_t byte_off)
{
    return *reinterpret_cast<const E *>(reinterpret_cast<const char *>(base) + byte_off);
}

template <int NBLK>
__device__ __forceinline__ void load_hdrs(hdr_st<NBLK> &G, uint32_t rb, const uint2 *__restrict__ H,
                                          const uint32_t *__restrict__ tstart, int lane)
{
    constexpr int TPR = 16 * NBLK, RPP = 64 / TPR;
#pragma unroll
    for (int p = 0; p < 4 * NBLK; ++p) {
        const uint32_t rec = rb + p * RPP + lane / TPR;
        G.hp[p] = ld_off(H, rec * (uint32_t)sizeof(uint2));
        G.t0[p] = ld_off(tstart, rec * (uint32_t)sizeof(uint32_t));
    }
    if constexpr (NBLK > 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) G.hg[q] = ld_off(H, (rb + 4 * q + (lane >> 4)) * (uint32_t)sizeof(uint2));
    }
}

// Nothing here uses a loaded value (only addresses), so the loads of the
// next batch stay in flight while the current one is applied: a lane with no
// tap left reads a padding tap from the tail of the tap array (index padidx,
// the plan's tap count), and the DCF of records past the item's end is masked
// when the batch is applied.
template <int NBLK, int NCB, bool PAIR, bool CT, typename TS, typename TT>
__device__ __forceinline__ void issue_batch(batch_st<NBLK, CT> &S, const hdr_st<NBLK> &G, uint32_t rb, uint32_t r1,
                                            const TT *__restrict__ T, uint32_t padidx, const TS *__restrict__ data,
                                            const float *__restrict__ dcf, uint32_t C, int ca, int cbb, int lane)
{
    constexpr int TPR = 16 * NBLK, RPP = 64 / TPR;
    const uint32_t u = (uint32_t)(lane % TPR);
#pragma unroll
    for (int p = 0; p < 4 * NBLK; ++p) {
        const uint32_t rec = rb + p * RPP + lane / TPR;
        const uint32_t n = rec < r1 ? G.hp[p].y : 0u;
        const uint32_t ti = u < n ? G.t0[p] + u : padidx;
        if constexpr (CT)
            S.tp[p] = ld_off(T, ti * (uint32_t)sizeof(TT));
        else
            S.tp[p] = T[ti];                    // 8-byte taps may pass 4 GiB
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint2 h = NBLK == 1 ? G.hp[q] : G.hg[NBLK == 1 ? 0 : q];
        S.w[q] = ld_off(dcf, h.x * (uint32_t)sizeof(float));
        const TS *row = data + (uint64_t)h.x * C;
        if constexpr (NCB == 4 && PAIR) {
            const float4 v = ld_pair(row, ca);
            S.a[q] = make_float2(v.x, v.y);
            S.b[q] = make_float2(v.z, v.w);
        } else {
            S.a[q] = ld_coil(row, ca);
            if constexpr (NCB == 4) S.b[q] = ld_coil(row, cbb);
        }
    }
}

template <int NBLK, int NCB, bool CT, bool FULL>
__device__ __forceinline__ void apply_batch(f4 (&acc)[NBLK][NCB], const batch_st<NBLK, CT> &S, float *tile,
                                            const float *slut, bool v0ok, bool v1ok, uint32_t rb, uint32_t r1,
                                            int lane)
{
    constexpr int TPR = 16 * NBLK, RPP = 64 / TPR;
    constexpr int ROW = 16 * NBLK + 1;          // cells (and the padding taps' row) per record column
    const int j = lane & 15, k = lane >> 4;
    // A tile [record][cell]: zero, then every tap writes its LUT weight
    // (the tile is allocated in whole multiples of 64 floats, so the zeroing
    // needs no bounds test and no branch)
#pragma unroll
    for (int z0 = 0; z0 < RB * ROW; z0 += 64) tile[z0 + lane] = 0.f;
#pragma unroll
    for (int p = 0; p < 4 * NBLK; ++p) {
        // taps of one record have distinct cells; padding taps all write 0
        // to the spare column cells_max, which no output reads
        const int rho = p * RPP + lane / TPR;
        tile[rho * ROW + tap_cell(S.tp[p])] = tap_weight(S.tp[p], slut);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rec = 4 * q + k;
        // B = dcf * sample (GridLut.cpp:109); padding coil columns and
        // records past the item's end get 0
        // (a record past the item's end is another unit's sample: its B row
        // is selected to 0, not multiplied by 0, so a NaN or Inf there
        // cannot reach this unit)
        const bool live = rb + (uint32_t)rec < r1;
        const float w = S.w[q];
        const float w0 = FULL || v0ok ? w : 0.f;
        float bv[NCB];
        if constexpr (NCB == 4) {
            const float w1 = FULL || v1ok ? w : 0.f;
            bv[0] = __fmul_rn(w0, S.a[q].x); bv[1] = __fmul_rn(w0, S.a[q].y);
            bv[2] = __fmul_rn(w1, S.b[q].x); bv[3] = __fmul_rn(w1, S.b[q].y);
        } else if constexpr (NCB == 2) {
            bv[0] = __fmul_rn(w0, S.a[q].x); bv[1] = __fmul_rn(w0, S.a[q].y);
        } else {
            bv[0] = __fmul_rn(w0, (j & 1) ? S.a[q].y : S.a[q].x);
        }
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) bv[cb] = live ? bv[cb] : 0.f;
#pragma unroll
        for (int b = 0; b < NBLK; ++b) {
            const float aw = tile[rec * ROW + 16 * b + j];
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
                acc[b][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aw, bv[cb], acc[b][cb], 0, 0, 0);
        }
    }
}

// NBLK blocks of 16 cells; NCB component blocks (chunks of 8 NCB coils).
// PAIR: C even, so lane j's two coils (2j, 2j+1 of a 32-coil chunk) are one
// aligned 16-byte (8-byte for FP16) load.
// CT: the plan's taps are compact (uint16_t: cell | LUT index << 6) and the
// weights come from the LUT staged in LDS (entry L is 0, for padding taps).
template <int NBLK, int NCB, bool PAIR, bool CT, bool FULL, typename TS>
__global__ void __launch_bounds__(64) k_grid_mfma(const fr_item *__restrict__ items, int64_t nitems,
                                                  const fr_unit *__restrict__ units, const uint2 *__restrict__ H,
                                                  const uint32_t *__restrict__ tstart,
                                                  const void *__restrict__ Tv, const TS *__restrict__ data,
                                                  const float *__restrict__ dcf, int C, int N, size_t cells,
                                                  float2 *__restrict__ out, float2 *__restrict__ partial,
                                                  int64_t nslots, int cells_max, const float *__restrict__ lut,
                                                  int L, int64_t nfront, uint32_t padidx)
{
    using TT = typename std::conditional<CT, uint16_t, uint2>::type;
    const TT *__restrict__ T = reinterpret_cast<const TT *>(Tv);
    __shared__ float slut[CT ? FR_CT_MAX_L + 1 : 1];
    if constexpr (CT) {
        for (int q = threadIdx.x; q <= L; q += 64) slut[q] = q < L ? lut[q] : 0.f;
        wave_sync();
    }
    // LDS: the A tile during the record loop, then the write-out transpose
    constexpr int CS = 16 * NBLK + 1;           // cells per component row of the transpose
    constexpr int TILE = (RB * (16 * NBLK + 1) + 63) / 64 * 64, OSM = 16 * NCB * CS;
    __shared__ float smem[(TILE > OSM ? TILE : OSM)];
    __shared__ uint32_t goff[16 * NBLK];        // grid offset of each unit cell
    float *tile = smem, *osm = smem;
    const int lane = threadIdx.x;
    const int j = lane & 15;                    // A row (cell in block) / B column (component)
    const int k = lane >> 4;                    // A column / B row (record in group)
    const int chunk = 8 * NCB;                  // coils per chunk
    const int c0 = blockIdx.y * chunk;
    const int nc = min(chunk, C - c0);
    // lane j's coils (NCB = 4: 2j, 2j+1; 2: j; 1: j/2), clamped to a valid
    // coil for the load and masked through the DCF
    const int cl0 = NCB == 4 ? 2 * j : (NCB == 2 ? j : j >> 1);
    const bool v0ok = cl0 < nc, v1ok = NCB == 4 && cl0 + 1 < nc;
    // PAIR (C and so nc even): the pair's first coil stays even and the pair
    // inside the chunk, ca + 1 <= c0 + nc - 1 < C
    const int ca = PAIR ? c0 + min(cl0, nc - 2) : c0 + min(cl0, nc - 1), cbb = c0 + min(cl0 + 1, nc - 1);

    // Items in the order of k_grid_items (fr_grid.hip): the heaviest units'
    // items first, one per workgroup; the rest, in Morton order, IPW
    // consecutive items per workgroup (one LUT staging for all of them),
    // remapped so that each XCD (workgroups are dealt to the 8 XCDs
    // round-robin) takes runs of XCD_RUN consecutive items.  The hardware
    // dispatches workgroups in order, so each XCD's running items stay a
    // compact patch of neighbouring units that share samples in its L2 (a
    // persistent static schedule drifts apart and measured slower).
    const int64_t bid = blockIdx.x;
    int64_t i0;
    int nown;
    if (bid < nfront) {
        i0 = bid;
        nown = 1;
    } else {
        constexpr int IPW = items_per_wg<NBLK>();
        constexpr int64_t XR = XCD_RUN / IPW;   // groups per XCD run
        const int64_t rr = bid - nfront, gg = rr / 8;
        i0 = nfront + ((gg / XR) * (8 * XR) + (bid % 8) * XR + gg % XR) * IPW;
        nown = IPW;
    }
#pragma unroll 1
    for (int own = 0; own < nown; ++own) {
        const int64_t i = i0 + own;
        if (i >= nitems) break;
        const fr_item it = items[i];
        const fr_unit U = units[it.unit];
        const int ux = U.ext & 255, uy = (U.ext >> 8) & 255, uz = U.ext >> 16;
        const int ncell = ux * uy * uz;

        f4 acc[NBLK][NCB];
#pragma unroll
        for (int b = 0; b < NBLK; ++b)
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) acc[b][cb] = f4{0.f, 0.f, 0.f, 0.f};

        if (it.r0 < it.r1) {
            // Software pipeline over batches of RB records: while batch b is
            // applied, batch b+1's taps, DCFs and sample rows and batch b+2's
            // headers are in flight.  Records past the item's end read valid
            // (padded) headers and are masked (padding taps, zero DCF).
            const uint32_t r1 = it.r1;
            uint32_t r = it.r0;
            hdr_st<NBLK> G0, G1;
            batch_st<NBLK, CT> S0, S1;
            load_hdrs<NBLK>(G0, r, H, tstart, lane);
            load_hdrs<NBLK>(G1, r + RB, H, tstart, lane);
            issue_batch<NBLK, NCB, PAIR, CT>(S0, G0, r, r1, T, padidx, data, dcf, (uint32_t)C, ca, cbb, lane);
            while (true) {
                load_hdrs<NBLK>(G0, r + 2 * RB, H, tstart, lane);
                issue_batch<NBLK, NCB, PAIR, CT>(S1, G1, r + RB, r1, T, padidx, data, dcf, (uint32_t)C, ca, cbb,
                                                 lane);
                // keep the prefetch ahead of the compute (the scheduler would
                // otherwise sink each load to its use and wait on it there)
                __builtin_amdgcn_sched_barrier(0);
                apply_batch<NBLK, NCB, CT, FULL>(acc, S0, tile, slut, v0ok, v1ok, r, r1, lane);
                r += RB;
                if (r >= r1) break;
                load_hdrs<NBLK>(G1, r + 2 * RB, H, tstart, lane);
                issue_batch<NBLK, NCB, PAIR, CT>(S0, G0, r + RB, r1, T, padidx, data, dcf, (uint32_t)C, ca, cbb,
                                                 lane);
                __builtin_amdgcn_sched_barrier(0);
                apply_batch<NBLK, NCB, CT, FULL>(acc, S1, tile, slut, v0ok, v1ok, r, r1, lane);
                r += RB;
                if (r >= r1) break;
            }
        }

        // Write-out.  Accumulator register r of block b holds cell 16 b + 4 k
        // + r, component NCB j + cb (coil c = re 2c, im 2c + 1).  They go to
        // LDS as [component][cell], then consecutive lanes store consecutive
        // cells of one coil, so each store covers whole rows of the unit (ux
        // contiguous cells) instead of one 8-byte value in each of 32 lines.
        // When the unit's cell count divides 64 (the usual full units of 16,
        // 32 or 64 cells) every lane keeps one cell for the whole write-out
        // and steps through the coils; otherwise cell offsets go through LDS.
        const bool lane_cell = (64 % ncell) == 0;
        wave_sync();
        if (!lane_cell) {
            for (int q = lane; q < ncell; q += 64) {
                const int lx = q % ux, ly = (q / ux) % uy, lz = q / (ux * uy);
                goff[q] = (uint32_t)(((size_t)(U.oz + lz) * N + (U.oy + ly)) * N + (U.ox + lx));
            }
        }
#pragma unroll
        for (int b = 0; b < NBLK; ++b)
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int r = 0; r < 4; ++r) osm[(NCB * j + cb) * CS + 16 * b + 4 * k + r] = acc[b][cb][r];
        wave_sync();
        // each grid value is written once and not read again by this pass: a
        // streaming store keeps the sample rows in L2
        auto put = [&](int cl, int q, size_t go) {
            const float2 v = make_float2(osm[(2 * cl) * CS + q], osm[(2 * cl + 1) * CS + q]);
            const int c = c0 + cl;
            if (it.slot < 0) {
                __builtin_nontemporal_store(__builtin_bit_cast(unsigned long long, v),
                                            reinterpret_cast<unsigned long long *>(out + (size_t)c * cells + go));
            } else {
                const size_t slot = ((size_t)(c >> 5) * nslots + it.slot) * 32 + (c & 31);
                partial[slot * cells_max + q] = v;
            }
        };
        if (lane_cell) {
            const int sh = __builtin_ctz((unsigned)ncell);     // ncell is a power of two here
            const int q = lane & (ncell - 1);
            const int lx = q % ux, ly = (q / ux) % uy, lz = q / (ux * uy);
            const size_t go = ((size_t)(U.oz + lz) * N + (U.oy + ly)) * N + (U.ox + lx);
            for (int cl = lane >> sh; cl < nc; cl += 64 >> sh) put(cl, q, go);
        } else {
            const int pairs = nc * ncell;
            for (int t = lane; t < pairs; t += 64) {
                const int cl = t / ncell;
                const int q = t - cl * ncell;
                put(cl, q, goff[q]);
            }
        }
        wave_sync();   // the next item reuses the LDS tile
    }
}

template <int NBLK, int NCB, bool PAIR, bool CT, bool FULL, typename TS>
fr_status launch_mfma_ct(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial)
{
    fr_ctx *ctx = p->ctx;
    // the kernel addresses headers, tap starts, compact taps and DCFs by 32-bit byte offsets
    const uint64_t lim = uint64_t(1) << 32;
    FR_CHECK(ctx, (uint64_t)(p->nrecords + 64) * sizeof(uint2) < lim &&
                      (!CT || (uint64_t)(p->ntaps + 576) * sizeof(uint16_t) < lim) &&
                      (uint64_t)s->M * sizeof(float) < lim,
             FR_ERR_UNSUPPORTED, "matrix-core gridding: plan of %lld records / %lld taps exceeds 4 GiB arrays "
             "(use accumulate 2)", (long long)p->nrecords, (long long)p->ntaps);
    const size_t cells = fr_plan_cells(p->dim, p->N, p->frames);
    const int chunks = (s->C + 8 * NCB - 1) / (8 * NCB);
    const int64_t light = p->nitems - p->nitems_front;
    constexpr int IPW = items_per_wg<NBLK>();
    const int64_t groups = (light + IPW - 1) / IPW, span = 8 * (XCD_RUN / IPW);
    const int64_t W = p->nitems_front + span * ((groups + span - 1) / span);
    hipLaunchKernelGGL((k_grid_mfma<NBLK, NCB, PAIR, CT, FULL, TS>), dim3((unsigned)W, chunks), dim3(64), 0, ctx->stream,
                       p->items, p->nitems, p->units, p->hdrs, p->tstart, (const void *)p->taps, (const TS *)s->data,
                       s->dcf,
                       s->C, (int)p->N, cells, out, partial, p->nslots, p->geom.cells_max, p->lut,
                       p->L, p->nitems_front, (uint32_t)p->ntaps);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

template <int NBLK, int NCB, bool PAIR, bool FULL, typename TS>
fr_status launch_mfma(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial)
{
    return p->compact_taps ? launch_mfma_ct<NBLK, NCB, PAIR, true, FULL, TS>(p, s, out, partial)
                           : launch_mfma_ct<NBLK, NCB, PAIR, false, FULL, TS>(p, s, out, partial);
}

// FULL: every chunk holds 8 NCB coils, so no lane's coil is masked
template <int NBLK, typename TS>
fr_status launch_ncb(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial)
{
    if (s->C <= 8) return launch_mfma<NBLK, 1, false, false, TS>(p, s, out, partial);
    if (s->C <= 16) return launch_mfma<NBLK, 2, false, false, TS>(p, s, out, partial);
    if (s->C % 32 == 0) return launch_mfma<NBLK, 4, true, true, TS>(p, s, out, partial);
    if (s->C % 2 == 0) return launch_mfma<NBLK, 4, true, false, TS>(p, s, out, partial);
    return launch_mfma<NBLK, 4, false, false, TS>(p, s, out, partial);
}

template <typename TS>
fr_status launch_nblk(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial)
{
    const int cm = p->geom.cells_max;
    if (cm <= 16) return launch_ncb<1, TS>(p, s, out, partial);
    if (cm <= 32) return launch_ncb<2, TS>(p, s, out, partial);
    return launch_ncb<4, TS>(p, s, out, partial);
}

}  // namespace

fr_status fr_launch_grid_mfma(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial)
{
    return s->type == FR_SAMPLES_C64 ? launch_nblk<float2>(p, s, out, partial)
                                     : launch_nblk<__half2>(p, s, out, partial);
}
