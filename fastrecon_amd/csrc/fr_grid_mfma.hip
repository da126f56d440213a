// Gridding on the matrix cores (accumulate mode FR_ACC_MFMA).
//
// Same work items, records and deterministic slice combine as the LDS
// kernels (fr_grid.hip), different arithmetic: a unit of at most 64 cells is
// cut into blocks of 16 cells, and its records are taken four at a time.  For
// each block the four records' contributions are one dense product
//
//     acc[16 cells][16 components] += A[16 cells][4 records] * B[4 records][16 components]
//
// on v_mfma_f32_16x16x4_f32, with
//   A[i][k] = LUT weight of record k's sample at cell i (0 where the sample
//             has no tap), computed on the fly with the reference arithmetic
//             of fr_taps.h: the tap set {cells with sqrtf(dz^2+dy^2+dx^2) <
//             hw} equals the reference's clamped box test (GridLut.cpp:99-
//             136; the box bounds c +/- hw are exact or only move across
//             cells the distance test rejects), and the LUT index is the
//             reference's roundf((dk / hw) * (L - 1));
//   B[k][j]  = dcf * sample component j of record k (GridLut.cpp:109).
// Every product w * (dcf * s) is the reference's; only the order of the sum
// differs (the matrix core adds four products at a time), so results agree
// with the CPU loop to FP32 rounding (relative L2 ~1e-7), and are
// deterministic.  No tap stream is read: a pass streams the record headers,
// the sample rows and coordinates, and writes the grid once.
//
// Lane layout of v_mfma_f32_16x16x4_f32 (wave64): lane l supplies A[l % 16][l
// / 16] and B[l / 16][l % 16]; accumulator register r of lane l holds
// D[4 (l / 16) + r][l % 16].  So lane l works on record l / 16 of the group:
// it loads that record's sample coordinates (for A) and its component slice
// l % 16 (for B).  Components: with NCB component blocks per 16 lanes, lane j
// holds components [NCB j, NCB j + NCB) of a chunk of 8 NCB coils (NCB = 4:
// coils 2j and 2j+1; NCB = 2: coil j; NCB = 1: coil j/2, re or im).
//
// Non-finite sample values differ from the reference in one respect: a
// sample's row meets every cell of the blocks it touches (with weight 0), so
// a NaN or Inf spreads to those cells.
#include "fr_internal.h"
#include "fr_taps.h"

#include <algorithm>
#include <cmath>

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

// Weight A[cell][record] of one (cell, sample) pair: the reference tap test
// and LUT entry.  The index comes from the fast square root (within one of
// the reference index) corrected by the exact thresholds B[m] <= d2 < B[m+1].
__device__ __forceinline__ float pair_weight(float fx, float fy, float fz, float px, float py, float pz, bool ok,
                                             const float2 *pairs, const float *slut, float dacc, float cidx, int L)
{
    const float dz = __fsub_rn(fz, pz);
    const float dy = __fsub_rn(fy, py);
    const float dx = __fsub_rn(fx, px);
    const float d2 = fr_tap_dk2(__fmul_rn(dz, dz), __fmul_rn(dy, dy), dx);   // (dz^2 + dy^2) + dx^2
    const bool tap = ok && d2 < dacc;                                        // sqrtf(d2) < hw
    const float t = __builtin_amdgcn_sqrtf(d2);
    int ka = (int)__builtin_fmaf(t, cidx, 0.5f);
    ka = min(ka, L);
    const float2 bb = pairs[ka];
    int ki = ka - (d2 < bb.x ? 1 : 0) + (d2 >= bb.y ? 1 : 0);
    ki = min(max(ki, 0), L - 1);                // (never binds for a tap; keeps the read in bounds)
    const float w = slut[ki];                   // computed for every pair: no branch
    return tap ? w : 0.f;
}

// One group of four records as lane (j, k) sees it: record k's sample
// coordinates and DCF, and the sample's component slice of lane j.
template <int NCB>
struct group_regs {
    float px, py, pz, w;
    float2 a, b;                                // coil slots (NCB = 4: two coils; else one)
};

template <int NCB, int ZM, bool PAIR, typename TS>
__device__ __forceinline__ void load_group(group_regs<NCB> &G, uint32_t s, const float *__restrict__ tx,
                                           const float *__restrict__ ty, const float *__restrict__ tz, int frames,
                                           const TS *__restrict__ data, const float *__restrict__ dcf, uint32_t C,
                                           int ca, int cbb)
{
    G.px = tx[s];
    G.py = ty[s];
    if constexpr (ZM == 1) G.pz = tz[s];
    else if constexpr (ZM == 2) G.pz = (float)min(max((int)tz[s], 0), frames - 1);
    else G.pz = 0.f;
    G.w = dcf[s];
    const TS *row = data + (uint64_t)s * C;
    if constexpr (NCB == 4 && PAIR) {
        // coils ca, ca + 1 in one 16-byte load (C even, ca even: aligned)
        const float4 v = ld_pair(row, ca);
        G.a = make_float2(v.x, v.y);
        G.b = make_float2(v.z, v.w);
    } else {
        G.a = ld_coil(row, ca);
        if constexpr (NCB == 4) G.b = ld_coil(row, cbb);
    }
}

template <int NBLK, int NCB>
__device__ __forceinline__ void apply_group(f4 (&acc)[NBLK][NCB], const group_regs<NCB> &G, bool ok,
                                            const float (&fx)[NBLK], const float (&fy)[NBLK],
                                            const float (&fz)[NBLK], const bool (&cok)[NBLK], bool v0ok, bool v1ok,
                                            int j, const float2 *pairs, const float *slut, float dacc, float cidx,
                                            int L)
{
    // B = dcf * sample (GridLut.cpp:109).  Padding columns (coils past the
    // chunk) and records past the item's end get a zero DCF here (their
    // loads read valid rows); the records also get zero weights in A.
    const float w0 = ok && v0ok ? G.w : 0.f;
    float bv[NCB];
    if constexpr (NCB == 4) {
        const float w1 = ok && v1ok ? G.w : 0.f;
        bv[0] = __fmul_rn(w0, G.a.x); bv[1] = __fmul_rn(w0, G.a.y);
        bv[2] = __fmul_rn(w1, G.b.x); bv[3] = __fmul_rn(w1, G.b.y);
    } else if constexpr (NCB == 2) {
        bv[0] = __fmul_rn(w0, G.a.x); bv[1] = __fmul_rn(w0, G.a.y);
    } else {
        bv[0] = __fmul_rn(w0, (j & 1) ? G.a.y : G.a.x);
    }
#pragma unroll
    for (int b = 0; b < NBLK; ++b) {
        const float aw = pair_weight(fx[b], fy[b], fz[b], G.px, G.py, G.pz, ok && cok[b], pairs, slut, dacc, cidx, L);
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[b][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aw, bv[cb], acc[b][cb], 0, 0, 0);
    }
}

constexpr int BG = 4;          // groups of four records per prefetched batch
// the pipeline reads headers up to 8 BG + 15 records past a live record; the
// plan pads H by 64 records
static_assert(8 * BG + 15 < 64, "record header padding");

// ZM: 0 = 2D, 1 = 3D, 2 = batched 2D frames (third trajectory row = frame).
// PAIR: C even, so lane j's two coils (2j, 2j+1 of a 32-coil chunk) are one
// aligned 16-byte (8-byte for FP16) load.
template <int NBLK, int NCB, int ZM, bool PAIR, typename TS>
__global__ void __launch_bounds__(64) k_grid_mfma(const fr_item *__restrict__ items, int64_t nitems,
                                                  const fr_unit *__restrict__ units, const uint2 *__restrict__ H,
                                                  const float *__restrict__ traj, int64_t M, int frames,
                                                  const TS *__restrict__ data, const float *__restrict__ dcf, int C,
                                                  int N, size_t cells, float2 *__restrict__ out,
                                                  float2 *__restrict__ partial, int64_t nslots, int cells_max,
                                                  const float *__restrict__ lut, const float2 *__restrict__ gpairs,
                                                  int L, float dacc, float cidx)
{
    // LDS: threshold pairs [L + 1] | LUT [L], staged once per wave
    extern __shared__ float2 smem[];
    float2 *pairs = smem;
    float *slut = reinterpret_cast<float *>(smem + L + 1);
    const int lane = threadIdx.x;
    for (int q = lane; q <= L; q += 64) pairs[q] = gpairs[q];
    for (int q = lane; q < L; q += 64) slut[q] = lut[q];
    __syncthreads();

    const int j = lane & 15;                    // A row (cell in block) / B column (component)
    const int k = lane >> 4;                    // A column / B row (record in group)
    const int chunk = 8 * NCB;                  // coils per chunk
    const int c0 = blockIdx.y * chunk;
    const int nc = min(chunk, C - c0);
    const float *tx = traj, *ty = traj + M, *tz = traj + 2 * M;
    // lane j's coils (NCB = 4: 2j, 2j+1; 2: j; 1: j/2), clamped to a valid
    // coil for the load and masked through the DCF
    const int cl0 = NCB == 4 ? 2 * j : (NCB == 2 ? j : j >> 1);
    const bool v0ok = cl0 < nc, v1ok = NCB == 4 && cl0 + 1 < nc;
    // PAIR (C and so nc even): the pair's first coil stays even and the pair
    // inside the chunk, ca + 1 <= c0 + nc - 1 < C
    const int ca = PAIR ? c0 + min(cl0, nc - 2) : c0 + min(cl0, nc - 1), cbb = c0 + min(cl0 + 1, nc - 1);

    // Persistent waves.  Round t covers items [t W, t W + W), W = gridDim.x;
    // workgroups are dealt to the 8 XCDs round-robin, so XCD x's waves take
    // the contiguous range x W/8 .. (x+1) W/8 - 1 of each round: neighbouring
    // units (which share samples) run on one XCD and share its L2.
    const int64_t W = gridDim.x;
    const int64_t slot_in_round = (int64_t)(blockIdx.x % 8) * (W / 8) + blockIdx.x / 8;
    for (int64_t i = slot_in_round; i < nitems; i += W) {
        const fr_item it = items[i];
        const fr_unit U = units[it.unit];
        const int ux = U.ext & 255, uy = (U.ext >> 8) & 255, uz = U.ext >> 16;
        const int ncell = ux * uy * uz;

        float fx[NBLK], fy[NBLK], fz[NBLK];
        bool cok[NBLK];
#pragma unroll
        for (int b = 0; b < NBLK; ++b) {
            const int q = 16 * b + j;
            cok[b] = q < ncell;
            const int qq = cok[b] ? q : 0;
            fx[b] = (float)(U.ox + qq % ux);
            fy[b] = (float)(U.oy + (qq / ux) % uy);
            fz[b] = (float)(U.oz + qq / (ux * uy));
        }

        f4 acc[NBLK][NCB];
#pragma unroll
        for (int b = 0; b < NBLK; ++b)
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) acc[b][cb] = f4{0.f, 0.f, 0.f, 0.f};

        if (it.r0 < it.r1) {
            // Software pipeline over batches of 16 records (4 groups): while
            // batch b is applied, batch b+1's coordinates and sample rows
            // (64 rows per wave) and batch b+2's sample indices are in
            // flight.  Records past the item's end read valid (padded)
            // headers and are masked; groups wholly past it are skipped.
            const uint32_t r1 = it.r1;
            uint32_t r = it.r0;
            group_regs<NCB> A[BG], B[BG];
            uint32_t hn[BG], h2[BG];
#pragma unroll
            for (int q = 0; q < BG; ++q) hn[q] = H[r + 4 * q + k].x;
#pragma unroll
            for (int q = 0; q < BG; ++q)
                load_group<NCB, ZM, PAIR>(A[q], hn[q], tx, ty, tz, frames, data, dcf, (uint32_t)C, ca, cbb);
#pragma unroll
            for (int q = 0; q < BG; ++q) hn[q] = H[r + 4 * BG + 4 * q + k].x;
            while (true) {
#pragma unroll
                for (int q = 0; q < BG; ++q) h2[q] = H[r + 8 * BG + 4 * q + k].x;
#pragma unroll
                for (int q = 0; q < BG; ++q)
                    load_group<NCB, ZM, PAIR>(B[q], hn[q], tx, ty, tz, frames, data, dcf, (uint32_t)C, ca, cbb);
                // keep the prefetch ahead of the compute (the scheduler would
                // otherwise sink each load to its use and wait on it there)
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < BG; ++q)
                    apply_group<NBLK, NCB>(acc, A[q], r + 4 * q + k < r1, fx, fy, fz, cok, v0ok, v1ok, j, pairs,
                                           slut, dacc, cidx, L);
                r += 4 * BG;
                if (r >= r1) break;
#pragma unroll
                for (int q = 0; q < BG; ++q) hn[q] = H[r + 8 * BG + 4 * q + k].x;
#pragma unroll
                for (int q = 0; q < BG; ++q)
                    load_group<NCB, ZM, PAIR>(A[q], h2[q], tx, ty, tz, frames, data, dcf, (uint32_t)C, ca, cbb);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < BG; ++q)
                    apply_group<NBLK, NCB>(acc, B[q], r + 4 * q + k < r1, fx, fy, fz, cok, v0ok, v1ok, j, pairs,
                                           slut, dacc, cidx, L);
                r += 4 * BG;
                if (r >= r1) break;
            }
        }

        // accumulator register r of block b: cell 16 b + 4 k + r, component slice j
#pragma unroll
        for (int b = 0; b < NBLK; ++b) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = 16 * b + 4 * k + r;
                // NCB = 1: lanes j (re) and j ^ 1 (im) hold one coil
                const float other = NCB == 1 ? __shfl_xor(acc[b][0][r], 1) : 0.f;
                if (q >= ncell) continue;
                const int lx = q % ux, ly = (q / ux) % uy, lz = q / (ux * uy);
                const size_t gi = ((size_t)(U.oz + lz) * N + (U.oy + ly)) * N + (U.ox + lx);
                auto put = [&](int cl, float re, float im) {
                    const int c = c0 + cl;
                    if (it.slot < 0) {
                        out[(size_t)c * cells + gi] = make_float2(re, im);
                    } else {
                        const size_t slot = ((size_t)(c >> 5) * nslots + it.slot) * 32 + (c & 31);
                        partial[slot * cells_max + q] = make_float2(re, im);
                    }
                };
                if constexpr (NCB == 4) {
                    if (v0ok) put(cl0, acc[b][0][r], acc[b][1][r]);
                    if (v1ok) put(cl0 + 1, acc[b][2][r], acc[b][3][r]);
                } else if constexpr (NCB == 2) {
                    if (v0ok) put(cl0, acc[b][0][r], acc[b][1][r]);
                } else {
                    if ((j & 1) == 0 && v0ok) put(cl0, acc[b][0][r], other);
                }
            }
        }
    }
}

template <int NBLK, int NCB, int ZM, bool PAIR, typename TS>
fr_status launch_mfma(fr_plan *p, const fr_samples *s, float2 *out)
{
    fr_ctx *ctx = p->ctx;
    const size_t cells = fr_plan_cells(p->dim, p->N, p->frames);
    const int chunks = (s->C + 8 * NCB - 1) / (8 * NCB);
    // persistent waves: a multiple of 8 (one share per XCD), at most one per item
    const int64_t want = (int64_t)std::max(1, ctx->cu_count) * FR_MFMA_WAVES_PER_CU;
    const int64_t W = std::max<int64_t>(8, std::min<int64_t>(want, (p->nitems + 7) / 8 * 8));
    const size_t lds = sizeof(float2) * (p->L + 1) + sizeof(float) * p->L;
    hipLaunchKernelGGL((k_grid_mfma<NBLK, NCB, ZM, PAIR, TS>), dim3((unsigned)W, chunks), dim3(64), lds, ctx->stream,
                       p->items, p->nitems, p->units, p->hdrs, s->traj, s->M, p->frames, (const TS *)s->data, s->dcf,
                       s->C, (int)p->N, cells, out, (float2 *)p->partial, p->nslots, p->geom.cells_max, p->lut,
                       p->mfma_pairs, p->L, p->mfma_dacc, p->mfma_cidx);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

template <int NBLK, int NCB, bool PAIR, typename TS>
fr_status launch_zm(fr_plan *p, const fr_samples *s, float2 *out)
{
    if (p->frames > 0) return launch_mfma<NBLK, NCB, 2, PAIR, TS>(p, s, out);
    if (p->dim == 3) return launch_mfma<NBLK, NCB, 1, PAIR, TS>(p, s, out);
    return launch_mfma<NBLK, NCB, 0, PAIR, TS>(p, s, out);
}

template <int NBLK, typename TS>
fr_status launch_ncb(fr_plan *p, const fr_samples *s, float2 *out)
{
    if (s->C <= 8) return launch_zm<NBLK, 1, false, TS>(p, s, out);
    if (s->C <= 16) return launch_zm<NBLK, 2, false, TS>(p, s, out);
    if (s->C % 2 == 0) return launch_zm<NBLK, 4, true, TS>(p, s, out);
    return launch_zm<NBLK, 4, false, TS>(p, s, out);
}

template <typename TS>
fr_status launch_nblk(fr_plan *p, const fr_samples *s, float2 *out)
{
    const int cm = p->geom.cells_max;
    if (cm <= 16) return launch_ncb<1, TS>(p, s, out);
    if (cm <= 32) return launch_ncb<2, TS>(p, s, out);
    return launch_ncb<4, TS>(p, s, out);
}

}  // namespace

fr_status fr_launch_grid_mfma(fr_plan *p, const fr_samples *s, float2 *out)
{
    return s->type == FR_SAMPLES_C64 ? launch_nblk<float2>(p, s, out) : launch_nblk<__half2>(p, s, out);
}
