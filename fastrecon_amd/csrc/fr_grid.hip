// Gridding execute: one wave per work unit, all coils (in chunks of 32) of
// a unit accumulated in that wave's slice of LDS.
//
// Replaces cuGridLut::griddingChannel / gridding_kernel
// (ReconData/cuGridLut.cu:231-309) and the CPU loop GridLut::griddingChannel
// (GridLut.cpp:80-147).
//
// Mapping (gfx950, wave64):
//   * samples are sample-major [M][C] complex, so the (up to) 32 coils of a
//     sample are one contiguous 256-byte read: lane p holds component p
//     (coil p/2, re or im), scaled by the DCF exactly like the reference's
//     `dcf * sample`;
//   * the unit's accumulator is [cell][2*coils + 1 pad] FP32 in LDS; a tap is
//     one multiply w * d and one LDS add into [cell][p] (consecutive words,
//     bank-conflict free);
//   * record headers are fetched 64 at a time with one vector load and read
//     with readlane; taps are fetched 64 at a time into VGPRs one chunk
//     ahead, staged through a 512-byte LDS buffer and read by broadcast; the
//     next record's sample data is requested before the current record's
//     taps are applied.  This keeps several HBM requests in flight per wave.
// Each cell belongs to exactly one unit, its wave walks the unit's records
// in ascending sample order, so every cell sums its contributions in the
// reference CPU order with the same FP32 operations: deterministic, no
// global atomics, bit-identical to the CPU path.  A unit's cells are written
// once, coalesced along x.
#include "fr_internal.h"

#include <algorithm>

namespace {

template <typename TS>
__device__ __forceinline__ float load_comp(const TS *__restrict__ data, uint32_t s, int C, int lanecomp);

template <>
__device__ __forceinline__ float load_comp<float2>(const float2 *__restrict__ data, uint32_t s, int C, int lanecomp)
{
    return reinterpret_cast<const float *>(data)[(size_t)s * C * 2 + lanecomp];
}

template <>
__device__ __forceinline__ float load_comp<__half2>(const __half2 *__restrict__ data, uint32_t s, int C, int lanecomp)
{
    return __half2float(reinterpret_cast<const __half *>(data)[(size_t)s * C * 2 + lanecomp]);
}

__device__ __forceinline__ uint32_t rl(uint32_t v, int lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

// A zero the compiler cannot prove uniform.  Indexing with it turns a load
// at a wave-uniform address into a vector load: vector loads complete in
// order and are waited on one at a time (vmcnt), whereas scalar loads may
// complete out of order, so waiting on any one of them drains them all and
// every prefetch in flight.
__device__ __forceinline__ uint32_t opaque_zero()
{
    uint32_t z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

constexpr int BATCH = 16;          // records per batch
constexpr int64_t XCD_RUN = 128;   // consecutive items one XCD takes in turn
constexpr int ACC_STRIDE = FR_ACC_STRIDE;   // floats per accumulator row (64 components + pad)
constexpr int STAGE = 4;           // 64-tap chunks staged per batch (256 taps)

// One batch of records [base, base + k) with taps from tb; its loads are
// issued by prep() and consumed by compute() one batch later.
template <typename TS>
struct batch_regs {
    float d[BATCH];                // lane p: component p of each record's sample
    float w;                       // lane j: DCF of record j
    uint32_t nv;                   // lane j: tap count of record j (a multiple of 4)
    uint2 ch[STAGE];               // lane l: tap cb + 64 q + l
    uint32_t cb, tb;
    int k;
};

template <typename TS>
__device__ __forceinline__ void prep(batch_regs<TS> &B, uint2 &h, uint32_t &nb, uint32_t &ntb, uint32_t re,
                                     const uint2 *__restrict__ H, const uint2 *__restrict__ T,
                                     const TS *__restrict__ data, const float *__restrict__ dcf, int C, int lc,
                                     uint32_t vz, int lane)
{
    // h holds the headers of records nb .. nb + 63 (lane l: record nb + l).
    // The batch takes the longest run of records (at most BATCH) whose taps
    // fit the stage: an inclusive prefix sum of the tap counts over lanes
    // 0 .. BATCH-1 (within one 16-lane row), then a ballot.
    const uint32_t cb = ntb & ~63u;
    const bool live = lane < BATCH && nb + (uint32_t)lane < re;
    uint32_t pre = live ? h.y : 0u;
#pragma unroll
    for (int d = 1; d < BATCH; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)pre, d, BATCH);
        if ((lane & (BATCH - 1)) >= d) pre += t;
    }
    // records carry at most 64 taps and the stage holds 64 STAGE, so k >= 1
    const bool fits = live && (ntb - cb) + pre <= 64u * STAGE;
    const int k = __popcll(__ballot(fits));
    const uint32_t te = ntb + (k > 0 ? rl(pre, k - 1) : 0u);
    B.k = k;
    B.cb = cb;
    B.tb = ntb;
    B.nv = h.y;                                // tap counts of the batch's records
#pragma clang loop unroll(full)
    for (int j = 0; j < BATCH; ++j) {          // unconditional: past-the-end records read a valid row
        const uint32_t s = rl(h.x, j < k ? j : 0);
        B.d[j] = load_comp<TS>(data, s, C, lc);
    }
    B.w = dcf[h.x + vz];
#pragma clang loop unroll(full)
    for (int q = 0; q < STAGE; ++q) B.ch[q] = T[cb + 64 * q + lane];
    nb += k;
    ntb = te;
    h = H[nb + lane];                          // next batch's headers (H is padded)
}

// Apply the taps of batch B.  A record's taps are distinct cells (padding
// taps go to the spare row), so each group's accumulator words are read
// together, updated and written back; LDS keeps a wave's accesses in order,
// so the next group and the next record see these writes.  Tap cells are
// byte offsets of the cell's accumulator row.
template <int ACC, typename TS>
__device__ __forceinline__ void compute(const batch_regs<TS> &B, uint2 *stage, char *accb, uint32_t lane4)
{
#pragma clang loop unroll(full)
    for (int q = 0; q < STAGE; ++q) stage[64 * q + (lane4 >> 2)] = B.ch[q];
    uint32_t off = B.tb - B.cb;
#pragma clang loop unroll(full)
    for (int j = 0; j < BATCH; ++j) {
        const uint32_t n = j < B.k ? rl(B.nv, j) : 0u;
        const float w = __uint_as_float(rl(__float_as_uint(B.w), j));
        const float d = __fmul_rn(w, B.d[j]);              // dcf * sample (GridLut.cpp:109)
        const uint2 *tp = stage + off;
        // records average under two groups: a plain loop (no unrolled body
        // plus remainder) keeps the per-record set-up short
#pragma clang loop unroll(disable)
        for (uint32_t q = 0; q < n; q += FR_TAP_GROUP) {
            float *p[FR_TAP_GROUP];
            float v[FR_TAP_GROUP], x[FR_TAP_GROUP];
#pragma unroll
            for (int u = 0; u < FR_TAP_GROUP; ++u) {
                const uint2 a = tp[q + u];
                p[u] = reinterpret_cast<float *>(accb + (a.x + lane4));
                v[u] = __fmul_rn(__uint_as_float(a.y), d);
            }
            if constexpr (ACC == 0) {
#pragma unroll
                for (int u = 0; u < FR_TAP_GROUP; ++u) x[u] = *p[u];
#pragma unroll
                for (int u = 0; u < FR_TAP_GROUP; ++u) *p[u] = __fadd_rn(x[u], v[u]);
            } else {
#pragma unroll
                for (int u = 0; u < FR_TAP_GROUP; ++u) atomicAdd(p[u], v[u]);   // LDS float atomic (slower)
            }
        }
        off += n;
    }
}

template <int ACC, typename TS>
__global__ void __launch_bounds__(64) k_grid_items(const fr_item *__restrict__ items, int64_t nitems,
                                                   const fr_unit *__restrict__ units,
                                                   const uint2 *__restrict__ H, const uint2 *__restrict__ T,
                                                   const TS *__restrict__ data, const float *__restrict__ dcf,
                                                   int C, int stride, int acc_floats, int N, size_t cells,
                                                   float2 *__restrict__ out, float2 *__restrict__ partial,
                                                   int64_t nslots, int cells_max, int64_t nfront)
{
    // LDS: accumulator rows [cells_max + 1][stride] (the last absorbs padding taps) | tap stage
    extern __shared__ float lds[];
    const int lane = threadIdx.x;
    // Items of the heaviest units run first, one per workgroup.  Workgroups
    // are dealt to the 8 XCDs round-robin; the remaining (Morton-ordered)
    // items are remapped so that each XCD takes runs of XCD_RUN consecutive
    // items, i.e. neighbouring units that share samples in its L2, while the
    // runs rotate over the XCDs to keep their loads even.
    const int64_t bid = blockIdx.x;
    int64_t i;
    if (bid < nfront) {
        i = bid;
    } else {
        const int64_t r = bid - nfront;
        const int64_t g = r / 8;
        const int64_t li = (g / XCD_RUN) * (8 * XCD_RUN) + (bid % 8) * XCD_RUN + g % XCD_RUN;
        i = nfront + li;
    }
    if (i >= nitems) return;
    float *acc = lds;
    uint2 *stage = reinterpret_cast<uint2 *>(lds + acc_floats);

    const fr_item it = items[i];
    const fr_unit U = units[it.unit];
    const int ux = U.ext & 255, uy = (U.ext >> 8) & 255, uz = U.ext >> 16;
    const int ncell = ux * uy * uz;
    for (int q = lane; q < ncell * ACC_STRIDE; q += 64) acc[q] = 0.f;

    const int c0 = blockIdx.y * 32;
    const int nc = min(32, C - c0);
    // lanes past the chunk's 2*nc components read a valid component and
    // accumulate into their own columns of the row, which are never read
    const int lc = lane < 2 * nc ? 2 * c0 + lane : 2 * c0;
    const uint32_t vz = opaque_zero();
    char *accb = reinterpret_cast<char *>(acc);
    const uint32_t lane4 = 4u * lane;

    const uint32_t re = it.r1;
    uint32_t nb = it.r0, ntb = it.t0;
    if (nb < re) {
        // Software pipeline, unrolled twice so buffers A and B never copy:
        // batch B's loads are in flight while batch A is applied.
        uint2 h = H[nb + lane];
        batch_regs<TS> A, B;
        prep(A, h, nb, ntb, re, H, T, data, dcf, C, lc, vz, lane);
        while (true) {
            prep(B, h, nb, ntb, re, H, T, data, dcf, C, lc, vz, lane);
            compute<ACC>(A, stage, accb, lane4);
            if (B.k == 0) break;
            prep(A, h, nb, ntb, re, H, T, data, dcf, C, lc, vz, lane);
            compute<ACC>(B, stage, accb, lane4);
            if (A.k == 0) break;
        }
    }

    if (it.slot < 0) {
        // coil-major grids [C][N^dim], x fastest
        for (int q = lane; q < nc * ncell; q += 64) {
            const int c = q / ncell;
            const int cell = q - c * ncell;
            const int lx = cell % ux;
            const int ly = (cell / ux) % uy;
            const int lz = cell / (ux * uy);
            const size_t g = ((size_t)(U.oz + lz) * N + (U.oy + ly)) * N + (U.ox + lx);
            out[(size_t)(c0 + c) * cells + g] = make_float2(acc[cell * ACC_STRIDE + 2 * c], acc[cell * ACC_STRIDE + 2 * c + 1]);
        }
    } else {
        float2 *pp = partial + ((size_t)blockIdx.y * nslots + it.slot) * (size_t)(32 * cells_max);
        for (int q = lane; q < nc * ncell; q += 64) {
            const int c = q / ncell;
            const int cell = q - c * ncell;
            pp[(size_t)c * cells_max + cell] = make_float2(acc[cell * ACC_STRIDE + 2 * c], acc[cell * ACC_STRIDE + 2 * c + 1]);
        }
    }
}


// ------------------------------------------------------------------------
// Gather kernel (FR_ACC_GATHER).  Same items and summation order, different
// data flow: an item's records are processed in batches of GB; the batch's
// sample rows, scaled by the DCF (d' = dcf * s, GridLut.cpp:109), are staged
// in LDS as rows of 64 components, and its taps come sorted by cell.  The
// unit's (at most 32) cells live in VGPRs, one per cell, indexed by the
// unrolled cell loop, so a tap is one LDS read of the sample's row at the
// wave's 64 components, one multiply and one add into a register: half the
// LDS traffic of a read-add-write and no LDS write per tap.  Within a cell
// the taps are in record (= ascending sample) order and the batches are in
// order, so every cell sums in the reference order (exact mode: bitwise).
// Padding taps read the batch's zero row with weight 0 (exact no-ops).
constexpr int GB = FR_GATHER_BATCH;
constexpr int GCELLS = FR_GATHER_CELLS;

template <typename TS>
struct gbatch {
    float d[GB];                   // lane p: component p of record j's sample
    float w;                       // lane j: DCF of record j
    uint32_t k;                    // records
};

template <typename TS>
__device__ __forceinline__ void gprep(gbatch<TS> &B, const fr_batch *__restrict__ batches, uint32_t bi, uint32_t nb,
                                      const uint2 *__restrict__ H, const TS *__restrict__ data,
                                      const float *__restrict__ dcf, int C, int lc, uint32_t vz, int lane)
{
    const bool ok = bi < nb;
    const fr_batch Bd = batches[ok ? bi : 0];
    const uint32_t k = ok ? Bd.nrec : 0u;
    B.k = k;
    const uint2 h = H[Bd.r0 + (lane < GB ? lane : 0) + vz];      // lane j: header of record j (H is padded)
#pragma clang loop unroll(full)
    for (int j = 0; j < GB; ++j) {
        const uint32_t sidx = rl(h.x, (uint32_t)j < k ? j : 0);
        B.d[j] = load_comp<TS>(data, sidx, C, lc);
    }
    B.w = dcf[h.x];
}

// rows of batch B -> LDS (row j at j * 64 floats)
template <typename TS>
__device__ __forceinline__ void gstage(const gbatch<TS> &B, float *rows, int lane)
{
#pragma clang loop unroll(full)
    for (int j = 0; j < GB; ++j) {
        const float w = __uint_as_float(rl(__float_as_uint(B.w), j));
        if ((uint32_t)j < B.k) rows[j * 64 + lane] = __fmul_rn(w, B.d[j]);   // dcf * sample
    }
}

__device__ __forceinline__ void gtaps(float (&acc)[GCELLS], const fr_batch &Bd, const uint2 *__restrict__ T,
                                      const char *rowsb, uint32_t lane4)
{
    uint64_t t = Bd.toff;
#pragma clang loop unroll(full)
    for (int c = 0; c < GCELLS; ++c) {
        const uint32_t n = (Bd.cnt[c >> 2] >> (8 * (c & 3))) & 255u;
        const uint2 *tp = T + t;
        for (uint32_t q = 0; q < n; q += 4) {
            float v[4], x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint2 a = tp[q + u];
                x[u] = *reinterpret_cast<const float *>(rowsb + a.x + lane4);
                v[u] = __uint_as_float(a.y);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[c] = __fadd_rn(acc[c], __fmul_rn(v[u], x[u]));
        }
        t += n;
    }
}

template <typename TS>
__global__ void __launch_bounds__(64) k_grid_gather(const fr_item *__restrict__ items, int64_t nitems,
                                                    const fr_unit *__restrict__ units,
                                                    const fr_batch *__restrict__ batches,
                                                    const uint2 *__restrict__ H, const uint2 *__restrict__ T,
                                                    const TS *__restrict__ data, const float *__restrict__ dcf, int C,
                                                    int N, size_t cells, float2 *__restrict__ out,
                                                    float2 *__restrict__ partial, int64_t nslots, int cells_max,
                                                    int64_t nfront)
{
    // LDS: rows [GB + 1][64] (row GB stays zero); reused as [32][ACC_STRIDE] for the write-out
    extern __shared__ float lds[];
    const int lane = threadIdx.x;
    const int64_t bid = blockIdx.x;
    int64_t i;
    if (bid < nfront) {
        i = bid;
    } else {
        const int64_t r = bid - nfront;
        const int64_t g = r / 8;
        const int64_t li = (g / XCD_RUN) * (8 * XCD_RUN) + (bid % 8) * XCD_RUN + g % XCD_RUN;
        i = nfront + li;
    }
    if (i >= nitems) return;
    const fr_item it = items[i];
    const fr_unit U = units[it.unit];
    const int ux = U.ext & 255, uy = (U.ext >> 8) & 255, uz = U.ext >> 16;
    const int ncell = ux * uy * uz;
    const int c0 = blockIdx.y * 32;
    const int nc = min(32, C - c0);
    const int lc = lane < 2 * nc ? 2 * c0 + lane : 2 * c0;
    const uint32_t vz = opaque_zero();
    const uint32_t lane4 = 4u * lane;
    lds[GB * 64 + lane] = 0.f;
    float acc[GCELLS];
#pragma unroll
    for (int c = 0; c < GCELLS; ++c) acc[c] = 0.f;

    const uint32_t nb = (it.r1 - it.r0 + GB - 1) / GB;
    const uint32_t b0 = it.t0;
    const char *rowsb = reinterpret_cast<const char *>(lds);
    if (nb > 0) {
        gbatch<TS> A, B;
        gprep(A, batches, b0, b0 + nb, H, data, dcf, C, lc, vz, lane);
        for (uint32_t bi = b0;; bi += 2) {
            // batch bi (in A): stage, prefetch bi + 1 into B, apply
            gstage(A, lds, lane);
            gprep(B, batches, bi + 1, b0 + nb, H, data, dcf, C, lc, vz, lane);
            gtaps(acc, batches[bi], T, rowsb, lane4);
            if (bi + 1 >= b0 + nb) break;
            gstage(B, lds, lane);
            gprep(A, batches, bi + 2, b0 + nb, H, data, dcf, C, lc, vz, lane);
            gtaps(acc, batches[bi + 1], T, rowsb, lane4);
            if (bi + 2 >= b0 + nb) break;
        }
    }

    // accumulators -> LDS rows [cell][ACC_STRIDE], then coalesced writes
    __syncthreads();
#pragma unroll
    for (int c = 0; c < GCELLS; ++c) lds[c * ACC_STRIDE + lane] = acc[c];
    __syncthreads();
    const float *a = lds;
    if (it.slot < 0) {
        for (int q = lane; q < nc * ncell; q += 64) {
            const int c = q / ncell;
            const int cell = q - c * ncell;
            const int lx = cell % ux;
            const int ly = (cell / ux) % uy;
            const int lz = cell / (ux * uy);
            const size_t g = ((size_t)(U.oz + lz) * N + (U.oy + ly)) * N + (U.ox + lx);
            out[(size_t)(c0 + c) * cells + g] = make_float2(a[cell * ACC_STRIDE + 2 * c], a[cell * ACC_STRIDE + 2 * c + 1]);
        }
    } else {
        float2 *pp = partial + ((size_t)blockIdx.y * nslots + it.slot) * (size_t)(32 * cells_max);
        for (int q = lane; q < nc * ncell; q += 64) {
            const int c = q / ncell;
            const int cell = q - c * ncell;
            pp[(size_t)c * cells_max + cell] = make_float2(a[cell * ACC_STRIDE + 2 * c], a[cell * ACC_STRIDE + 2 * c + 1]);
        }
    }
}

// Heavy units: sum their slices' partial boxes in slice order, write the grid.
__global__ void k_combine(const fr_heavy *__restrict__ heavy, const fr_unit *__restrict__ units,
                          const float2 *__restrict__ partial, int64_t nslots, int cells_max, int C, int N, size_t cells,
                          float2 *__restrict__ out)
{
    const fr_heavy h = heavy[blockIdx.x];
    const fr_unit U = units[h.unit];
    const int ux = U.ext & 255, uy = (U.ext >> 8) & 255, uz = U.ext >> 16;
    const int ncell = ux * uy * uz;
    const int c0 = blockIdx.y * 32;
    const int nc = min(32, C - c0);
    const size_t slot_stride = (size_t)32 * cells_max;
    const float2 *base = partial + ((size_t)blockIdx.y * nslots + h.slot0) * slot_stride;
    for (int q = threadIdx.x; q < nc * ncell; q += blockDim.x) {
        const int c = q / ncell;
        const int cell = q - c * ncell;
        float2 a = make_float2(0.f, 0.f);
        for (uint32_t k = 0; k < h.nitems; ++k) {
            const float2 v = base[k * slot_stride + (size_t)c * cells_max + cell];
            a.x = __fadd_rn(a.x, v.x);
            a.y = __fadd_rn(a.y, v.y);
        }
        const int lx = cell % ux;
        const int ly = (cell / ux) % uy;
        const int lz = cell / (ux * uy);
        const size_t g = ((size_t)(U.oz + lz) * N + (U.oy + ly)) * N + (U.ox + lx);
        out[(size_t)(c0 + c) * cells + g] = a;
    }
}

template <int ACC, typename TS>
fr_status launch(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial)
{
    fr_ctx *ctx = p->ctx;
    const fr_geom &g = p->geom;
    const int stride = ACC_STRIDE;
    // the tap stage behind the accumulator must be 16-byte aligned (64-bit LDS reads)
    const int acc_floats = ((g.cells_max + 1) * stride + 3) & ~3;
    const size_t lds = (size_t)acc_floats * sizeof(float) + sizeof(uint2) * 64 * STAGE;
    const int chunks = (s->C + 31) / 32;
    const size_t cells = fr_plan_cells(p->dim, p->N, p->frames);
    const int64_t light = p->nitems - p->nitems_front;
    const int64_t span = 8 * XCD_RUN;
    dim3 grid((unsigned)(p->nitems_front + span * ((light + span - 1) / span)), chunks);
    if constexpr (ACC == FR_ACC_GATHER) {
        const size_t glds = sizeof(float) * (size_t)(GB + 1) * 64;
        hipLaunchKernelGGL(k_grid_gather<TS>, grid, dim3(64), glds, ctx->stream, p->items, p->nitems, p->units,
                           p->batches, p->hdrs, p->taps, (const TS *)s->data, s->dcf, s->C, (int)p->N, cells, out,
                           partial, p->nslots, g.cells_max, p->nitems_front);
    } else {
        auto kern = k_grid_items<ACC == FR_ACC_LDS_ATOMIC ? 1 : 0, TS>;
        FR_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(kern, grid, dim3(64), lds, ctx->stream, p->items, p->nitems, p->units, p->hdrs, p->taps,
                           (const TS *)s->data, s->dcf, s->C, stride, acc_floats, (int)p->N, cells, out,
                           partial, p->nslots, g.cells_max, p->nitems_front);
    }
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

}  // namespace

size_t fr_grid_scratch_bytes(const fr_plan *p, int C)
{
    const int chunks = (C + 31) / 32;
    return (size_t)chunks * p->nslots * 32 * p->geom.cells_max * sizeof(float2);
}

fr_status fr_launch_grid(fr_plan *p, const fr_samples *s, float2 *out, int force_mode, float2 *scratch,
                         size_t scratch_bytes, const fr_samples *dcf)
{
    fr_ctx *ctx = p->ctx;
    const fr_geom &g = p->geom;
    // heavy-unit slices: partial boxes [32-coil chunk][slot][32 coils][cells_max],
    // in the caller's scratch or in the plan's own buffer (grown on demand)
    const int chunks = (s->C + 31) / 32;
    const size_t need = fr_grid_scratch_bytes(p, s->C);
    float2 *partial = nullptr;
    if (scratch) {
        FR_CHECK(ctx, need <= scratch_bytes, FR_ERR_UNSUPPORTED,
                 "this plan's heavy-unit slices need %zu bytes of scratch for %d coils; the scratch holds %zu",
                 need, s->C, scratch_bytes);
        partial = scratch;
    } else {
        if (need > p->partial_bytes) {
            FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
            hipFree(p->partial);
            p->partial = nullptr;
            p->partial_bytes = 0;
            fr_status e = fr_alloc(ctx, (void **)&p->partial, need);
            if (e) return e;
            p->partial_bytes = need;
        }
        partial = (float2 *)p->partial;
    }
    if (p->nitems == 0) return FR_OK;
    const int mode = force_mode ? force_mode : p->accumulate;
    FR_CHECK(ctx, !dcf || mode == FR_ACC_POS, FR_ERR_STATE,
             "another handle's DCF applies to position-weight gridding only");
    const bool c64 = s->type == FR_SAMPLES_C64;
    fr_status e;
    switch (mode) {
    case FR_ACC_MFMA:
        e = fr_launch_grid_mfma(p, s, out, partial);
        break;
    case FR_ACC_POS:
        e = fr_launch_grid_pos(p, s, out, partial, 1, dcf);
        break;
    case FR_ACC_GATHER:
        e = c64 ? launch<FR_ACC_GATHER, float2>(p, s, out, partial) : launch<FR_ACC_GATHER, __half2>(p, s, out, partial);
        break;
    case FR_ACC_LDS_ATOMIC:
        e = c64 ? launch<FR_ACC_LDS_ATOMIC, float2>(p, s, out, partial)
                : launch<FR_ACC_LDS_ATOMIC, __half2>(p, s, out, partial);
        break;
    default:
        e = c64 ? launch<FR_ACC_LDS_RMW, float2>(p, s, out, partial) : launch<FR_ACC_LDS_RMW, __half2>(p, s, out, partial);
        break;
    }
    if (e) return e;
    if (p->nheavy > 0) {
        const size_t cells = fr_plan_cells(p->dim, p->N, p->frames);
        hipLaunchKernelGGL(k_combine, dim3((unsigned)p->nheavy, chunks), dim3(256), 0, ctx->stream, p->heavy,
                           p->units, (const float2 *)partial, p->nslots, g.cells_max, s->C, (int)p->N, cells, out);
        FR_HIP(ctx, hipGetLastError());
    }
    return FR_OK;
}

// Argument checks shared by the gridding entry points.
fr_status fr_grid_check(const fr_plan *p, const fr_samples *s, const fr_grids *out)
{
    fr_ctx *ctx = p->ctx;
    FR_CHECK(ctx, s->M == p->M && s->dim == p->dim, FR_ERR_INVALID,
             "samples (dim %d, M %lld) do not match the plan (dim %d, M %lld)", s->dim, (long long)s->M, p->dim,
             (long long)p->M);
    FR_CHECK(ctx, out->dim == p->dim && out->N == p->N && out->C == s->C && out->frames == p->frames,
             FR_ERR_INVALID,
             "grids (dim %d, N %u, C %d, frames %d) do not match plan/samples (dim %d, N %u, C %d, frames %d)",
             out->dim, out->N, out->C, out->frames, p->dim, p->N, s->C, p->frames);
    FR_CHECK(ctx, s->frames == p->frames, FR_ERR_INVALID, "samples and plan differ in frames");
    FR_CHECK(ctx, s->has_dcf, FR_ERR_STATE, "DCF not set");
    FR_CHECK(ctx, p->geom.cells_max <= 64, FR_ERR_INVALID, "plan unit box exceeds 64 cells");
    return FR_OK;
}

extern "C" fr_status fr_grid_execute(fr_plan *p, const fr_samples *s, fr_grids *out)
{
    if (!p || !s || !out) return FR_ERR_INVALID;
    fr_status e = fr_grid_check(p, s, out);
    if (e) return e;
    return fr_launch_grid(p, s, out->data, 0);
}
