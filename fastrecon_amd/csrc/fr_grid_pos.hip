// Gridding on the matrix cores with tap weights computed from the sample
// positions (accumulate mode FR_ACC_POS, the default for plans that are not
// exact-order).
//
// A unit is a box of at most 16 cells (4 x 2 x 2 by default, the most
// compact 16-cell box: a radius-2 kernel ball meets about 8 such boxes per
// sample, against 10 for 8 x 2 x 1).  The plan lists, per unit, the samples
// that have at least one accepted tap in it (ascending sample order, padded
// to a multiple of 4 with a null sample).  Four records at a time are one
// dense product per 16 components
//
//     acc[16 cells][16 components] += A[16 cells][4 records] * B[4 records][16 components]
//
// on v_mfma_f32_16x16x4_f32, with
//   A[i][k] = LUT weight of record k's sample at cell i, 0 where the sample
//             has no tap there: computed here, lane (i, k) evaluating the
//             reference tap test and LUT index for its own cell;
//   B[k][j] = dcf * sample component j of record k (GridLut.cpp:109); the
//             DCF travels in the .w of the record's position entry, which
//             k_pos_dcf refreshes from the samples before every pass.
//
// Tap test and LUT index (GridLut.cpp:112-136) without the reference's
// correctly rounded square root and division: the squared distance d2 is
// formed with the reference's FP32 operations in its order ((dz*dz + dy*dy)
// + dx*dx); the tap is accepted iff d2 < dacc, the least d2 whose sqrtf is >=
// hw; the reference index roundf(sqrtf(d2) / hw * (L - 1)) is monotone in d2,
// so with B[m] = least d2 whose index is >= m (computed on the host with the
// reference arithmetic) the index is the m with B[m] <= d2 < B[m + 1].  An
// estimate e = floor(sqrt~(d2) * (L - 1) / hw + 0.49) from the hardware's
// 1-ulp square root is the index or one below it, so one comparison with
// B[e + 1] selects LUT[e] or LUT[e + 1] from the table entry (B[e + 1],
// LUT[e], LUT[e + 1]).  Every tap set and LUT index is the reference's; only
// the order of the sum over samples differs (four products per matrix
// instruction), so results agree with the CPU loop to FP32 rounding and are
// deterministic.
//
// Cells a sample does not tap get weight 0, so a NaN or Inf sample value
// spreads to the other cells of the units it touches (0 * NaN), unlike the
// reference, which leaves those cells alone.
//
// Lane layout of v_mfma_f32_16x16x4_f32 (wave64): lane l supplies A[l % 16][l
// / 16] and B[l / 16][l % 16]; accumulator register r of lane l holds
// D[4 (l / 16) + r][l % 16].  Lane l evaluates cell l % 16 for record l / 16
// of a group, and holds component slice l % 16: with NCB component blocks,
// components [NCB j, NCB j + NCB) of a chunk of 8 NCB coils.
#include "fr_internal.h"

#include <algorithm>
#include <type_traits>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int RB = 16;               // records per batch: four groups of four
constexpr int PW = FR_POS_WAVES;     // waves per workgroup
constexpr int CS = 17;               // LDS row stride of the write-out transpose (16 cells + 1)
// the pipeline reads record indices up to 3 RB - 1 past a live record; the
// plan pads the record array by 64 entries
static_assert(3 * RB - 1 < 64, "record padding");

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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

// element at a 32-bit byte offset from a uniform base (the record, position
// and DCF arrays stay below 4 GiB: at most 2^27 samples)
template <typename E>
__device__ __forceinline__ E ld_off(const E *__restrict__ base, uint32_t byte_off)
{
    return *reinterpret_cast<const E *>(reinterpret_cast<const char *>(base) + byte_off);
}

// One batch of RB records in flight: lane (j, k) holds group q's record
// 4 q + k: its position, DCF and the lane's component slice of its row.
struct batch_st {
    float4 p[4];
    float w[4];
    float2 a[4], b[4];
};

__device__ __forceinline__ void load_ids(uint32_t (&id)[4], const uint32_t *__restrict__ recs, uint32_t rb, int k)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) id[q] = ld_off(recs, (rb + 4 * q + k) * (uint32_t)sizeof(uint32_t));
}

// The position entry carries the sample's DCF in .w (refreshed from the
// samples before every pass), so a record costs one 16-byte gather for
// position and weight.
template <int NCB, bool PAIR, typename TS>
__device__ __forceinline__ void issue_batch(batch_st &S, const uint32_t (&id)[4], const float4 *__restrict__ pos,
                                            const TS *__restrict__ data, uint32_t C, int ca, int cbb)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t s = id[q];
        S.p[q] = ld_off(pos, s * (uint32_t)sizeof(float4));
        S.w[q] = S.p[q].w;
        const TS *row = data + (uint64_t)s * C;
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

// Weight of the lane's cell for a sample at p: the reference tap test and
// LUT index through the selection table (file comment).
__device__ __forceinline__ float tap_weight(const float4 &p, const float4 *stbl, float fx, float fy, float fz,
                                            float dacc, float cidx, float lmax)
{
    // d2 with the reference's operations and order (GridLut.cpp:120-131)
    const float dz = __fsub_rn(fz, p.z);
    const float dy = __fsub_rn(fy, p.y);
    const float dx = __fsub_rn(fx, p.x);
    const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dz, dz), __fmul_rn(dy, dy)), __fmul_rn(dx, dx));
    // index estimate from the 1-ulp square root, then the exact selection
    const float est = fminf(__builtin_fmaf(__builtin_amdgcn_sqrtf(d2), cidx, 0.49f), lmax);
    const float4 e = stbl[(int)est];
    const float wsel = d2 >= e.x ? e.z : e.y;
    return d2 < dacc ? wsel : 0.f;                    // tap test sqrtf(d2) < hw (GridLut.cpp:133)
}

// B = dcf * sample (GridLut.cpp:109); padding coil columns get 0
template <int NCB, bool FULL>
__device__ __forceinline__ void b_values(float (&bv)[NCB], float w, float2 a, float2 b, bool v0ok, bool v1ok, int j)
{
    const float w0 = FULL || v0ok ? w : 0.f;
    if constexpr (NCB == 4) {
        const float w1 = FULL || v1ok ? w : 0.f;
        bv[0] = __fmul_rn(w0, a.x); bv[1] = __fmul_rn(w0, a.y);
        bv[2] = __fmul_rn(w1, b.x); bv[3] = __fmul_rn(w1, b.y);
    } else if constexpr (NCB == 2) {
        bv[0] = __fmul_rn(w0, a.x); bv[1] = __fmul_rn(w0, a.y);
    } else {
        bv[0] = __fmul_rn(w0, (j & 1) ? a.y : a.x);
    }
}

// Apply the four groups of one batch.  The null sample (index M) sits
// far outside the grid with DCF 0 and a zero row, so padding records add
// exact zeros.  The four weight computations are independent (their table
// reads overlap); the matrix instructions then cycle through the NCB
// accumulators.
template <int NCB, bool FULL>
__device__ __forceinline__ void apply_batch(f4 (&acc)[NCB], const batch_st &S, const float4 *stbl, float fx,
                                            float fy, float fz, float dacc, float cidx, float lmax, bool v0ok,
                                            bool v1ok, int j)
{
    float aw[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) aw[q] = tap_weight(S.p[q], stbl, fx, fy, fz, dacc, cidx, lmax);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float bv[NCB];
        b_values<NCB, FULL>(bv, S.w[q], S.a[q], S.b[q], v0ok, v1ok, j);
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(aw[q], bv[cb], acc[cb], 0, 0, 0);
    }
}

// Lane j's cell of an item's unit box (t0 = ox | oy << 16, pad = oz | (ux |
// uy << 5 | uz << 10) << 16): cell index j, x fastest; lanes past the box's
// cells sit far away, so every tap test rejects them.
struct item_box {
    int ox, oy, oz, ux, uy, uz, ncell;
};

__device__ __forceinline__ item_box unpack_box(uint32_t b0, uint32_t b1)
{
    item_box b;
    b.ox = (int)(b0 & 0xffffu);
    b.oy = (int)(b0 >> 16);
    b.oz = (int)(b1 & 0xffffu);
    b.ux = (int)((b1 >> 16) & 31u);
    b.uy = (int)((b1 >> 21) & 31u);
    b.uz = (int)((b1 >> 26) & 31u);
    b.ncell = b.ux * b.uy * b.uz;
    return b;
}

__device__ __forceinline__ void lane_cell(const item_box &b, int j, float &fx, float &fy, float &fz)
{
    fx = fy = fz = -3e30f;
    if (j < b.ncell) {
        fx = (float)(b.ox + j % b.ux);
        fy = (float)(b.oy + (j / b.ux) % b.uy);
        fz = (float)(b.oz + j / (b.ux * b.uy));
    }
}

// Grid the records [r, r1) of one item into acc: software pipeline over
// batches of RB records -- while batch b is applied, batch b+1's positions,
// DCFs and rows and batch b+2's sample indices are in flight (past r1 they
// read the next records or the 64 padding records at the end).
template <int NCB, bool PAIR, bool FULL, typename TS>
__device__ __forceinline__ void grid_records(f4 (&acc)[NCB], uint32_t r, uint32_t r1,
                                             const uint32_t *__restrict__ recs, const float4 *__restrict__ pos,
                                             const TS *__restrict__ data, uint32_t C,
                                             int ca, int cbb, const float4 *stbl, float fx, float fy, float fz,
                                             float dacc, float cidx, float lmax, bool v0ok, bool v1ok, int j, int k)
{
    uint32_t id0[4], id1[4];
    batch_st S0, S1;
    load_ids(id0, recs, r, k);
    load_ids(id1, recs, r + RB, k);
    issue_batch<NCB, PAIR>(S0, id0, pos, data, C, ca, cbb);
    while (true) {
        load_ids(id0, recs, r + 2 * RB, k);
        issue_batch<NCB, PAIR>(S1, id1, pos, data, C, ca, cbb);
        __builtin_amdgcn_sched_barrier(0);
        apply_batch<NCB, FULL>(acc, S0, stbl, fx, fy, fz, dacc, cidx, lmax, v0ok, v1ok, j);
        r += RB;
        if (r >= r1) break;
        load_ids(id1, recs, r + 2 * RB, k);
        issue_batch<NCB, PAIR>(S0, id0, pos, data, C, ca, cbb);
        __builtin_amdgcn_sched_barrier(0);
        apply_batch<NCB, FULL>(acc, S1, stbl, fx, fy, fz, dacc, cidx, lmax, v0ok, v1ok, j);
        r += RB;
        if (r >= r1) break;
    }
}

// grid_records over [r, r1) with a switch at rmid (a multiple of RB): the
// records up to rmid belong to one unit, the rest to the next; at rmid the
// pipeline keeps running and on_mid() moves the lane to the next unit's
// cell (fx, fy, fz) and accumulators.
template <int NCB, bool PAIR, bool FULL, typename TS, typename Mid>
__device__ __forceinline__ void grid_records_mid(f4 (&acc)[NCB], uint32_t r, uint32_t rmid, uint32_t r1,
                                                 const uint32_t *__restrict__ recs, const float4 *__restrict__ pos,
                                                 const TS *__restrict__ data, uint32_t C, int ca, int cbb,
                                                 const float4 *stbl, float &fx, float &fy, float &fz, float dacc,
                                                 float cidx, float lmax, bool v0ok, bool v1ok, int j, int k, Mid on_mid)
{
    uint32_t id0[4], id1[4];
    batch_st S0, S1;
    load_ids(id0, recs, r, k);
    load_ids(id1, recs, r + RB, k);
    issue_batch<NCB, PAIR>(S0, id0, pos, data, C, ca, cbb);
    while (true) {
        load_ids(id0, recs, r + 2 * RB, k);
        issue_batch<NCB, PAIR>(S1, id1, pos, data, C, ca, cbb);
        __builtin_amdgcn_sched_barrier(0);
        apply_batch<NCB, FULL>(acc, S0, stbl, fx, fy, fz, dacc, cidx, lmax, v0ok, v1ok, j);
        r += RB;
        if (r == rmid) on_mid();
        if (r >= r1) break;
        load_ids(id1, recs, r + 2 * RB, k);
        issue_batch<NCB, PAIR>(S0, id0, pos, data, C, ca, cbb);
        __builtin_amdgcn_sched_barrier(0);
        apply_batch<NCB, FULL>(acc, S1, stbl, fx, fy, fz, dacc, cidx, lmax, v0ok, v1ok, j);
        r += RB;
        if (r == rmid) on_mid();
        if (r >= r1) break;
    }
}

// NCB component blocks (chunks of 8 NCB coils).  PAIR: C even, so lane j's
// two coils (2j, 2j+1 of a 32-coil chunk) are one aligned 16-byte (8-byte
// for FP16) load.  FULL: every chunk holds 8 NCB coils.
//
// Three kinds of workgroup:
//  * front: one heavy-unit item per wave (a whole unit, or a slice written
//    to its scratch slot), written out by the wave;
//  * tile (when the plan has tiles): the PW waves grid the PW x-adjacent
//    units of a tile, stage their sums in LDS and store the tile in whole
//    rows of PW * bx cells (128 bytes for 4-cell-wide units);
//  * otherwise a run of ipw consecutive light items per wave, as one
//    record stream, each unit written out by its wave.
template <int NCB, bool PAIR, bool FULL, typename TS, int OCC>
__global__ void __launch_bounds__(64 * PW) __attribute__((amdgpu_waves_per_eu(OCC, 8))) k_grid_pos(
    const fr_item *__restrict__ items, int64_t nitems, const fr_tile *__restrict__ tiles, int64_t ntiles,
    int lbx, int lby, int lbz, const uint32_t *__restrict__ recs, const float4 *__restrict__ pos,
    const float4 *__restrict__ tbl, int L, float dacc, float cidx, const TS *__restrict__ data,
    int C, int N, size_t cells, float2 *__restrict__ out, float2 *__restrict__ partial,
    int64_t nslots, int cells_max, int64_t nfront, int64_t front_wg, int ipw, int64_t xr, int nt, int tw,
    int osm_floats, int64_t tstride, int fsub)
{
    extern __shared__ float4 smem_pos[];
    float4 *stbl = smem_pos;                                         // L entries
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    float *osm = reinterpret_cast<float *>(smem_pos + L) + wv * osm_floats;
    for (int q = threadIdx.x; q < L; q += 64 * PW) stbl[q] = tbl[q];
    __syncthreads();

    const int j = lane & 15;                    // A row (cell) / B column (component slice)
    const int k = lane >> 4;                    // A column / B row (record in group)
    const int chunk = 8 * NCB;                  // coils per chunk
    const int c0 = blockIdx.y * chunk;
    const int nc = min(chunk, C - c0);
    const int cl0 = NCB == 4 ? 2 * j : (NCB == 2 ? j : j >> 1);
    const bool v0ok = cl0 < nc, v1ok = NCB == 4 && cl0 + 1 < nc;
    const int ca = PAIR ? c0 + min(cl0, nc - 2) : c0 + min(cl0, nc - 1), cbb = c0 + min(cl0 + 1, nc - 1);
    const float lmax = (float)(L - 1);

    f4 acc[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[cb] = f4{0.f, 0.f, 0.f, 0.f};

    // Launch order: the front items first; then tiles (or runs of light
    // items) in Morton order, remapped so that each XCD (workgroups are dealt
    // to the 8 XCDs round-robin) takes runs of xr consecutive workgroups, a
    // compact patch of neighbouring units sharing samples in its L2.
    const int64_t bid = blockIdx.x;
    const int64_t rr = bid - front_wg, gg = rr / 8;
    const int64_t grp = (gg / xr) * (8 * xr) + (bid % 8) * xr + gg % xr;

    if (bid >= front_wg && ntiles > 0 && tw == PW) {
        // ---------------------------------------------------------- tile
        if (grp >= ntiles) return;              // the whole workgroup
        const fr_tile T = tiles[grp];
        // (a select, not T.item[wv]: a per-lane index would put T in scratch)
        const int32_t myitem = wv == 0 ? T.item[0] : wv == 1 ? T.item[1] : wv == 2 ? T.item[2] : T.item[3];
        const int bx = 1 << lbx, by = 1 << lby;
        const int W = PW * bx, rows = by << lbz;    // tile row length, rows per coil
        float *tb = reinterpret_cast<float *>(smem_pos + L);
        const int cstride = rows * W * 2 + 4;       // floats per coil (+4: fewer bank conflicts)
        if (myitem >= 0) {
            const fr_item it = items[myitem];
            float fx, fy, fz;
            lane_cell(unpack_box(it.t0, it.pad), j, fx, fy, fz);
            if (it.r0 < it.r1)
                grid_records<NCB, PAIR, FULL>(acc, it.r0, it.r1, recs, pos, data, (uint32_t)C, ca, cbb, stbl, fx,
                                              fy, fz, dacc, cidx, lmax, v0ok, v1ok, j, k);
            // register r holds cell q = 4 k + r of the unit, component NCB j + cb
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int q = 4 * k + r;
                    const int cx = q & (bx - 1), cy = (q >> lbx) & (by - 1), cz = q >> (lbx + lby);
                    const int comp = NCB * j + cb;
                    tb[(comp >> 1) * cstride + ((cz * by + cy) * W + wv * bx + cx) * 2 + (comp & 1)] = acc[cb][r];
                }
        }
        __syncthreads();
        // 16 bytes (two cells) per thread and step, whole rows; cells of
        // units without an item here are left alone
        const int ox = (int)(T.box0 & 0xffffu), oy = (int)(T.box0 >> 16), oz = (int)T.box1;
        const int lhalf = lbx + __builtin_ctz((unsigned)PW) - 1;       // log2(W / 2)
        const int lrows = lby + lbz;
        const int total = nc << (lrows + lhalf);
        for (int ci = threadIdx.x; ci < total; ci += 64 * PW) {
            const int ch = ci & ((1 << lhalf) - 1);
            const int line = ci >> lhalf;
            const int coil = line >> lrows, row = line & (rows - 1);
            const int x = 2 * ch;
            const int slot = x >> lbx;
            const int32_t owner = slot == 0 ? T.item[0] : slot == 1 ? T.item[1] : slot == 2 ? T.item[2] : T.item[3];
            if (owner < 0) continue;
            const int cz = row >> lby, cy = row & (by - 1);
            const f4 v = *reinterpret_cast<const f4 *>(tb + coil * cstride + (row * W + x) * 2);
            f4 *dst = reinterpret_cast<f4 *>(out + (size_t)(c0 + coil) * cells +
                                             ((size_t)(oz + cz) * N + (oy + cy)) * N + (ox + x));
            if (nt)
                __builtin_nontemporal_store(v, dst);
            else
                *dst = v;
        }
        return;
    }

    // Write the accumulators of item it (unit box b) through an LDS
    // transpose: register r holds cell 4 k + r, component NCB j + cb (coil
    // c: re 2c, im 2c + 1); consecutive lanes then store consecutive cells
    // of one coil.
    auto write_out = [&](const fr_item &it, const item_box &b) {
        wave_sync();
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) osm[(NCB * j + cb) * CS + 4 * k + r] = acc[cb][r];
        wave_sync();
        auto put = [&](int cl, int q, size_t go) {
            const float2 v = make_float2(osm[(2 * cl) * CS + q], osm[(2 * cl + 1) * CS + q]);
            const int c = c0 + cl;
            if (it.slot < 0) {
                if (nt)
                    __builtin_nontemporal_store(__builtin_bit_cast(unsigned long long, v),
                                                reinterpret_cast<unsigned long long *>(out + (size_t)c * cells + go));
                else
                    out[(size_t)c * cells + go] = v;
            } else {
                const size_t slot = ((size_t)(c >> 5) * nslots + it.slot) * 32 + (c & 31);
                partial[slot * cells_max + q] = v;
            }
        };
        if ((16 % b.ncell) == 0) {
            // the cell count divides 16 (full units): each lane keeps one cell
            const int sh = __builtin_ctz((unsigned)b.ncell);
            const int q = lane & (b.ncell - 1);
            const size_t go =
                ((size_t)(b.oz + q / (b.ux * b.uy)) * N + (b.oy + (q / b.ux) % b.uy)) * N + (b.ox + q % b.ux);
            for (int cl = lane >> sh; cl < nc; cl += 64 >> sh) put(cl, q, go);
        } else {
            for (int t = lane; t < nc * b.ncell; t += 64) {
                const int cl = t / b.ncell;
                const int q = t - cl * b.ncell;
                const size_t go =
                    ((size_t)(b.oz + q / (b.ux * b.uy)) * N + (b.oy + (q / b.ux) % b.uy)) * N + (b.ox + q % b.ux);
                put(cl, q, go);
            }
        }
        wave_sync();   // the next item reuses the transpose buffer
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[cb] = f4{0.f, 0.f, 0.f, 0.f};
    };

    // ----------------------------------------------------------- pairs
    // One wave grids the two x-adjacent units of a pair one after the other
    // (the first unit's sums held in registers) and writes them out
    // together: rows of 2 bx cells (64 bytes for 4-cell-wide units) instead
    // of bx.  Components go through the LDS transpose in steps of 16 coils
    // (8 when a chunk has 8): region cell x + 2 bx (y + by z) of the pair's
    // 2 bx x by x bz box, unit A at x < bx, unit B at x >= bx.
    if (bid >= front_wg && ntiles > 0) {
        const int64_t pidx = fr_sampled_tile(grp * PW + wv, tstride);   // (tstride 1: every pair)
        if (pidx >= ntiles) return;
        const fr_tile T = tiles[pidx];
        const int32_t ia = T.item[0], ib = T.item[1];
        if (ia < 0 && ib < 0) return;
        // a unit alone (its partner is a front unit): the single write-out
        if (ia < 0 || ib < 0) {
            const fr_item it = items[ia >= 0 ? ia : ib];
            const item_box b = unpack_box(it.t0, it.pad);
            float fx, fy, fz;
            lane_cell(b, j, fx, fy, fz);
            if (it.r0 < it.r1)
                grid_records<NCB, PAIR, FULL>(acc, it.r0, it.r1, recs, pos, data, (uint32_t)C, ca, cbb, stbl, fx, fy,
                                              fz, dacc, cidx, lmax, v0ok, v1ok, j, k);
            write_out(it, b);
            return;
        }
        constexpr int RS = 33;                        // LDS row: 32 region cells + 1
        const int bxs = 1 << lbx, lb2 = lbx + 1;
        constexpr int steps = NCB == 4 ? 2 : 1;
        constexpr int ncs = NCB == 1 ? 8 : 16;        // coils per step
        // lane (j, k): register r holds cell q = 4 k + r of a unit; unit A's
        // cells go to region x < bx, B's to x >= bx
        auto stage = [&](f4 a, int cb, int xoff) {
            const int comp = NCB == 1 ? j : 2 * j + (cb & 1);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = 4 * k + r;
                osm[comp * RS + (q & (bxs - 1)) + ((q >> lbx) << lb2) + xoff] = a[r];
            }
        };
        const fr_item ita = items[ia], itb = items[ib];
        const item_box ba = unpack_box(ita.t0, ita.pad), bb = unpack_box(itb.t0, itb.pad);
        float fx, fy, fz;
        lane_cell(ba, j, fx, fy, fz);
        // at the end of A: all its component slices into the transpose
        // buffer (step h at offset h 32 RS); then B's cell and fresh
        // accumulators
        auto to_b = [&]() {
            wave_sync();
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) stage(acc[cb], cb, (cb >> 1) * 32 * RS);
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) acc[cb] = f4{0.f, 0.f, 0.f, 0.f};
            lane_cell(bb, j, fx, fy, fz);
        };
        if (ita.r0 < ita.r1 && itb.r0 < itb.r1 && itb.r0 == ita.r1) {
            // A's and B's records are consecutive: one pipelined stream
            grid_records_mid<NCB, PAIR, FULL>(acc, ita.r0, ita.r1, itb.r1, recs, pos, data, (uint32_t)C, ca, cbb, stbl,
                                              fx, fy, fz, dacc, cidx, lmax, v0ok, v1ok, j, k, to_b);
        } else {
            if (ita.r0 < ita.r1)
                grid_records<NCB, PAIR, FULL>(acc, ita.r0, ita.r1, recs, pos, data, (uint32_t)C, ca, cbb, stbl, fx,
                                              fy, fz, dacc, cidx, lmax, v0ok, v1ok, j, k);
            to_b();
            if (itb.r0 < itb.r1)
                grid_records<NCB, PAIR, FULL>(acc, itb.r0, itb.r1, recs, pos, data, (uint32_t)C, ca, cbb, stbl, fx,
                                              fy, fz, dacc, cidx, lmax, v0ok, v1ok, j, k);
        }
        const int cell = lane & 31;
        const int x = cell & ((2 << lbx) - 1), row = cell >> lb2;
        const size_t go = ((size_t)(ba.oz + (row >> lby)) * N + (ba.oy + (row & ((1 << lby) - 1)))) * N + (ba.ox + x);
        wave_sync();
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) stage(acc[cb], cb, (cb >> 1) * 32 * RS + bxs);
        wave_sync();
#pragma unroll
        for (int h = 0; h < steps; ++h) {
            const float *os = osm + h * 32 * RS;
#pragma unroll
            for (int t = 0; t < ncs / 2; ++t) {
                const int cs = (lane >> 5) + 2 * t;
                const int cl = NCB == 4 ? 2 * cs + h : cs;
                if (cl < nc) {
                    const float2 v = make_float2(os[(2 * cs) * RS + cell], os[(2 * cs + 1) * RS + cell]);
                    float2 *d = out + (size_t)(c0 + cl) * cells + go;
                    if (nt)
                        __builtin_nontemporal_store(__builtin_bit_cast(unsigned long long, v),
                                                    reinterpret_cast<unsigned long long *>(d));
                    else
                        *d = v;
                }
            }
        }
        return;
    }

    int64_t i, iend;
    const int64_t fw = bid * PW + wv;
    if (bid < front_wg) {
        // fsub > 1 (a sample of the grid): wave fw takes sub-slice fw % fsub
        // of front item fw / fsub
        i = fsub > 1 ? fw / fsub : fw;
        iend = min(i + 1, nfront);
    } else {
        i = nfront + (grp * PW + wv) * ipw;
        iend = min(i + ipw, nitems);
    }
    if (i >= iend) return;

    // current item; empty items (units without records) write zeros
    fr_item cur = items[i];
    if (bid < front_wg && fsub > 1) {
        // only the front units of every tstride-th rank; the sub-slice's
        // records (a multiple of FR_POS_BATCH, possibly none) go to
        // virtual slot fw
        if (cur.unit % (uint32_t)tstride != 0) return;
        const uint32_t sub = (uint32_t)(fw % fsub), len = cur.r1 - cur.r0;
        const uint32_t sl = ((len + fsub - 1) / fsub + (FR_POS_BATCH - 1)) & ~(uint32_t)(FR_POS_BATCH - 1);
        const uint32_t a = min(cur.r0 + sub * sl, cur.r1);
        cur.r0 = a;
        cur.r1 = min(a + sl, cur.r1);
        cur.slot = (int32_t)fw;
    }
    item_box box = unpack_box(cur.t0, cur.pad);
    while (cur.r0 == cur.r1) {
        write_out(cur, box);
        if (++i >= iend) return;
        cur = items[i];
        box = unpack_box(cur.t0, cur.pad);
    }
    float fx, fy, fz;
    lane_cell(box, j, fx, fy, fz);

    // At the end of the current item: write it out, then move to the next
    // item with records (writing zeros for empty ones).  False when the run
    // is done.
    auto next_item = [&]() -> bool {
        write_out(cur, box);
        while (++i < iend) {
            cur = items[i];
            box = unpack_box(cur.t0, cur.pad);
            if (cur.r0 < cur.r1) {
                lane_cell(box, j, fx, fy, fz);
                return true;
            }
            write_out(cur, box);
        }
        return false;
    };

    // the run's items as one record stream (their records are consecutive)
    uint32_t r = cur.r0;
    uint32_t id0[4], id1[4];
    batch_st S0, S1;
    load_ids(id0, recs, r, k);
    load_ids(id1, recs, r + RB, k);
    issue_batch<NCB, PAIR>(S0, id0, pos, data, (uint32_t)C, ca, cbb);
    while (true) {
        load_ids(id0, recs, r + 2 * RB, k);
        issue_batch<NCB, PAIR>(S1, id1, pos, data, (uint32_t)C, ca, cbb);
        __builtin_amdgcn_sched_barrier(0);
        apply_batch<NCB, FULL>(acc, S0, stbl, fx, fy, fz, dacc, cidx, lmax, v0ok, v1ok, j);
        r += RB;
        if (r == cur.r1 && !next_item()) break;
        load_ids(id1, recs, r + 2 * RB, k);
        issue_batch<NCB, PAIR>(S0, id0, pos, data, (uint32_t)C, ca, cbb);
        __builtin_amdgcn_sched_barrier(0);
        apply_batch<NCB, FULL>(acc, S1, stbl, fx, fy, fz, dacc, cidx, lmax, v0ok, v1ok, j);
        r += RB;
        if (r == cur.r1 && !next_item()) break;
    }
}

template <int NCB, bool PAIR, bool FULL, typename TS, int OCC>
fr_status launch(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial, int64_t tstride, int fsub)
{
    fr_ctx *ctx = p->ctx;
    const size_t cells = fr_plan_cells(p->dim, p->N, p->frames);
    const int chunks = (s->C + 8 * NCB - 1) / (8 * NCB);
    const int ipw = std::max(1, p->pos_ipw);
    const int64_t per_wg = (int64_t)PW * ipw;
    // tstride > 1 (pair plans only): the sampled pairs (fr_sampled_tile), and
    // the front items cut into fsub sub-slices each (one wave per sub-slice)
    const int64_t front_wg = (p->nitems_front * (tstride > 1 ? fsub : 1) + PW - 1) / PW;
    const int64_t light = p->nitems - p->nitems_front;
    const int64_t npairs = fr_sampled_tiles(p->ntiles, tstride);
    // light work: tiles of four (one per workgroup), pairs (PW per
    // workgroup) or runs of ipw items per wave
    const bool fours = p->ntiles > 0 && p->tile_units == PW;
    const int64_t groups = p->ntiles > 0 ? (fours ? p->ntiles : (npairs + PW - 1) / PW)
                                         : (light + per_wg - 1) / per_wg;
    const int64_t xr = std::max<int64_t>(1, p->xcd_run / (p->ntiles > 0 ? (fours ? PW : 2 * PW) : per_wg));
    const int64_t span = 8 * xr;
    const int64_t W = front_wg + span * ((groups + span - 1) / span);
    FR_CHECK(ctx, W < (int64_t(1) << 31), FR_ERR_UNSUPPORTED, "gridding launch of %lld workgroups", (long long)W);
    // per wave: the write-out transpose (16 cells x 16 NCB components, or a
    // pair's 32 cells x 32 components)
    const int osm_floats = std::max(16 * NCB * CS, (NCB == 4 ? 64 : 32) * 33);
    const size_t lds = sizeof(float4) * p->L + sizeof(float) * PW * osm_floats;
    auto kern = k_grid_pos<NCB, PAIR, FULL, TS, OCC>;
    FR_HIP(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int lbx = __builtin_ctz((unsigned)p->geom.bx), lby = __builtin_ctz((unsigned)p->geom.by),
              lbz = __builtin_ctz((unsigned)p->geom.bz);
    hipLaunchKernelGGL(kern, dim3((unsigned)W, chunks), dim3(64 * PW), lds, ctx->stream, p->items, p->nitems,
                       p->tiles, p->ntiles, lbx, lby, lbz, p->recs, p->pos, p->tsel, p->L, p->mfma_dacc, p->mfma_cidx,
                       (const TS *)s->data, s->C, (int)p->N, cells, out, partial, p->nslots,
                       p->geom.cells_max, p->nitems_front, front_wg, ipw, xr, p->pos_stores == 1 ? 1 : 0,
                       p->tile_units, osm_floats, tstride, tstride > 1 ? fsub : 1);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

template <typename TS, int OCC>
fr_status launch_ncb(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial, int64_t ts, int fs)
{
    if (s->C <= 8) return launch<1, false, false, TS, OCC>(p, s, out, partial, ts, fs);
    if (s->C <= 16) return launch<2, false, false, TS, OCC>(p, s, out, partial, ts, fs);
    if (s->C % 32 == 0) return launch<4, true, true, TS, OCC>(p, s, out, partial, ts, fs);
    if (s->C % 2 == 0) return launch<4, true, false, TS, OCC>(p, s, out, partial, ts, fs);
    return launch<4, false, false, TS, OCC>(p, s, out, partial, ts, fs);
}

// pos[i].w <- the samples' DCF (the null sample keeps 0)
__global__ void k_pos_dcf(float4 *__restrict__ pos, const float *__restrict__ dcf, int64_t M)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) pos[i].w = dcf[i];
}

template <typename TS>
fr_status launch_occ(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial, int64_t ts, int fs)
{
    return p->pos_waves == 4 ? launch_ncb<TS, 4>(p, s, out, partial, ts, fs)
                             : launch_ncb<TS, 1>(p, s, out, partial, ts, fs);
}

}  // namespace

fr_status fr_launch_grid_pos(fr_plan *p, const fr_samples *s, float2 *out, float2 *partial, int64_t tile_stride,
                             const fr_samples *dcf, int front_sub)
{
    FR_CHECK(p->ctx, p->recs && p->pos && p->tsel && p->geom.cells_max <= 16 && p->L <= FR_POS_MAX_L,
             FR_ERR_STATE, "position-weight gridding needs a plan built for it (accumulate 6)");
    FR_CHECK(p->ctx, tile_stride == 1 || (tile_stride > 1 && p->ntiles > 0 && p->tile_units == 2), FR_ERR_STATE,
             "a tile stride needs a plan with unit pairs");
    FR_CHECK(p->ctx, tile_stride == 1 || (front_sub >= 1 && s->C == 1), FR_ERR_INVALID,
             "a sample of the grid is single-channel, with at least one sub-slice per front item");
    const fr_samples *w = dcf ? dcf : s;
    FR_CHECK(p->ctx, w->M == s->M, FR_ERR_INVALID, "the DCF's samples handle has another sample count");
    // the DCF into the positions' .w when the copy there is not this
    // handle's current DCF (another samples handle, or a changed DCF), and
    // always under graph capture: a replay may follow DCF changes that no
    // host call sees, so the captured pass carries its copy and the next
    // eager pass refreshes again
    fr_ctx *ctx = p->ctx;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FR_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    const bool capturing = cs != hipStreamCaptureStatusNone;
    if (p->M > 0 && (capturing || p->pos_dcf_uid != w->uid || p->pos_dcf_version != w->dcf_version)) {
        hipLaunchKernelGGL(k_pos_dcf, dim3((unsigned)((p->M + 255) / 256)), dim3(256), 0, ctx->stream, p->pos,
                           w->dcf, p->M);
        FR_HIP(ctx, hipGetLastError());
        p->pos_dcf_uid = capturing ? 0 : w->uid;
        p->pos_dcf_version = w->dcf_version;
    }
    return s->type == FR_SAMPLES_C64 ? launch_occ<float2>(p, s, out, partial, tile_stride, front_sub)
                                     : launch_occ<__half2>(p, s, out, partial, tile_stride, front_sub);
}
