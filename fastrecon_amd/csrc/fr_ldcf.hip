// Learned density compensation (net-new; the reference has no DCF
// estimation): per-sample trajectory features and an 8-64-64-1 SiLU MLP with
// BF16 weight matrices and FP32 accumulation, trained and applied on the
// device.
//
// Features (one row of 8 floats per sample):
//   u = (t - (N-1)/2) / (N/2) for the normalised point t, rho = |u| N/2
//   (the radius |k| in grid cells),
//   f0 = log1p(rho)/5 (log-compressed so that the first cells around the
//   k-space centre are resolved),
//   f1 = atan2(u_y, u_x)/pi, f2 = acos(u_z/|u|)/pi (0.5 in 2D),
//   f3..f5 = log1p(n_r)/8 with n_r the number of other samples whose nearest
//   cell lies within Chebyshev distance r = 1, 2, 4 of this sample's cell,
//   f6 = min(rho_max - rho, 4)/4, the distance in cells to the outermost
//   sample's shell (rho_max = the largest rho of all samples): the outermost
//   shells (a spoke's end samples) carry the largest weights and differ by
//   one cell, which f0 resolves only to 0.002,
//   f7 = log1p(c_b / V_b)/8, the samples per cell of this sample's radius
//   shell: b = floor(2 rho), c_b the samples (this one included) with that
//   b, V_b the cells of the shell b/2 <= rho < b/2 + 1/2 ((4 pi/3)(hi^3 -
//   lo^3) in 3D, pi (hi^2 - lo^2) in 2D).  A smooth density for radial-type
//   trajectories, where the cell counts near the centre vary from cell to
//   cell with the spoke crossings.
// The counts come from a cell histogram (binned by tile on grids of at most
// 8192 tiles, else one atomic per sample), its summed-area table and a pass
// over the cells that stores each cell's three box sums (r = 1, 2, 4; 8
// corner reads each in 3D, neighbouring cells sharing them) in one uint4,
// so a sample reads one 16-byte cell record whatever the density; rho_max
// (one atomic maximum per wave) and the shell counts (per workgroup in LDS)
// come from the histogram's first pass.
//
// The DCF head: w = exp(mu + sigma y) for the network output y, divided by
// 1 + n_k for a count feature k = 3..5 when the network carries that base
// (fr_mlp_set_base; exp(mu + sigma y - 8 f_k), the network then learns the
// DCF relative to the inverse local density).
//
// MLP layer 2 (64 x 64) runs on MFMA (v_mfma_f32_16x16x32_bf16): a wave
// takes 16 samples, computes h1 = silu(x W1 + b1) for them in VGPRs already
// in the B-operand layout (lane l: sample l%16, hidden units 32ks+8(l/16)+j),
// and multiplies by W2^T held as eight A fragments; the D fragments (lane l:
// sample l%16, outputs 16mt+4(l/16)+i) feed SiLU and the 64 -> 1 layer, whose
// four partial sums per sample are added across lane groups in a fixed order.
// Training recomputes this forward, back-propagates in FP32 and accumulates
// each wave's weight gradient in registers; per-wave partials are summed in
// wave order by a second kernel (deterministic), then Adam updates the FP32
// master weights and refreshes the BF16 copies.
#include "fr_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <initializer_list>

namespace {

constexpr int NF = 8;
constexpr int H = 64;
constexpr int FSTRIDE = 8;   // floats per feature row
constexpr int O_W1 = 0, O_B1 = NF * H, O_W2 = O_B1 + H, O_B2 = O_W2 + H * H, O_W3 = O_B2 + H, O_B3 = O_W3 + H;
constexpr int NPARAM = O_B3 + 1;
constexpr int SLOT = NPARAM + 1;   // per-wave partial: gradient + loss

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

__host__ __device__ inline uint32_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;   // round to nearest even (finite values)
}

typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;

// two floats to BF16, a in the low half: one v_cvt_pk_bf16_f32 (round to
// nearest even, the same bits as f2bf for finite values)
__device__ inline uint32_t pack_bf16(float a, float b)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

__device__ inline float bf2f(uint32_t b) { return __uint_as_float(b << 16); }

// SiLU and its derivative with the hardware reciprocal (v_rcp_f32, 1 ulp):
// the network's BF16 weights and hidden layer set its precision, so an IEEE
// division (about ten instructions, 32 per sample) buys nothing.  Training and
// inference use the same functions.
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ inline float silu(float z) { return z * sigmoid(z); }

__device__ inline float dsilu(float z)
{
    const float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// ------------------------------------------------------------- features
// Cell histogram of the samples' nearest cells, for grids too large for the
// binned histogram below (more than 8192 tiles).  Radial-type trajectories
// cross the k-space centre once per spoke, so the few cells around N/2 take
// thousands of hits that serialise as same-address atomics; each workgroup
// counts the box of HR cells around the centre in LDS over its samples
// (grid-stride) and adds the box to memory once.  Integer counts: the
// result does not depend on order.
constexpr int HR = 3;                        // central box half-width (cells)
constexpr int HB = 2 * HR + 1;

// FP32 radius in cells of the normalised point (same operations as
// sample_features)
__device__ inline float rho_of(const float (&t)[3], int dim, int N)
{
    const float cf = 0.5f * (float)(N - 1), hN = 0.5f * (float)N;
    float u[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < dim; ++d) u[d] = __fdiv_rn(__fsub_rn(t[d], cf), hN);
    const float r2 = dim == 3 ? __fadd_rn(__fadd_rn(__fmul_rn(u[0], u[0]), __fmul_rn(u[1], u[1])), __fmul_rn(u[2], u[2]))
                              : __fadd_rn(__fmul_rn(u[0], u[0]), __fmul_rn(u[1], u[1]));
    return __fmul_rn(sqrtf(r2), hN);
}
__device__ inline float sample_rho(const float *__restrict__ traj, int64_t M, int dim, int N, int64_t s)
{
    float t[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < dim; ++d) t[d] = traj[d * M + s];
    return rho_of(t, dim, N);
}

// radius shell of a sample: floor(2 rho), clamped to the nb shells
__device__ inline int shell_of(float rho, int nb) { return min(max((int)floorf(2.f * rho), 0), nb - 1); }

// hist[cells] (one past the histogram) receives rho_max as float bits
// (non-negative floats order like their bits), hist[cells + 1 + b] the
// samples of radius shell b (b < nb); with lds_shells, counted per
// workgroup in dynamic LDS (nb words) and added once.
__global__ void __launch_bounds__(256) k_cell_hist(const float *__restrict__ traj, int64_t M, int dim, int N,
                                                   uint32_t *__restrict__ hist, int nb, int lds_shells)
{
    __shared__ uint32_t hot[HB * HB * HB];
    extern __shared__ uint32_t shells[];
    const int nhot = dim == 3 ? HB * HB * HB : HB * HB;
    for (int i = threadIdx.x; i < nhot; i += blockDim.x) hot[i] = 0u;
    if (lds_shells)
        for (int i = threadIdx.x; i < nb; i += blockDim.x) shells[i] = 0u;
    __syncthreads();
    const int lo = N / 2 - HR;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const size_t ncells = dim == 3 ? (size_t)N * N * N : (size_t)N * N;
    uint32_t *gshell = hist + ncells + 1;
    float rmax = 0.f;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < M; s += stride) {
        const float rho = sample_rho(traj, M, dim, N, s);
        rmax = fmaxf(rmax, rho);
        const int b = shell_of(rho, nb);
        if (lds_shells)
            atomicAdd(&shells[b], 1u);
        else
            atomicAdd(&gshell[b], 1u);
        size_t cell = 0;
        int hidx = 0;
        bool central = true;
        for (int d = dim - 1; d >= 0; --d) {
            int c = (int)floorf(traj[d * M + s] + 0.5f);
            c = min(max(c, 0), N - 1);
            cell = cell * N + c;
            const int h = c - lo;
            central = central && h >= 0 && h < HB;
            hidx = hidx * HB + h;
        }
        if (central)
            atomicAdd(&hot[hidx], 1u);
        else
            atomicAdd(&hist[cell], 1u);
    }
    // wave maximum, then one atomic per wave
    for (int m = 32; m >= 1; m >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, m));
    if ((threadIdx.x & 63) == 0) atomicMax(&hist[ncells], __float_as_uint(rmax));
    __syncthreads();
    for (int i = threadIdx.x; i < nhot; i += blockDim.x) {
        const uint32_t v = hot[i];
        if (v == 0u) continue;
        size_t cell = 0;
        int r = i;
        int c[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            c[d] = lo + r % HB;
            r /= HB;
        }
        for (int d = dim - 1; d >= 0; --d) cell = cell * N + c[d];
        atomicAdd(&hist[cell], v);
    }
    if (lds_shells)
        for (int i = threadIdx.x; i < nb; i += blockDim.x)
            if (shells[i]) atomicAdd(&gshell[i], shells[i]);
}

// Binned histogram (grids of at most 8192 tiles of T^dim cells): the same
// counts without one global atomic per sample.  k_bin_count counts samples
// per tile (and shells, rho_max) in LDS, adds the tile totals once per
// workgroup and keeps each sample's key (tile << 16 | cell within the tile);
// k_bin_scan turns the totals into bucket offsets and the tiles' chunk
// lists; k_bin_scatter writes each sample's cell within its tile (16 bits)
// into its tile's bucket, slots reserved per workgroup; k_bin_hist counts
// one chunk of BIN_CHUNK bucket entries in LDS: the first chunk of every
// tile writes the tile's cells, a second launch adds the further chunks of
// the dense tiles (the few around the k-space centre) with atomics.  Integer
// counts: the order within a bucket does not matter.
// Tile totals and bucket cursors are kept in BIN_SETS copies, workgroup b
// using copy b % BIN_SETS (the XCD it runs on): every workgroup touches the
// tiles its spokes cross, about 750 on the 256^3 kooshball of 51200 spokes,
// so one copy would take some 300 atomics per address from all XCDs; the
// scan gives each copy its own run of the tile's bucket.
constexpr uint32_t BIN_CHUNK = 8192;
constexpr int BIN_SETS = 8;
struct bins_view {
    int tb, nt;            // tile side 2^tb cells; nt tiles per axis
    int64_t NT;            // tiles
    uint32_t *total, *start, *cursor, *more;   // [SETS][NT], [NT + 1], [SETS][NT], [NT + 1]
    uint32_t *key;         // [M]
    uint16_t *bucket;      // [M]
};

__device__ inline uint32_t key_of(const int (&c)[3], int dim, int tb, int nt)
{
    const int m = (1 << tb) - 1;
    const uint32_t tile = dim == 3 ? (uint32_t)(((c[2] >> tb) * nt + (c[1] >> tb)) * nt + (c[0] >> tb))
                                   : (uint32_t)((c[1] >> tb) * nt + (c[0] >> tb));
    const uint32_t in = dim == 3 ? (uint32_t)(((c[2] & m) << (2 * tb)) | ((c[1] & m) << tb) | (c[0] & m))
                                 : (uint32_t)(((c[1] & m) << tb) | (c[0] & m));
    return tile << 16 | in;
}

// One LDS atomic per run of equal tiles in consecutive lanes (consecutive
// samples of a spoke share a tile for about T samples): every lane of the
// wave calls this together; a lane with tile < 0 adds nothing and ends the
// run before it.  Returns the counter's value before this lane's unit.
__device__ inline uint32_t wave_run_add(uint32_t *ctr, int tile)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(tile, 1);
    const uint64_t bounds = __ballot(lane == 0 || prev != tile);
    const uint64_t upto = lane == 63 ? ~0ull : (2ull << lane) - 1;   // lanes 0..lane
    const int head = 63 - __clzll(bounds & upto);
    const uint64_t after = bounds & ~upto;
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    uint32_t base = 0;
    if (lane == head && tile >= 0) base = atomicAdd(&ctr[tile], (uint32_t)(end - lane));
    return __shfl(base, head) + (uint32_t)(lane - head);
}

// per-tile sample counts into B.total, each sample's key, shells and
// rho_max into hist's tail (as k_cell_hist); dynamic LDS: NT tile counters,
// then nb shell counters when lds_shells.  The loop runs whole waves (the
// stride is a multiple of 64) so the lanes of a wave meet in wave_run_add.
__global__ void __launch_bounds__(256) k_bin_count(const float *__restrict__ traj, int64_t M, int dim, int N,
                                                   bins_view B, uint32_t *__restrict__ hist, int nb, int lds_shells)
{
    extern __shared__ uint32_t lds[];
    uint32_t *tcnt = lds, *shells = lds + B.NT;
    for (int64_t i = threadIdx.x; i < B.NT; i += blockDim.x) tcnt[i] = 0u;
    if (lds_shells)
        for (int i = threadIdx.x; i < nb; i += blockDim.x) shells[i] = 0u;
    __syncthreads();
    const size_t ncells = dim == 3 ? (size_t)N * N * N : (size_t)N * N;
    uint32_t *gshell = hist + ncells + 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    float rmax = 0.f;
    // the next step's coordinates are loaded before this step's counting
    float nt[3] = {0.f, 0.f, 0.f};
    auto load = [&](int64_t s) {
        if (s < M)
            for (int d = 0; d < dim; ++d) nt[d] = traj[d * M + s];
    };
    int64_t s0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
    load(s0 + lane);
    for (; s0 < M; s0 += stride) {
        const int64_t s = s0 + lane;
        const float t[3] = {nt[0], nt[1], nt[2]};
        load(s + stride);
        int tile = -1;
        if (s < M) {
            const float rho = rho_of(t, dim, N);
            rmax = fmaxf(rmax, rho);
            const int b = shell_of(rho, nb);
            if (lds_shells)
                atomicAdd(&shells[b], 1u);
            else
                atomicAdd(&gshell[b], 1u);
            int c[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) c[d] = min(max((int)floorf(t[d] + 0.5f), 0), N - 1);
            const uint32_t k = key_of(c, dim, B.tb, B.nt);
            B.key[s] = k;
            tile = (int)(k >> 16);
        }
        wave_run_add(tcnt, tile);
    }
    for (int m = 32; m >= 1; m >>= 1) rmax = fmaxf(rmax, __shfl_xor(rmax, m));
    if ((threadIdx.x & 63) == 0) atomicMax(&hist[ncells], __float_as_uint(rmax));
    __syncthreads();
    uint32_t *total = B.total + (int64_t)(blockIdx.x % BIN_SETS) * B.NT;
    for (int64_t i = threadIdx.x; i < B.NT; i += blockDim.x)
        if (tcnt[i]) atomicAdd(&total[i], tcnt[i]);
    if (lds_shells)
        for (int i = threadIdx.x; i < nb; i += blockDim.x)
            if (shells[i]) atomicAdd(&gshell[i], shells[i]);
}

// exclusive scans of the NT (<= 8192) tile totals (summed over the copies)
// into start[0..NT] and of the tiles' chunks beyond the first into
// more[0..NT]; cursor[c][t] = start[t] + the totals of copies before c.
// One 1024-thread workgroup, 8 tiles per thread at most; each wave scans its
// lanes' sums with shuffles, then the 16 wave totals are scanned in LDS.
__global__ void __launch_bounds__(1024) k_bin_scan(bins_view B)
{
    __shared__ uint32_t wsum[16], wmore[16];
    const int64_t per = (B.NT + 1023) / 1024;
    const int64_t i0 = min((int64_t)threadIdx.x * per, B.NT), i1 = min(i0 + per, B.NT);
    uint32_t sum = 0, extra = 0;
    for (int64_t i = i0; i < i1; ++i) {
        uint32_t t = 0;
        for (int c = 0; c < BIN_SETS; ++c) t += B.total[c * B.NT + i];
        sum += t;
        extra += t > BIN_CHUNK ? (t - 1) / BIN_CHUNK : 0u;
    }
    // inclusive wave scans
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t is = sum, im = extra;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(is, d), v = __shfl_up(im, d);
        if (lane >= d) {
            is += u;
            im += v;
        }
    }
    if (lane == 63) {
        wsum[wave] = is;
        wmore[wave] = im;
    }
    __syncthreads();
    uint32_t run = is - sum, runm = im - extra;     // exclusive within the wave
    for (int w = 0; w < wave; ++w) {
        run += wsum[w];
        runm += wmore[w];
    }
    if (threadIdx.x == blockDim.x - 1) {
        B.start[B.NT] = run + sum;
        B.more[B.NT] = runm + extra;
    }
    for (int64_t i = i0; i < i1; ++i) {
        B.start[i] = run;
        B.more[i] = runm;
        uint32_t t = 0;
        for (int c = 0; c < BIN_SETS; ++c) {
            const uint32_t v = B.total[c * B.NT + i];
            B.cursor[c * B.NT + i] = run + t;
            t += v;
        }
        run += t;
        runm += t > BIN_CHUNK ? (t - 1) / BIN_CHUNK : 0u;
    }
}

// each sample's cell within its tile into the tile's bucket: the workgroup
// counts its samples per tile, reserves that many slots per tile at once,
// then hands them out in LDS (whole waves per step, as k_bin_count)
__global__ void __launch_bounds__(256) k_bin_scatter(int64_t M, bins_view B)
{
    extern __shared__ uint32_t tcnt[];
    for (int64_t i = threadIdx.x; i < B.NT; i += blockDim.x) tcnt[i] = 0u;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t first = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
    for (int64_t s0 = first; s0 < M; s0 += stride) {
        const int64_t s = s0 + (threadIdx.x & 63);
        wave_run_add(tcnt, s < M ? (int)(B.key[s] >> 16) : -1);
    }
    __syncthreads();
    // this workgroup's copy: the one k_bin_count added its totals to (same
    // grid, same samples per workgroup)
    uint32_t *cursor = B.cursor + (int64_t)(blockIdx.x % BIN_SETS) * B.NT;
    for (int64_t i = threadIdx.x; i < B.NT; i += blockDim.x)
        if (tcnt[i]) tcnt[i] = atomicAdd(&cursor[i], tcnt[i]);
    __syncthreads();
    for (int64_t s0 = first; s0 < M; s0 += stride) {
        const int64_t s = s0 + (threadIdx.x & 63);
        const uint32_t k = s < M ? B.key[s] : ~0u;
        const uint32_t slot = wave_run_add(tcnt, s < M ? (int)(k >> 16) : -1);
        if (s < M) B.bucket[slot] = (uint16_t)(k & 0xffffu);
    }
}

// one chunk of a tile's bucket counted in LDS (T^dim counters, dynamic).
// extra == 0: block b takes the first BIN_CHUNK entries of tile b and
// writes every cell of the tile inside the grid; extra == 1: block b takes
// the b-th chunk beyond the first (tile found in B.more) and adds its
// non-zero counts to cells the first launch wrote
__global__ void __launch_bounds__(256) k_bin_hist(bins_view B, int dim, int N, uint32_t *__restrict__ hist, int extra)
{
    extern __shared__ uint32_t h[];
    int64_t tile = blockIdx.x;
    uint32_t j = 0;
    if (extra) {
        const uint32_t item = blockIdx.x;
        if (item >= B.more[B.NT]) return;
        // the last tile t with more[t] <= item (more is non-decreasing and
        // more[NT] > item, so more[t + 1] > item: t has the item)
        int64_t lo = 0, hi = B.NT;   // more[lo] <= item < more[hi]
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) / 2;
            if (B.more[mid] <= item) lo = mid;
            else hi = mid;
        }
        tile = lo;
        j = item - B.more[lo] + 1;
    }
    const int tb = B.tb, T = 1 << tb;
    const int cellsT = dim == 3 ? T * T * T : T * T;
    for (int i = threadIdx.x; i < cellsT; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const uint32_t end = B.start[tile + 1];
    const uint32_t e0 = B.start[tile] + j * BIN_CHUNK;
    const uint32_t e1 = min(end, e0 + BIN_CHUNK);
    for (uint32_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) atomicAdd(&h[B.bucket[e]], 1u);
    __syncthreads();
    const int tx = (int)(tile % B.nt), ty = (int)((tile / B.nt) % B.nt), tz = dim == 3 ? (int)(tile / ((int64_t)B.nt * B.nt)) : 0;
    for (int i = threadIdx.x; i < cellsT; i += blockDim.x) {
        const int x = (tx << tb) + (i & (T - 1)), y = (ty << tb) + ((i >> tb) & (T - 1)), z = (tz << tb) + (i >> (2 * tb));
        if (x < N && y < N && (dim == 2 || z < N)) {
            const size_t cell = ((size_t)(dim == 3 ? z : 0) * N + y) * N + x;
            if (!extra)
                hist[cell] = h[i];
            else if (h[i])
                atomicAdd(&hist[cell], h[i]);
        }
    }
}

// inclusive prefix sum along one axis; axis stride `st`, lines enumerated by
// the other two coordinates
__global__ void k_prefix_axis(uint32_t *__restrict__ a, int N, int dim, int axis)
{
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t lines = dim == 3 ? (int64_t)N * N : N;
    if (line >= lines) return;
    size_t base, st;
    if (dim == 2) {
        if (axis == 0) { base = (size_t)line * N; st = 1; }
        else { base = (size_t)line; st = N; }
    } else {
        const int64_t p = line % N, q = line / N;
        if (axis == 0) { base = ((size_t)q * N + p) * N; st = 1; }
        else if (axis == 1) { base = (size_t)q * N * N + p; st = N; }
        else { base = (size_t)q * N + p; st = (size_t)N * N; }
    }
    // 16 loads in flight, then their running sums (a store to a[] may alias a
    // later load, so one load and one store per step would serialise on the
    // memory latency)
    constexpr int K = 16;
    uint32_t run = 0;
    int i = 0;
    for (; i + K <= N; i += K) {
        uint32_t v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = a[base + (size_t)(i + k) * st];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            run += v[k];
            a[base + (size_t)(i + k) * st] = run;
        }
    }
    for (; i < N; ++i) {
        run += a[base + (size_t)i * st];
        a[base + (size_t)i * st] = run;
    }
}

// inclusive prefix sum along x (contiguous lines of N): one wave per line,
// 256 cells per step (four per lane, then a wave scan of the lane totals);
// integer sums, so the same values as a sequential scan
__global__ void __launch_bounds__(256) k_prefix_x(uint32_t *__restrict__ a, int N, int64_t lines)
{
    const int64_t line = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (line >= lines) return;                  // whole waves
    uint32_t *row = a + (size_t)line * N;
    uint32_t carry = 0;
    for (int x0 = 0; x0 < N; x0 += 256) {
        uint32_t v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = x0 + 4 * lane + k;
            v[k] = x < N ? row[x] : 0u;
        }
        v[1] += v[0];
        v[2] += v[1];
        v[3] += v[2];
        uint32_t t = v[3];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t u = __shfl_up(t, d);
            if (lane >= d) t += u;
        }
        const uint32_t base = carry + t - v[3];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = x0 + 4 * lane + k;
            if (x < N) row[x] = base + v[k];
        }
        carry += __shfl(t, 63);
    }
}

// the count of the box of Chebyshev radius r around cell (x, y, z) (clipped
// to the grid) from the summed-area table S by inclusion-exclusion of its
// eight corners, corners outside the table (index -1) reading as zero;
// 32-bit cell indices (cells < 2^31)
__device__ inline uint32_t row_box(const uint32_t *__restrict__ S, int dim, int N, int x, int y, int z, int r)
{
    const int xl = max(x - r, 0) - 1, xh = min(x + r, N - 1);
    const int yl = max(y - r, 0) - 1, yh = min(y + r, N - 1);
    auto at = [&](int zz, int yy) -> uint32_t {      // S(xh) - S(xl) on row (yy, zz)
        if (zz < 0 || yy < 0) return 0u;
        const uint32_t *row = S + ((uint32_t)zz * (uint32_t)N + (uint32_t)yy) * (uint32_t)N;
        return row[xh] - (xl >= 0 ? row[xl] : 0u);
    };
    if (dim == 2) return at(0, yh) - at(0, yl);
    const int zl = max(z - r, 0) - 1, zh = min(z + r, N - 1);
    return (at(zh, yh) - at(zh, yl)) - (at(zl, yh) - at(zl, yl));
}

// per cell: the sample counts of the boxes of Chebyshev radius 1, 2 and 4
// around it (.x, .y, .z; .w = 0), from the summed-area table S; one
// workgroup per x-row y + N z.  A cell whose radius-1 box is empty holds no
// sample, so no feature reads it: its record is not written (about half of
// a kooshball's 256^3 cells, the corners outside the k-space ball).
// Workgroups go round-robin to the eight XCDs, so workgroup b takes row
// (b % 8) per + b / 8: each XCD walks its own run of consecutive rows, whose
// corner rows its L2 then holds.  Unsigned sums wrap, so the grouping of the
// eight corner terms does not change the count.
__global__ void __launch_bounds__(256) k_box_counts(const uint32_t *__restrict__ S, int dim, int N, uint32_t lines,
                                                    uint32_t per, uint4 *__restrict__ B)
{
    const uint32_t rowi = (blockIdx.x % 8u) * per + blockIdx.x / 8u;
    if (rowi >= lines) return;
    const int y = (int)(rowi % (uint32_t)N), z = (int)(rowi / (uint32_t)N);
    uint4 *out = B + (size_t)rowi * N;
    for (int x = threadIdx.x; x < N; x += blockDim.x) {
        const uint32_t n1 = row_box(S, dim, N, x, y, z, 1);
        if (n1 != 0u) out[x] = make_uint4(n1, row_box(S, dim, N, x, y, z, 2), row_box(S, dim, N, x, y, z, 4), 0u);
    }
}

// The count tables of one trajectory: S the summed-area table of the cell
// counts followed by rho_max (float bits) and the nb shell counts; B the
// cells' box counts (k_box_counts).
struct counts_view {
    const uint32_t *S;
    const uint4 *B;
    int nb;
};

// cells of radius shell b (b/2 <= rho < b/2 + 1/2), FP32
__device__ inline float shell_cells(int b, int dim)
{
    const float lo = 0.5f * (float)b, hi = __fadd_rn(lo, 0.5f);
    if (dim == 3)
        return __fmul_rn(4.18879032f, __fsub_rn(__fmul_rn(__fmul_rn(hi, hi), hi), __fmul_rn(__fmul_rn(lo, lo), lo)));
    return __fmul_rn(3.14159274f, __fsub_rn(__fmul_rn(hi, hi), __fmul_rn(lo, lo)));
}

// the eight features of sample s
__device__ inline void sample_features(const float *__restrict__ traj, int64_t M, int dim, int N, const counts_view &K,
                                       int64_t s, float (&f)[FSTRIDE])
{
    const float cf = 0.5f * (float)(N - 1), hN = 0.5f * (float)N;
    float u[3] = {0.f, 0.f, 0.f};
    int c[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
        const float t = traj[d * M + s];
        u[d] = __fdiv_rn(__fsub_rn(t, cf), hN);
        c[d] = min(max((int)floorf(t + 0.5f), 0), N - 1);
    }
    const float r2 = dim == 3 ? __fadd_rn(__fadd_rn(__fmul_rn(u[0], u[0]), __fmul_rn(u[1], u[1])), __fmul_rn(u[2], u[2]))
                              : __fadd_rn(__fmul_rn(u[0], u[0]), __fmul_rn(u[1], u[1]));
    const float r = sqrtf(r2);
    const float inv_pi = 0.318309886183790671f;
    f[0] = log1pf(r * hN) * 0.2f;             // radius in cells, log-compressed
    f[1] = atan2f(u[1], u[0]) * inv_pi;
    f[2] = (dim == 3 && r > 0.f) ? acosf(fminf(fmaxf(u[2] / r, -1.f), 1.f)) * inv_pi : 0.5f;
    const size_t cell = ((size_t)c[2] * N + c[1]) * N + c[0];
    const uint4 n = K.B[cell];
    f[3] = log1pf((float)(n.x - 1u)) * 0.125f;
    f[4] = log1pf((float)(n.y - 1u)) * 0.125f;
    f[5] = log1pf((float)(n.z - 1u)) * 0.125f;
    const size_t ncells = dim == 3 ? (size_t)N * N * N : (size_t)N * N;
    const float rho_max = __uint_as_float(K.S[ncells]);
    const float rho = __fmul_rn(r, hN);
    f[6] = fminf(__fsub_rn(rho_max, rho), 4.f) * 0.25f;
    const int b = shell_of(rho, K.nb);
    f[7] = log1pf(__fdiv_rn((float)K.S[ncells + 1 + b], shell_cells(b, dim))) * 0.125f;
}

__global__ void k_features(const float *__restrict__ traj, int64_t M, int dim, int N, counts_view K,
                           float *__restrict__ feat)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= M) return;
    float f[FSTRIDE];
    sample_features(traj, M, dim, N, K, s, f);
    float4 *o = reinterpret_cast<float4 *>(feat + s * FSTRIDE);
    o[0] = make_float4(f[0], f[1], f[2], f[3]);
    o[1] = make_float4(f[4], f[5], f[6], f[7]);
}

// ------------------------------------------------------------------- MLP
struct mlp_dev {
    const float *p;          // FP32 master parameters (biases are read from here)
    const uint16_t *w1;      // [6][64] BF16
    const uint16_t *w2;      // [64 in][64 out] BF16
    const uint16_t *w3;      // [64] BF16
};

// W2^T as A fragments: fragment (mt, ks), lane l holds W2[k][m] for
// m = 16 mt + l%16, k = 32 ks + 8 (l/16) + j, j = 0..7.
__device__ inline void load_a(const mlp_dev &W, int lane, u32x4 (&a)[4][2])
{
    const int m0 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            uint32_t w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = W.w2[(32 * ks + 8 * g + j) * H + 16 * mt + m0];
            a[mt][ks] = u32x4{w[0] | (w[1] << 16), w[2] | (w[3] << 16), w[4] | (w[5] << 16), w[6] | (w[7] << 16)};
        }
}

// Forward for the 16 samples of one wave.  Lane l: sample l%16.
// Outputs: z1[ks][j] (unit 32ks+8g+j), h1b (BF16-rounded h1 as float),
// z2[mt][i] (output 16mt+4g+i), h2, y (identical in the 4 lanes of a sample).
struct fwd_state {
    float z1[2][8], h1[2][8];
    float z2[4][4], h2[4][4];
    float y;
};

// Where the forward reads the layer-1 weights, the biases and w3: straight
// from the device parameters, or from a workgroup's LDS copy (same values).
struct wsrc_global {
    const mlp_dev &W;
    __device__ float w1(int i, int k) const { return bf2f(W.w1[i * H + k]); }
    __device__ float b1(int k) const { return W.p[O_B1 + k]; }
    __device__ float b2(int o) const { return W.p[O_B2 + o]; }
    __device__ float w3(int o) const { return bf2f(W.w3[o]); }
    __device__ float b3() const { return W.p[O_B3]; }
};

struct lds_weights {
    float w1[NF][H];
    float b1[H], b2[H], w3[H];
    float b3;
};

struct wsrc_lds {
    const lds_weights &L;
    __device__ float w1(int i, int k) const { return L.w1[i][k]; }
    __device__ float b1(int k) const { return L.b1[k]; }
    __device__ float b2(int o) const { return L.b2[o]; }
    __device__ float w3(int o) const { return L.w3[o]; }
    __device__ float b3() const { return L.b3; }
};

__device__ inline void stage_weights(const mlp_dev &W, lds_weights &L)
{
    for (int q = threadIdx.x; q < NF * H; q += blockDim.x) L.w1[q / H][q % H] = bf2f(W.w1[q]);
    for (int q = threadIdx.x; q < H; q += blockDim.x) {
        L.b1[q] = W.p[O_B1 + q];
        L.b2[q] = W.p[O_B2 + q];
        L.w3[q] = bf2f(W.w3[q]);
    }
    if (threadIdx.x == 0) L.b3 = W.p[O_B3];
}

template <class WS>
__device__ inline void mlp_forward16(const WS &W, const u32x4 (&a)[4][2], const float *x, int lane, fwd_state &F)
{
    const int g = lane >> 4;
    u32x4 b[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        float h[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 32 * ks + 8 * g + j;
            float z = W.b1(k);
#pragma unroll
            for (int i = 0; i < NF; ++i) z = fmaf(x[i], W.w1(i, k), z);
            F.z1[ks][j] = z;
            h[j] = silu(z);
        }
        uint32_t hb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            hb[j] = pack_bf16(h[2 * j], h[2 * j + 1]);
            F.h1[ks][2 * j] = bf2f(hb[j] & 0xffffu);
            F.h1[ks][2 * j + 1] = bf2f(hb[j] >> 16);
        }
        b[ks] = u32x4{hb[0], hb[1], hb[2], hb[3]};
    }
    float part = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[mt][0]),
                                                     __builtin_bit_cast(bf16x8, b[0]), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[mt][1]),
                                                     __builtin_bit_cast(bf16x8, b[1]), acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int o = 16 * mt + 4 * g + i;
            const float z = __fadd_rn(acc[i], W.b2(o));
            F.z2[mt][i] = z;
            F.h2[mt][i] = silu(z);
            part = fmaf(F.h2[mt][i], W.w3(o), part);
        }
    }
    // the four lane groups hold the partial sums of the same sample
    const float p1 = __fadd_rn(part, __shfl_xor(part, 16));
    const float p2 = __fadd_rn(p1, __shfl_xor(p1, 32));
    F.y = __fadd_rn(p2, W.b3());
}

// ------------------------------------------------------------- inference
// Inference runs both dense layers on MFMA.  Layer 1 (FP32 16x16x4, two K
// steps of four features): A = W1^T fragments (lane l: W1[4s + l/16][16mt +
// l%16]), B = the features (lane l: feature 4s + l/16 of sample l%16), C =
// b1; the D fragments leave lane l holding z1 of hidden units 16mt +
// 4(l/16) + r of sample l%16, r = 0..3.  Layer 2 contracts over the hidden
// units in the order that puts exactly those sixteen units into lane l's
// BF16 B operand: K position 32ks + 8(l/16) + j holds unit hidden_of(...)
// = 16(2ks + j/4) + 4(l/16) + j%4, and W2's A fragments are loaded in the
// same order.  Against the training forward (mlp_forward16, sequential FMAs
// in layer 1) only the FP32 summation order of z1 differs.
__device__ inline int hidden_of(int k)
{
    const int ks = k >> 5, g = (k >> 3) & 3, j = k & 7;
    return 16 * (2 * ks + (j >> 2)) + 4 * g + (j & 3);
}

struct infer_frags {
    u32x4 a2[4][2];   // W2^T, K positions in hidden_of order
    float a1[4][2];   // W1^T
};

__device__ inline void load_infer_frags(const mlp_dev &W, int lane, infer_frags &F)
{
    const int m0 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            uint32_t w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = W.w2[hidden_of(32 * ks + 8 * g + j) * H + 16 * mt + m0];
            F.a2[mt][ks] = u32x4{w[0] | (w[1] << 16), w[2] | (w[3] << 16), w[4] | (w[5] << 16), w[6] | (w[7] << 16)};
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) F.a1[mt][s] = bf2f(W.w1[(4 * s + g) * H + 16 * mt + m0]);
    }
}

// y of the 16 samples of one wave (identical in the four lanes of a
// sample); x0, x1: features g and 4 + g of sample l%16 (g = l/16)
template <class WS>
__device__ inline float mlp_infer16(const WS &W, const infer_frags &F, float x0, float x1, int lane)
{
    const int g = lane >> 4;
    uint32_t hb[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int u = 16 * mt + 4 * g;
        f32x4 acc = {W.b1(u), W.b1(u + 1), W.b1(u + 2), W.b1(u + 3)};
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(F.a1[mt][0], x0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(F.a1[mt][1], x1, acc, 0, 0, 0);
        hb[mt][0] = pack_bf16(silu(acc[0]), silu(acc[1]));
        hb[mt][1] = pack_bf16(silu(acc[2]), silu(acc[3]));
    }
    u32x4 b[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
        b[ks] = u32x4{hb[2 * ks][0], hb[2 * ks][1], hb[2 * ks + 1][0], hb[2 * ks + 1][1]};
    float part = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, F.a2[mt][0]),
                                                     __builtin_bit_cast(bf16x8, b[0]), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, F.a2[mt][1]),
                                                     __builtin_bit_cast(bf16x8, b[1]), acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int o = 16 * mt + 4 * g + i;
            part = fmaf(silu(__fadd_rn(acc[i], W.b2(o))), W.w3(o), part);
        }
    }
    const float p1 = __fadd_rn(part, __shfl_xor(part, 16));
    const float p2 = __fadd_rn(p1, __shfl_xor(p1, 32));
    return __fadd_rn(p2, W.b3());
}

__device__ inline void load_x(const float *__restrict__ feat, int64_t row, float *x)
{
    const float4 *r = reinterpret_cast<const float4 *>(feat + row * FSTRIDE);
    const float4 a = r[0], b = r[1];
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
}

// mode 0: y; mode 1: the DCF exp(mu + sigma y - lb), lb = 8 x_k for a
// network with base feature k (log(1 + n_k)), else 0
__device__ inline float mlp_out(float y, int mode, float mu, float sigma, float lb)
{
    return mode == 0 ? y : __expf(__fsub_rn(__fadd_rn(mu, __fmul_rn(sigma, y)), lb));
}

// y for every sample; mode 0 writes y, mode 1 the DCF (mlp_out; base: the
// base feature's index, -1 for none)
__global__ void __launch_bounds__(256) k_mlp_forward(mlp_dev W, const float *__restrict__ feat, int64_t M, int mode,
                                                     float mu, float sigma, int base, float *__restrict__ out)
{
    __shared__ lds_weights LW;
    stage_weights(W, LW);
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    infer_frags F;
    load_infer_frags(W, lane, F);
    __syncthreads();
    const wsrc_lds ws{LW};
    const int64_t groups = (M + 15) / 16;
    for (int64_t gi = wave; gi < groups; gi += nwaves) {
        const int64_t s = gi * 16 + (lane & 15);
        const float *row = feat + (s < M ? s : M - 1) * FSTRIDE;
        const float y = mlp_infer16(ws, F, row[g], row[4 + g], lane);
        if (s < M && lane < 16) out[s] = mlp_out(y, mode, mu, sigma, base >= 0 ? 8.f * row[base] : 0.f);
    }
}

// Features and forward in one pass (no feature rows in memory): each wave
// takes 64 samples, lane l computes the features of sample l into LDS, then
// the wave runs the network over them 16 at a time.  Same arithmetic as
// k_features followed by k_mlp_forward.
__global__ void __launch_bounds__(256) k_ldcf_dcf(mlp_dev W, const float *__restrict__ traj, int64_t M, int dim, int N,
                                                  counts_view K, int mode, float mu, float sigma, int base,
                                                  float *__restrict__ out)
{
    __shared__ lds_weights LW;
    __shared__ float xs[4][64][NF + 1];
    stage_weights(W, LW);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    infer_frags F;
    load_infer_frags(W, lane, F);
    __syncthreads();
    const wsrc_lds ws{LW};
    const int64_t groups = (M + 63) / 64;
    for (int64_t gi = wave; gi < groups; gi += nwaves) {
        const int64_t s0 = gi * 64;
        {
            float f[FSTRIDE];
            sample_features(traj, M, dim, N, K, min(s0 + lane, M - 1), f);
#pragma unroll
            for (int i = 0; i < NF; ++i) xs[wv][lane][i] = f[i];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
#pragma unroll 1
        for (int q = 0; q < 4; ++q) {
            const int r = 16 * q + (lane & 15);
            const float y = mlp_infer16(ws, F, xs[wv][r][g], xs[wv][r][4 + g], lane);
            const int64_t s = s0 + r;
            if (s < M && lane < 16) out[s] = mlp_out(y, mode, mu, sigma, base >= 0 ? 8.f * xs[wv][r][base] : 0.f);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Training pass over a batch: forward, MSE loss against t (normalised
// targets), back-propagation; each wave writes its gradient partial
// (NPARAM floats) and loss partial to its slot.  Every wave runs the same
// number of 16-sample groups (groups past the batch are masked), so the
// block barriers match.
constexpr int TB_WAVES = 4;
struct wave_lds {
    float x[16][NF];
    float h1[16][H];
    float h2[16][H];
    float dz2[16][H];
    float dz1[16][H];
    float dy[16];
};

__global__ void __launch_bounds__(256) k_mlp_backward(mlp_dev W, const float *__restrict__ feat,
                                                      const float *__restrict__ targ, const int64_t *__restrict__ idx,
                                                      int64_t B, int iters, float *__restrict__ partial)
{
    __shared__ wave_lds L[TB_WAVES];
    __shared__ float w2f[H][H + 1];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int64_t wave = (int64_t)blockIdx.x * TB_WAVES + wv;
    const int64_t nwaves = (int64_t)gridDim.x * TB_WAVES;
    for (int q = threadIdx.x; q < H * H; q += blockDim.x) w2f[q / H][q % H] = bf2f(W.w2[q]);
    u32x4 a[4][2];
    load_a(W, lane, a);
    __syncthreads();

    const int g = lane >> 4, sl = lane & 15;
    float gW2[H];
#pragma unroll
    for (int o = 0; o < H; ++o) gW2[o] = 0.f;
    float gW1[NF];
#pragma unroll
    for (int i = 0; i < NF; ++i) gW1[i] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, gl = 0.f;
    const float inv2B = 2.0f / (float)B;
    wave_lds &S = L[wv];

    for (int it = 0; it < iters; ++it) {
        const int64_t gi = wave + (int64_t)it * nwaves;
        const int64_t b = gi * 16 + sl;
        const bool valid = b < B;
        const int64_t row = valid ? (idx ? idx[b] : b) : (idx ? idx[0] : 0);
        float x[NF];
        load_x(feat, row, x);
        fwd_state F;
        mlp_forward16(wsrc_global{W}, a, x, lane, F);
        const float r = valid ? __fsub_rn(F.y, targ[row]) : 0.f;
        const float dy = __fmul_rn(inv2B, r);
        // dz2 for the lane's outputs; stage the per-sample rows in LDS
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int o = 16 * mt + 4 * g + i;
                S.h2[sl][o] = F.h2[mt][i];
                S.dz2[sl][o] = __fmul_rn(__fmul_rn(dy, bf2f(W.w3[o])), dsilu(F.z2[mt][i]));
            }
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int j = 0; j < 8; ++j) S.h1[sl][32 * ks + 8 * g + j] = F.h1[ks][j];
        if (g == 0) {
#pragma unroll
            for (int i = 0; i < NF; ++i) S.x[sl][i] = x[i];
            S.dy[sl] = dy;
            gl = __fadd_rn(gl, __fmul_rn(r, r));
        }
        __syncthreads();
        // dh1 = dz2 W2^T for the lane's units, then dz1
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 32 * ks + 8 * g + j;
                float dh = 0.f;
                for (int o = 0; o < H; ++o) dh = __fadd_rn(dh, __fmul_rn(S.dz2[sl][o], w2f[k][o]));
                S.dz1[sl][k] = __fmul_rn(dh, dsilu(F.z1[ks][j]));
            }
        __syncthreads();
        // lane l owns row l of W2 (input unit l), column l of W1, and index l
        // of b1, b2, w3
        for (int s2 = 0; s2 < 16; ++s2) {
            const float h1s = S.h1[s2][lane];
#pragma unroll
            for (int o = 0; o < H; ++o) gW2[o] = __fadd_rn(gW2[o], __fmul_rn(h1s, S.dz2[s2][o]));
            const float d1 = S.dz1[s2][lane];
#pragma unroll
            for (int i = 0; i < NF; ++i) gW1[i] = __fadd_rn(gW1[i], __fmul_rn(S.x[s2][i], d1));
            gb1 = __fadd_rn(gb1, d1);
            gb2 = __fadd_rn(gb2, S.dz2[s2][lane]);
            gw3 = __fadd_rn(gw3, __fmul_rn(S.h2[s2][lane], S.dy[s2]));
            gb3 = __fadd_rn(gb3, S.dy[s2]);
        }
        __syncthreads();
    }
    // loss partial: lanes 0..15 hold one sample each
    float l1 = gl;
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) l1 = __fadd_rn(l1, __shfl_xor(l1, m));
    float *P = partial + (size_t)wave * SLOT;
#pragma unroll
    for (int i = 0; i < NF; ++i) P[O_W1 + i * H + lane] = gW1[i];
    P[O_B1 + lane] = gb1;
#pragma unroll
    for (int o = 0; o < H; ++o) P[O_W2 + lane * H + o] = gW2[o];
    P[O_B2 + lane] = gb2;
    P[O_W3 + lane] = gw3;
    if (lane == 0) {
        P[O_B3] = gb3;
        P[NPARAM] = l1;
    }
}

__global__ void k_reduce_slots(const float *__restrict__ partial, int64_t nslots, float *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= SLOT) return;
    float a = 0.f;
    for (int64_t w = 0; w < nslots; ++w) a = __fadd_rn(a, partial[(size_t)w * SLOT + p]);
    out[p] = a;
}

__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v, const float *__restrict__ g,
                       float lr, float b1, float b2, float eps, float c1, float c2, uint16_t *__restrict__ w1,
                       uint16_t *__restrict__ w2, uint16_t *__restrict__ w3)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (i < O_B1) w1[i - O_W1] = (uint16_t)f2bf(pi);
    else if (i >= O_W2 && i < O_B2) w2[i - O_W2] = (uint16_t)f2bf(pi);
    else if (i >= O_W3 && i < O_B3) w3[i - O_W3] = (uint16_t)f2bf(pi);
}

__global__ void k_refresh_bf16(const float *__restrict__ p, uint16_t *__restrict__ w1, uint16_t *__restrict__ w2,
                               uint16_t *__restrict__ w3)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float pi = p[i];
    if (i < O_B1) w1[i - O_W1] = (uint16_t)f2bf(pi);
    else if (i >= O_W2 && i < O_B2) w2[i - O_W2] = (uint16_t)f2bf(pi);
    else if (i >= O_W3 && i < O_B3) w3[i - O_W3] = (uint16_t)f2bf(pi);
}

}  // namespace

struct fr_mlp {
    fr_ctx *ctx = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *w1 = nullptr, *w2 = nullptr, *w3 = nullptr;
    float *partial = nullptr;
    int64_t partial_slots = 0;
    int64_t step = 0;
    float mu = 0.f, sigma = 1.f;
    int base = -1;                // base feature of the DCF head (fr_mlp_set_base), -1: none
    uint32_t *hist = nullptr;     // fr_ldcf_dcf's count tables (counts_layout), kept between calls
    size_t hist_bytes = 0;
};

namespace {
mlp_dev view(const fr_mlp *m) { return mlp_dev{m->p, m->w1, m->w2, m->w3}; }
}

extern "C" int fr_mlp_param_count(void) { return NPARAM; }

// The count tables of a trajectory in one buffer: the histogram /
// summed-area table (cells uint32), rho_max, the shell counts (nb uint32;
// radii up to N/2 sqrt(dim) in half cells, plus margin); then from the next
// 16-byte boundary the cells' box counts (uint4 each); then, for a binned
// histogram, the tile totals (BIN_SETS copies), starts, cursors (BIN_SETS
// copies) and chunk offsets, the M sample keys and the M-entry bucket.
// Zeroed before each use: everything up to the box counts (zero_bytes), or,
// binned, rho_max, the shells and the tile totals (the tiles write every
// cell).
constexpr int64_t MAX_TILES = 8192;
struct counts_layout {
    size_t cells, zero_bytes, box_off, bins_off, bytes;
    int nb;
    bool binned;
    int tb, nt;            // binned: tile side 2^tb, nt tiles per axis
    int64_t NT, M;
};
static counts_layout layout_of(int dim, unsigned N, int64_t M)
{
    counts_layout L;
    L.cells = fr_cells(dim, N);
    L.nb = (int)std::ceil((double)N * std::sqrt((double)dim)) + 2;
    L.zero_bytes = (L.cells + 1 + (size_t)L.nb) * sizeof(uint32_t);
    L.box_off = (L.zero_bytes + 15) & ~(size_t)15;
    L.bins_off = L.box_off + L.cells * sizeof(uint4);
    // tiles of 16^3 cells in 3D, 64^2 or 128^2 in 2D (at most 64 KB of LDS
    // counters per tile), at most MAX_TILES of them; bucket slots are 32-bit
    L.binned = false;
    L.tb = L.nt = 0;
    L.NT = 0;
    L.M = M;
    for (int tb : dim == 3 ? std::initializer_list<int>{4} : std::initializer_list<int>{6, 7}) {
        if (M <= 0 || M > (int64_t)UINT32_MAX) break;
        const int nt = (int)((N + (1u << tb) - 1) >> tb);
        const int64_t NT = dim == 3 ? (int64_t)nt * nt * nt : (int64_t)nt * nt;
        if (NT <= MAX_TILES) {
            L.binned = true;
            L.tb = tb;
            L.nt = nt;
            L.NT = NT;
            break;
        }
    }
    L.bytes = L.bins_off + (L.binned ? ((2 * (BIN_SETS + 1) * (size_t)L.NT + 2 + (size_t)M) * sizeof(uint32_t) +
                                        (size_t)M * sizeof(uint16_t))
                                     : 0);
    return L;
}
static bins_view bins_of(uint32_t *counts, const counts_layout &L)
{
    uint32_t *total = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(counts) + L.bins_off);
    uint32_t *start = total + BIN_SETS * L.NT, *cursor = start + L.NT + 1;
    uint32_t *more = cursor + BIN_SETS * L.NT, *key = more + L.NT + 1;
    return bins_view{L.tb, L.nt, L.NT, total, start, cursor, more, key, reinterpret_cast<uint16_t *>(key + L.M)};
}
// zero what the tables' passes accumulate into
static hipError_t zero_counts(hipStream_t q, uint32_t *counts, const counts_layout &L)
{
    if (!L.binned) return hipMemsetAsync(counts, 0, L.zero_bytes, q);
    hipError_t e = hipMemsetAsync(counts + L.cells, 0, (1 + (size_t)L.nb) * sizeof(uint32_t), q);
    if (e == hipSuccess)
        e = hipMemsetAsync(bins_of(counts, L).total, 0, (size_t)BIN_SETS * L.NT * sizeof(uint32_t), q);
    return e;
}
static counts_view view_of(const uint32_t *counts, const counts_layout &L)
{
    return counts_view{counts, reinterpret_cast<const uint4 *>(reinterpret_cast<const char *>(counts) + L.box_off),
                       L.nb};
}

// shells counted in LDS up to this many (larger 2D grids count in memory)
constexpr int LDS_SHELLS = 8192;

// zeroed tables -> per-cell and per-shell counts -> summed-area table (in
// place) -> the cells' box counts
static void launch_counts(hipStream_t q, const fr_samples *s, unsigned N, uint32_t *counts, const counts_layout &L)
{
    // about 32 samples per thread: few enough workgroups that the central
    // box is flushed a few thousand times, not once per spoke
    const unsigned wgs = std::max(1u, std::min(nblk(s->M, 256), nblk(s->M, 256 * 32)));
    const int lds = L.nb <= LDS_SHELLS;
    if (L.binned) {
        const bins_view B = bins_of(counts, L);
        const size_t tile_lds = sizeof(uint32_t) * (size_t)L.NT;
        const size_t cell_lds = sizeof(uint32_t) << (s->dim * L.tb);
        hipLaunchKernelGGL(k_bin_count, dim3(wgs), dim3(256), tile_lds + (lds ? sizeof(uint32_t) * L.nb : 0), q,
                           s->traj, s->M, s->dim, (int)N, B, counts, L.nb, lds);
        hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, q, B);
        hipLaunchKernelGGL(k_bin_scatter, dim3(wgs), dim3(256), tile_lds, q, s->M, B);
        hipLaunchKernelGGL(k_bin_hist, dim3((unsigned)L.NT), dim3(256), cell_lds, q, B, s->dim, (int)N, counts, 0);
        // chunks beyond the tiles' first: at most M / BIN_CHUNK
        const unsigned more = (unsigned)(s->M / BIN_CHUNK);
        if (more > 0)
            hipLaunchKernelGGL(k_bin_hist, dim3(more), dim3(256), cell_lds, q, B, s->dim, (int)N, counts, 1);
    } else {
        hipLaunchKernelGGL(k_cell_hist, dim3(wgs), dim3(256), lds ? sizeof(uint32_t) * L.nb : 0, q, s->traj, s->M,
                           s->dim, (int)N, counts, L.nb, lds);
    }
    const int64_t lines = s->dim == 3 ? (int64_t)N * N : N;
    hipLaunchKernelGGL(k_prefix_x, dim3(nblk(lines, 4)), dim3(256), 0, q, counts, (int)N, lines);
    for (int ax = 1; ax < s->dim; ++ax)
        hipLaunchKernelGGL(k_prefix_axis, dim3(nblk(lines, 256)), dim3(256), 0, q, counts, (int)N, s->dim, ax);
    const uint32_t per = (uint32_t)((lines + 7) / 8);
    hipLaunchKernelGGL(k_box_counts, dim3(8 * per), dim3(256), 0, q, counts, s->dim, (int)N, (uint32_t)lines, per,
                       const_cast<uint4 *>(view_of(counts, L).B));
}

extern "C" fr_status fr_ldcf_features(fr_ctx *ctx, const fr_samples *s, unsigned N, float *feat_dev)
{
    if (!ctx || !s || !feat_dev) return FR_ERR_INVALID;
    FR_CHECK(ctx, s->has_traj, FR_ERR_STATE, "trajectory not set");
    FR_CHECK(ctx, s->frames == 0, FR_ERR_UNSUPPORTED, "features of batched frames: compute them per frame");
    FR_CHECK(ctx, N >= 2 && (s->dim == 2 || s->dim == 3), FR_ERR_INVALID, "bad grid size %u / dim %d", N, s->dim);
    FR_CHECK(ctx, fr_cells(s->dim, N) < (size_t(1) << 31), FR_ERR_UNSUPPORTED,
             "learned-DCF features on %u^%d cells (at most 2^31 cells)", N, s->dim);
    FR_EXTENT(ctx, feat_dev, sizeof(float) * FSTRIDE * (size_t)s->M, "fr_ldcf_features rows");
    const counts_layout L = layout_of(s->dim, N, s->M);
    uint32_t *hist = nullptr;
    fr_status st = fr_alloc(ctx, (void **)&hist, L.bytes);
    if (st) return st;
    hipStream_t q = ctx->stream;
    hipError_t e = zero_counts(q, hist, L);
    if (e == hipSuccess && s->M > 0) {
        launch_counts(q, s, N, hist, L);
        hipLaunchKernelGGL(k_features, dim3(nblk(s->M, 256)), dim3(256), 0, q, s->traj, s->M, s->dim, (int)N,
                           view_of(hist, L), feat_dev);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(q);
    hipFree(hist);
    FR_CHECK(ctx, e == hipSuccess, FR_ERR_HIP, "features: %s", hipGetErrorString(e));
    return FR_OK;
}

extern "C" fr_status fr_mlp_create(fr_ctx *ctx, fr_mlp **out)
{
    if (!ctx || !out) return FR_ERR_INVALID;
    fr_mlp *m = new fr_mlp;
    m->ctx = ctx;
    fr_status st;
    if ((st = fr_alloc(ctx, (void **)&m->p, NPARAM * 4)) || (st = fr_alloc(ctx, (void **)&m->m, NPARAM * 4)) ||
        (st = fr_alloc(ctx, (void **)&m->v, NPARAM * 4)) || (st = fr_alloc(ctx, (void **)&m->grad, SLOT * 4)) ||
        (st = fr_alloc(ctx, (void **)&m->w1, NF * H * 2)) || (st = fr_alloc(ctx, (void **)&m->w2, H * H * 2)) ||
        (st = fr_alloc(ctx, (void **)&m->w3, H * 2))) {
        fr_mlp_destroy(m);
        return st;
    }
    FR_HIP(ctx, hipMemsetAsync(m->p, 0, NPARAM * 4, ctx->stream));
    FR_HIP(ctx, hipMemsetAsync(m->m, 0, NPARAM * 4, ctx->stream));
    FR_HIP(ctx, hipMemsetAsync(m->v, 0, NPARAM * 4, ctx->stream));
    hipLaunchKernelGGL(k_refresh_bf16, dim3(nblk(NPARAM, 256)), dim3(256), 0, ctx->stream, m->p, m->w1, m->w2, m->w3);
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = m;
    return FR_OK;
}

extern "C" void fr_mlp_destroy(fr_mlp *m)
{
    if (!m) return;
    if (m->ctx) hipStreamSynchronize(m->ctx->stream);
    hipFree(m->p); hipFree(m->m); hipFree(m->v); hipFree(m->grad);
    hipFree(m->w1); hipFree(m->w2); hipFree(m->w3); hipFree(m->partial);
    hipFree(m->hist);
    delete m;
}

extern "C" fr_status fr_mlp_set_params(fr_mlp *m, const float *params, float mu, float sigma)
{
    if (!m || !params) return FR_ERR_INVALID;
    fr_ctx *ctx = m->ctx;
    FR_CHECK(ctx, std::isfinite(mu) && std::isfinite(sigma) && sigma > 0.f, FR_ERR_INVALID,
             "output normalisation must be finite with sigma > 0");
    FR_HIP(ctx, hipMemcpyAsync(m->p, params, NPARAM * 4, hipMemcpyHostToDevice, ctx->stream));
    FR_HIP(ctx, hipMemsetAsync(m->m, 0, NPARAM * 4, ctx->stream));
    FR_HIP(ctx, hipMemsetAsync(m->v, 0, NPARAM * 4, ctx->stream));
    hipLaunchKernelGGL(k_refresh_bf16, dim3(nblk(NPARAM, 256)), dim3(256), 0, ctx->stream, m->p, m->w1, m->w2, m->w3);
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    m->mu = mu;
    m->sigma = sigma;
    m->step = 0;
    return FR_OK;
}

extern "C" fr_status fr_mlp_set_base(fr_mlp *m, int feature)
{
    if (!m) return FR_ERR_INVALID;
    FR_CHECK(m->ctx, feature == -1 || (feature >= 3 && feature <= 5), FR_ERR_INVALID,
             "the DCF head's base is a box-count feature (3, 4 or 5) or -1, not %d", feature);
    m->base = feature;
    return FR_OK;
}

extern "C" int fr_mlp_get_base(const fr_mlp *m) { return m ? m->base : -1; }

extern "C" fr_status fr_mlp_get_params(fr_mlp *m, float *params, float *mu, float *sigma)
{
    if (!m || !params) return FR_ERR_INVALID;
    fr_ctx *ctx = m->ctx;
    FR_HIP(ctx, hipMemcpyAsync(params, m->p, NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream));
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (mu) *mu = m->mu;
    if (sigma) *sigma = m->sigma;
    return FR_OK;
}

static unsigned fwd_blocks(fr_ctx *ctx, int64_t M)
{
    const int64_t groups = (M + 15) / 16;
    const int64_t waves = std::min<int64_t>(groups, (int64_t)std::max(ctx->cu_count, 1) * 32);
    return (unsigned)std::max<int64_t>(1, (waves + 3) / 4);
}

extern "C" fr_status fr_mlp_forward(fr_mlp *m, const float *feat_dev, int64_t M, float *y_dev)
{
    if (!m || !feat_dev || !y_dev || M < 0) return FR_ERR_INVALID;
    fr_ctx *ctx = m->ctx;
    if (M == 0) return FR_OK;
    FR_EXTENT(ctx, feat_dev, sizeof(float) * FSTRIDE * (size_t)M, "fr_mlp_forward feature rows");
    FR_EXTENT(ctx, y_dev, sizeof(float) * (size_t)M, "fr_mlp_forward output");
    hipLaunchKernelGGL(k_mlp_forward, dim3(fwd_blocks(ctx, M)), dim3(256), 0, ctx->stream, view(m), feat_dev, M, 0,
                       0.f, 1.f, -1, y_dev);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_ldcf_dcf(fr_mlp *m, fr_samples *s, unsigned N)
{
    if (!m || !s) return FR_ERR_INVALID;
    fr_ctx *ctx = m->ctx;
    FR_CHECK(ctx, s->has_traj, FR_ERR_STATE, "trajectory not set");
    FR_CHECK(ctx, s->frames == 0, FR_ERR_UNSUPPORTED, "learned DCF of batched frames: compute it per frame");
    FR_CHECK(ctx, N >= 2 && (s->dim == 2 || s->dim == 3), FR_ERR_INVALID, "bad grid size %u / dim %d", N, s->dim);
    FR_CHECK(ctx, fr_cells(s->dim, N) < (size_t(1) << 31), FR_ERR_UNSUPPORTED,
             "learned-DCF features on %u^%d cells (at most 2^31 cells)", N, s->dim);
    const counts_layout L = layout_of(s->dim, N, s->M);
    if (L.bytes > m->hist_bytes) {
        FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        hipFree(m->hist);
        m->hist = nullptr;
        m->hist_bytes = 0;
        fr_status st = fr_alloc(ctx, (void **)&m->hist, L.bytes);
        if (st) return st;
        m->hist_bytes = L.bytes;
    }
    if (s->M > 0) {
        FR_HIP(ctx, zero_counts(ctx->stream, m->hist, L));
        launch_counts(ctx->stream, s, N, m->hist, L);
        const int64_t groups = (s->M + 63) / 64;
        const int64_t waves = std::min<int64_t>(groups, (int64_t)std::max(ctx->cu_count, 1) * 32);
        hipLaunchKernelGGL(k_ldcf_dcf, dim3((unsigned)std::max<int64_t>(1, (waves + 3) / 4)), dim3(256), 0,
                           ctx->stream, view(m), s->traj, s->M, s->dim, (int)N, view_of(m->hist, L), 1, m->mu,
                           m->sigma, m->base, s->dcf);
    }
    FR_HIP(ctx, hipGetLastError());
    s->has_dcf = true;
    fr_dcf_changed(s);
    return FR_OK;
}

extern "C" fr_status fr_mlp_dcf(fr_mlp *m, const float *feat_dev, fr_samples *s)
{
    if (!m || !feat_dev || !s) return FR_ERR_INVALID;
    fr_ctx *ctx = m->ctx;
    FR_EXTENT(ctx, feat_dev, sizeof(float) * FSTRIDE * (size_t)s->M, "fr_mlp_dcf feature rows");
    if (s->M > 0)
        hipLaunchKernelGGL(k_mlp_forward, dim3(fwd_blocks(ctx, s->M)), dim3(256), 0, ctx->stream, view(m), feat_dev,
                           s->M, 1, m->mu, m->sigma, m->base, s->dcf);
    FR_HIP(ctx, hipGetLastError());
    s->has_dcf = true;
    fr_dcf_changed(s);
    return FR_OK;
}

// gradient (and loss) of the batch into m->grad; shared by train and grad
static fr_status mlp_backward(fr_mlp *m, const float *feat_dev, const float *targ_dev, const int64_t *idx_dev,
                              int64_t B)
{
    fr_ctx *ctx = m->ctx;
    FR_CHECK(ctx, B > 0 && feat_dev && targ_dev, FR_ERR_INVALID, "empty training batch");
    const int64_t groups = (B + 15) / 16;
    const int64_t maxw = (int64_t)std::max(ctx->cu_count, 1) * 4;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((groups + TB_WAVES - 1) / TB_WAVES, maxw));
    const int64_t nwaves = blocks * TB_WAVES;
    const int iters = (int)((groups + nwaves - 1) / nwaves);
    if (nwaves > m->partial_slots) {
        FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        hipFree(m->partial);
        m->partial = nullptr;
        m->partial_slots = 0;
        fr_status st = fr_alloc(ctx, (void **)&m->partial, (size_t)nwaves * SLOT * 4);
        if (st) return st;
        m->partial_slots = nwaves;
    }
    hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)blocks), dim3(64 * TB_WAVES), 0, ctx->stream, view(m), feat_dev,
                       targ_dev, idx_dev, B, iters, m->partial);
    hipLaunchKernelGGL(k_reduce_slots, dim3(nblk(SLOT, 256)), dim3(256), 0, ctx->stream, m->partial, nwaves, m->grad);
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

extern "C" fr_status fr_mlp_grad(fr_mlp *m, const float *feat_dev, const float *targ_dev, const int64_t *idx_dev,
                                 int64_t B, float *grad_host, float *loss)
{
    if (!m || !grad_host) return FR_ERR_INVALID;
    fr_ctx *ctx = m->ctx;
    fr_status st = mlp_backward(m, feat_dev, targ_dev, idx_dev, B);
    if (st) return st;
    float tmp[SLOT];
    FR_HIP(ctx, hipMemcpyAsync(tmp, m->grad, SLOT * 4, hipMemcpyDeviceToHost, ctx->stream));
    FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(grad_host, tmp, NPARAM * 4);
    if (loss) *loss = tmp[NPARAM] / (float)B;
    return FR_OK;
}

extern "C" fr_status fr_mlp_train_step(fr_mlp *m, const float *feat_dev, const float *targ_dev,
                                       const int64_t *idx_dev, int64_t B, float lr, float *loss)
{
    if (!m) return FR_ERR_INVALID;
    fr_ctx *ctx = m->ctx;
    FR_CHECK(ctx, lr > 0.f && std::isfinite(lr), FR_ERR_INVALID, "learning rate must be positive");
    fr_status st = mlp_backward(m, feat_dev, targ_dev, idx_dev, B);
    if (st) return st;
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(NPARAM, 256)), dim3(256), 0, ctx->stream, m->p, m->m, m->v, m->grad, lr, b1,
                       b2, eps, c1, c2, m->w1, m->w2, m->w3);
    FR_HIP(ctx, hipGetLastError());
    if (loss) {
        float l;
        FR_HIP(ctx, hipMemcpyAsync(&l, m->grad + NPARAM, 4, hipMemcpyDeviceToHost, ctx->stream));
        FR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        *loss = l / (float)B;
    }
    return FR_OK;
}
