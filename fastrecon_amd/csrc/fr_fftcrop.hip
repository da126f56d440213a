// Cropped FFT + fftshift + sum-of-squares in three (3D) or two (2D) passes.
//
// The reconstruction needs, per coil, the unnormalised sign +1 FFT of the
// grid (ImageFilter/FFT.cpp:40-51,82-84), fftshifted (ImageFilter.cpp:231-
// 280), cropped to rc at offset (N - rc)/2 and combined as sqrt(sum_c |v_c|^2)
// (ImageFilter.cpp:159-209, Main.cpp:175-181).  Only rc of the N outputs of
// each 1D transform reach the crop, so each pass transforms rows along one
// axis, keeps those outputs and writes them transposed for the next axis:
//
//   3D  G[c][z][y][x]  -x->  B1[c][z][kx][y]  -y->  B2[c][ky][kx][z]  -z->  sum over c of |.|^2
//   2D  G[c][f][y][x]  -x->  B1[c][f][kx][y]  -y->  sum over c of |.|^2  (f = frame)
//
// k indices are crop positions (0 .. rc-1, already fftshifted).  Each pass
// reads its input once and writes half (rc = N/2) of it; the last pass sums
// the coils in order, fl(re*re) + fl(im*im) per coil, then sqrtf and the
// optional deapodisation divide: the arithmetic of k_sos (fr_image.hip).
//
// Row transform: N-point Stockham radix-4 (N = 4^m, or 2 * 4^m with a final
// radix-2 step; N = 384 takes radices 6, 8, 8), one wave per row, twiddles exp(+2 pi i k / N) from an FP64
// table rounded to FP32.  The result differs from rocFFT's in rounding
// (relative ~1e-7), well inside the 1e-5 image bar.
#include "fr_internal.h"

#include <cmath>
#include <vector>

namespace {

constexpr int TR = 16;          // rows per workgroup tile of the x and y passes
constexpr int TS = 8;           // rows (output columns) per workgroup tile of the last pass: twice the
                                // workgroups of a 16-row tile (the pass has only rc^2 / TS tiles)
// waves per workgroup, each taking TR / NW rows of the tile (8 waves for
// N = 256 measured slower than 4: 2.26 vs 2.16 ms for 32 coils at 256^3)
#ifndef FR_FFT_WAVES
#define FR_FFT_WAVES 4
#endif
template <int N> constexpr int waves() { return N >= 512 ? 4 : FR_FFT_WAVES; }
static_assert(TR % (2 * FR_FFT_WAVES) == 0 && TS % (2 * FR_FFT_WAVES) == 0 && 16 % TS == 0,
              "an even number of rows per wave (the passes load two rows ahead, the last pass alternates two buffers)");

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__fsub_rn(__fmul_rn(a.x, b.x), __fmul_rn(a.y, b.y)),
                       __fadd_rn(__fmul_rn(a.x, b.y), __fmul_rn(a.y, b.x)));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(__fadd_rn(a.x, b.x), __fadd_rn(a.y, b.y)); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(__fsub_rn(a.x, b.x), __fsub_rn(a.y, b.y)); }

// Input slots per lane: b4 sets of 4 (the first radix-4 iteration's inputs;
// N = 384 keeps its 6 first-stage inputs in the first 6 slots)
template <int N> constexpr int b4() { return (N / 4 + 63) / 64; }
// Outputs per lane after a row transform, and which output k lane l holds in
// v[q]: k = l + 64 q, or for N = 384 (last stage radix 8 over 48 lanes)
// k = l + 48 q on lanes 0 .. 47.
template <int N> constexpr int nv() { return N == 384 ? 8 : N / 64; }
template <int N> __device__ __forceinline__ int out_k(int lane, int q)
{
    if constexpr (N == 384) return lane < 48 ? lane + 48 * q : -1;
    else return lane + 64 * q;
}

// The first radix-4 iteration's inputs of the row at `in` (unit stride):
// lane l, butterfly b takes elements j + r N/4, j = l + 64 b.  Loaded one row
// ahead so that the global-memory latency overlaps the previous row's work.
template <int N>
__device__ __forceinline__ void load_row(const float2 *__restrict__ in, float2 (&x0)[b4<N>()][4], int lane)
{
    if constexpr (N == 384) {
        // radix-6 first stage: lane l takes elements l + 64 r, r = 0 .. 5
#pragma unroll
        for (int r = 0; r < 6; ++r) x0[r >> 2][r & 3] = in[lane + 64 * r];
        x0[1][2] = x0[1][3] = make_float2(0.f, 0.f);
    } else {
#pragma unroll
        for (int b = 0; b < b4<N>(); ++b) {
            const int j = lane + 64 * b;
#pragma unroll
            for (int r = 0; r < 4; ++r) x0[b][r] = in[j + r * (N / 4)];
        }
    }
}

// Natural-order N-point transform of a row whose first-iteration inputs are
// in x0 (load_row).  After the call, lane l holds output k = out_k<N>(l, q)
// in v[q] (q = 0 .. nv<N>() - 1).  buf: 2N float2 of this wave's LDS; tw: N twiddles
// exp(s 2 pi i k / N).
__device__ __forceinline__ float2 mul_si(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }  // a (s i)

// 3-point DFT with w = exp(s 2 pi i / 3), h = s sqrt(3) / 2
__device__ __forceinline__ void dft3(float2 b0, float2 b1, float2 b2, float h, float2 &X0, float2 &X1, float2 &X2)
{
    const float2 sum = cadd(b1, b2), dif = csub(b1, b2);
    const float2 m = make_float2(__fsub_rn(b0.x, __fmul_rn(0.5f, sum.x)), __fsub_rn(b0.y, __fmul_rn(0.5f, sum.y)));
    const float2 rot = make_float2(-__fmul_rn(h, dif.y), __fmul_rn(h, dif.x));      // i h (b1 - b2)
    X0 = cadd(b0, sum);
    X1 = cadd(m, rot);
    X2 = csub(m, rot);
}

// 4-point DFT with w = exp(s 2 pi i / 4) = s i
__device__ __forceinline__ void dft4(float2 a0, float2 a1, float2 a2, float2 a3, float s, float2 (&X)[4])
{
    const float2 a = cadd(a0, a2), bb = csub(a0, a2), c = cadd(a1, a3), d = mul_si(csub(a1, a3), s);
    X[0] = cadd(a, c);
    X[1] = cadd(bb, d);
    X[2] = csub(a, c);
    X[3] = csub(bb, d);
}

// 6-point DFT, w6 = exp(s 2 pi i / 6): n = n2 + 3 n1, k = k1 + 2 k2; size-2
// DFTs over n1, twiddles w6^(n2 k1), size-3 DFTs over n2
__device__ __forceinline__ void dft6(float2 (&x)[6], float s)
{
    const float h = 0.866025403784438647f * s;                       // s sqrt(3) / 2
    float2 e[3], o[3];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
        e[n2] = cadd(x[n2], x[n2 + 3]);
        o[n2] = csub(x[n2], x[n2 + 3]);
    }
    // o1 *= w6 = (1/2, h); o2 *= w6^2 = (-1/2, h)
    const float2 o1 = make_float2(__fsub_rn(__fmul_rn(0.5f, o[1].x), __fmul_rn(h, o[1].y)),
                                  __fadd_rn(__fmul_rn(0.5f, o[1].y), __fmul_rn(h, o[1].x)));
    const float2 o2 = make_float2(__fsub_rn(__fmul_rn(-0.5f, o[2].x), __fmul_rn(h, o[2].y)),
                                  __fadd_rn(__fmul_rn(-0.5f, o[2].y), __fmul_rn(h, o[2].x)));
    dft3(e[0], e[1], e[2], h, x[0], x[2], x[4]);
    dft3(o[0], o1, o2, h, x[1], x[3], x[5]);
}

// 8-point DFT, w8 = exp(s 2 pi i / 8): even and odd 4-point DFTs, then
// X[k] = E[k] + w8^k O[k], X[k + 4] = E[k] - w8^k O[k]
__device__ __forceinline__ void dft8(float2 (&x)[8], float s)
{
    const float c = 0.707106781186547524f;                            // sqrt(1/2)
    float2 E[4], O[4];
    dft4(x[0], x[2], x[4], x[6], s, E);
    dft4(x[1], x[3], x[5], x[7], s, O);
    const float sc = s * c;
    // w8 = (c, s c), w8^2 = s i, w8^3 = (-c, s c)
    const float2 t1 = make_float2(__fsub_rn(__fmul_rn(c, O[1].x), __fmul_rn(sc, O[1].y)),
                                  __fadd_rn(__fmul_rn(c, O[1].y), __fmul_rn(sc, O[1].x)));
    const float2 t2 = mul_si(O[2], s);
    const float2 t3 = make_float2(__fsub_rn(__fmul_rn(-c, O[3].x), __fmul_rn(sc, O[3].y)),
                                  __fadd_rn(__fmul_rn(-c, O[3].y), __fmul_rn(sc, O[3].x)));
    x[0] = cadd(E[0], O[0]);
    x[4] = csub(E[0], O[0]);
    x[1] = cadd(E[1], t1);
    x[5] = csub(E[1], t1);
    x[2] = cadd(E[2], t2);
    x[6] = csub(E[2], t2);
    x[3] = cadd(E[3], t3);
    x[7] = csub(E[3], t3);
}

// N = 384 = 6 8 8 (Stockham): a radix-6 stage (Ns = 1, 64 butterflies, one
// per lane, inputs from x0), a radix-8 stage (Ns = 6, 48 butterflies) and a
// final radix-8 stage (Ns = 48, 48 butterflies) whose outputs j + 48 r are
// lane j's v[r] (lanes 48 .. 63 hold none).  Twiddles of a radix-R stage:
// exp(s 2 pi i r (j mod Ns) / (R Ns)) = tw[r (j mod Ns) N / (R Ns)].
__device__ __forceinline__ void fft_row384(const float2 (&x0)[b4<384>()][4], float2 (&v)[8], float2 *buf,
                                           const float2 *tw, int lane, float s)
{
    float2 *p = buf, *q = buf + 384;
    {
        float2 x[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) x[r] = x0[r >> 2][r & 3];
        dft6(x, s);
#pragma unroll
        for (int r = 0; r < 6; ++r) p[6 * lane + r] = x[r];
    }
    if (lane < 48) {
        const int jm = lane % 6;
        float2 x[8];
        x[0] = p[lane];
#pragma unroll
        for (int r = 1; r < 8; ++r) x[r] = cmul(p[lane + 48 * r], tw[8 * r * jm]);
        dft8(x, s);
        const int idx = (lane / 6) * 48 + jm;
#pragma unroll
        for (int r = 0; r < 8; ++r) q[idx + 6 * r] = x[r];
    }
    if (lane < 48) {
        v[0] = q[lane];
#pragma unroll
        for (int r = 1; r < 8; ++r) v[r] = cmul(q[lane + 48 * r], tw[r * lane]);
        dft8(v, s);
    } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = make_float2(0.f, 0.f);
    }
}

template <int N>
__device__ __forceinline__ void fft_row(const float2 (&x0)[b4<N>()][4], float2 (&v)[nv<N>()], float2 *buf,
                                        const float2 *tw, int lane, float s)
{
    if constexpr (N == 384) {
        fft_row384(x0, v, buf, tw, lane, s);
    } else {
        constexpr int Q = N / 64;                    // values per lane
        constexpr int B4 = N / 4 / 64;               // radix-4 butterflies per lane
        float2 *src = buf, *dst = buf + N;
        int Ns = 1;
        bool first = true;
    #pragma unroll
        for (int it = 0; it < 8; ++it) {
            if (Ns * 4 > N) break;
    #pragma unroll
            for (int b = 0; b < B4; ++b) {
                const int j = lane + 64 * b;
                float2 x[4];
    #pragma unroll
                for (int r = 0; r < 4; ++r) x[r] = first ? x0[b][r] : src[j + r * (N / 4)];
                const int jm = j % Ns;
                const int t = jm * (N / (4 * Ns));
    #pragma unroll
                for (int r = 1; r < 4; ++r) x[r] = cmul(x[r], tw[(t * r) & (N - 1)]);
                const float2 a = cadd(x[0], x[2]), bb = csub(x[0], x[2]), c = cadd(x[1], x[3]);
                const float2 dd = csub(x[1], x[3]);
                const float2 d = make_float2(-s * dd.y, s * dd.x);          // (x1 - x3) * (s i)
                const int idx = (j / Ns) * Ns * 4 + jm;
                dst[idx] = cadd(a, c);
                dst[idx + Ns] = cadd(bb, d);
                dst[idx + 2 * Ns] = csub(a, c);
                dst[idx + 3 * Ns] = csub(bb, d);
            }
            first = false;
            float2 *tmp = src; src = dst; dst = tmp;
            Ns *= 4;
        }
        if (Ns < N) {
            // N = 2 * 4^m: one radix-2 iteration with Ns = N / 2 (output index = j)
    #pragma unroll
            for (int b = 0; b < N / 2 / 64; ++b) {
                const int j = lane + 64 * b;
                const float2 x0 = src[j];
                const float2 x1 = cmul(src[j + N / 2], tw[j]);
                v[b] = cadd(x0, x1);
                v[b + N / 128] = csub(x0, x1);
            }
        } else {
    #pragma unroll
            for (int q = 0; q < Q; ++q) v[q] = src[lane + 64 * q];
        }
    }
}

// crop position of output k (or -1): shifted crop k' = (k - N/2 - x0) mod N
__device__ __forceinline__ int crop_pos(int k, int N, int rc, int shift)
{
    const int x0 = (N - rc) / 2;
    int kp = shift ? k - N / 2 - x0 : k - x0;
    kp = kp < 0 ? kp + N : kp;
    kp = kp >= N ? kp - N : kp;
    return kp < rc ? kp : -1;
}

// Pass over tiles of TR rows: row t of tile T starts at in + rowA(T) + t *
// rowStride; kept output k' of row t goes to out[outA(T) + k' * outK + t].
// Layout parameters: tile T = (a, b, c0) over dims (na, nb, nt = rows / TR).
struct pass_desc {
    int64_t na, nb, nt;                          // tile grid
    int64_t inA, inB, inT, rowStride;            // in offset = a inA + b inB + (t0 + t) rowStride (t0 = tt * TR)
    int64_t outA, outB, outT, outK;              // out offset = a outA + b outB + tt*TR*outT + k' outK + t outT
};

template <int N, int NW = waves<N>(), int WGT = 64 * waves<N>()>
__global__ void __launch_bounds__(WGT) k_fft_pass(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                  const float2 *__restrict__ gtw, pass_desc d, int rc, int shift,
                                                  float s)
{
    __shared__ float2 tw[N];
    __shared__ float2 fbuf[NW][2 * N];
    extern __shared__ float2 otile[];           // [rc][TR]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < N; i += WGT) tw[i] = gtw[i];
    __syncthreads();
    const int64_t T = blockIdx.x;
    const int64_t tt = T % d.nt, b = (T / d.nt) % d.nb, a = T / (d.nt * d.nb);
    const float2 *base = in + a * d.inA + b * d.inB + tt * TR * d.rowStride;
    // rows w, w + NW, .. of the tile, each wave's rows loaded two ahead
    float2 xa[b4<N>()][4], xb[b4<N>()][4], xc[b4<N>()][4];
    load_row<N>(base + (int64_t)w * d.rowStride, xa, lane);
    load_row<N>(base + (int64_t)(w + NW) * d.rowStride, xb, lane);
#pragma unroll
    for (int i = 0; i < TR / NW; ++i) {
        const int t = w + NW * i;
        float2 (&cur)[b4<N>()][4] = i % 3 == 0 ? xa : (i % 3 == 1 ? xb : xc);
        float2 (&nxt)[b4<N>()][4] = (i + 2) % 3 == 0 ? xa : ((i + 2) % 3 == 1 ? xb : xc);
        if (i + 2 < TR / NW) load_row<N>(base + (int64_t)(t + 2 * NW) * d.rowStride, nxt, lane);
        float2 v[nv<N>()];
        fft_row<N>(cur, v, fbuf[w], tw, lane, s);
#pragma unroll
        for (int q = 0; q < nv<N>(); ++q) {
            const int k = out_k<N>(lane, q);
            const int kp = k < 0 ? -1 : crop_pos(k, N, rc, shift);
            if (kp >= 0) otile[kp * TR + t] = v[q];
        }
    }
    __syncthreads();
    float2 *ob = out + a * d.outA + b * d.outB + tt * TR * d.outT;
    for (int i = threadIdx.x; i < rc * TR; i += WGT) {
        const int kp = i / TR, t = i % TR;
        ob[(int64_t)kp * d.outK + (int64_t)t * d.outT] = otile[i];
    }
}

// Last pass with the coil sum: tile (a = outer slice, tt = TS rows), coils
// c = 0 .. C-1 in order; row t of coil c at in + c inC + a inA + (tt TS + t)
// rowStride.  acc[k'][t] += fl(re^2) + fl(im^2); then sqrt, deapod, store to
// img[a outA + k' outK + tt TS + t].
template <int N, int NW = waves<N>(), int WGT = 64 * waves<N>()>
__global__ void __launch_bounds__(WGT) k_fft_sos(const float2 *__restrict__ in, float *__restrict__ img,
                                                 const float2 *__restrict__ gtw, int64_t na, int64_t nt, int C,
                                                 int64_t inC, int64_t inA, int64_t rowStride, int64_t outA,
                                                 int64_t outK, int rc, int shift, float s,
                                                 const float *__restrict__ deapod, int64_t dmod)
{
    __shared__ float2 tw[N];
    __shared__ float2 fbuf[NW][2 * N];
    extern __shared__ float acc[];               // [rc][TS]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < N; i += WGT) tw[i] = gtw[i];
    for (int i = threadIdx.x; i < rc * TS; i += WGT) acc[i] = 0.f;
    __syncthreads();
    const int64_t T = blockIdx.x;
    const int64_t tt = T % nt, a = T / nt;
    // rows (c, w + NW i) in coil order, each wave's next row loaded ahead
    const float2 *tbase = in + a * inA + tt * TS * rowStride;
    float2 xa[b4<N>()][4], xb[b4<N>()][4];
    load_row<N>(tbase + (int64_t)w * rowStride, xa, lane);
#pragma unroll 1
    for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int i = 0; i < TS / NW; ++i) {
            const int t = w + NW * i;
            float2 (&cur)[b4<N>()][4] = (i & 1) ? xb : xa;
            float2 (&nxt)[b4<N>()][4] = (i & 1) ? xa : xb;
            // next row: (c, t + NW), or (c + 1, w) after the last of this coil
            if (i + 1 < TS / NW)
                load_row<N>(tbase + c * inC + (int64_t)(t + NW) * rowStride, nxt, lane);
            else if (c + 1 < C)
                load_row<N>(tbase + (c + 1) * inC + (int64_t)w * rowStride, nxt, lane);
            float2 v[nv<N>()];
            fft_row<N>(cur, v, fbuf[w], tw, lane, s);
#pragma unroll
            for (int q = 0; q < nv<N>(); ++q) {
                const int k = out_k<N>(lane, q);
                const int kp = k < 0 ? -1 : crop_pos(k, N, rc, shift);
                if (kp >= 0) {
                    float &x = acc[kp * TS + t];
                    x = __fadd_rn(x, __fadd_rn(__fmul_rn(v[q].x, v[q].x), __fmul_rn(v[q].y, v[q].y)));
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < rc * TS; i += WGT) {
        const int kp = i / TS, t = i % TS;
        const int64_t o = a * outA + (int64_t)kp * outK + tt * TS + t;
        float m = sqrtf(acc[i]);
        if (deapod) m = __fdiv_rn(m, deapod[o % dmod]);
        img[o] = m;
    }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <int N>
fr_status run_passes(fr_fft *f, const fr_grids *g, int rcx, int rcy, int rcz, int shift, const float *deapod,
                     float *out)
{
    fr_ctx *ctx = f->ctx;
    const int C = g->C;
    const float s = (float)f->sign;
    constexpr int WGT = 64 * waves<N>();
    const size_t lds_row = sizeof(float2) * 16;
    if (g->dim == 3) {
        // B1[c][z][kx][y], B2[c][ky][kx][z]
        const int64_t n1 = (int64_t)C * N * rcx * N, n2 = (int64_t)C * rcy * rcx * N;
        fr_status e = fr_fft_workspace(f, sizeof(float2) * (size_t)(n1 + n2));
        if (e) return e;
        float2 *B1 = (float2 *)f->crop_ws, *B2 = B1 + n1;
        // x: tiles (c, z, y-block); rows along x; out B1
        pass_desc d1{C, N, N / TR, (int64_t)N * N * N, (int64_t)N * N, 0, N,
                     (int64_t)N * rcx * N, (int64_t)rcx * N, 1, N};
        hipLaunchKernelGGL(k_fft_pass<N>, dim3((unsigned)(C * N * (N / TR))), dim3(WGT), lds_row * rcx, ctx->stream,
                           g->data, B1, f->tw, d1, rcx, shift, s);
        // y: tiles (c, kx, z-block); row (c, z, kx) of B1; out B2[c][ky][kx][z]
        pass_desc d2{C, rcx, N / TR, (int64_t)N * rcx * N, N, 0, (int64_t)rcx * N,
                     (int64_t)rcy * rcx * N, N, 1, (int64_t)rcx * N};
        hipLaunchKernelGGL(k_fft_pass<N>, dim3((unsigned)(C * rcx * (N / TR))), dim3(WGT), lds_row * rcy,
                           ctx->stream, B1, B2, f->tw, d2, rcy, shift, s);
        // z + coil sum: tiles (ky, kx-block); row (c, ky, kx) of B2; img[kz][ky][kx]
        hipLaunchKernelGGL(k_fft_sos<N>, dim3((unsigned)(rcy * (rcx / TS))), dim3(WGT), sizeof(float) * TS * rcz,
                           ctx->stream, B2, out, f->tw, (int64_t)rcy, (int64_t)(rcx / TS), C,
                           (int64_t)rcy * rcx * N, (int64_t)rcx * N, (int64_t)N, (int64_t)rcx,
                           (int64_t)rcy * rcx, rcz, shift, s, deapod, (int64_t)rcz * rcy * rcx);
    } else {
        const int F = std::max(g->frames, 1);
        const int64_t n1 = (int64_t)C * F * rcx * N;
        fr_status e = fr_fft_workspace(f, sizeof(float2) * (size_t)n1);
        if (e) return e;
        float2 *B1 = (float2 *)f->crop_ws;
        // x: tiles (c, f, y-block); out B1[c][f][kx][y]
        pass_desc d1{C, F, N / TR, (int64_t)F * N * N, (int64_t)N * N, 0, N,
                     (int64_t)F * rcx * N, (int64_t)rcx * N, 1, N};
        hipLaunchKernelGGL(k_fft_pass<N>, dim3((unsigned)(C * F * (N / TR))), dim3(WGT), lds_row * rcx, ctx->stream,
                           g->data, B1, f->tw, d1, rcx, shift, s);
        // y + coil sum: tiles (f, kx-block); row (c, f, kx) of B1; img[f][ky][kx]
        // (the deapodisation map is one rcy x rcx image shared by the frames)
        hipLaunchKernelGGL(k_fft_sos<N>, dim3((unsigned)(F * (rcx / TS))), dim3(WGT), sizeof(float) * TS * rcy,
                           ctx->stream, B1, out, f->tw, (int64_t)F, (int64_t)(rcx / TS), C, (int64_t)F * rcx * N,
                           (int64_t)rcx * N, (int64_t)N, (int64_t)rcy * rcx, (int64_t)rcx, rcy, shift, s, deapod,
                           (int64_t)rcy * rcx);
    }
    FR_HIP(ctx, hipGetLastError());
    return FR_OK;
}

}  // namespace

fr_status fr_fft_workspace(fr_fft *f, size_t bytes)
{
    if (bytes <= f->crop_ws_bytes) return FR_OK;
    FR_HIP(f->ctx, hipStreamSynchronize(f->ctx->stream));
    (void)hipFree(f->crop_ws);
    f->crop_ws = nullptr;
    f->crop_ws_bytes = 0;
    fr_status e = fr_alloc(f->ctx, &f->crop_ws, bytes);
    if (e) return e;
    f->crop_ws_bytes = bytes;
    return FR_OK;
}

bool fr_fft_sos_fused(const fr_fft *f, const fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz)
{
    (void)rcz;
    return (f->N == 256 || f->N == 384 || f->N == 512) && g->N == f->N && rcx % 16 == 0 && rcx >= 16 && rcy >= 1 &&
           (g->dim == 2 || rcz >= 1) && (int64_t)g->C * std::max(g->frames, 1) == f->batch;
}

// Twiddles exp(sign 2 pi i k / N), k < N, FP64-evaluated, uploaded once per
// plan (the first fr_fft_sos call; run it before capturing a graph).
fr_status fr_fft_tw(fr_fft *f)
{
    if (f->tw) return FR_OK;
    std::vector<float2> tw(f->N);
    for (unsigned k = 0; k < f->N; ++k) {
        const double a = 2.0 * M_PI * (double)k / (double)f->N * (double)f->sign;
        tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2 *d = nullptr;
    fr_status e = fr_alloc(f->ctx, (void **)&d, sizeof(float2) * f->N);
    if (e) return e;
    hipError_t h = hipMemcpyAsync(d, tw.data(), sizeof(float2) * f->N, hipMemcpyHostToDevice, f->ctx->stream);
    if (h == hipSuccess) h = hipStreamSynchronize(f->ctx->stream);
    if (h != hipSuccess) {
        (void)hipFree(d);
        return fr_fail(f->ctx, FR_ERR_HIP, "twiddle upload: %s", hipGetErrorString(h));
    }
    f->tw = d;
    return FR_OK;
}

extern "C" fr_status fr_fft_sos(fr_fft *f, fr_grids *g, unsigned rcx, unsigned rcy, unsigned rcz, int apply_shift,
                                const float *deapod_dev, float *out_dev)
{
    if (!f || !g || !out_dev) return FR_ERR_INVALID;
    fr_ctx *ctx = f->ctx;
    if (g->dim == 2) rcz = g->frames > 0 ? (unsigned)g->frames : 1u;
    FR_CHECK(ctx, g->dim == f->dim && g->N == f->N && (int64_t)g->C * std::max(g->frames, 1) == f->batch,
             FR_ERR_INVALID, "grids (dim %d, N %u, C %d) do not match the FFT plan (dim %d, N %u, batch %d)", g->dim,
             g->N, g->C, f->dim, f->N, f->batch);
    FR_CHECK(ctx, rcx >= 1 && rcy >= 1 && rcz >= 1 && rcx <= g->N && rcy <= g->N && (g->dim == 2 || rcz <= g->N),
             FR_ERR_INVALID, "crop %ux%ux%u outside grid %u", rcx, rcy, rcz, g->N);
    FR_CHECK(ctx, !apply_shift || g->N % 2 == 0, FR_ERR_UNSUPPORTED, "fftshift needs an even grid size");
    FR_EXTENT(ctx, out_dev, sizeof(float) * rcx * rcy * rcz, "fr_fft_sos image");
    FR_EXTENT(ctx, deapod_dev, sizeof(float) * rcx * rcy * (g->dim == 3 ? rcz : 1u), "fr_fft_sos deapodisation map");
    if (!fr_fft_sos_fused(f, g, rcx, rcy, rcz)) {
        // other sizes: the full batched FFT in place, then the fused epilogue
        fr_status e = fr_fft_execute(f, g);
        if (!e) e = fr_sos(ctx, g, rcx, rcy, rcz, apply_shift, deapod_dev, out_dev);
        return e;
    }
    fr_status e = fr_fft_tw(f);
    if (e) return e;
    if (f->N == 256) return run_passes<256>(f, g, (int)rcx, (int)rcy, (int)rcz, apply_shift, deapod_dev, out_dev);
    if (f->N == 384) return run_passes<384>(f, g, (int)rcx, (int)rcy, (int)rcz, apply_shift, deapod_dev, out_dev);
    return run_passes<512>(f, g, (int)rcx, (int)rcy, (int)rcz, apply_shift, deapod_dev, out_dev);
}
