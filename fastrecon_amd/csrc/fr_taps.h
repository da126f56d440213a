// Per-sample tap enumeration with the reference CPU arithmetic.
//
// For sample position c (normalised grid units) the reference
// (ReconData/GridLut.cpp:99-136) visits the clamped box
//   start = max(ceil(c - hw), 0), end = min(floor(c + hw), N - 1)
// in z, y, x order, computes dk = sqrtf((dz*dz + dy*dy) + dx*dx) with
// dx = float(x) - cx etc., accepts dk < hw and takes the LUT entry
// ki = (int)roundf((dk / hw) * float(L - 1)).  The device code below keeps
// every one of those float operations (IEEE sqrt and division, no
// contraction: the library is built with -ffp-contract=off) so that tap
// sets and LUT indices are identical to the CPU's.  Note: on gfx950 plain
// sqrtf() lowers to v_sqrt_f32 plus the residual-based correction to the
// correctly rounded result, whereas __fsqrt_rn() lowers to the bare 1-ulp
// v_sqrt_f32; only sqrtf() is used here.
#pragma once

#include <hip/hip_runtime.h>

struct fr_box {
    int s[3], e[3];
    float c[3];
};

// For batched 2D frames (frames > 0) the third trajectory row holds the
// sample's frame index f: the box's z range is exactly [f, f] and its z
// centre is f, so dz = 0 and the taps are the 2D ones of that frame.
__device__ __forceinline__ void fr_sample_box(int dim, int N, const float *__restrict__ traj, int64_t M,
                                              int64_t s, float hw, fr_box &b, int frames = 0)
{
#pragma unroll
    for (int i = 0; i < 3; ++i) { b.c[i] = 0.f; b.s[i] = 0; b.e[i] = 0; }
    for (int i = 0; i < dim; ++i) {
        float c = traj[(int64_t)i * M + s];
        b.c[i] = c;
        int st = (int)ceilf(c - hw);
        int en = (int)floorf(c + hw);
        b.s[i] = st < 0 ? 0 : st;
        b.e[i] = en > N - 1 ? N - 1 : en;
    }
    if (frames > 0) {
        const float f = traj[(int64_t)2 * M + s];
        const int fi = min(max((int)f, 0), frames - 1);
        b.c[2] = (float)fi;
        b.s[2] = fi;
        b.e[2] = fi;
    }
}

__device__ __forceinline__ float fr_tap_dk2(float dz2, float dy2, float dx)
{
    // (dz2 + dy2) + dx*dx, the reference's left-to-right order
    return __fadd_rn(__fadd_rn(dz2, dy2), __fmul_rn(dx, dx));
}

__device__ __forceinline__ int fr_tap_index(float dk, float hw, int L)
{
    float q = __fdiv_rn(dk, hw);
    float r = __fmul_rn(q, (float)(L - 1));
    return (int)roundf(r);
}

// Calls f(x, y, z, ki) for every accepted tap of box b restricted to the
// sub-box [lo, hi] (inclusive), in z, y, x order.
template <typename F>
__device__ __forceinline__ void fr_box_taps(const fr_box &b, const int lo[3], const int hi[3], float hw,
                                            int L, F &&f)
{
    for (int z = lo[2]; z <= hi[2]; ++z) {
        float dz = __fsub_rn((float)z, b.c[2]);
        float dz2 = __fmul_rn(dz, dz);
        for (int y = lo[1]; y <= hi[1]; ++y) {
            float dy = __fsub_rn((float)y, b.c[1]);
            float dy2 = __fmul_rn(dy, dy);
            for (int x = lo[0]; x <= hi[0]; ++x) {
                float dx = __fsub_rn((float)x, b.c[0]);
                float dk = sqrtf(fr_tap_dk2(dz2, dy2, dx));
                if (dk < hw) f(x, y, z, fr_tap_index(dk, hw, L));
            }
        }
    }
}
