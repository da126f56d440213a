// Host-side pieces of the reference's gridding setup that the backend needs:
// the Kaiser-Bessel LUT and the trajectory normalisation.  Both are
// evaluated on the host in the reference's exact float/double arithmetic
// (the library is compiled with -ffp-contract=off).
#include "fr_internal.h"

#include <cmath>

namespace {

// I0(x) by its power series sum_k ((x/2)^k / k!)^2 in double; converges in
// < 40 terms for the arguments used by the LUT (|x| <= ~12).
double bessel_i0(double x)
{
    const double h2 = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        term *= h2 / ((double)k * (double)k);
        sum += term;
        if (term < sum * 1e-18) break;
    }
    return sum;
}

}  // namespace

// ConvKernel::ConvKernel (ReconData/ConvKernel.cpp:7-30).
extern "C" fr_status fr_kb_lut(float kWidth, float og, int L, float *lut, float *beta_out)
{
    if (!lut || L < 2 || !(kWidth > 0.f) || !(og > 0.f)) return FR_ERR_INVALID;
    const float w = kWidth, a = og;
    // :14  w*w/(a*a) in float, the (a - 0.5) factors in double
    const float q = (w * w) / (a * a);
    const double am = (double)a - 0.5;
    const float beta = (float)(M_PI * std::sqrt((double)q * am * am - 0.8));
    const float dk = (float)((double)w / 2.0 / (double)(L - 1));       // :16
    float k0 = 1.f;
    for (int i = 0; i < L; ++i) {
        const float k = dk * (float)i;                                   // :20
        const float t = (2.0f * k) / w;                                  // :21
        const double x = (double)beta * std::sqrt(1.0 - (double)t * (double)t);
        const float v = (float)(bessel_i0(x) / (double)w);               // :23
        if (i == 0) k0 = v;                                              // :25
        lut[i] = v / k0;                                                 // :26
    }
    if (beta_out) *beta_out = beta;
    return FR_OK;
}

// storeTrajComponent bounds (ReconData.cpp:56-81) + normalizeTraj /
// transformLocalTraj (ReconData.cpp:201-225): every component is mapped
// with component 0's bounds, x <- (x - min0) * float(N - 1) / (max0 - min0).
extern "C" fr_status fr_traj_normalize(float *traj, int64_t M, int stride, int dim, unsigned N, float *bounds)
{
    if (!traj || M <= 0 || dim < 1 || stride < dim || N < 2) return FR_ERR_INVALID;
    float b[8];
    for (int d = 0; d < dim && d < 4; ++d) {
        float lo = traj[d], hi = traj[d];
        for (int64_t s = 1; s < M; ++s) {
            const float v = traj[s * stride + d];
            if (v < lo) lo = v;
            if (!(v < hi)) hi = v;
        }
        b[2 * d] = lo;
        b[2 * d + 1] = hi;
    }
    const float tr = -b[0];
    const float scale = (float)(N - 1u) / (b[1] - b[0]);
    for (int d = 0; d < dim && d < 4; ++d) {
        b[2 * d] = (b[2 * d] + tr) * scale;
        b[2 * d + 1] = (b[2 * d + 1] + tr) * scale;
    }
    for (int64_t s = 0; s < M; ++s)
        for (int d = 0; d < dim; ++d) traj[s * stride + d] = (traj[s * stride + d] + tr) * scale;
    if (bounds)
        for (int d = 0; d < 2 * dim && d < 8; ++d) bounds[d] = b[d];
    return FR_OK;
}

// Normalisation with given (translation, scale), as transformLocalTraj.
extern "C" fr_status fr_traj_transform(float *traj, int64_t M, int stride, int dim, float tr, float scale)
{
    if (!traj || M < 0 || dim < 1 || stride < dim) return FR_ERR_INVALID;
    for (int64_t s = 0; s < M; ++s)
        for (int d = 0; d < dim; ++d) traj[s * stride + d] = (traj[s * stride + d] + tr) * scale;
    return FR_OK;
}
