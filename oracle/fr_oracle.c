/*
 * fr_oracle.c -- CPU oracle (test infrastructure only; see fr_oracle.h).
 *
 * Build: oracle/build_oracle.sh (gcc -O2 -fopenmp -ffp-contract=off, x86-64
 * baseline, no -march) -- the reference is built with -O3 and no -march
 * (CMakeLists.txt:5,15), so float arithmetic is plain SSE2 with no FMA.
 * Every float operation below keeps the reference's evaluation order and
 * float/double promotions; comments give the reference line it mirrors.
 */
#include "fr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ I0 */
/* Power series I0(x) = sum_k ((x/2)^k / k!)^2, summed to convergence in
 * double.  For the arguments used here (|x| <= ~12) it converges in < 40
 * terms with relative error ~1e-16. */
double or_i0(double x)
{
    double h = 0.5 * x, h2 = h * h;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        term *= h2 / ((double)k * (double)k);
        sum += term;
        if (term < sum * 1e-18) break;
    }
    return sum;
}

/* ------------------------------------------------------------ KB LUT */
/* ConvKernel::ConvKernel (ReconData/ConvKernel.cpp:7-30). */
void or_kb_lut(float kWidth, float og, int L, float *lut, float *beta_out)
{
    float w = kWidth;
    float a = og;
    /* :14  float beta = M_PI * sqrt(w*w/(a*a) * (a-0.5) * (a-0.5) - 0.8);
     * w*w and a*a are float products, the quotient is float; (a-0.5) is
     * double, so the rest is evaluated in double and rounded into beta. */
    float q = (w * w) / (a * a);
    double am = (double)a - 0.5;
    float beta = (float)(M_PI * sqrt((double)q * am * am - 0.8));
    /* :16  float dk = w / 2.0 / (m_length - 1);  (double, then float) */
    float dk = (float)((double)w / 2.0 / (double)(L - 1));
    float kernel0 = 1.0f;
    for (int i = 0; i < L; ++i) {
        /* :20  float k = dk * i; */
        float k = dk * (float)i;
        /* :21  double x = beta * sqrt(1 - std::pow(2 * k / w, 2));
         * 2*k/w is float; pow promotes to double (exact square). */
        float t = (2.0f * k) / w;
        double td = (double)t;
        double x = (double)beta * sqrt(1.0 - td * td);
        /* :23  data[i] = I0(x) / w;   (double quotient stored as float) */
        float v = (float)(or_i0(x) / (double)w);
        /* :25-26 kernel0 = data[0]; data[i] /= kernel0; (float) */
        if (i == 0) kernel0 = v;
        lut[i] = v / kernel0;
    }
    if (beta_out) *beta_out = beta;
}

/* ------------------------------------------------------ trajectories */
/* storeTrajComponent (ReconData.cpp:56-81): bounds via minmax_element. */
void or_traj_bounds(const float *traj, int64_t M, int stride, int dim, float *bounds)
{
    for (int d = 0; d < dim; ++d) {
        float lo = traj[d], hi = traj[d];
        for (int64_t s = 1; s < M; ++s) {
            float v = traj[s * stride + d];
            if (v < lo) lo = v;          /* minmax_element: first minimum   */
            if (!(v < hi)) hi = v;       /* and last maximum (same value)   */
        }
        bounds[2 * d] = lo;
        bounds[2 * d + 1] = hi;
    }
}

/* normalizeTraj + transformLocalTraj (ReconData.cpp:201-225). */
void or_normalize_traj(float *traj, int64_t M, int stride, int dim, float *bounds, unsigned size)
{
    /* :215-216  tr = -bounds0.first; scale = (size-1) / (b0.second - b0.first)
     * (size - 1) is unsigned, converted to float for the float division. */
    float tr = -bounds[0];
    float scale = (float)(size - 1u) / (bounds[1] - bounds[0]);
    for (int d = 0; d < dim; ++d) {                 /* :218-222 */
        bounds[2 * d] = (bounds[2 * d] + tr) * scale;
        bounds[2 * d + 1] = (bounds[2 * d + 1] + tr) * scale;
    }
    for (int64_t s = 0; s < M; ++s)                 /* :201-209 */
        for (int d = 0; d < dim; ++d)
            traj[s * stride + d] = (traj[s * stride + d] + tr) * scale;
}

/* ----------------------------------------------------------- gridding */
/* Per-sample box and per-tap decision exactly as GridLut.cpp:99-136.
 * Calls `emit(cell, ki)` for every accepted tap in z, y, x order. */
typedef struct {
    int start[3], end[3];
    float center[3];
} box_t;

static void sample_box(int dim, unsigned N, const float *p, float kHW, box_t *b)
{
    for (int i = 0; i < 3; ++i) { b->center[i] = 0.0f; b->start[i] = 0; b->end[i] = 0; }
    for (int i = 0; i < dim; ++i) {                 /* :101-106 */
        b->center[i] = p[i];
        int st = (int)ceil(b->center[i] - kHW);
        int en = (int)floor(b->center[i] + kHW);
        st = (int)fmax((double)st, 0.0);
        en = (int)fmin((double)en, (double)(N - 1u));
        b->start[i] = st;
        b->end[i] = en;
    }
}

static inline int tap_index(float dk, float kHW, int L)
{
    /* :136  (int)std::round(dk / kHW * (klength - 1))  -- float ops, roundf */
    float q = dk / kHW;
    float r = q * (float)(L - 1);
    return (int)roundf(r);
}

void or_grid_f32(int dim, unsigned N, int64_t M, int C,
                 const float *traj, int stride, const float *dcf,
                 const float *data, float kWidth, const float *lut, int L,
                 float *out, int nthreads)
{
    const float kHW = kWidth / 2.0f;                     /* :86 */
    const size_t cells = (dim == 3) ? (size_t)N * N * N : (size_t)N * N;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int c = 0; c < C; ++c) {
        float *o = out + 2 * cells * (size_t)c;
        const float *kd = data + 2 * (size_t)M * (size_t)c;
        memset(o, 0, 2 * cells * sizeof(float));
        for (int64_t s = 0; s < M; ++s) {
            box_t b;
            sample_box(dim, N, traj + s * stride, kHW, &b);
            /* :109  data = dcf * sample  -> (dcf*re, dcf*im) */
            float dre = dcf[s] * kd[2 * s];
            float dim_ = dcf[s] * kd[2 * s + 1];
            for (int z = b.start[2]; z <= b.end[2]; ++z) {
                float dz = (float)z - b.center[2];
                float dz2 = dz * dz;
                for (int y = b.start[1]; y <= b.end[1]; ++y) {
                    float dy = (float)y - b.center[1];
                    float dy2 = dy * dy;
                    for (int x = b.start[0]; x <= b.end[0]; ++x) {
                        float dx = (float)x - b.center[0];
                        float dk = sqrtf((dz2 + dy2) + dx * dx);   /* :131 */
                        if (dk < kHW) {
                            int ki = tap_index(dk, kHW, L);
                            float w = lut[ki];
                            size_t cell = (size_t)z * N * N + (size_t)y * N + (size_t)x;
                            /* :137  *itOut += LUT[ki] * data */
                            float tre = w * dre;
                            float tim = w * dim_;
                            o[2 * cell] = o[2 * cell] + tre;
                            o[2 * cell + 1] = o[2 * cell + 1] + tim;
                        }
                    }
                }
            }
        }
    }
}

void or_grid_f64(int dim, unsigned N, int64_t M, int C,
                 const float *traj, int stride, const float *dcf,
                 const float *data, float kWidth, const float *lut, int L,
                 double *out, int nthreads)
{
    const float kHW = kWidth / 2.0f;
    const size_t cells = (dim == 3) ? (size_t)N * N * N : (size_t)N * N;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int c = 0; c < C; ++c) {
        double *o = out + 2 * cells * (size_t)c;
        const float *kd = data + 2 * (size_t)M * (size_t)c;
        memset(o, 0, 2 * cells * sizeof(double));
        for (int64_t s = 0; s < M; ++s) {
            box_t b;
            sample_box(dim, N, traj + s * stride, kHW, &b);
            double dre = (double)dcf[s] * (double)kd[2 * s];
            double dim_ = (double)dcf[s] * (double)kd[2 * s + 1];
            for (int z = b.start[2]; z <= b.end[2]; ++z) {
                float dz = (float)z - b.center[2];
                float dz2 = dz * dz;
                for (int y = b.start[1]; y <= b.end[1]; ++y) {
                    float dy = (float)y - b.center[1];
                    float dy2 = dy * dy;
                    for (int x = b.start[0]; x <= b.end[0]; ++x) {
                        float dx = (float)x - b.center[0];
                        float dk = sqrtf((dz2 + dy2) + dx * dx);
                        if (dk < kHW) {
                            double w = (double)lut[tap_index(dk, kHW, L)];
                            size_t cell = (size_t)z * N * N + (size_t)y * N + (size_t)x;
                            o[2 * cell] += w * dre;
                            o[2 * cell + 1] += w * dim_;
                        }
                    }
                }
            }
        }
    }
}

void or_degrid_f64(int dim, unsigned N, int64_t M, int C,
                   const float *traj, int stride,
                   const double *grid, float kWidth, const float *lut, int L,
                   double *out)
{
    const float kHW = kWidth / 2.0f;
    const size_t cells = (dim == 3) ? (size_t)N * N * N : (size_t)N * N;
    for (int c = 0; c < C; ++c) {
        const double *g = grid + 2 * cells * (size_t)c;
        double *o = out + 2 * (size_t)M * (size_t)c;
        for (int64_t s = 0; s < M; ++s) {
            box_t b;
            sample_box(dim, N, traj + s * stride, kHW, &b);
            double are = 0.0, aim = 0.0;
            for (int z = b.start[2]; z <= b.end[2]; ++z) {
                float dz = (float)z - b.center[2];
                float dz2 = dz * dz;
                for (int y = b.start[1]; y <= b.end[1]; ++y) {
                    float dy = (float)y - b.center[1];
                    float dy2 = dy * dy;
                    for (int x = b.start[0]; x <= b.end[0]; ++x) {
                        float dx = (float)x - b.center[0];
                        float dk = sqrtf((dz2 + dy2) + dx * dx);
                        if (dk < kHW) {
                            double w = (double)lut[tap_index(dk, kHW, L)];
                            size_t cell = (size_t)z * N * N + (size_t)y * N + (size_t)x;
                            are += w * g[2 * cell];
                            aim += w * g[2 * cell + 1];
                        }
                    }
                }
            }
            o[2 * s] = are;
            o[2 * s + 1] = aim;
        }
    }
}

/* The same degridding sum in FP32, accumulated in the tap order above
 * (z, y, x), one rounding per multiply and per add: the arithmetic of the
 * device degridding kernels (net-new: the reference has no degridding). */
void or_degrid_f32(int dim, unsigned N, int64_t M, int C,
                   const float *traj, int stride,
                   const float *grid, float kWidth, const float *lut, int L,
                   float *out)
{
    const float kHW = kWidth / 2.0f;
    const size_t cells = (dim == 3) ? (size_t)N * N * N : (size_t)N * N;
    for (int c = 0; c < C; ++c) {
        const float *g = grid + 2 * cells * (size_t)c;
        float *o = out + 2 * (size_t)M * (size_t)c;
        for (int64_t s = 0; s < M; ++s) {
            box_t b;
            sample_box(dim, N, traj + s * stride, kHW, &b);
            float are = 0.0f, aim = 0.0f;
            for (int z = b.start[2]; z <= b.end[2]; ++z) {
                float dz = (float)z - b.center[2];
                float dz2 = dz * dz;
                for (int y = b.start[1]; y <= b.end[1]; ++y) {
                    float dy = (float)y - b.center[1];
                    float dy2 = dy * dy;
                    for (int x = b.start[0]; x <= b.end[0]; ++x) {
                        float dx = (float)x - b.center[0];
                        float dk = sqrtf((dz2 + dy2) + dx * dx);
                        if (dk < kHW) {
                            float w = lut[tap_index(dk, kHW, L)];
                            size_t cell = (size_t)z * N * N + (size_t)y * N + (size_t)x;
                            are = are + w * g[2 * cell];
                            aim = aim + w * g[2 * cell + 1];
                        }
                    }
                }
            }
            o[2 * s] = are;
            o[2 * s + 1] = aim;
        }
    }
}

/* ---------------------------------------------------------------- FFT */
/* Mixed-radix decimation-in-time FFT in double, unnormalised,
 * X[k] = sum_n x[n] exp(sign * 2 pi i n k / n_len). */
typedef struct { double re, im; } cplx;

static void fft_rec(const cplx *in, size_t is, cplx *out, size_t n, int sign, cplx *tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    size_t p = 0;
    for (size_t f = 2; f * f <= n; ++f) if (n % f == 0) { p = f; break; }
    if (p == 0) p = n;                               /* n prime: naive DFT */
    size_t m = n / p;
    for (size_t r = 0; r < p; ++r)
        fft_rec(in + r * is, is * p, out + r * m, m, sign, tmp);
    const double base = (double)sign * 2.0 * M_PI / (double)n;
    for (size_t k = 0; k < m; ++k) {
        for (size_t r = 0; r < p; ++r) {
            double ang = base * (double)(r * k);
            double c = cos(ang), s = sin(ang);
            cplx y = out[r * m + k];
            tmp[r].re = y.re * c - y.im * s;
            tmp[r].im = y.re * s + y.im * c;
        }
        for (size_t q = 0; q < p; ++q) {
            double sre = 0.0, sim = 0.0;
            for (size_t r = 0; r < p; ++r) {
                size_t e = (r * q) % p;
                double ang = (double)sign * 2.0 * M_PI * (double)e / (double)p;
                double c = cos(ang), s = sin(ang);
                sre += tmp[r].re * c - tmp[r].im * s;
                sim += tmp[r].re * s + tmp[r].im * c;
            }
            out[k + q * m].re = sre;
            out[k + q * m].im = sim;
        }
    }
}

static void fft_lines(cplx *a, size_t n, size_t stride, size_t count_outer, size_t outer_stride,
                      size_t count_inner, size_t inner_stride, int sign)
{
    cplx *line = (cplx *)malloc(n * sizeof(cplx));
    cplx *res = (cplx *)malloc(n * sizeof(cplx));
    cplx *tmp = (cplx *)malloc(n * sizeof(cplx));
    for (size_t o = 0; o < count_outer; ++o)
        for (size_t i = 0; i < count_inner; ++i) {
            cplx *base = a + o * outer_stride + i * inner_stride;
            for (size_t j = 0; j < n; ++j) line[j] = base[j * stride];
            fft_rec(line, 1, res, n, sign, tmp);
            for (size_t j = 0; j < n; ++j) base[j * stride] = res[j];
        }
    free(line); free(res); free(tmp);
}

void or_fft_c2c(int dim, unsigned N, int batch, double *data, int sign)
{
    size_t n = N;
    size_t cells = (dim == 3) ? n * n * n : n * n;
    for (int b = 0; b < batch; ++b) {
        cplx *a = (cplx *)(data + 2 * cells * (size_t)b);
        if (dim == 2) {
            fft_lines(a, n, 1, n, n, 1, 0, sign);          /* along x */
            fft_lines(a, n, n, 1, 0, n, 1, sign);          /* along y */
        } else {
            fft_lines(a, n, 1, n * n, n, 1, 0, sign);      /* along x */
            fft_lines(a, n, n, n, n * n, n, 1, sign);      /* along y */
            fft_lines(a, n, n * n, 1, 0, n * n, 1, sign);  /* along z */
        }
    }
}

/* ----------------------------------------------------------- fftshift */
/* ImageFilter::fftShift2 / fftShift3 (ImageFilter.cpp:231-280): swap
 * element (z,y,x) with (z+N/2, (y+N/2)%N, (x+N/2)%N) for z < N/2. */
#define DEFINE_SHIFT(NAME, T)                                                     \
void NAME(int dim, unsigned N, int batch, T *data)                               \
{                                                                                 \
    size_t n = N, h = N / 2;                                                      \
    size_t cells = (dim == 3) ? n * n * n : n * n;                                \
    for (int b = 0; b < batch; ++b) {                                             \
        T *d = data + 2 * cells * (size_t)b;                                      \
        if (dim == 2) {                                                           \
            for (size_t y = 0; y < h; ++y) {                                      \
                size_t y1 = y + h;                                                \
                for (size_t x = 0; x < n; ++x) {                                  \
                    size_t x1 = x < h ? x + h : x - h;                            \
                    size_t i = y * n + x, j = y1 * n + x1;                        \
                    T t0 = d[2 * i], t1 = d[2 * i + 1];                           \
                    d[2 * i] = d[2 * j]; d[2 * i + 1] = d[2 * j + 1];             \
                    d[2 * j] = t0; d[2 * j + 1] = t1;                             \
                }                                                                 \
            }                                                                     \
        } else {                                                                  \
            for (size_t z = 0; z < h; ++z) {                                      \
                size_t z1 = z + h;                                                \
                for (size_t y = 0; y < n; ++y) {                                  \
                    size_t y1 = y < h ? y + h : y - h;                            \
                    for (size_t x = 0; x < n; ++x) {                              \
                        size_t x1 = x < h ? x + h : x - h;                        \
                        size_t i = z * n * n + y * n + x;                         \
                        size_t j = z1 * n * n + y1 * n + x1;                      \
                        T t0 = d[2 * i], t1 = d[2 * i + 1];                       \
                        d[2 * i] = d[2 * j]; d[2 * i + 1] = d[2 * j + 1];         \
                        d[2 * j] = t0; d[2 * j + 1] = t1;                         \
                    }                                                             \
                }                                                                 \
            }                                                                     \
        }                                                                         \
    }                                                                             \
}
DEFINE_SHIFT(or_fftshift_f32, float)
DEFINE_SHIFT(or_fftshift_f64, double)

/* ---------------------------------------------------------------- SOS */
/* ImageFilter::SOS (ImageFilter.cpp:159-209): crop offsets
 * x0 = (N - rc)/2 (unsigned); out += v * conj(v) over coils in order, whose
 * real part is fl(fl(re*re) + fl(im*im)); then sqrt and |.| (Main.cpp:177). */
void or_sos_f32(int dim, unsigned N, int C, const float *in,
                unsigned rcx, unsigned rcy, unsigned rcz, float *out)
{
    size_t n = N;
    size_t cells = (dim == 3) ? n * n * n : n * n;
    unsigned nz = (dim == 3) ? N : 1u;
    if (dim == 2) rcz = 1u;
    unsigned x0 = (N - rcx) / 2u, y0 = (N - rcy) / 2u, z0 = (nz - rcz) / 2u;
    size_t rc = (size_t)rcx * rcy * rcz;
    for (size_t i = 0; i < rc; ++i) out[i] = 0.0f;
    for (int c = 0; c < C; ++c) {
        const float *d = in + 2 * cells * (size_t)c;
        for (unsigned z = 0; z < rcz; ++z)
            for (unsigned y = 0; y < rcy; ++y)
                for (unsigned x = 0; x < rcx; ++x) {
                    size_t i = (size_t)(z + z0) * n * n + (size_t)(y + y0) * n + (x + x0);
                    float re = d[2 * i], im = d[2 * i + 1];
                    float p = re * re + im * im;
                    size_t o = ((size_t)z * rcy + y) * rcx + x;
                    out[o] = out[o] + p;
                }
    }
    for (size_t i = 0; i < rc; ++i) out[i] = sqrtf(out[i]);
}

void or_sos_f64(int dim, unsigned N, int C, const double *in,
                unsigned rcx, unsigned rcy, unsigned rcz, double *out)
{
    size_t n = N;
    size_t cells = (dim == 3) ? n * n * n : n * n;
    unsigned nz = (dim == 3) ? N : 1u;
    if (dim == 2) rcz = 1u;
    unsigned x0 = (N - rcx) / 2u, y0 = (N - rcy) / 2u, z0 = (nz - rcz) / 2u;
    size_t rc = (size_t)rcx * rcy * rcz;
    for (size_t i = 0; i < rc; ++i) out[i] = 0.0;
    for (int c = 0; c < C; ++c) {
        const double *d = in + 2 * cells * (size_t)c;
        for (unsigned z = 0; z < rcz; ++z)
            for (unsigned y = 0; y < rcy; ++y)
                for (unsigned x = 0; x < rcx; ++x) {
                    size_t i = (size_t)(z + z0) * n * n + (size_t)(y + y0) * n + (x + x0);
                    size_t o = ((size_t)z * rcy + y) * rcx + x;
                    out[o] += d[2 * i] * d[2 * i] + d[2 * i + 1] * d[2 * i + 1];
                }
    }
    for (size_t i = 0; i < rc; ++i) out[i] = sqrt(out[i]);
}

/* -------------------------------------------------------- Pipe-Menon */
/* w <- w / (G^H G w): grid the real weights with the LUT kernel (same taps
 * as gridding), interpolate back with the same taps, divide.  FP64. */
void or_pipe_menon_f64(int dim, unsigned N, int64_t M, const float *traj, int stride,
                       float kWidth, const float *lut, int L, int iters, double *w)
{
    const size_t cells = (dim == 3) ? (size_t)N * N * N : (size_t)N * N;
    const float kHW = kWidth / 2.0f;
    double *g = (double *)malloc(2 * cells * sizeof(double));
    double *d = (double *)malloc(2 * (size_t)M * sizeof(double));
    for (int it = 0; it < iters; ++it) {
        memset(g, 0, 2 * cells * sizeof(double));
        for (int64_t s = 0; s < M; ++s) {
            box_t b;
            sample_box(dim, N, traj + s * stride, kHW, &b);
            for (int z = b.start[2]; z <= b.end[2]; ++z) {
                float dz = (float)z - b.center[2];
                float dz2 = dz * dz;
                for (int y = b.start[1]; y <= b.end[1]; ++y) {
                    float dy = (float)y - b.center[1];
                    float dy2 = dy * dy;
                    for (int x = b.start[0]; x <= b.end[0]; ++x) {
                        float dx = (float)x - b.center[0];
                        float dk = sqrtf((dz2 + dy2) + dx * dx);
                        if (dk < kHW) {
                            size_t cell = (size_t)z * N * N + (size_t)y * N + (size_t)x;
                            g[2 * cell] += (double)lut[tap_index(dk, kHW, L)] * w[s];
                        }
                    }
                }
            }
        }
        or_degrid_f64(dim, N, M, 1, traj, stride, g, kWidth, lut, L, d);
        for (int64_t s = 0; s < M; ++s) {
            double den = d[2 * s];
            w[s] = den > 0.0 ? w[s] / den : 0.0;
        }
    }
    free(g); free(d);
}

/* ImageFilter::lowFilter (ImageFilter.cpp:56-93). */
void or_lowfilter_f32(int dim, unsigned N, int C, float *data, int res)
{
    unsigned nx = N, ny = N, nz = (dim == 3) ? N : 1u;
    unsigned x0 = nx / 2u, y0 = ny / 2u, z0 = nz / 2u;
    float att = (float)(2.0 * res * res / 4.0);
    float coeff[2000];
    for (int r = 0; r < 2000; ++r) coeff[r] = expf((float)(-r / 100.0));
    size_t cells = (size_t)nx * ny * nz;
    for (int c = 0; c < C; ++c) {
        float *d = data + 2 * cells * (size_t)c;
        size_t i = 0;
        for (unsigned long z = 0; z < nz; ++z) {
            int r1 = (int)((z - z0) * (z - z0));
            for (unsigned long y = 0; y < ny; ++y) {
                int r2 = (int)((y - y0) * (y - y0)) + r1;
                for (unsigned long x = 0; x < nx; ++x, ++i) {
                    int r = (int)((x - x0) * (x - x0) + (unsigned long)r2);
                    int idx = (int)((double)((float)r / att) * 100.0);
                    if (idx >= 2000) {
                        d[2 * i] = 0.0f;
                        d[2 * i + 1] = 0.0f;
                    } else {
                        d[2 * i] *= coeff[idx];
                        d[2 * i + 1] *= coeff[idx];
                    }
                }
            }
        }
    }
}

/* ImageFilter::normalize (ImageFilter.cpp:95-118). */
void or_normalize_f32(int dim, unsigned N, int C, float *data)
{
    size_t n = N;
    size_t cells = (dim == 3) ? n * n * n : n * n;
    float *mag = (float *)calloc(cells, sizeof(float));
    for (int c = 0; c < C; ++c) {
        const float *d = data + 2 * cells * (size_t)c;
        for (size_t i = 0; i < cells; ++i) {
            float a = hypotf(d[2 * i], d[2 * i + 1]);
            mag[i] = mag[i] + a * a;
        }
    }
    for (int c = 0; c < C; ++c) {
        float *d = data + 2 * cells * (size_t)c;
        for (size_t i = 0; i < cells; ++i) {
            float s = sqrtf(mag[i]);
            d[2 * i] = d[2 * i] / s;
            d[2 * i + 1] = d[2 * i + 1] / s;
        }
    }
    free(mag);
}

/* ImageFilter::SOS(map, size) (ImageFilter.cpp:170-209). */
void or_sos_map_f32(int dim, unsigned N, int C, const float *in, const float *map,
                    unsigned rcx, unsigned rcy, unsigned rcz, float *out)
{
    size_t n = N;
    size_t cells = (dim == 3) ? n * n * n : n * n;
    unsigned nz = (dim == 3) ? N : 1u;
    if (dim == 2) rcz = 1u;
    unsigned x0 = (N - rcx) / 2u, y0 = (N - rcy) / 2u, z0 = (nz - rcz) / 2u;
    size_t rc = (size_t)rcx * rcy * rcz;
    for (size_t i = 0; i < 2 * rc; ++i) out[i] = 0.0f;
    for (int c = 0; c < C; ++c) {
        const float *d = in + 2 * cells * (size_t)c;
        const float *m = map + 2 * cells * (size_t)c;
        for (unsigned z = 0; z < rcz; ++z)
            for (unsigned y = 0; y < rcy; ++y)
                for (unsigned x = 0; x < rcx; ++x) {
                    size_t i = (size_t)(z + z0) * n * n + (size_t)(y + y0) * n + (x + x0);
                    float a = d[2 * i], b = d[2 * i + 1];
                    float mc = m[2 * i], md = -m[2 * i + 1];
                    float re = a * mc - b * md;
                    float im = a * md + b * mc;
                    size_t o = ((size_t)z * rcy + y) * rcx + x;
                    out[2 * o] = out[2 * o] + re;
                    out[2 * o + 1] = out[2 * o + 1] + im;
                }
    }
}
