/*
 * sct_oracle.c -- CPU restatement of the segmentctscans hot path.
 * TEST INFRASTRUCTURE ONLY (the checker); see sct_oracle.h.
 *
 * Built by oracle/Makefile into oracle/_build/libsct_oracle.so with
 * -ffp-contract=off, so every fused multiply-add is the explicit fma() call
 * written in the source and matches the HIP kernels' explicit __fmaf_rn.
 */
#include "sct_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ---- layout helpers, include/layout_cylinder.h:81-94 ------------------- */
static inline uint32_t row_begin(const or_layout *m, uint32_t col) { return m->start_row[col]; }
static inline uint32_t row_end(const or_layout *m, uint32_t col)
{
    return m->start_row[col] + (m->base[col + 1] - m->base[col]) / m->nstride;
}
/* Line index of (row, col): base[col] is a multiple of nstride, so
 * offset(n, row, col) = line_of(row, col) * nstride + n. */
static inline size_t line_of(const or_layout *m, uint32_t row, uint32_t col)
{
    return (size_t)(m->base[col] / m->nstride) + (row - m->start_row[col]);
}
/* Neighbour predicates, potts-min-partition.cpp:128,131,390,393. */
static inline int has_next_row(const or_layout *m, uint32_t row, uint32_t col)
{
    return row_end(m, col) - 1 > row;
}
static inline int has_prev_row(const or_layout *m, uint32_t row, uint32_t col)
{
    return row_begin(m, col) < row;
}
static inline int has_next_col(const or_layout *m, uint32_t row, uint32_t col)
{
    return m->col_end - 1 > col && row_begin(m, col + 1) <= row && row < row_end(m, col + 1);
}
static inline int has_prev_col(const or_layout *m, uint32_t row, uint32_t col)
{
    return m->col_begin < col && row_begin(m, col - 1) <= row && row < row_end(m, col - 1);
}
/* field.h:37-40, 128-byte alignment for 4-byte elements. */
static inline uint32_t stride32(uint32_t n) { return ((n + 31u) / 32u) * 32u; }

/* ---- cylinder constructor, layout_cylinder.h:152-195 ------------------- */
int or_cylinder(uint32_t n_images, uint32_t height, uint32_t width,
                double xc, double yc, double r, uint32_t nstride,
                uint32_t *start_row, uint32_t *base,
                uint32_t *col_begin, uint32_t *col_end, uint64_t *data_size)
{
    const double r2 = r * r;
    uint32_t off = 0, cb = ~0u, ce = 0;
    for (uint32_t col = 0; col < width; ++col) {
        double d2 = r2 - ((double)col - xc) * ((double)col - xc);
        base[col] = off;
        double r0 = 0.0, r1 = 0.0;
        if (0 < d2) {
            double d = sqrt(d2);
            double c = ceil(yc - d);
            if (yc - d == c) c += 1.0;
            double f = floor(yc + d);
            if (yc + d == f) f -= 1.0;
            r0 = c > 0.0 ? c : 0.0;
            r1 = f < (double)(height - 1) ? f : (double)(height - 1);
        }
        if (0 < d2 && r0 <= r1) {
            uint32_t rr0 = (uint32_t)r0, rr1 = (uint32_t)r1;
            if (cb == ~0u) { cb = col; ce = col + 1; } else { ++ce; }
            start_row[col] = rr0;
            off += (rr1 - rr0 + 1) * nstride;
        } else {
            start_row[col] = ~0u;
        }
    }
    base[width] = off;
    uint64_t ds = 0;
    for (uint32_t col = cb; col < ce; ++col)
        ds += start_row[col] + (base[col + 1] - base[col]) / nstride - start_row[col];
    *col_begin = cb;
    *col_end = ce;
    *data_size = ds * n_images;
    return 0;
}

/* ---- gradient / divergence (predicate check) --------------------------- */
void or_grad_d(const or_layout *m, const double *v, double *p0, double *p1, double *p2)
{
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                size_t o = line_of(m, row, col) * m->nstride + n;
                p0[o] = (n + 1 < m->n_images) ? v[o + 1] - v[o] : 0.0;
                p1[o] = has_next_row(m, row, col) ? v[o + m->nstride] - v[o] : 0.0;
                p2[o] = has_next_col(m, row, col)
                    ? v[line_of(m, row, col + 1) * m->nstride + n] - v[o] : 0.0;
            }
}

void or_div_d(const or_layout *m, const double *p0, const double *p1, const double *p2, double *d)
{
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                size_t o = line_of(m, row, col) * m->nstride + n;
                double s = 0;
                if (n > 0) s -= p0[o - 1];
                if (n + 1 < m->n_images) s += p0[o];
                if (has_prev_row(m, row, col)) s -= p1[o - m->nstride];
                if (has_next_row(m, row, col)) s += p1[o];
                if (has_prev_col(m, row, col)) s -= p2[line_of(m, row, col - 1) * m->nstride + n];
                if (has_next_col(m, row, col)) s += p2[o];
                d[o] = s;
            }
}

/* ---- reduced-precision storage --------------------------------------------
 * Round-to-nearest-even float -> bfloat16 / IEEE half -> float, written on
 * the bit patterns (the HIP kernels use the same rounding).  Only the FP32
 * mirror rounds; the FP64 oracle ignores the storage mode. */
static inline float bf16_round(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) {
        x |= 0x00400000u;                      /* quiet NaN */
    } else {
        x += 0x7fffu + ((x >> 16) & 1u);
    }
    x &= 0xffff0000u;
    memcpy(&f, &x, 4);
    return f;
}

static inline uint16_t f32_to_f16_bits(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);          /* rounds to infinity */
    if (ax < 0x38800000u) {                                             /* half subnormal or zero */
        const int e = (int)(ax >> 23);
        const int shift = 126 - e;
        if (e == 0 || shift > 24) return (uint16_t)sign;
        const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
        uint32_t q = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (uint16_t)(sign | q);
    }
    uint32_t r = ax - 0x38000000u;                                       /* rebias 127 -> 15 */
    r += 0xfffu + ((r >> 13) & 1u);
    return (uint16_t)(sign | (r >> 13));
}

static inline float f16_bits_to_f32(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, mant = h & 0x3ffu, x;
    if (e == 0x1f) {
        x = sign | 0x7f800000u | (mant << 13);
    } else if (e == 0) {
        if (mant == 0) {
            x = sign;
        } else {                                                        /* normalise the subnormal */
            int k = 0;
            while (!(mant & 0x400u)) { mant <<= 1; ++k; }
            x = sign | ((uint32_t)(113 - k) << 23) | ((mant & 0x3ffu) << 13);
        }
    } else {
        x = sign | ((e + 112u) << 23) | (mant << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

static inline float f16_round(float f) { return f16_bits_to_f32(f32_to_f16_bits(f)); }

static inline float store_round_f(float f, int storage)
{
    return storage == 1 ? bf16_round(f) : (storage == 2 ? f16_round(f) : f);
}

float or_bf16_round(float f) { return bf16_round(f); }
float or_f16_round(float f) { return f16_round(f); }

/* ---- precision-generic solvers ----------------------------------------- */
#define REAL double
#define SFX _d
#define STORE_ROUND(x, storage) ((void)(storage), (x))
#define REAL_MAX DBL_MAX
#define FMA fma
#define SQRT sqrt
#define FABS fabs
#define FMAX fmax
#define FMIN fmin
#include "oracle_body.h"
#undef REAL
#undef SFX
#undef REAL_MAX
#undef FMA
#undef SQRT
#undef FABS
#undef FMAX
#undef FMIN
#undef STORE_ROUND

#define REAL float
#define STORE_ROUND(x, storage) store_round_f((x), (storage))
#define SFX _f
#define REAL_MAX FLT_MAX
#define FMA fmaf
#define SQRT sqrtf
#define FABS fabsf
#define FMAX fmaxf
#define FMIN fminf
#include "oracle_body.h"

/* ---- reductions ---------------------------------------------------------- */

/* calc_hist.cpp:59-63 with delta = 0:
 *   accept rs <= x && x <= re; t = (x - rs) / (re - rs);
 *   bin = clamp(int(round(nbins * t - 0.5f)), 0, nbins - 1)
 * nbins * t - 0.5f is written as one fma (the contraction a CUDA compiler
 * makes by default; exact anyway for power-of-two nbins). */
static inline int32_t hist_bin(float x, float rs, float re, uint32_t nbins)
{
    float t = (x - rs) / (re - rs);
    int32_t b = (int32_t)roundf(fmaf((float)nbins, t, -0.5f));
    int32_t hi = (int32_t)nbins - 1;
    return b < 0 ? 0 : (b > hi ? hi : b);
}

void or_calc_hist(const or_layout *m, const float *u, float rs, float re,
                  uint32_t nbins, uint32_t *counts)
{
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                float x = u[line_of(m, row, col) * m->nstride + n];
                if (rs <= x && x <= re) counts[hist_bin(x, rs, re, nbins)] += 1;
            }
}

void or_min_max(const or_layout *m, const float *u, float *mn, float *mx)
{
    float a = FLT_MAX, b = -FLT_MAX;
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                float x = u[line_of(m, row, col) * m->nstride + n];
                a = (x < a) ? x : a;
                b = (b < x) ? x : b;
            }
    *mn = a;
    *mx = b;
}

void or_grad_min_max(const or_layout *m, const float *v, float *mn, float *mx)
{
    float a = FLT_MAX, b = -FLT_MAX;
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                float x = v[line_of(m, row, col) * m->nstride + n];
                if (0 < x) {
                    a = (x < a) ? x : a;
                    b = (b < x) ? x : b;
                }
            }
    *mn = a;
    *mx = b;
}

void or_grad_magn(const or_layout *m, const float *u, float *v)
{
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                size_t o = line_of(m, row, col) * m->nstride + n;
                if (!(n + 1 < m->n_images) || !has_next_row(m, row, col) || !has_next_col(m, row, col) ||
                    !(n > 0) || !has_prev_row(m, row, col) || !has_prev_col(m, row, col)) {
                    v[o] = -1.0f;
                } else {
                    float q1 = u[o + 1] - u[o - 1];
                    float q2 = u[o + m->nstride] - u[o - m->nstride];
                    float q3 = u[line_of(m, row, col + 1) * m->nstride + n]
                             - u[line_of(m, row, col - 1) * m->nstride + n];
                    v[o] = 0.5f * sqrtf(fmaf(q3, q3, fmaf(q2, q2, q1 * q1)));
                }
            }
}

void or_uv_hist(const or_layout *m, const float *u, const float *v,
                float u0, float u1, float v0, float v1, uint32_t nbins,
                uint32_t stride, uint32_t *counts)
{
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                size_t o = line_of(m, row, col) * m->nstride + n;
                float x = u[o], y = v[o];
                if (!(u0 <= x && x <= u1 && v0 <= y && y <= v1)) continue;
                int32_t xb = hist_bin(x, u0, u1, nbins);
                int32_t yb = hist_bin(y, v0, v1, nbins);
                counts[(size_t)yb * stride + (size_t)xb] += 1;
            }
}
