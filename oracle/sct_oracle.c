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
    /* NaN: quiet, with the top 9 payload bits kept (what the hardware
     * conversion of the solver's half storage does) */
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x1ffu));
    if (ax == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
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

/* ---- FP64 duality gap of the Potts saddle problem ---------------------------
 * Not in the reference; the definition the library's gap follows:
 *   L(V, Xi) = <V, G> + <grad V, Xi>,  G_i = lambda (u - c_i)^2,
 *   V in the simplex, Xi in K = {|xi_i - xi_j| <= 1, i < j} per voxel.
 *   primal E(V)  = <V, G> + sum_x h_K(grad V(x)),
 *   dual   D(Xi) = sum_x min_i (G_i - (div Xi)_i),
 * with h_K the support function of K.  By conic duality h_K(p) is the
 * cheapest flow on the complete graph with node supplies p_i:
 *   h_K(p) = min sum_{i<j} |f_ij|  s.t.  p_i = sum_{j>i} f_ij - sum_{j<i} f_ji.
 * K = 2: |p_0|.  K = 3: the Fermat-Weber cost of the points {0, p_0, -p_1}.
 * K >= 4: iteratively reweighted least squares (conductances = current flow
 * magnitudes, node K-1 grounded), 256 steps from unit conductances, taking the
 * cheapest iterate; every iterate is a feasible flow, so the value is an
 * upper bound of h_K and E an upper bound of the primal energy. */
#define OR_IRLS_STEPS 256

static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static double fermat3(const double *p0, const double *p1)
{
    /* points A = 0, B = p0, C = -p1 */
    double s[3], cr[3];
    for (int a = 0; a < 3; ++a) s[a] = p0[a] + p1[a];
    const double ab = norm3(p0), ac = norm3(p1), bc = norm3(s);
    /* angle >= 120 degrees at a vertex: the median is that vertex */
    if (-dot3(p0, p1) <= -0.5 * ab * ac) return ab + ac;          /* at A: <B, C> */
    if (dot3(p0, s) <= -0.5 * ab * bc) return ab + bc;            /* at B: <A-B, C-B> */
    if (dot3(p1, s) <= -0.5 * ac * bc) return ac + bc;            /* at C: <A-C, B-C> */
    cr[0] = p0[1] * p1[2] - p0[2] * p1[1];
    cr[1] = p0[2] * p1[0] - p0[0] * p1[2];
    cr[2] = p0[0] * p1[1] - p0[1] * p1[0];
    return sqrt(0.5 * (ab * ab + ac * ac + bc * bc) + sqrt(3.0) * norm3(cr));
}

static double irls_flow(int K, const double (*p)[3])
{
    double c[15], L[25], ch[25], phi[6][3], best = INFINITY, scale = 0.0;
    int np = K * (K - 1) / 2, M = K - 1;
    for (int i = 0; i < K; ++i) scale = fmax(scale, norm3(p[i]));
    if (scale == 0.0) return 0.0;
    for (int e = 0; e < np; ++e) c[e] = 1.0;
    for (int step = 0; step < OR_IRLS_STEPS; ++step) {
        memset(L, 0, sizeof(L));
        for (int i = 0, e = 0; i < K; ++i)
            for (int j = i + 1; j < K; ++j, ++e) {
                if (i < M) L[i * M + i] += c[e];
                if (j < M) L[j * M + j] += c[e];
                if (i < M && j < M) { L[i * M + j] -= c[e]; L[j * M + i] -= c[e]; }
            }
        /* Cholesky of the grounded Laplacian */
        for (int r = 0; r < M; ++r)
            for (int q = 0; q <= r; ++q) {
                double acc = L[r * M + q];
                for (int t = 0; t < q; ++t) acc -= ch[r * M + t] * ch[q * M + t];
                ch[r * M + q] = (r == q) ? sqrt(acc) : acc / ch[q * M + q];
            }
        for (int a = 0; a < 3; ++a) {
            double y[6];
            for (int r = 0; r < M; ++r) {
                double acc = p[r][a];
                for (int t = 0; t < r; ++t) acc -= ch[r * M + t] * y[t];
                y[r] = acc / ch[r * M + r];
            }
            for (int r = M - 1; r >= 0; --r) {
                double acc = y[r];
                for (int t = r + 1; t < M; ++t) acc -= ch[t * M + r] * phi[t][a];
                phi[r][a] = acc / ch[r * M + r];
            }
            phi[M][a] = 0.0;
        }
        double cost = 0.0;
        for (int i = 0, e = 0; i < K; ++i)
            for (int j = i + 1; j < K; ++j, ++e) {
                double f[3];
                for (int a = 0; a < 3; ++a) f[a] = c[e] * (phi[i][a] - phi[j][a]);
                const double fn = norm3(f);
                cost += fn;
                c[e] = fmax(fn, 1e-12 * scale);
            }
        if (cost < best) best = cost;
    }
    return best;
}

double or_support_d(int K, const double *p)
{
    const double (*pp)[3] = (const double (*)[3])p;
    if (K == 2) return norm3(pp[0]);
    if (K == 3) return fermat3(pp[0], pp[1]);
    return irls_flow(K, pp);
}

double or_support_irls_d(int K, const double *p) { return irls_flow(K, (const double (*)[3])p); }

int or_potts_gap_d(const or_layout *m, const float *u, int K, const float *c, float lambda,
                   const double *V, const double *xi, double *primal, double *dual)
{
    if (K < 2 || K > 6) return 1;
    const size_t span = m->base[m->width];
    double E = 0.0, D = 0.0;
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                const size_t o = line_of(m, row, col) * m->nstride + n;
                const int nx[3] = {n + 1 < m->n_images, has_next_row(m, row, col), has_next_col(m, row, col)};
                const int px[3] = {n > 0, has_prev_row(m, row, col), has_prev_col(m, row, col)};
                const size_t on[3] = {o + 1, o + m->nstride, nx[2] ? line_of(m, row, col + 1) * m->nstride + n : 0};
                const size_t op[3] = {o - 1, o - m->nstride, px[2] ? line_of(m, row, col - 1) * m->nstride + n : 0};
                double v[6], p[6][3], dv[6], g[6];
                double vk = 1.0;
                for (int i = 0; i < K - 1; ++i) { v[i] = V[i * span + o]; vk -= v[i]; }
                v[K - 1] = vk;
                for (int a = 0; a < 3; ++a) {
                    double sk = 0.0;
                    for (int i = 0; i < K - 1; ++i) {
                        p[i][a] = nx[a] ? V[i * span + on[a]] - v[i] : 0.0;
                        sk -= p[i][a];
                    }
                    p[K - 1][a] = sk;
                }
                for (int i = 0; i < K; ++i) dv[i] = 0.0;
                for (int a = 0; a < 3; ++a) {
                    double own_k = 0.0, prev_k = 0.0;
                    for (int i = 0; i < K - 1; ++i) {
                        const double own = xi[(3 * i + a) * span + o];
                        const double prev = px[a] ? xi[(3 * i + a) * span + op[a]] : 0.0;
                        if (nx[a]) dv[i] += own;
                        if (px[a]) dv[i] -= prev;
                        own_k -= own;
                        prev_k -= prev;
                    }
                    if (nx[a]) dv[K - 1] += own_k;
                    if (px[a]) dv[K - 1] -= prev_k;
                }
                double lin = 0.0, mn = INFINITY;
                for (int i = 0; i < K; ++i) {
                    const double d = (double)u[o] - (double)c[i];
                    g[i] = (double)lambda * d * d;
                    lin += v[i] * g[i];
                    mn = fmin(mn, g[i] - dv[i]);
                }
                E += lin + or_support_d(K, &p[0][0]);
                D += mn;
            }
    *primal = E;
    *dual = D;
    return 0;
}

/* ---- FP64 adjoint of the unrolled Potts iteration ---------------------------
 * Not in the reference.  Loss after `iters` iterations from the start state
 * (V = V-bar = one-hot of init, Xi = 0):
 *   L = 1/(N K) sum_x sum_{i<K} (V_i(x) - Y_i(x))^2,  Y one-hot of target,
 * N in-mask voxels, V_{K-1} = 1 - sum of the others.  Gradients with respect
 * to c, lambda and u flow through the data term G_i = lambda (u - c_i)^2; the
 * nearest-label start is piecewise constant (gradient 0 almost everywhere).
 * Backward through one iteration (s_t -> s_{t+1}):
 *   V+ = Pi_S(q), q = V + tau1 (div Xi+ - G): Jacobian of Pi_S = diag(a) - a a^T/|a|
 *        on the active set a (q_i > theta);
 *   Xi+ = Pi_K(x), x = Xi + tau2 grad V-bar: Dykstra replayed with the same
 *        sweeps, each pair step back-propagated from its h (excess map
 *        f(h) = h (1 - 1/|h|) for |h| > 1, Jacobian (1 - 1/|h|) I + h h^T/|h|^3);
 *   div and grad are each other's negative adjoints on the masked domain.
 * States are checkpointed every ck iterations and each segment recomputed. */
typedef struct {
    const or_layout *m;
    int K;
    double c[8], lambda, tau1, tau2, tol;
    const double *u;
    size_t span;
} adj_t;

static void adj_nbrs(const or_layout *m, uint32_t n, uint32_t row, uint32_t col, size_t o, int nx[3], int px[3],
                     size_t on[3], size_t op[3])
{
    nx[0] = n + 1 < m->n_images; nx[1] = has_next_row(m, row, col); nx[2] = has_next_col(m, row, col);
    px[0] = n > 0; px[1] = has_prev_row(m, row, col); px[2] = has_prev_col(m, row, col);
    on[0] = o + 1; on[1] = o + m->nstride; on[2] = nx[2] ? line_of(m, row, col + 1) * m->nstride + n : 0;
    op[0] = o - 1; op[1] = o - m->nstride; op[2] = px[2] ? line_of(m, row, col - 1) * m->nstride + n : 0;
}

/* Dykstra with the oracle's semantics (20 sweeps, stop when the largest
 * correction norm of a sweep is below tol), recording h of every pair step. */
static int adj_dykstra(int K, double q[8][3], double tol, double hist[20][15][3])
{
    double r[15][3] = {{0}};
    int sweep;
    for (sweep = 0; sweep < 20; ++sweep) {
        double mx = -DBL_MAX;
        for (int i1 = 0, p = 0; i1 < K - 1; ++i1)
            for (int i2 = i1 + 1; i2 < K; ++i2, ++p) {
                double h[3], e[3], d[3];
                for (int a = 0; a < 3; ++a) hist[sweep][p][a] = h[a] = (q[i2][a] - q[i1][a]) + r[p][a];
                double l = sqrt(fma(h[2], h[2], fma(h[1], h[1], h[0] * h[0])));
                if (l > 1.0) { double s = (l - 1.0) / l; for (int a = 0; a < 3; ++a) e[a] = s * h[a]; }
                else e[0] = e[1] = e[2] = 0.0;
                for (int a = 0; a < 3; ++a) d[a] = 0.5 * (e[a] - r[p][a]);
                double dn = sqrt(fma(d[2], d[2], fma(d[1], d[1], d[0] * d[0])));
                mx = (dn < mx) ? mx : dn;
                for (int a = 0; a < 3; ++a) { q[i1][a] += d[a]; q[i2][a] -= d[a]; r[p][a] = e[a]; }
            }
        if (tol > mx) { ++sweep; break; }
    }
    return sweep;
}

/* g: adjoint of the output q on entry, of the input q on return. */
static void adj_dykstra_back(int K, int sweeps, double hist[20][15][3], double g[8][3])
{
    double gr[15][3] = {{0}};
    for (int s = sweeps - 1; s >= 0; --s)
        for (int i1 = K - 2, p = K * (K - 1) / 2 - 1; i1 >= 0; --i1)
            for (int i2 = K - 1; i2 > i1; --i2, --p) {
                const double *h = hist[s][p];
                double gd[3], ge[3], gh[3] = {0, 0, 0};
                for (int a = 0; a < 3; ++a) { gd[a] = g[i1][a] - g[i2][a]; ge[a] = 0.5 * gd[a] + gr[p][a]; }
                double l = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
                if (l > 1.0) {
                    double hg = h[0] * ge[0] + h[1] * ge[1] + h[2] * ge[2], il = 1.0 / l;
                    for (int a = 0; a < 3; ++a) gh[a] = (1.0 - il) * ge[a] + hg * il * il * il * h[a];
                }
                for (int a = 0; a < 3; ++a) {
                    g[i1][a] -= gh[a];
                    g[i2][a] += gh[a];
                    gr[p][a] = -0.5 * gd[a] + gh[a];
                }
            }
}

/* Pi_S of q in place; act[i] = 1 on the active set. */
static void adj_simplex(int K, double *q, int *act)
{
    double u[8];
    for (int i = 0; i < K; ++i) u[i] = q[i];
    sort_desc_d(K, u);
    double sum = 0, rho_sum = 0;
    int rho = 0;
    for (int i = 0; i < K; ++i) {
        sum += u[i];
        if (u[i] > (sum - 1.0) / (double)(i + 1)) { rho = i + 1; rho_sum = sum; }
    }
    double theta = (rho_sum - 1.0) / (double)rho;
    for (int i = 0; i < K; ++i) {
        double a = q[i] - theta;
        act[i] = a > 0.0;
        q[i] = a < 0.0 ? 0.0 : a;
    }
}

static void adj_x(const adj_t *A, uint32_t n, uint32_t row, uint32_t col, const double *Vb, const double *Xi,
                  double x[8][3])
{
    const or_layout *m = A->m;
    const int K = A->K;
    size_t o = line_of(m, row, col) * m->nstride + n, on[3], op[3];
    int nx[3], px[3];
    adj_nbrs(m, n, row, col, o, nx, px, on, op);
    double vb[8], vk = 1.0;
    for (int i = 0; i < K - 1; ++i) { vb[i] = Vb[i * A->span + o]; vk -= vb[i]; }
    vb[K - 1] = vk;
    for (int a = 0; a < 3; ++a) {
        double p[8], xk = 0.0;
        if (nx[a]) {
            double vka = 1.0;
            for (int i = 0; i < K - 1; ++i) { double t = Vb[i * A->span + on[a]]; p[i] = t; vka -= t; }
            p[K - 1] = vka;
            for (int i = 0; i < K; ++i) p[i] -= vb[i];
        } else {
            for (int i = 0; i < K; ++i) p[i] = 0.0;
        }
        for (int i = 0; i < K - 1; ++i) { double xo = Xi[(3 * i + a) * A->span + o]; xk -= xo; x[i][a] = fma(p[i], A->tau2, xo); }
        x[K - 1][a] = fma(p[K - 1], A->tau2, xk);
    }
}

/* q of the V step at a voxel (K components), from V and Xi+. */
static void adj_qv(const adj_t *A, uint32_t n, uint32_t row, uint32_t col, const double *V, const double *Xi1,
                   double *q, double *d_out)
{
    const or_layout *m = A->m;
    const int K = A->K;
    size_t o = line_of(m, row, col) * m->nstride + n, on[3], op[3];
    int nx[3], px[3];
    adj_nbrs(m, n, row, col, o, nx, px, on, op);
    double xo[8][3], xp[8][3], v[8], vk = 1.0;
    for (int a = 0; a < 3; ++a) {
        double xk = 0.0, xkp = 0.0;
        for (int i = 0; i < K - 1; ++i) {
            xo[i][a] = Xi1[(3 * i + a) * A->span + o]; xk -= xo[i][a];
            xp[i][a] = px[a] ? Xi1[(3 * i + a) * A->span + op[a]] : 0.0; xkp -= xp[i][a];
        }
        xo[K - 1][a] = xk; xp[K - 1][a] = xkp;
    }
    for (int i = 0; i < K - 1; ++i) { v[i] = V[i * A->span + o]; vk -= v[i]; }
    v[K - 1] = vk;
    const double uu = A->u[o];
    for (int i = 0; i < K; ++i) {
        double div = 0.0;
        for (int a = 0; a < 3; ++a) { if (px[a]) div -= xp[i][a]; if (nx[a]) div += xo[i][a]; }
        double d = uu - A->c[i];
        d_out[i] = d;
        q[i] = fma(A->tau1, fma(-A->lambda, d * d, div), v[i]);
    }
}

#define ADJ_FOR_VOXELS(m)                                                        \
    _Pragma("omp parallel for schedule(dynamic, 8)")                            \
    for (uint32_t col = (m)->col_begin; col < (m)->col_end; ++col)              \
        for (uint32_t row = row_begin((m), col); row < row_end((m), col); ++row) \
            for (uint32_t n = 0; n < (m)->n_images; ++n)

/* One forward iteration s -> s1 (arrays of 10(K-1)... words: V | Vb | Xi). */
static void adj_forward(const adj_t *A, const double *s, double *s1)
{
    const or_layout *m = A->m;
    const int K = A->K;
    const size_t S = A->span;
    const double *V = s, *Vb = s + (K - 1) * S, *Xi = s + 2 * (K - 1) * S;
    double *V1 = s1, *Vb1 = s1 + (K - 1) * S, *Xi1 = s1 + 2 * (K - 1) * S;
    ADJ_FOR_VOXELS(m) {
        double x[8][3], hist[20][15][3];
        size_t o = line_of(m, row, col) * m->nstride + n;
        adj_x(A, n, row, col, Vb, Xi, x);
        adj_dykstra(K, x, A->tol, hist);
        for (int i = 0; i < K - 1; ++i) for (int a = 0; a < 3; ++a) Xi1[(3 * i + a) * S + o] = x[i][a];
    }
    ADJ_FOR_VOXELS(m) {
        double q[8], d[8], v[8];
        int act[8];
        size_t o = line_of(m, row, col) * m->nstride + n;
        adj_qv(A, n, row, col, V, Xi1, q, d);
        adj_simplex(K, q, act);
        for (int i = 0; i < K - 1; ++i) v[i] = V[i * S + o];
        for (int i = 0; i < K - 1; ++i) { Vb1[i * S + o] = 2.0 * q[i] - v[i]; V1[i * S + o] = q[i]; }
    }
}

static void adj_start(const adj_t *A, const int8_t *init, double *s)
{
    const or_layout *m = A->m;
    const int K = A->K;
    const size_t S = A->span;
    memset(s, 0, sizeof(double) * 5 * (K - 1) * S);
    ADJ_FOR_VOXELS(m) {
        size_t o = line_of(m, row, col) * m->nstride + n;
        int cl = 0;
        if (init) cl = init[o];
        else {
            double best = DBL_MAX;
            for (int k = 0; k < K; ++k) { double dd = fabs(A->c[k] - A->u[o]); if (dd < best) { best = dd; cl = k; } }
        }
        for (int i = 0; i < K - 1; ++i) s[i * S + o] = s[(K - 1 + i) * S + o] = (i == cl) ? 1.0 : 0.0;
    }
}

static double adj_loss(const adj_t *A, const double *s, const float *target, double *gV)
{
    const or_layout *m = A->m;
    const int K = A->K;
    const size_t S = A->span;
    double L = 0.0, N = 0.0;
    for (uint32_t col = m->col_begin; col < m->col_end; ++col) N += (double)(row_end(m, col) - row_begin(m, col));
    N *= m->n_images;
    const double w = 1.0 / (N * K);
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : L)
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                size_t o = line_of(m, row, col) * m->nstride + n;
                int y = (int)target[o];
                double vk = 1.0, e[8];
                for (int i = 0; i < K - 1; ++i) { double v = s[i * S + o]; vk -= v; e[i] = v - (i == y ? 1.0 : 0.0); }
                e[K - 1] = vk - (K - 1 == y ? 1.0 : 0.0);
                for (int i = 0; i < K; ++i) L += w * e[i] * e[i];
                if (gV) for (int i = 0; i < K - 1; ++i) gV[i * S + o] = 2.0 * w * (e[i] - e[K - 1]);
            }
    return L;
}

static int adj_setup(adj_t *A, const or_layout *m, const double *u, int K, const double *c, double lambda,
                     double tau1, double tau2, double tol)
{
    if (K < 2 || K > 6) return 1;
    A->m = m; A->K = K; A->u = u; A->lambda = lambda; A->tau1 = tau1; A->tau2 = tau2; A->tol = tol;
    for (int i = 0; i < K; ++i) A->c[i] = c[i];
    A->span = (size_t)m->base[m->width];
    return 0;
}

double or_adj_loss_d(const or_layout *m, const double *u, const float *target, int K, const double *c,
                     double lambda, double tau1, double tau2, int iters, const int8_t *init, double tol)
{
    adj_t A;
    if (adj_setup(&A, m, u, K, c, lambda, tau1, tau2, tol)) return NAN;
    const size_t W = 5 * (size_t)(K - 1) * A.span;
    double *s = (double *)malloc(sizeof(double) * W), *s1 = (double *)malloc(sizeof(double) * W);
    if (!s || !s1) { free(s); free(s1); return NAN; }
    adj_start(&A, init, s);
    memset(s1, 0, sizeof(double) * W);
    for (int it = 0; it < iters; ++it) { adj_forward(&A, s, s1); double *t = s; s = s1; s1 = t; }
    double L = adj_loss(&A, s, target, NULL);
    free(s); free(s1);
    return L;
}

/* Returns the loss; grad_c[K], *grad_lambda and grad_u[span] (optional). */
double or_adj_grad_d(const or_layout *m, const double *u, const float *target, int K, const double *c,
                     double lambda, double tau1, double tau2, int iters, const int8_t *init, double tol,
                     int ck, double *grad_c, double *grad_lambda, double *grad_u)
{
    adj_t A;
    if (adj_setup(&A, m, u, K, c, lambda, tau1, tau2, tol) || iters < 1 || ck < 1) return NAN;
    const size_t S = A.span, W = 5 * (size_t)(K - 1) * S;
    const int nck = (iters + ck - 1) / ck;
    double *cks = (double *)malloc(sizeof(double) * W * (size_t)nck);
    double *seg = (double *)malloc(sizeof(double) * W * (size_t)(ck + 1));
    double *gs = (double *)calloc(W, sizeof(double));          /* gV | gVb | gXi of the current state */
    double *GQ = (double *)calloc((size_t)K * S, sizeof(double));
    double *GP = (double *)calloc((size_t)3 * K * S, sizeof(double));
    double *gu = (double *)calloc(S, sizeof(double));
    if (!cks || !seg || !gs || !GQ || !GP || !gu) { free(cks); free(seg); free(gs); free(GQ); free(GP); free(gu); return NAN; }
    /* forward with checkpoints at t = 0, ck, 2 ck, ... */
    adj_start(&A, init, cks);
    double *cur = seg, *nxt = seg + W;
    memcpy(cur, cks, sizeof(double) * W);
    for (int it = 0; it < iters; ++it) {
        adj_forward(&A, cur, nxt);
        double *t = cur; cur = nxt; nxt = t;
        if ((it + 1) % ck == 0 && (it + 1) / ck < nck) memcpy(cks + W * (size_t)((it + 1) / ck), cur, sizeof(double) * W);
    }
    const double loss = adj_loss(&A, cur, target, gs);
    double gl = 0.0, gc[8] = {0};
    double *gV = gs, *gVb = gs + (K - 1) * S, *gXi = gs + 2 * (K - 1) * S;
    for (int k = nck - 1; k >= 0; --k) {
        const int t0 = k * ck, t1 = (t0 + ck < iters) ? t0 + ck : iters;
        memcpy(seg, cks + W * (size_t)k, sizeof(double) * W);
        for (int j = 0; j < t1 - t0; ++j) adj_forward(&A, seg + W * (size_t)j, seg + W * (size_t)(j + 1));
        for (int t = t1 - 1; t >= t0; --t) {
            const double *st = seg + W * (size_t)(t - t0), *st1 = st + W;
            const double *V = st, *Vb = st + (K - 1) * S, *Xi = st + 2 * (K - 1) * S, *Xi1 = st1 + 2 * (K - 1) * S;
            double gl_t = 0.0, gc_t[8] = {0};
            /* A: through V+ = Pi_S(q) and V-bar+ = 2 V+ - V */
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : gl_t)
            for (uint32_t col = m->col_begin; col < m->col_end; ++col) {
                double gcl[8] = {0};
                for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
                    for (uint32_t n = 0; n < m->n_images; ++n) {
                        size_t o = line_of(m, row, col) * m->nstride + n;
                        double q[8], d[8], ge[8], gq[8];
                        int act[8], na = 0;
                        adj_qv(&A, n, row, col, V, Xi1, q, d);
                        adj_simplex(K, q, act);
                        double mean = 0.0;
                        for (int i = 0; i < K; ++i) {
                            ge[i] = (i < K - 1) ? gV[i * S + o] + 2.0 * gVb[i * S + o] : 0.0;
                            if (act[i]) { mean += ge[i]; ++na; }
                        }
                        mean /= (double)na;
                        for (int i = 0; i < K; ++i) { gq[i] = act[i] ? ge[i] - mean : 0.0; GQ[i * S + o] = gq[i]; }
                        for (int i = 0; i < K - 1; ++i) gV[i * S + o] = gq[i] - gq[K - 1] - gVb[i * S + o];
                        for (int i = 0; i < K; ++i) {
                            const double gG = -A.tau1 * gq[i];
                            gl_t += gG * d[i] * d[i];
                            gcl[i] += gG * (-2.0 * A.lambda * d[i]);
                            gu[o] += gG * 2.0 * A.lambda * d[i];
                        }
                    }
#pragma omp critical
                for (int i = 0; i < K; ++i) gc_t[i] += gcl[i];
            }
            gl += gl_t;
            for (int i = 0; i < K; ++i) gc[i] += gc_t[i];
            /* B: through Xi+ = Pi_K(Xi + tau2 grad V-bar) */
            ADJ_FOR_VOXELS(m) {
                size_t o = line_of(m, row, col) * m->nstride + n, on[3], op[3];
                int nx[3], px[3];
                adj_nbrs(m, n, row, col, o, nx, px, on, op);
                double g[8][3], full[8][3], x[8][3], hist[20][15][3];
                for (int i = 0; i < K; ++i)
                    for (int a = 0; a < 3; ++a)
                        full[i][a] = nx[a] ? -A.tau1 * (GQ[i * S + on[a]] - GQ[i * S + o]) : 0.0;
                for (int a = 0; a < 3; ++a) {
                    for (int i = 0; i < K - 1; ++i) g[i][a] = gXi[(3 * i + a) * S + o] + full[i][a] - full[K - 1][a];
                    g[K - 1][a] = 0.0;
                }
                adj_x(&A, n, row, col, Vb, Xi, x);
                const int sw = adj_dykstra(K, x, A.tol, hist);
                adj_dykstra_back(K, sw, hist, g);
                for (int i = 0; i < K - 1; ++i)
                    for (int a = 0; a < 3; ++a) gXi[(3 * i + a) * S + o] = g[i][a] - g[K - 1][a];
                for (int i = 0; i < K; ++i)
                    for (int a = 0; a < 3; ++a) GP[(3 * i + a) * S + o] = A.tau2 * g[i][a];
            }
            /* C: V-bar enters through grad V-bar: adjoint -div */
            ADJ_FOR_VOXELS(m) {
                size_t o = line_of(m, row, col) * m->nstride + n, on[3], op[3];
                int nx[3], px[3];
                adj_nbrs(m, n, row, col, o, nx, px, on, op);
                double full[8];
                for (int i = 0; i < K; ++i) {
                    double dv = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        if (nx[a]) dv += GP[(3 * i + a) * S + o];
                        if (px[a]) dv -= GP[(3 * i + a) * S + op[a]];
                    }
                    full[i] = -dv;
                }
                for (int i = 0; i < K - 1; ++i) gVb[i * S + o] = full[i] - full[K - 1];
            }
        }
    }
    if (grad_c) for (int i = 0; i < K; ++i) grad_c[i] = gc[i];
    if (grad_lambda) *grad_lambda = gl;
    if (grad_u) memcpy(grad_u, gu, sizeof(double) * S);
    free(cks); free(seg); free(gs); free(GQ); free(GP); free(gu);
    return loss;
}

/* ---- ingest: transpose/convert_to_fp.cpp:11-46, find_intensity_range.cpp:22-110 */
void or_u16_hist(const or_layout *m, const uint16_t *u, uint64_t *hist)
{
    memset(hist, 0, sizeof(uint64_t) * 65536);
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) ++hist[u[line_of(m, row, col) * m->nstride + n]];
}

void or_find_intensity_range(const or_layout *m, const uint16_t *u, double tail, uint16_t *lo, uint16_t *hi)
{
    uint64_t *h = (uint64_t *)malloc(sizeof(uint64_t) * 65536), total = 0, cum = 0;
    or_u16_hist(m, u, h);
    for (int i = 0; i < 65536; ++i) total += h[i];
    const double pl = tail * (double)total, ph = (1.0 - tail) * (double)total;
    size_t l = 0, r = 65536;
    for (size_t i = 0; i < 65536; ++i) {
        cum += h[i];
        if (pl >= (double)cum) l = i;
        if (ph <= (double)cum) { r = i; break; }
    }
    *lo = (uint16_t)l;
    *hi = (uint16_t)r;
    free(h);
}

void or_convert_to_fp(const or_layout *m, const void *in, int is_u16, float lo, float hi, float *out)
{
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                size_t o = line_of(m, row, col) * m->nstride + n;
                float v = is_u16 ? (float)((const uint16_t *)in)[o] : ((const float *)in)[o];
                out[o] = v <= lo ? 0.0f : (v >= hi ? 1.0f : (v - lo) / (hi - lo));
            }
}
