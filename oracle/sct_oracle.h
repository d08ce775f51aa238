/*
 * sct_oracle.h -- CPU restatement of the segmentctscans volumetric hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked by the
 * pytest suite, by the package's smoke check and by the CPU-baseline leg of
 * the benchmark.  It is never part of the product path, which lives in the
 * HIP library under segmentctscans_amd/ and fails loudly when that library
 * is missing.
 *
 * Every solver comes in two precisions built from one body (oracle_body.h):
 *   *_d : FP64 oracle.  The reference update in double precision.
 *   *_f : FP32 mirror.  The same update in float with the operation order and
 *         fused-multiply-add choices the HIP kernels use, so HIP results can
 *         be compared with it bit for bit.
 *
 * Reference files followed (paths relative to the reference repository):
 *   include/layout_cylinder.h:81-195      cylinder mapping, row ranges
 *   include/field.h:37-45                 128-byte stride rounding
 *   potts-model/potts-min-partition.cpp   init_v :56-94, grad + Dykstra :96-247,
 *                                         div + simplex :249-489, update :491-538,
 *                                         slab/halo driver :540-679
 *   totalvarmin/chambolle.cpp:33-276      Chambolle ROF (div, grad, update)
 *   calc_hist/calc_hist.cpp:23-138        histogram
 *   min_max/min_max.cpp:18-81             min / max
 *   grad-magn/grad_magn.cpp:17-79         gradient magnitude
 *   uv-hist/uv-hist-kernel.cpp:31-198     2-D (u, |grad u|) histogram
 *
 * Pinning of each function (what checks it other than the HIP path) is listed
 * in the header comment of oracle_body.h next to the function.
 */
#ifndef SCT_ORACLE_H
#define SCT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_layout {
    uint32_t n_images, height, width; /* extents: n (fast), rows, cols (slow) */
    uint32_t nstride;                 /* floats per (row, col) line            */
    uint32_t col_begin, col_end;      /* live column range                     */
    const uint32_t *start_row;        /* [width]                               */
    const uint32_t *base;             /* [width + 1], base[width] = span       */
} or_layout;

/* layout_cylinder.h:152-195 (cylinder constructor, FP64 edge rule). */
int or_cylinder(uint32_t n_images, uint32_t height, uint32_t width,
                double xc, double yc, double r, uint32_t nstride,
                uint32_t *start_row, uint32_t *base,
                uint32_t *col_begin, uint32_t *col_end, uint64_t *data_size);

/* Forward gradient / backward divergence with the reference's neighbour
 * predicates (potts-min-partition.cpp:125-133 and :387-395), used to check
 * adjointness <grad v, p> = -<v, div p>.  p is 3 span-sized arrays. */
void or_grad_d(const or_layout *m, const double *v, double *p0, double *p1, double *p2);
void or_div_d(const or_layout *m, const double *p0, const double *p1, const double *p2, double *d);

/* Single-voxel projections, exported for closed-form / brute-force checks. */
void or_proj_simplex_d(int K, double *q);
void or_proj_simplex_f(int K, float *q);
/* q: K x 3, row-major (label, axis).  Returns sweeps performed. */
int or_dykstra_d(int K, double *q, int max_sweeps, double tol);
int or_dykstra_f(int K, float *q, int max_sweeps, double tol);

/* Potts solver, potts-min-partition.cpp:540-679.
 *   u       input image in the cylinder layout (span floats)
 *   res     output, span floats: c[argmax label] on every in-mask voxel
 *   V_out   optional, (K-1) x span   final V        (single slab only)
 *   xi_out  optional, 3(K-1) x span  final Xi       (single slab only)
 *   dq      optional, [iters]        max_x sum_k (dV_k)^2 per iteration
 *                                    (last slab when several)
 *   slab_w  0: one slab; else the n-width of a slab (halo = iters)
 * Returns 0, or 1 when the slab is too narrow for the halo (the reference's
 * out-of-memory return), or 2 on invalid arguments. */
int or_potts_d(const or_layout *m, const float *u, int K, const float *c,
               float lambda, float tau1, float tau2, int iters, int slab_w,
               float *res, double *V_out, double *xi_out, double *dq, int storage, int nthreads);
/* storage: 0 FP32; 1 Xi and V-bar stored as bfloat16; 2 as IEEE half: both
 * are rounded (to nearest even) where they are stored and every later read,
 * the V step of the same iteration included, sees the stored values. */
int or_potts_f(const or_layout *m, const float *u, int K, const float *c,
               float lambda, float tau1, float tau2, int iters, int slab_w,
               float *res, float *V_out, float *xi_out, float *dq, int storage, int nthreads);
/* FP64 duality gap of the Potts saddle problem (no counterpart in the
 * reference; definition in sct_oracle.c).  V: (K-1) x span, xi: 3(K-1) x span,
 * component-major in the input layout. */
double or_support_d(int K, const double *p);        /* p: K x 3, sum_i p_i = 0 */
double or_support_irls_d(int K, const double *p);   /* the K >= 4 method, any K >= 2 */
int or_potts_gap_d(const or_layout *m, const float *u, int K, const float *c, float lambda,
                   const double *V, const double *xi, double *primal, double *dual);

/* FP64 adjoint of the unrolled iteration (no counterpart in the reference;
 * definition in sct_oracle.c).  u: double [span]; target: label index per
 * voxel [span]; init: start labels [span] or NULL for the nearest label;
 * tol: Dykstra stopping tolerance (1e-6 in the solver). */
double or_adj_loss_d(const or_layout *m, const double *u, const float *target, int K, const double *c,
                     double lambda, double tau1, double tau2, int iters, const int8_t *init, double tol);
double or_adj_grad_d(const or_layout *m, const double *u, const float *target, int K, const double *c,
                     double lambda, double tau1, double tau2, int iters, const int8_t *init, double tol,
                     int ck, double *grad_c, double *grad_lambda, double *grad_u);

/* Ingest (transpose/): u16 histogram [65536], tail cut, conversion to float. */
void or_u16_hist(const or_layout *m, const uint16_t *u, uint64_t *hist);
void or_find_intensity_range(const or_layout *m, const uint16_t *u, double tail, uint16_t *lo, uint16_t *hi);
void or_convert_to_fp(const or_layout *m, const void *in, int is_u16, float lo, float hi, float *out);

float or_bf16_round(float f);
float or_f16_round(float f);

/* Chambolle ROF, chambolle.cpp:225-276.  g is overwritten with the result.
 * xi_out optional (3 x span), dq optional [iters]. */
int or_chambolle_d(const or_layout *m, double *g, float lambda, float tau, int iters,
                   double *xi_out, double *dq, int nthreads);
int or_chambolle_f(const or_layout *m, float *g, float lambda, float tau, int iters,
                   float *xi_out, float *dq, int nthreads);

/* calc_hist.cpp:23-80 with dither 0; counts[nbins] are incremented. */
void or_calc_hist(const or_layout *m, const float *u, float rs, float re,
                  uint32_t nbins, uint32_t *counts);
/* min_max.cpp:18-81 */
void or_min_max(const or_layout *m, const float *u, float *mn, float *mx);
/* uv-hist-kernel.cpp:31-71: min/max over strictly positive values */
void or_grad_min_max(const or_layout *m, const float *v, float *mn, float *mx);
/* grad_magn.cpp:17-58 */
void or_grad_magn(const or_layout *m, const float *u, float *v);
/* uv-hist-kernel.cpp:101-198; counts[(nbins-1)*stride + nbins] incremented */
void or_uv_hist(const or_layout *m, const float *u, const float *v,
                float u0, float u1, float v0, float v1, uint32_t nbins,
                uint32_t stride, uint32_t *counts);

#ifdef __cplusplus
}
#endif
#endif
