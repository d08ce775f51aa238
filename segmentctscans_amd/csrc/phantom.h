/*
 * phantom.h -- seeded synthetic micro-CT phantom in the cylinder layout.
 *
 * Synthetic-input tooling for the tests and the benchmark; it has no
 * counterpart in the reference (which only reads scanned TIFF stacks).
 *
 * Construction (all steps deterministic for a given seed, independent of the
 * thread count):
 *   1. dense box (n, rows, cols), 0 outside the cylinder mask, c_matrix inside;
 *   2. random sequential addition of non-overlapping spheres, radius
 *      U[rmin, rmax], centre uniform in the cylinder with the sphere fully
 *      inside radially, until the summed sphere volume reaches
 *      porosity * cylinder volume or max_attempts is spent; each sphere takes
 *      c_phase2 with probability p_phase2, else c_phase1;
 *   3. separable Gaussian blur (sigma, truncation radius, clamped edges);
 *   4. additive N(0, noise_sd^2) on in-mask voxels;
 *   5. u = (x - lo) / (hi - lo), stored in the cylinder layout.
 * Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by
 * the seed, with explicit Box-Muller for normals.
 */
#ifndef SCT_PHANTOM_H
#define SCT_PHANTOM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sct_phantom_params {
    uint32_t n_images, height, width;   /* n (cylinder axis), rows, cols */
    double xc, yc, r;                   /* mask circle in (col, row)     */
    double rmin, rmax;                  /* sphere radius range (voxels)  */
    double porosity;                    /* target sphere volume fraction */
    double c_matrix, c_phase1, c_phase2, p_phase2;
    double blur_sigma;
    int32_t blur_radius;
    double noise_sd;
    double lo, hi;                      /* normalisation                 */
    uint64_t seed;
    int64_t max_attempts;
} sct_phantom_params;

typedef struct sct_phantom_stats {
    uint32_t n_spheres;
    double porosity;          /* voxel fraction of in-mask voxels in spheres */
} sct_phantom_stats;

/* Fills the defaults of the benchmark phantoms for an N^2 x n box:
 * cylinder radius 0.45 N centred at (N-1)/2, glass 0.0339, brine 0.0409,
 * CO2 0.0527 (p = 0.5), 35% porosity, sigma 1 (radius 4), noise 0.008,
 * lo/hi = 0.019182531 / 0.075767376. */
void sct_phantom_defaults(sct_phantom_params *p, uint32_t n_images, uint32_t side,
                          double rmin, double rmax, uint64_t seed);

/* u: span floats (start_row/base describe the layout, nstride its stride).
 * labels: optional span bytes, 0 matrix / 1 phase1 / 2 phase2 (pre-blur).
 * Returns 0 on success, nonzero on bad arguments or allocation failure. */
int sct_phantom_generate(const sct_phantom_params *p, const uint32_t *start_row,
                         const uint32_t *base, uint32_t nstride, uint32_t col_begin,
                         uint32_t col_end, float *u, uint8_t *labels,
                         sct_phantom_stats *stats, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
