/*
 * sct.h -- C ABI of the MI355X-native segmentctscans hot path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (SCT_OK on success); the message of the last failure on a context is
 * returned by sct_ctx_last_error().  The C++ header segmentctscans.hpp wraps
 * these calls in the reference's own C++ signatures and turns status codes
 * back into the reference's exceptions / null returns.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   include/cuda-context.h:11-37             context, upload/download, memory query
 *   include/cuda-details/device-field-impl.h DeviceField<T> (voxels + mapping tables)
 *   include/layout_cylinder.h:152-195        cylinder mapping constructor
 *   include/field.h:37-45                    get_stride / prev_stride
 *   potts-model/potts-min-partition.h:8-16   potts_min_partition<K>
 *   totalvarmin/chambolle.h:5-6              chambolle
 *   include/calc_hist.h:8-12                 calc_hist (both overloads)
 *   include/min_max.h:7                      min_max
 *   include/grad_magn.h:5-6                  grad_magn
 *   uv-hist/uv-hist.h:18-44                  uv_hist (both overloads)
 *   transpose/convert_to_fp.h:8-10           convert_to_fp<uint16_t | float>
 *   transpose/find_intensity_range.h:8-9     find_intensity_range
 *   include/field.h:92-153                   Field<T> .bin archive save / load
 *
 * Threading: one HIP stream per context; a context is not thread-safe.
 * Every call that returns host values synchronises the stream, as the
 * reference does; sct_grad_magn is asynchronous like grad_magn.cpp:60-79.
 */
#ifndef SCT_H
#define SCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SCT_OK          0
#define SCT_ERR_INVALID 1  /* bad argument (reference: validation_error)            */
#define SCT_ERR_HIP     2  /* HIP runtime failure (reference: cuda::cuda_error)      */
#define SCT_ERR_SIZE    3  /* host/device span mismatch (cuda-context.cpp:69-71)     */
#define SCT_ERR_NOMEM   4  /* slab cannot hold the halo: the reference's null return
                              (potts-min-partition.cpp:571-577)                      */

/* ---- element types ------------------------------------------------------ */
#define SCT_F32 0
#define SCT_U16 1

typedef struct sct_ctx sct_ctx;
typedef struct sct_field sct_field;
typedef struct sct_potts_solver sct_potts_solver;

/* Cylinder layout (layout_cylinder::mapping, layout_cylinder.h:29-112).
 * offset(n, row, col) = base[col] + (row - start_row[col]) * nstride + n. */
typedef struct sct_layout {
    uint32_t n_images, height, width;   /* extents: n (contiguous), rows, cols */
    uint32_t nstride;                   /* elements per (row, col) line         */
    uint32_t col_begin, col_end;
    const uint32_t *start_row;          /* [width]                              */
    const uint32_t *base;               /* [width + 1]; base[width] = span      */
} sct_layout;

/* ---- host-side layout helpers (no device needed) ----------------------- */

/* layout_cylinder::mapping_storage(extents, xc, yc, r, nstride),
 * layout_cylinder.h:152-195.  start_row[width], base[width + 1]. */
int sct_cylinder_layout(uint32_t n_images, uint32_t height, uint32_t width,
                        double xc, double yc, double r, uint32_t nstride,
                        uint32_t *start_row, uint32_t *base,
                        uint32_t *col_begin, uint32_t *col_end, uint64_t *data_size);
/* Field<float>::get_stride / prev_stride, field.h:37-45 (128-byte rows). */
uint32_t sct_get_stride(uint32_t n);
uint32_t sct_prev_stride(uint32_t n);

/* ---- context: cuda-context.h:11-14, cuda-context.cpp:14-20 ------------- */
int  sct_ctx_create(int device, sct_ctx **out);
void sct_ctx_destroy(sct_ctx *ctx);
const char *sct_ctx_last_error(const sct_ctx *ctx);
/* The context's hipStream_t (for callers that time or order work on it). */
void *sct_ctx_stream(sct_ctx *ctx);
int  sct_ctx_synchronize(sct_ctx *ctx);
/* device_memory(), cuda-context.cpp:22-41: free + pool (reserved - used). */
int  sct_device_memory(sct_ctx *ctx, size_t *available, size_t *total);

/* ---- fields: cuda-context.h:24-31, device-field-impl.h:14-54 ----------- */
/* DeviceField(mapping): uninitialised voxels. */
int  sct_field_create(sct_ctx *ctx, const sct_layout *layout, int dtype, sct_field **out);
/* upload(ctx, mdspan): copies span elements from host. */
int  sct_upload(sct_ctx *ctx, const sct_layout *layout, int dtype, const void *host,
                sct_field **out);
/* download(ctx, field, mdspan): span_elems must equal the field's span
 * (SCT_ERR_SIZE otherwise, as cuda-context.cpp:69-71 throws); synchronises. */
int  sct_download(sct_ctx *ctx, const sct_field *field, void *host, size_t span_elems);
void sct_field_destroy(sct_field *field);
void *sct_field_data(sct_field *field);       /* device pointer to the voxels   */
size_t sct_field_span(const sct_field *field); /* required_span_size()          */
uint64_t sct_field_data_size(const sct_field *field); /* in-mask voxels         */

/* ---- Potts: potts-min-partition.h:8-16 ---------------------------------- */
#define SCT_STORAGE_F32  0
#define SCT_STORAGE_BF16 1   /* Xi and V-bar stored as bfloat16, arithmetic FP32 */
#define SCT_STORAGE_F16  2   /* Xi and V-bar stored as IEEE half                 */

#define SCT_KERNEL_AUTO    0  /* the two-pass kernels (fastest measured on MI355X) */
#define SCT_KERNEL_TWOPASS 1  /* reference-shaped: Xi pass, then V pass       */
#define SCT_KERNEL_FUSED   2  /* one pass per iteration, col-marching LDS tiles */

typedef struct sct_potts_opts {
    uint32_t size;          /* sizeof(sct_potts_opts), for forward compatibility  */
    int32_t  storage;       /* SCT_STORAGE_*                                      */
    int32_t  kernel;        /* SCT_KERNEL_*                                       */
    int32_t  slab_w;        /* >0: force this slab width (halo = iters)           */
    int64_t  mem_available; /* >=0: pretend device_memory().available is this     */
    float   *dq_trace;      /* optional host [iters]: max_x sum_k (dV_k)^2 per
                               iteration (last slab when several)                 */
    double  *gap_trace;     /* optional host [gap_capacity]: relative primal-dual
                               gap (E - D) / E in FP64 after every gap_every-th
                               iteration (extension; last slab when several)      */
    int32_t  gap_every;     /* iterations between gap evaluations (0: never)      */
    int32_t  gap_capacity;  /* entries of gap_trace; sct_potts_min_partition
                               takes 0 as iters / gap_every                       */
    double   gap_tol;       /* > 0: sct_potts_min_partition stops a slab at the
                               first gap evaluation whose relative gap is below
                               gap_tol (needs gap_every > 0; extension, the
                               reference runs a fixed number of iterations)       */
    int32_t *iters_done;    /* optional host out: iterations performed (the most
                               over the slabs)                                    */
} sct_potts_opts;

void sct_potts_opts_init(sct_potts_opts *opts);

/* potts_min_partition<K>(ctx, g, cpu_map, ci, lambda, tau1, tau2, iters):
 * K in [2, 6]; *out receives a new field (same layout as g) holding c[argmax]
 * per voxel.  SCT_ERR_NOMEM when a slab cannot hold 2*iters halo planes. */
int sct_potts_min_partition(sct_ctx *ctx, const sct_field *g, int K, const float *c,
                            float lambda, float tau1, float tau2, int iters,
                            const sct_potts_opts *opts, sct_field **out);

/* Stepping interface over the same solver (single slab; extension): create
 * runs init_v (potts-min-partition.cpp:641-644), iterate runs more primal-dual
 * iterations, labels runs update_kernel (:491-538) into res. */
int  sct_potts_solver_create(sct_ctx *ctx, const sct_field *g, int K, const float *c,
                             float lambda, float tau1, float tau2,
                             const sct_potts_opts *opts, sct_potts_solver **out);
int  sct_potts_solver_iterate(sct_potts_solver *s, int iters);
/* Kernel launches per iteration: 2 for the two-pass kernels (Xi, then V),
 * 1 for the fused kernel. */
int  sct_potts_solver_kernels(const sct_potts_solver *s);
/* iterate, with HIP events on the library stream between the launches:
 * kernel_ms[i * kernels + j] receives the duration of launch j of iteration i
 * (host array of iters * sct_potts_solver_kernels(s)).  Synchronises.  No
 * counterpart in the reference, which has no kernel timing. */
int  sct_potts_solver_iterate_timed(sct_potts_solver *s, int iters, float *kernel_ms);
int  sct_potts_solver_labels(sct_potts_solver *s, sct_field *res);
/* FP64 primal energy E(V) and dual energy D(Xi) of the current state for the
 * saddle function <V, G> + <grad V, Xi>, G_i = lambda (u - c_i)^2; E >= D and
 * E - D -> 0 as the iteration converges.  Synchronises.  No counterpart in the
 * reference, which runs a fixed number of iterations. */
int  sct_potts_solver_gap(sct_potts_solver *s, double *primal, double *dual);
/* Downloads V ((K-1) x span) and Xi (3(K-1) x span), component-major, in the
 * input layout; either pointer may be NULL.  Synchronises. */
int  sct_potts_solver_state(sct_potts_solver *s, float *V, float *xi);
int  sct_potts_solver_dq(sct_potts_solver *s, float *dq, int n);
void sct_potts_solver_destroy(sct_potts_solver *s);

/* ---- Chambolle TV: chambolle.h:5-6 (g overwritten) ---------------------- */
int sct_chambolle(sct_ctx *ctx, sct_field *g, float lambda, float tau, int iters,
                  float *dq_trace);

/* ---- reductions ---------------------------------------------------------- */
/* min_max.h:7 */
int sct_min_max(sct_ctx *ctx, const sct_field *u, float *mn, float *mx);
/* calc_hist.h:11-12 with nbins widened from uint16_t to uint32_t (65536 bins
 * are not expressible in the reference's type).  counts[nbins] receives the
 * integer counts, density[nbins] (optional) count / (data_size * bin_width)
 * as calc_hist.cpp:131-137.  delta != 0 adds uniform dither in
 * [-delta/2, delta/2) from Philox4x32-10 keyed by seed. */
int sct_calc_hist(sct_ctx *ctx, const sct_field *u, float range_start, float range_end,
                  uint32_t nbins, float delta, uint64_t seed, uint32_t *counts,
                  double *density);
/* sct_calc_hist without allocating.  No counterpart in the reference: for
 * callers that cannot allocate device or pinned memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as sct_calc_hist, for nbins up to
 * SCT_NOALLOC_MAX_BINS (SCT_ERR_INVALID beyond).  Its device temporaries
 * live in a fixed module-scope buffer of 17 MiB per GPU, held until the
 * process ends; it calls no hipMalloc, hipFree or hipHostMalloc and copies
 * the counts straight into `counts`.  Every call on one GPU, from any context
 * or thread, shares that buffer: calls that overlap on one GPU race.
 * Synchronises. */
#define SCT_NOALLOC_MAX_BINS 262144u
int sct_calc_hist_noalloc(sct_ctx *ctx, const sct_field *u, float range_start, float range_end,
                          uint32_t nbins, float delta, uint64_t seed, uint32_t *counts, double *density);
/* calc_hist.h:8-9: range from min_max (calc_hist.cpp:83-92). */
int sct_calc_hist_auto(sct_ctx *ctx, const sct_field *u, uint32_t nbins, float delta,
                       uint64_t seed, uint32_t *counts, double *density,
                       float *range_start, float *range_end);
/* grad_magn.h:5-6; asynchronous. */
int sct_grad_magn(sct_ctx *ctx, const sct_field *u, sct_field **v);
/* uv-hist-kernel.cpp:31-71: min/max over strictly positive v. */
int sct_grad_min_max(sct_ctx *ctx, const sct_field *v, float *mn, float *mx);
/* uv-hist.h:37-44: counts[(nbins - 1) * stride + nbins], stride = nbins
 * rounded up to 32 (returned in *stride). */
int sct_uv_hist(sct_ctx *ctx, const sct_field *u, const sct_field *v,
                float u_start, float u_end, float v_start, float v_end,
                uint32_t nbins, uint32_t *counts, uint32_t *stride);
/* uv-hist.h:30-35: ranges from min_max(u) and grad_min_max(v);
 * ranges[4] = {u_start, u_end, v_start, v_end}. */
int sct_uv_hist_auto(sct_ctx *ctx, const sct_field *u, const sct_field *v, uint32_t nbins,
                     uint32_t *counts, uint32_t *stride, float *ranges);

/* ---- .bin archive (host only): Field<T>::save / load, field.h:92-153 -------
 * One MSB-first bit stream: the mapping header of layout_cylinder.h:233-318
 * (Exp-Golomb ue / se, io/writeprimitives.h:55-66, io/readprimitives.h:53-75),
 * lo and hi as 32-bit patterns for float fields, then the in-mask voxels in
 * (col, row, n) order as 16- or 32-bit values; the file is the stream's bytes,
 * the last zero-padded (transpose/transpose.cpp:303-308).  dtype: SCT_F32
 * (Field<float>) or SCT_U16 (Field<uint16_t>, no lo / hi). */
typedef struct sct_bin_info {
    uint32_t n_images, height, width, nstride, col_begin, col_end;
    uint64_t span;          /* elements of the field (base[width])             */
    uint64_t data_size;     /* in-mask voxels                                   */
    float    lo, hi;        /* float fields; 0 for u16                          */
    uint64_t header_bits;   /* stream bits before the first voxel               */
} sct_bin_info;

/* Parses the header of an archive in buf[len]; start_row[width] and
 * base[width + 1] (optional) receive the mapping tables.  SCT_ERR_SIZE when
 * buf is shorter than the archive it describes, SCT_ERR_INVALID when the
 * header is not a cylinder mapping. */
int sct_bin_parse(const void *buf, size_t len, int dtype, sct_bin_info *info, uint32_t *start_row,
                  uint32_t *base);
/* Reads the voxels of a parsed archive into data[span] (padding zero). */
int sct_bin_read(const void *buf, size_t len, int dtype, const sct_bin_info *info, void *data);
/* Upper bound of the bytes sct_bin_write needs. */
size_t sct_bin_bound(const sct_layout *layout, int dtype);
/* Writes the archive of data[span] (layout's mapping; lo, hi for float) into
 * buf[cap]; *written receives its bytes. */
int sct_bin_write(const sct_layout *layout, int dtype, float lo, float hi, const void *data, void *buf,
                  size_t cap, size_t *written);

/* ---- ingest: transpose/ ------------------------------------------------- */
/* convert_to_fp<ElementType>(ctx, in, lo, hi), convert_to_fp.h:8-10 with
 * ElementType uint16_t (SCT_U16) or float (SCT_F32): *out receives a new float
 * field, v <= lo -> 0, v >= hi -> 1, else (v - lo) / (hi - lo); padding 0.
 * Asynchronous, as the reference. */
int sct_convert_to_fp(sct_ctx *ctx, const sct_field *in, float lo, float hi, sct_field **out);
/* find_intensity_range(ctx, field, tail_thr), find_intensity_range.h:8-9: the
 * 65536-bin histogram of a u16 field and its tail cut; hist (optional)
 * receives the 65536 counts (64-bit; SCT_ERR_SIZE if 1/128 of the field's
 * lines hold 2^32 or more voxels).  Synchronises. */
int sct_find_intensity_range(sct_ctx *ctx, const sct_field *field, double tail_thr, uint16_t *lo,
                             uint16_t *hi, uint64_t *hist);

/* ---- batched Potts (no counterpart in the reference, whose driver solves one
 * volume per call, potts-model.cpp:164) ------------------------------------ */
typedef struct sct_potts_problem {
    const sct_field *g;     /* input field (F32); problems may share one      */
    int32_t K;              /* 2..6                                           */
    float   c[6];           /* normalised label intensities c[0..K-1]         */
    float   lambda;
} sct_potts_problem;

#define SCT_BATCH_GRAPH 1   /* capture all iterations into one hipGraph launch */
#define SCT_BATCH_AUTOTUNE 2 /* pick each K group's Xi workgroup tile (and, persistent, its
                               workgroups per CU) by timing (cached per context) */
#define SCT_BATCH_PERSISTENT 4 /* each K group's iterations in one persistent launch: tiles
                                  from one ordered ticket queue; a half-step's tiles start once
                                  every tile of the half-step before is complete */

/* Solves n problems, iters primal-dual iterations each, with common tau1 and
 * tau2 and storage (SCT_STORAGE_*): out[p] receives a new field of c[argmax]
 * labels for problem p, equal to what sct_potts_min_partition returns for it
 * (one slab).  Problems of one K share each Xi and V launch; *elapsed_ms
 * (optional) receives the event time of the iterations.  Synchronises. */
int sct_potts_batch(sct_ctx *ctx, int n, const sct_potts_problem *problems, float tau1, float tau2,
                    int iters, int storage, int flags, sct_field **out, float *elapsed_ms);
/* The Xi workgroup tiles (waves per workgroup: 4, 2 or 1; 0 for a K with no
 * problem) the last sct_potts_batch call on ctx used, for K = 2..6.  With
 * SCT_BATCH_AUTOTUNE the call times 3 iterations of each tile per K group
 * before the timed iterations and restarts the group from its initial state,
 * so the results do not depend on the choice. */
int sct_potts_batch_tiles(sct_ctx *ctx, int *waves);
/* Fixes the tile of each K (waves[K - 2] in {1, 2, 4}; 0 leaves K to the
 * default or the autotuner) for the following sct_potts_batch calls on ctx. */
int sct_potts_batch_set_tiles(sct_ctx *ctx, const int *waves);
/* Fixes the workgroups per CU of each K's persistent launches (per_cu[K - 2] in
 * 1..64; 0 leaves K to the default or the autotuner) for the following
 * sct_potts_batch calls on ctx.  A count above what a CU holds gives a grid
 * that is not resident all at once, which the launch handles (its waits only
 * involve workgroups that have started).  No counterpart in the reference. */
int sct_potts_batch_set_occupancy(sct_ctx *ctx, const int *per_cu);
/* The workgroups per CU of the persistent launches of the last
 * sct_potts_batch call on ctx, for K = 2..6 (0 without SCT_BATCH_PERSISTENT).
 * No counterpart in the reference. */
int sct_potts_batch_occupancy(sct_ctx *ctx, int *per_cu);

/* ---- adjoint (no counterpart in the reference) ---------------------------- */
/* Gradient of L = 1/(N K) sum_x sum_{i<K} (V_i(x) - [target(x) = i])^2, V after
 * iters primal-dual iterations from the nearest-label start (N in-mask voxels,
 * target: float field of label indices with the layout of g), with respect to
 * c[0..K-1], lambda and u, by reverse-mode differentiation of the unrolled
 * iteration (FP32).  States are checkpointed every checkpoint_every
 * iterations (0: round(sqrt(iters))) and the states between two checkpoints
 * recomputed segment by segment, the next segment on a second stream while
 * the current one is differentiated (memory: iters / checkpoint_every + 2
 * checkpoint_every states); checkpoint_every = 1 keeps all iters + 1 states
 * and recomputes none.  A state is 5 (K - 1) floats and one byte per padded
 * voxel.
 * *grad_u (optional) receives a new field; times_ms[2] (optional) the event
 * times of the forward pass and of the reverse pass.  Synchronises. */
int sct_potts_adjoint(sct_ctx *ctx, const sct_field *g, const sct_field *target, int K, const float *c,
                      float lambda, float tau1, float tau2, int iters, int checkpoint_every,
                      double *loss, double *grad_c, double *grad_lambda, sct_field **grad_u,
                      float *times_ms);

/* The same differentiation split into a forward and a backward call, for a
 * caller-defined loss.  sct_potts_tape_forward runs iters iterations from the
 * nearest-label start, keeps the checkpoints in *tape and, with V non-null,
 * stores K new float fields V[i] = V_i after the last iteration (V_{K-1} = 1
 * minus the stored components; the layout of g).  sct_potts_tape_backward
 * takes dL/dV_i (K float fields with the layout of g) and returns dL/dc[K],
 * dL/dlambda and (optional) a new field dL/du; it may be called any number of
 * times on one tape.  g must stay alive and unchanged while the tape lives.
 * *forward_ms / *backward_ms (optional): event times.  Both synchronise. */
typedef struct sct_potts_tape sct_potts_tape;
int sct_potts_tape_forward(sct_ctx *ctx, const sct_field *g, int K, const float *c, float lambda, float tau1,
                           float tau2, int iters, int checkpoint_every, sct_field **V, sct_potts_tape **tape,
                           float *forward_ms);
int sct_potts_tape_backward(sct_potts_tape *tape, const sct_field *const *dV, double *grad_c, double *grad_lambda,
                            sct_field **grad_u, float *backward_ms);
void sct_potts_tape_destroy(sct_potts_tape *tape);

/* ---- gmm-peaks: passes of the 1-D Gaussian mixture fit, gmm-peaks/gmm.h ----
 * Samples (doubles) live on the device; component k has weight pi[k], mean
 * mu[k] and variance var[k], density pi_k (2 pi)^-1/2 exp(-(x - mu_k)^2 /
 * (2 var_k)) / sqrt(|var_k|), K in [1, 16]. */
typedef struct sct_samples sct_samples;
int sct_samples_create(sct_ctx *ctx, const double *x, uint64_t n, sct_samples **out);
void sct_samples_destroy(sct_samples *s);
/* The fused E+M pass, gmm.h:145-201: responsibilities g_ik = density_k /
 * sum_j density_j; w[k] = sum_i g_ik, s1[k] = sum_i g_ik x_i, s2[k] =
 * sum_i g_ik x_i^2, *nll = -sum_i log sum_k density_k (a sample whose
 * densities all underflow counts 1/K to each component and -log DBL_MIN).
 * Deterministic: workgroup partial sums are added in a fixed order. */
int sct_gmm1d_pass(const sct_samples *s, int K, const double *pi, const double *mu, const double *var, double *w,
                   double *s1, double *s2, double *nll);
/* gmm.h:204-219: clusters[i] = the component of largest density (first on ties). */
int sct_gmm1d_assign(const sct_samples *s, int K, const double *pi, const double *mu, const double *var,
                     int32_t *clusters);

/* ---- bimm-peaks: the sample likelihood of the bivariate intensity /
 * gradient-magnitude mixture, bimm-peaks/puv.h:42-130 ------------------------
 * Samples (u, v) (doubles) live on the device.  The mixture has K in [1, 6]
 * pure phases with intensities I[K] and weights w[0..K-1], and the interfaces
 * of the pairs i < j (strict-upper order) with weights w[K..]; noise
 * deviation exp(log_sn), blur exp(log_sb), interface half-width exp(log_ds),
 * noise correlation rho < 1. */
typedef struct sct_uv_samples sct_uv_samples;
int sct_uv_samples_create(sct_ctx *ctx, const double *u, const double *v, uint64_t n, sct_uv_samples **out);
void sct_uv_samples_destroy(sct_uv_samples *s);
/* *nll = sum over samples [first, first + count) of -log full_p(u, v)
 * (bimm-peaks.cpp:458-472 before the division by the count), each interface
 * density's integral over t by adaptive Gauss-Kronrod to 1e-10 relative.
 * Deterministic.  Synchronises. */
int sct_bimm_nll(const sct_uv_samples *s, int K, const double *w, const double *I, double log_sn, double log_sb,
                 double log_ds, double rho, uint64_t first, uint64_t count, double *nll);

/* ---- diagnostics (no counterpart in the reference) ------------------------ */
/* Checks on the GPU that the Dykstra projection's range-specific correctly
 * rounded square root (x in [1, 2^64)) and quotient (b in [1, 2^64),
 * a in [2^-23, b)), the same quotient on the histograms' range (b in
 * [2^-32, 2^32], |a| in [b 2^-40, b]), and the simplex projection's quotients
 * by 3, 5 and 7 (any finite y other than -0), equal IEEE sqrtf and division
 * on n random operands; *bad_sqrt and *bad_div receive the numbers of
 * differing results. */
int sct_selftest_cr_math(sct_ctx *ctx, uint64_t n, uint64_t seed, uint64_t *bad_sqrt, uint64_t *bad_div);

/* Rounds x[n] to the stored element of Xi and V-bar for storage
 * SCT_STORAGE_BF16 or SCT_STORAGE_F16 with the solver's own device code;
 * out[n] receives the 16-bit patterns.  Synchronises. */
int sct_selftest_store_round(sct_ctx *ctx, int storage, const float *x, uint64_t n, uint16_t *out);

#ifdef __cplusplus
}
#endif
#endif
