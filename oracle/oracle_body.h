/*
 * oracle_body.h -- precision-generic body of the oracle (test infrastructure).
 *
 * Included twice by sct_oracle.c: REAL=double (suffix _d, the FP64 oracle)
 * and REAL=float (suffix _f, the FP32 mirror of the HIP kernels).  Both
 * follow the reference's operation order; where the reference leaves the
 * rounding to its compiler (FMA contraction) this file writes FMA explicitly
 * at the places a contracting compiler fuses a*b+c, and the HIP kernels do
 * the same, so the FP32 mirror and the HIP path agree bit for bit.
 * Two documented deviations from the reference's written order, shared with
 * the kernels:
 *   - Dykstra excess: s = (l - 1) / l; e = s * h   (reference: (l-1)*h/l,
 *     potts-min-partition.cpp:218) -- one division per violated pair
 *     instead of three;
 *   - Potts data term: q = fma(tau1, fma(-lambda, d*d, div), v)
 *     (reference: v + tau1*(div - lambda*sqr(u-c)), :467).
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   cylinder            brute-force disc membership, span/offset bijection
 *   grad/div predicates adjointness <grad v, p> = -<v, div p> (exact in
 *                       integers), constants annihilated
 *   proj_simplex        KKT conditions and a bisection solver
 *   dykstra             K=2 closed form; K>=3 10^4-sweep limit, feasibility
 *                       and variational inequality
 *   potts               invariants (sum V = 1, V >= 0, Xi feasible), constant
 *                       image fixed point, multi-slab == single-slab bitwise,
 *                       FP32 mirror vs FP64 label agreement
 *   chambolle           |xi| <= 1, constant fixed point, exact 1-D step-edge
 *                       ROF solution, lambda -> 0 identity
 *   calc_hist, uv_hist  exact known answers (bin edges, ramps, NaN, range)
 *   min_max, grad_magn  numpy minmax, linear ramps
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* ---- simplex projection, potts-min-partition.cpp:292-347 ------------- */
static inline void FN(sort_desc)(int K, REAL *u)
{
    /* Sorting networks of :251-290 (K = 2..5), insertion sort for K = 6
     * (:318-323).  ce(a, b): u[a] = max, u[b] = min (:303-309). */
#define CE(a, b) do { REAL x_ = u[a], y_ = u[b]; u[a] = FMAX(x_, y_); u[b] = FMIN(x_, y_); } while (0)
    switch (K) {
    case 2: CE(0, 1); break;
    case 3: CE(0, 2); CE(0, 1); CE(1, 2); break;
    case 4: CE(0, 2); CE(1, 3); CE(0, 1); CE(2, 3); CE(1, 2); break;
    case 5: CE(0, 3); CE(1, 4); CE(0, 2); CE(1, 3); CE(0, 1); CE(2, 4); CE(1, 2); CE(3, 4); CE(2, 3); break;
    default:
        for (int i = 1; i < K; ++i)
            for (int jj = 0; jj < i; ++jj) { int j = i - 1 - jj; CE(j, j + 1); }
    }
#undef CE
}

static inline void FN(proj_simplex_)(int K, REAL *q)
{
    REAL u[8];
    for (int i = 0; i < K; ++i) u[i] = q[i];
    FN(sort_desc)(K, u);
    REAL sum = 0, rho_sum = 0;
    int rho = 0;
    for (int i = 0; i < K; ++i) {            /* :330-338 */
        sum += u[i];
        REAL theta = (sum - (REAL)1) / (REAL)(i + 1);
        if (u[i] > theta) { rho = i + 1; rho_sum = sum; }
    }
    REAL theta = (rho_sum - (REAL)1) / (REAL)rho;   /* :340 */
    for (int i = 0; i < K; ++i) {            /* :342-344, std::max(a, 0) */
        REAL a = q[i] - theta;
        q[i] = (a < (REAL)0) ? (REAL)0 : a;
    }
}

void FN(or_proj_simplex)(int K, REAL *q) { FN(proj_simplex_)(K, q); }

/* ---- Dykstra projection onto K = {|q_i - q_j| <= 1}, :202-238 --------- */
static inline int FN(dykstra_)(int K, REAL (*q)[3], int max_sweeps, double tol)
{
    REAL r[15][3];
    int np = K * (K - 1) / 2, sweep;
    for (int p = 0; p < np; ++p) r[p][0] = r[p][1] = r[p][2] = 0;
    for (sweep = 0; sweep < max_sweeps; ++sweep) {
        REAL mx = -REAL_MAX;                  /* numeric_limits::lowest() */
        for (int i1 = 0; i1 < K - 1; ++i1) {
            for (int i2 = i1 + 1; i2 < K; ++i2) {
                int p = (2 * K - 1 - i1) * i1 / 2 + i2 - i1 - 1;   /* :209 */
                REAL h[3], e[3], d[3];
                for (int a = 0; a < 3; ++a) h[a] = (q[i2][a] - q[i1][a]) + r[p][a];
                REAL l = SQRT(FMA(h[2], h[2], FMA(h[1], h[1], h[0] * h[0])));
                if (l > (REAL)1) {
                    REAL s = (l - (REAL)1) / l;
                    for (int a = 0; a < 3; ++a) e[a] = s * h[a];
                } else {
                    e[0] = e[1] = e[2] = 0;
                }
                for (int a = 0; a < 3; ++a) d[a] = (REAL)0.5 * (e[a] - r[p][a]);
                REAL dn = SQRT(FMA(d[2], d[2], FMA(d[1], d[1], d[0] * d[0])));
                mx = (dn < mx) ? mx : dn;     /* cuda::std::max(dn, mx) */
                for (int a = 0; a < 3; ++a) { q[i1][a] += d[a]; q[i2][a] -= d[a]; }
                for (int a = 0; a < 3; ++a) r[p][a] = e[a];
            }
        }
        if (tol > (double)mx) { ++sweep; break; }
    }
    return sweep;
}

int FN(or_dykstra)(int K, REAL *q, int max_sweeps, double tol)
{
    return FN(dykstra_)(K, (REAL (*)[3])q, max_sweeps, tol);
}

/* ---- Potts solver ------------------------------------------------------ */

/* Per-solve geometry of one slab in the work layout. */
typedef struct {
    const or_layout *m;
    uint32_t s0, s1;     /* slab [s0, s1) along n */
    uint32_t wstride;    /* work-buffer line stride */
    size_t lines, wspan;
} FN(slab_t);

/* potts-min-partition.cpp:96-247 for one voxel (n, row, col). */
static inline void FN(potts_xi_voxel)(const FN(slab_t) *S, int K, uint32_t n, uint32_t row,
                                      uint32_t col, REAL tau2, REAL *xi, const REAL *vbar)
{
    const or_layout *m = S->m;
    size_t L = line_of(m, row, col);
    size_t w = L * S->wstride + (n - S->s0);
    size_t ws = S->wspan;
    int nx[3];
    size_t wn[3];
    nx[0] = n + 1 < S->s1;
    nx[1] = has_next_row(m, row, col);
    nx[2] = has_next_col(m, row, col);
    wn[0] = w + 1;
    wn[1] = w + S->wstride;
    wn[2] = nx[2] ? line_of(m, row, col + 1) * S->wstride + (n - S->s0) : 0;

    REAL vb[8], vbn[8][3], p[8][3], q[8][3], xo[8][3];
    REAL vk = 1;
    for (int i = 0; i < K - 1; ++i) { vb[i] = vbar[i * ws + w]; vk = vk - vb[i]; }
    vb[K - 1] = vk;
    for (int a = 0; a < 3; ++a) {
        if (nx[a]) {
            REAL vka = 1;
            for (int i = 0; i < K - 1; ++i) { vbn[i][a] = vbar[i * ws + wn[a]]; vka = vka - vbn[i][a]; }
            vbn[K - 1][a] = vka;
            for (int i = 0; i < K; ++i) p[i][a] = vbn[i][a] - vb[i];
        } else {
            for (int i = 0; i < K; ++i) p[i][a] = 0;
        }
    }
    for (int a = 0; a < 3; ++a) {
        REAL xk = 0;
        for (int i = 0; i < K - 1; ++i) { xo[i][a] = xi[(3 * i + a) * ws + w]; xk = xk - xo[i][a]; }
        xo[K - 1][a] = xk;
        for (int i = 0; i < K; ++i) q[i][a] = FMA(p[i][a], tau2, xo[i][a]);   /* :193-200 */
    }
    FN(dykstra_)(K, q, 20, 1e-6);
    for (int i = 0; i < K - 1; ++i)
        for (int a = 0; a < 3; ++a) xi[(3 * i + a) * ws + w] = q[i][a];       /* :239-243 */
}

/* potts-min-partition.cpp:349-488 for one voxel; returns sum_k (q_k - v_k)^2. */
static inline REAL FN(potts_v_voxel)(const FN(slab_t) *S, int K, uint32_t n, uint32_t row,
                                     uint32_t col, const float *u, const REAL *cc, REAL lambda,
                                     REAL tau1, const REAL *xi, REAL *V, REAL *vbar, int storage)
{
    const or_layout *m = S->m;
    size_t L = line_of(m, row, col);
    size_t w = L * S->wstride + (n - S->s0);
    size_t ws = S->wspan;
    int px[3], nx[3];
    size_t wp[3];
    px[0] = n > S->s0;
    px[1] = has_prev_row(m, row, col);
    px[2] = has_prev_col(m, row, col);
    nx[0] = n + 1 < S->s1;
    nx[1] = has_next_row(m, row, col);
    nx[2] = has_next_col(m, row, col);
    wp[0] = w - 1;
    wp[1] = w - S->wstride;
    wp[2] = px[2] ? line_of(m, row, col - 1) * S->wstride + (n - S->s0) : 0;

    REAL xo[8][3], xp[8][3], v[8], q[8];
    for (int a = 0; a < 3; ++a) {
        REAL xk = 0;
        for (int i = 0; i < K - 1; ++i) { xo[i][a] = xi[(3 * i + a) * ws + w]; xk = xk - xo[i][a]; }
        xo[K - 1][a] = xk;
        if (px[a]) {
            REAL xkp = 0;
            for (int i = 0; i < K - 1; ++i) { xp[i][a] = xi[(3 * i + a) * ws + wp[a]]; xkp = xkp - xp[i][a]; }
            xp[K - 1][a] = xkp;
        }
    }
    REAL vk = 1;
    for (int i = 0; i < K - 1; ++i) { v[i] = V[i * ws + w]; vk = vk - v[i]; }
    v[K - 1] = vk;
    REAL uu = (REAL)u[L * m->nstride + n];
    for (int i = 0; i < K; ++i) {
        REAL div = 0;                                   /* :425-457 */
        if (px[0]) div -= xp[i][0];
        if (nx[0]) div += xo[i][0];
        if (px[1]) div -= xp[i][1];
        if (nx[1]) div += xo[i][1];
        if (px[2]) div -= xp[i][2];
        if (nx[2]) div += xo[i][2];
        REAL d = uu - cc[i];
        q[i] = FMA(tau1, FMA(-lambda, d * d, div), v[i]);  /* :467 */
    }
    FN(proj_simplex_)(K, q);
    REAL dq = 0;
    for (int i = 0; i < K - 1; ++i) {
        vbar[i * ws + w] = STORE_ROUND((REAL)2 * q[i] - v[i], storage);   /* :471-472 */
        V[i * ws + w] = q[i];
    }
    for (int k = 0; k < K; ++k) { REAL d = q[k] - v[k]; dq = FMA(d, d, dq); }  /* :474-480 */
    return dq;
}

int FN(or_potts)(const or_layout *m, const float *u, int K, const float *c,
                 float lambda, float tau1, float tau2, int iters, int slab_w,
                 float *res, REAL *V_out, REAL *xi_out, REAL *dq, int storage, int nthreads)
{
    if (K < 2 || K > 6 || iters < 0 || storage < 0 || storage > 2) return 2;
    const uint32_t n = m->n_images;
    const int halo = iters;
    int single = (slab_w <= 0) || ((uint32_t)slab_w >= n);
    int sw = single ? (int)n : slab_w;
    int res_w = sw - 2 * halo;
    if (!single && res_w <= 0) return 1;                /* :571-577 */
    int n_slabs = single ? 1 : 1 + ((int)n - sw + res_w - 1) / res_w;   /* :627 */
    size_t lines = (size_t)m->base[m->width] / m->nstride;
    uint32_t wstride = stride32((uint32_t)sw);
    size_t wspan = lines * wstride;
    REAL *xi = (REAL *)calloc((size_t)3 * (K - 1) * wspan, sizeof(REAL));
    REAL *V = (REAL *)calloc((size_t)(K - 1) * wspan, sizeof(REAL));
    REAL *vbar = (REAL *)calloc((size_t)(K - 1) * wspan, sizeof(REAL));
    if (!xi || !V || !vbar) { free(xi); free(V); free(vbar); return 2; }
    REAL cc[8];
    for (int i = 0; i < K; ++i) cc[i] = (REAL)c[i];
    if (nthreads > 0) omp_set_num_threads(nthreads);

    for (int ks = 0; ks < n_slabs; ++ks) {
        FN(slab_t) S;
        S.m = m;
        S.s0 = (uint32_t)(ks * res_w);
        S.s1 = (uint32_t)((int)n < (int)S.s0 + sw ? (int)n : (int)S.s0 + sw);
        S.wstride = wstride; S.lines = lines; S.wspan = wspan;

        /* init_v_kernel, :56-94 */
#pragma omp parallel for schedule(dynamic, 16)
        for (uint32_t col = m->col_begin; col < m->col_end; ++col) {
            for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row) {
                size_t L = line_of(m, row, col);
                for (uint32_t nn = S.s0; nn < S.s1; ++nn) {
                    REAL x = (REAL)u[L * m->nstride + nn];
                    REAL best = REAL_MAX;
                    int cl = 0;
                    for (int k = 0; k < K; ++k) {
                        REAL dd = FABS(cc[k] - x);
                        if (dd < best) { best = dd; cl = k; }
                    }
                    size_t w = L * wstride + (nn - S.s0);
                    for (int i = 0; i < K - 1; ++i) V[i * wspan + w] = vbar[i * wspan + w] = (i == cl) ? 1 : 0;
                }
            }
        }
        memset(xi, 0, (size_t)3 * (K - 1) * wspan * sizeof(REAL));   /* :644 */

        for (int it = 0; it < iters; ++it) {                          /* :645-660 */
#pragma omp parallel for schedule(dynamic, 16)
            for (uint32_t col = m->col_begin; col < m->col_end; ++col)
                for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
                    for (uint32_t nn = S.s0; nn < S.s1; ++nn)
                        FN(potts_xi_voxel)(&S, K, nn, row, col, (REAL)tau2, xi, vbar);
            /* Reduced-precision storage: Xi+ is rounded where it is stored, and
             * the V step below reads the stored values. */
            if (storage) {
#pragma omp parallel for schedule(static)
                for (size_t k = 0; k < (size_t)3 * (K - 1) * wspan; ++k) xi[k] = STORE_ROUND(xi[k], storage);
            }
            REAL mx = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(max : mx)
            for (uint32_t col = m->col_begin; col < m->col_end; ++col)
                for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
                    for (uint32_t nn = S.s0; nn < S.s1; ++nn) {
                        REAL d = FN(potts_v_voxel)(&S, K, nn, row, col, u, cc, (REAL)lambda,
                                                   (REAL)tau1, xi, V, vbar, storage);
                        if (d > mx) mx = d;
                    }
            if (dq) dq[it] = mx;
        }
        /* update_kernel over the slab interior, :491-538 and :663-674 */
        uint32_t r0 = (S.s0 == 0) ? S.s0 : S.s0 + (uint32_t)halo;
        uint32_t r1 = (S.s1 == n) ? S.s1 : S.s1 - (uint32_t)halo;
#pragma omp parallel for schedule(dynamic, 16)
        for (uint32_t col = m->col_begin; col < m->col_end; ++col) {
            for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row) {
                size_t L = line_of(m, row, col);
                for (uint32_t nn = r0; nn < r1; ++nn) {
                    size_t w = L * wstride + (nn - S.s0);
                    REAL best = 0, vk = 1, vi;
                    int lab = 0;
                    for (int i = 0; i < K; ++i) {
                        if (i < K - 1) { vi = V[i * wspan + w]; vk = vk - vi; } else vi = vk;
                        if (vi > best) { best = vi; lab = i; }
                    }
                    res[L * m->nstride + nn] = c[lab];
                }
            }
        }
    }
    if (single) {
        /* Export the state in the input layout (work stride may differ). */
        for (size_t L = 0; L < lines; ++L)
            for (uint32_t nn = 0; nn < n; ++nn) {
                size_t w = L * wstride + nn, o = L * m->nstride + nn;
                size_t span = (size_t)m->base[m->width];
                if (V_out)
                    for (int i = 0; i < K - 1; ++i) V_out[i * span + o] = V[i * wspan + w];
                if (xi_out)
                    for (int j = 0; j < 3 * (K - 1); ++j) xi_out[j * span + o] = xi[j * wspan + w];
            }
    }
    free(xi); free(V); free(vbar);
    return 0;
}

/* ---- Chambolle ROF, chambolle.cpp:33-276 ------------------------------ */

static inline REAL FN(tv_div)(const or_layout *m, uint32_t n, uint32_t row, uint32_t col,
                              const REAL *x0, const REAL *x1, const REAL *x2)
{
    size_t o = line_of(m, row, col) * m->nstride + n;
    REAL d = 0;                                        /* :67-78 */
    if (n > 0) d -= x0[o - 1];
    if (n + 1 < m->n_images) d += x0[o];
    if (has_prev_row(m, row, col)) d -= x1[o - m->nstride];
    if (has_next_row(m, row, col)) d += x1[o];
    if (has_prev_col(m, row, col)) d -= x2[line_of(m, row, col - 1) * m->nstride + n];
    if (has_next_col(m, row, col)) d += x2[o];
    return d;
}

int FN(or_chambolle)(const or_layout *m, REAL *g, float lambda, float tau, int iters,
                     REAL *xi_out, REAL *dq, int nthreads)
{
    size_t span = (size_t)m->base[m->width];
    REAL *x0 = (REAL *)calloc(3 * span, sizeof(REAL));
    REAL *v = (REAL *)calloc(span, sizeof(REAL));
    if (!x0 || !v) { free(x0); free(v); return 2; }
    REAL *x1 = x0 + span, *x2 = x1 + span;
    const REAL lam = (REAL)lambda, ta = (REAL)tau;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    for (int it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(dynamic, 16)
        for (uint32_t col = m->col_begin; col < m->col_end; ++col)
            for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
                for (uint32_t n = 0; n < m->n_images; ++n)
                    v[line_of(m, row, col) * m->nstride + n] = FN(tv_div)(m, n, row, col, x0, x1, x2);
        REAL mx = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(max : mx)
        for (uint32_t col = m->col_begin; col < m->col_end; ++col)
            for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row) {
                size_t L = line_of(m, row, col);
                int nr = has_next_row(m, row, col), nc = has_next_col(m, row, col);
                size_t oc = nc ? line_of(m, row, col + 1) * m->nstride : 0;
                for (uint32_t n = 0; n < m->n_images; ++n) {
                    size_t o = L * m->nstride + n;
                    REAL g1 = 0, g2 = 0, g3 = 0;               /* :129-145 */
                    if (n + 1 < m->n_images) g1 = (v[o + 1] - v[o]) - (g[o + 1] - g[o]) / lam;
                    if (nr) g2 = (v[o + m->nstride] - v[o]) - (g[o + m->nstride] - g[o]) / lam;
                    if (nc) g3 = (v[oc + n] - v[o]) - (g[oc + n] - g[o]) / lam;
                    REAL den = FMA(ta, SQRT(FMA(g3, g3, FMA(g2, g2, g1 * g1))), (REAL)1);  /* :147 */
                    REAL q1 = FMA(ta, g1, x0[o]) / den;        /* :152-154 */
                    REAL q2 = FMA(ta, g2, x1[o]) / den;
                    REAL q3 = FMA(ta, g3, x2[o]) / den;
                    REAL d1 = x0[o] - q1, d2 = x1[o] - q2, d3 = x2[o] - q3;
                    REAL d = FMA(d3, d3, FMA(d2, d2, d1 * d1)); /* :156 */
                    if (d > mx) mx = d;
                    x0[o] = q1; x1[o] = q2; x2[o] = q3;
                }
            }
        if (dq) dq[it] = mx;
    }
    /* update_kernel :171-223: g = g - lambda * div xi */
#pragma omp parallel for schedule(dynamic, 16)
    for (uint32_t col = m->col_begin; col < m->col_end; ++col)
        for (uint32_t row = row_begin(m, col); row < row_end(m, col); ++row)
            for (uint32_t n = 0; n < m->n_images; ++n) {
                size_t o = line_of(m, row, col) * m->nstride + n;
                g[o] = FMA(-lam, FN(tv_div)(m, n, row, col, x0, x1, x2), g[o]);
            }
    if (xi_out) memcpy(xi_out, x0, 3 * span * sizeof(REAL));
    free(x0); free(v);
    return 0;
}

#undef FN
#undef CAT
#undef CAT_
