// chambolle.hip -- Chambolle (2004) dual projection for ROF / TV denoising
// (totalvarmin/chambolle.cpp).  g is overwritten with g - lambda div xi.
#include "sct_internal.hpp"

#include <algorithm>

namespace {

struct TvGeom {
    const LineNbr *nbr;
    uint32_t lines, len, chunks, stride;
    uint64_t span;
};

__device__ __forceinline__ bool tv_voxel(const TvGeom &g, uint32_t &line, uint32_t &n)
{
    const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    line = uint32_t(gw / g.chunks);
    n = uint32_t(gw % g.chunks) * kWave + (threadIdx.x % kWave);
    return line < g.lines && n < g.len;
}

// div xi in the order of chambolle.cpp:67-78.
__device__ __forceinline__ float tv_div(const TvGeom &g, const LineNbr &nb, uint32_t line, uint32_t n,
                                        const float *__restrict__ x0, const float *__restrict__ x1,
                                        const float *__restrict__ x2)
{
    const uint64_t o = uint64_t(line) * g.stride + n;
    float d = 0.0f;
    if (n > 0) d -= x0[o - 1];
    if (n + 1 < g.len) d += x0[o];
    if (nb.prow >= 0) d -= x1[uint64_t(nb.prow) * g.stride + n];
    if (nb.nrow >= 0) d += x1[o];
    if (nb.pcol >= 0) d -= x2[uint64_t(nb.pcol) * g.stride + n];
    if (nb.ncol >= 0) d += x2[o];
    return d;
}

// One Chambolle iteration in one pass (div_kernel then grad_kernel,
// chambolle.cpp:33-169).  The divergence of the previous xi is evaluated where
// the gradient needs it -- at the voxel, at n + 1 (the next lane, or for the
// last lane of a chunk directly), at the +row and at the +col neighbour -- so
// no divergence field is stored, and the new xi goes to a second buffer: per
// voxel and iteration, xi is read and written once and g read once (28 B
// instead of the three-kernel 48 B).  Every divergence is the same sequence of
// operations as chambolle.cpp:67-78 (tv_div), so the result is bitwise that of
// the reference's two kernels.  Each lane owns VPL consecutive voxels of a
// chunk of kWave * VPL, loaded and stored as one vector: VPL = 4 (2) needs
// nstride to be a multiple of 4 (2), so that the vectors are aligned and stay
// inside the line's storage.  The waves walk the lines in band order, every
// XCD one contiguous eighth of it, so the neighbour lines come from the L2
// that fetched them; the new xi is stored non-temporally (it is next read a
// whole sweep later).  The kernel below walks runs of rows (tv_step_kernel).
struct TvStep {
    const OrderedLine *ol;                     // the lines in band order
    uint32_t lines, len, chunks, stride, xg;   // chunks per line; xg: workgroups per XCD run
    const float *x0, *x1, *x2;                 // xi at the start of the iteration
    float *y0, *y1, *y2;                       // xi after it
    const float *g;
    float lambda, tau;
};

template <int N>
__device__ __forceinline__ void tv_ld(const float *p, float (&o)[N])
{
    if constexpr (N == 1) {
        o[0] = *p;
    } else {
        typedef float V __attribute__((ext_vector_type(N)));
        const V v = *reinterpret_cast<const V *>(p);
#pragma unroll
        for (int j = 0; j < N; ++j) o[j] = v[j];
    }
}

template <int N>
__device__ __forceinline__ void tv_st(float *p, const float (&x)[N], int nv)
{
    if constexpr (N == 1) {
        p[0] = x[0];
    } else {
        typedef float V __attribute__((ext_vector_type(N)));
        if (nv == N) {
            V v;
#pragma unroll
            for (int j = 0; j < N; ++j) v[j] = x[j];
            __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
        } else {
#pragma unroll
            for (int j = 0; j < N; ++j)
                if (j < nv) p[j] = x[j];
        }
    }
}

// div xi at one voxel from its six candidate terms, chambolle.cpp:67-78.
__device__ __forceinline__ float tv_div6(bool hpn, float x0m, bool hnn, float x0, bool hpr, float x1p, bool hnr,
                                         float x1, bool hpc, float x2p, bool hnc, float x2)
{
    float d = 0.0f;
    if (hpn) d -= x0m;
    if (hnn) d += x0;
    if (hpr) d -= x1p;
    if (hnr) d += x1;
    if (hpc) d -= x2p;
    if (hnc) d += x2;
    return d;
}

// Each wave walks a run of consecutive rows of one column (up to kMarchRows
// lines of the band order) for one chunk of n.  Going from a line to the next
// row reuses what the previous step loaded: the +row line's xi and g become
// the line's own, its divergence d_r becomes d_o, and the +col line's xi_row
// becomes the -row term of the next +col divergence, so a step loads only the
// +row and +col lines (9 rows of lanes; 16 without the reuse).
template <int VPL, bool TRACE>
__global__ __launch_bounds__(256) void tv_step_kernel(TvStep a, const uint2 *__restrict__ runs, uint32_t n_runs,
                                                      float *dq)
{
    const int lane = threadIdx.x % kWave;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x, a.xg);
    const uint64_t gw = (uint64_t(b) * blockDim.x + threadIdx.x) / kWave;
    float dd = 0.0f;
    if (gw < uint64_t(n_runs) * a.chunks) {   // uniform over the wave
        const uint2 run = runs[uint32_t(gw / a.chunks)];
        const uint32_t n0 = uint32_t(gw % a.chunks) * (kWave * VPL) + lane * VPL;
        const uint32_t len = a.len;
        const int nv = n0 < len ? int(min(uint32_t(VPL), len - n0)) : 0;
        const auto at = [&](int32_t l) { return uint64_t(l) * a.stride; };
        OrderedLine L = a.ol[run.x];
        uint64_t lo = at(L.line);

        // carried from line to line: xi and g of the line, xi_row of the
        // -row line of the +col line, div xi of the line
        float x0o[VPL], x1o[VPL], x2o[VPL], go[VPL], x1cpr[VPL], d_o[VPL];
        {
            float x1pr[VPL], x2pc[VPL];
#pragma unroll
            for (int j = 0; j < VPL; ++j) x0o[j] = x1o[j] = x2o[j] = go[j] = x1cpr[j] = x1pr[j] = x2pc[j] = 0.0f;
            if (nv) {
                tv_ld(a.x0 + lo + n0, x0o);
                tv_ld(a.x1 + lo + n0, x1o);
                tv_ld(a.x2 + lo + n0, x2o);
                tv_ld(a.g + lo + n0, go);
                if (L.prow >= 0) tv_ld(a.x1 + at(L.prow) + n0, x1pr);
                if (L.pcol >= 0) tv_ld(a.x2 + at(L.pcol) + n0, x2pc);
                if (L.ncol_prow >= 0) tv_ld(a.x1 + at(L.ncol_prow) + n0, x1cpr);
            }
            float x0om = __shfl_up(x0o[VPL - 1], 1, kWave);
            if (lane == 0 && nv && n0 > 0) x0om = a.x0[lo + n0 - 1];
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                const uint32_t n = n0 + j;
                d_o[j] = tv_div6(n > 0, j ? x0o[j - 1] : x0om, n + 1 < len, x0o[j], L.prow >= 0, x1pr[j], L.nrow >= 0,
                                 x1o[j], L.pcol >= 0, x2pc[j], L.ncol >= 0, x2o[j]);
            }
        }
        for (uint32_t k = 0; k < run.y; ++k) {
            const bool r_nrow = L.flags & kNrowHasNrow, r_ncol = L.flags & kNrowHasNcol;
            const bool c_nrow = L.flags & kNcolHasNrow, c_ncol = L.flags & kNcolHasNcol;
            float x0r[VPL], x1r[VPL], x2r[VPL], gr[VPL], x2rpc[VPL], x0c[VPL], x1c[VPL], x2c[VPL], gc[VPL];
#pragma unroll
            for (int j = 0; j < VPL; ++j)
                x0r[j] = x1r[j] = x2r[j] = gr[j] = x2rpc[j] = x0c[j] = x1c[j] = x2c[j] = gc[j] = 0.0f;
            if (nv) {
                if (L.nrow >= 0) {
                    const uint64_t lr = at(L.nrow);
                    tv_ld(a.x0 + lr + n0, x0r);
                    tv_ld(a.g + lr + n0, gr);
                    if (r_nrow) tv_ld(a.x1 + lr + n0, x1r);
                    if (r_ncol) tv_ld(a.x2 + lr + n0, x2r);
                    if (L.nrow_pcol >= 0) tv_ld(a.x2 + at(L.nrow_pcol) + n0, x2rpc);
                }
                if (L.ncol >= 0) {
                    const uint64_t lc = at(L.ncol);
                    tv_ld(a.x0 + lc + n0, x0c);
                    tv_ld(a.g + lc + n0, gc);
                    if (c_nrow) tv_ld(a.x1 + lc + n0, x1c);
                    if (c_ncol) tv_ld(a.x2 + lc + n0, x2c);
                }
            }
            float x0rm = __shfl_up(x0r[VPL - 1], 1, kWave);
            float x0cm = __shfl_up(x0c[VPL - 1], 1, kWave);
            if (lane == 0 && nv && n0 > 0) {
                if (L.nrow >= 0) x0rm = a.x0[at(L.nrow) + n0 - 1];
                if (L.ncol >= 0) x0cm = a.x0[at(L.ncol) + n0 - 1];
            }
            float d_r[VPL], d_c[VPL];
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                const uint32_t n = n0 + j;
                const bool hpn = n > 0, hnn = n + 1 < len;
                d_r[j] = tv_div6(hpn, j ? x0r[j - 1] : x0rm, hnn, x0r[j], true, x1o[j], r_nrow, x1r[j],
                                 L.nrow_pcol >= 0, x2rpc[j], r_ncol, x2r[j]);
                d_c[j] = tv_div6(hpn, j ? x0c[j - 1] : x0cm, hnn, x0c[j], L.ncol_prow >= 0, x1cpr[j], c_nrow, x1c[j],
                                 true, x2o[j], c_ncol, x2c[j]);
            }
            float d_next = __shfl_down(d_o[0], 1, kWave);
            float g_next = __shfl_down(go[0], 1, kWave);
            if (lane == kWave - 1 && nv == VPL && n0 + VPL < len) {
                const uint64_t m = lo + n0 + VPL;
                const float x1p = L.prow >= 0 ? a.x1[at(L.prow) + n0 + VPL] : 0.0f;
                const float x2p = L.pcol >= 0 ? a.x2[at(L.pcol) + n0 + VPL] : 0.0f;
                d_next = tv_div6(true, x0o[VPL - 1], n0 + VPL + 1 < len, a.x0[m], L.prow >= 0, x1p, L.nrow >= 0,
                                 a.x1[m], L.pcol >= 0, x2p, L.ncol >= 0, a.x2[m]);
                g_next = a.g[m];
            }
            float y0[VPL], y1[VPL], y2[VPL];
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                const uint32_t n = n0 + j;
                const float dn = j + 1 < VPL ? d_o[j + 1] : d_next, gn = j + 1 < VPL ? go[j + 1] : g_next;
                float g1 = 0.0f, g2 = 0.0f, g3 = 0.0f;
                if (n + 1 < len) g1 = (dn - d_o[j]) - (gn - go[j]) / a.lambda;
                if (L.nrow >= 0) g2 = (d_r[j] - d_o[j]) - (gr[j] - go[j]) / a.lambda;
                if (L.ncol >= 0) g3 = (d_c[j] - d_o[j]) - (gc[j] - go[j]) / a.lambda;
                const float den = __fmaf_rn(a.tau, sqrtf(__fmaf_rn(g3, g3, __fmaf_rn(g2, g2, g1 * g1))), 1.0f);
                y0[j] = __fmaf_rn(a.tau, g1, x0o[j]) / den;
                y1[j] = __fmaf_rn(a.tau, g2, x1o[j]) / den;
                y2[j] = __fmaf_rn(a.tau, g3, x2o[j]) / den;
                if (TRACE && j < nv) {
                    const float e1 = x0o[j] - y0[j], e2 = x1o[j] - y1[j], e3 = x2o[j] - y2[j];
                    dd = fmaxf(dd, __fmaf_rn(e3, e3, __fmaf_rn(e2, e2, e1 * e1)));
                }
            }
            if (nv) {
                tv_st(a.y0 + lo + n0, y0, nv);
                tv_st(a.y1 + lo + n0, y1, nv);
                tv_st(a.y2 + lo + n0, y2, nv);
            }
            if (k + 1 < run.y) {   // the next line of the run is this line's +row line
                L = a.ol[run.x + k + 1];
                lo = at(L.line);
#pragma unroll
                for (int j = 0; j < VPL; ++j) {
                    x0o[j] = x0r[j];
                    x1o[j] = x1r[j];
                    x2o[j] = x2r[j];
                    go[j] = gr[j];
                    x1cpr[j] = x1c[j];
                    d_o[j] = d_r[j];
                }
            }
        }
    }
    if constexpr (TRACE) {
        __shared__ float red[4];
        dd = wave_max(dd);
        if ((threadIdx.x % kWave) == 0) red[threadIdx.x / kWave] = dd;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = red[0];
            for (int i = 1; i < int(blockDim.x / kWave); ++i) m = fmaxf(m, red[i]);
            atomic_max_nonneg(dq, m);
        }
    }
}

// update_kernel, chambolle.cpp:171-223: g = g - lambda * div xi.
__global__ __launch_bounds__(256) void tv_update_kernel(TvGeom g, float *__restrict__ gio, const float *__restrict__ xi,
                                                        float lambda)
{
    uint32_t line, n;
    if (!tv_voxel(g, line, n)) return;
    const LineNbr nb = g.nbr[line];
    const uint64_t o = uint64_t(line) * g.stride + n;
    gio[o] = __fmaf_rn(-lambda, tv_div(g, nb, line, n, xi, xi + g.span, xi + 2 * g.span), gio[o]);
}

}  // namespace

template <int VPL>
static void launch_tv_step(TvStep a, const Geometry &geo, float *dq, hipStream_t st)
{
    const uint64_t waves = uint64_t(geo.n_runs) * a.chunks;
    const uint32_t grid = uint32_t((waves + 3) / 4);
    a.xg = std::max(1u, grid / kXcdCount);   // one contiguous eighth of the band order per XCD
    if (dq)
        tv_step_kernel<VPL, true><<<grid, 256, 0, st>>>(a, geo.d_runs, geo.n_runs, dq);
    else
        tv_step_kernel<VPL, false><<<grid, 256, 0, st>>>(a, geo.d_runs, geo.n_runs, nullptr);
}

// chambolle(), chambolle.cpp:225-276.  xi starts at 0 in both buffers; the
// iterations alternate between them and the last one written feeds the final
// update.  With dq_trace, entry i receives max over voxels of |xi_i+1 - xi_i|^2.
extern "C" int sct_chambolle(sct_ctx *ctx, sct_field *gf, float lambda, float tau, int iters, float *dq_trace)
{
    if (!ctx || !gf) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (gf->dtype != SCT_F32) return sct_fail(ctx, SCT_ERR_INVALID, "chambolle needs a float field");
    if (iters < 0) return sct_fail(ctx, SCT_ERR_INVALID, "iters must be >= 0");
    const Geometry &geo = gf->geo;
    TvGeom g;
    g.nbr = geo.d_nbr;
    g.lines = geo.lines;
    g.len = geo.n_images;
    g.chunks = (g.len + kWave - 1) / kWave;
    g.stride = geo.nstride;
    g.span = geo.span();
    if (!g.lines || !g.len) return SCT_OK;
    // voxels per lane: the widest of 4, 2, 1 that divides nstride
    const int vpl = g.stride % 4 == 0 ? 4 : g.stride % 2 == 0 ? 2 : 1;
    float *xi = nullptr, *d_dq = nullptr;
    int rc = sct_alloc(ctx, (void **)&xi, sizeof(float) * 6 * g.span);
    if (!rc && dq_trace) rc = sct_alloc(ctx, (void **)&d_dq, sizeof(float) * size_t(std::max(iters, 1)));
    auto cleanup = [&] {
        (void)hipStreamSynchronize(ctx->stream);
        sct_free(ctx, xi);
        sct_free(ctx, d_dq);
    };
    if (rc) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemsetAsync(xi, 0, sizeof(float) * 6 * g.span, ctx->stream);
    if (e == hipSuccess && d_dq) e = hipMemsetAsync(d_dq, 0, sizeof(float) * size_t(std::max(iters, 1)), ctx->stream);
    float *gp = (float *)gf->data;
    float *buf[2] = {xi, xi + 3 * g.span};
    TvStep s;
    s.ol = geo.d_oline;
    s.lines = g.lines;
    s.len = g.len;
    s.chunks = (g.len + kWave * vpl - 1) / (kWave * vpl);
    s.stride = g.stride;
    s.g = gp;
    s.lambda = lambda;
    s.tau = tau;
    for (int it = 0; e == hipSuccess && it < iters; ++it) {
        const float *x = buf[it & 1];
        float *y = buf[(it + 1) & 1];
        s.x0 = x;
        s.x1 = x + g.span;
        s.x2 = x + 2 * g.span;
        s.y0 = y;
        s.y1 = y + g.span;
        s.y2 = y + 2 * g.span;
        float *dq = d_dq ? d_dq + it : nullptr;
        if (vpl == 4)
            launch_tv_step<4>(s, geo, dq, ctx->stream);
        else if (vpl == 2)
            launch_tv_step<2>(s, geo, dq, ctx->stream);
        else
            launch_tv_step<1>(s, geo, dq, ctx->stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        const uint32_t grid = uint32_t((uint64_t(g.lines) * g.chunks + 3) / 4);
        tv_update_kernel<<<grid, 256, 0, ctx->stream>>>(g, gp, buf[iters & 1], lambda);
        e = hipGetLastError();
    }
    if (e == hipSuccess && d_dq)
        e = hipMemcpyAsync(dq_trace, d_dq, sizeof(float) * size_t(iters), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "chambolle");
    return SCT_OK;
}
