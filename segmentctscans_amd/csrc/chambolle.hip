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

// div_kernel, chambolle.cpp:33-84.
__global__ __launch_bounds__(256) void tv_div_kernel(TvGeom g, const float *__restrict__ xi, float *__restrict__ v)
{
    uint32_t line, n;
    if (!tv_voxel(g, line, n)) return;
    const LineNbr nb = g.nbr[line];
    v[uint64_t(line) * g.stride + n] = tv_div(g, nb, line, n, xi, xi + g.span, xi + 2 * g.span);
}

// grad_kernel, chambolle.cpp:86-169.
template <bool TRACE>
__global__ __launch_bounds__(256) void tv_grad_kernel(TvGeom g, const float *__restrict__ gin, float *__restrict__ xi,
                                                      const float *__restrict__ v, float lambda, float tau, float *dq)
{
    uint32_t line, n;
    float dd = 0.0f;
    if (tv_voxel(g, line, n)) {
        const LineNbr nb = g.nbr[line];
        const uint64_t o = uint64_t(line) * g.stride + n;
        const float vo = v[o], go = gin[o];
        float g1 = 0.0f, g2 = 0.0f, g3 = 0.0f;
        if (n + 1 < g.len) g1 = (v[o + 1] - vo) - (gin[o + 1] - go) / lambda;
        if (nb.nrow >= 0) {
            const uint64_t on = uint64_t(nb.nrow) * g.stride + n;
            g2 = (v[on] - vo) - (gin[on] - go) / lambda;
        }
        if (nb.ncol >= 0) {
            const uint64_t on = uint64_t(nb.ncol) * g.stride + n;
            g3 = (v[on] - vo) - (gin[on] - go) / lambda;
        }
        const float den = __fmaf_rn(tau, sqrtf(__fmaf_rn(g3, g3, __fmaf_rn(g2, g2, g1 * g1))), 1.0f);
        float *x0 = xi, *x1 = xi + g.span, *x2 = xi + 2 * g.span;
        const float q1 = x0[o], q2 = x1[o], q3 = x2[o];
        const float n1 = __fmaf_rn(tau, g1, q1) / den;
        const float n2 = __fmaf_rn(tau, g2, q2) / den;
        const float n3 = __fmaf_rn(tau, g3, q3) / den;
        const float d1 = q1 - n1, d2 = q2 - n2, d3 = q3 - n3;
        dd = __fmaf_rn(d3, d3, __fmaf_rn(d2, d2, d1 * d1));
        x0[o] = n1;
        x1[o] = n2;
        x2[o] = n3;
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

// chambolle(), chambolle.cpp:225-276.
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
    float *xi = nullptr, *v = nullptr, *d_dq = nullptr;
    int rc = sct_alloc(ctx, (void **)&xi, sizeof(float) * 3 * g.span);
    if (!rc) rc = sct_alloc(ctx, (void **)&v, sizeof(float) * g.span);
    if (!rc && dq_trace) rc = sct_alloc(ctx, (void **)&d_dq, sizeof(float) * size_t(std::max(iters, 1)));
    auto cleanup = [&] {
        (void)hipStreamSynchronize(ctx->stream);
        sct_free(ctx, xi);
        sct_free(ctx, v);
        sct_free(ctx, d_dq);
    };
    if (rc) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemsetAsync(xi, 0, sizeof(float) * 3 * g.span, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(v, 0, sizeof(float) * g.span, ctx->stream);
    if (e == hipSuccess && d_dq) e = hipMemsetAsync(d_dq, 0, sizeof(float) * size_t(std::max(iters, 1)), ctx->stream);
    const uint32_t grid = uint32_t((uint64_t(g.lines) * g.chunks + 3) / 4);
    float *gp = (float *)gf->data;
    for (int it = 0; e == hipSuccess && it < iters; ++it) {
        tv_div_kernel<<<grid, 256, 0, ctx->stream>>>(g, xi, v);
        if (d_dq)
            tv_grad_kernel<true><<<grid, 256, 0, ctx->stream>>>(g, gp, xi, v, lambda, tau, d_dq + it);
        else
            tv_grad_kernel<false><<<grid, 256, 0, ctx->stream>>>(g, gp, xi, v, lambda, tau, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        tv_update_kernel<<<grid, 256, 0, ctx->stream>>>(g, gp, xi, lambda);
        e = hipGetLastError();
    }
    if (e == hipSuccess && d_dq)
        e = hipMemcpyAsync(dq_trace, d_dq, sizeof(float) * size_t(iters), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "chambolle");
    return SCT_OK;
}
