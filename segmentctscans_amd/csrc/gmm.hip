// gmm.hip -- the sample passes of gmm-peaks' one-dimensional Gaussian mixture
// fit (gmm-peaks/gmm.h:145-226): the fused E+M pass and the hard assignment,
// over samples resident on the device.  The EM iteration around them is host
// logic (segmentctscans_amd/gmm.py).
//
// Component k of the mixture has weight pi_k, mean mu_k and variance var_k;
// its density at x is pi_k (2 pi)^(-1/2) exp(-(x - mu_k)^2 / (2 var_k)) /
// sqrt(|var_k|) (gmm.h:80, :160-163, with the 1 x 1 covariance's inverse
// 1 / var_k and determinant var_k).
#include "sct_internal.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

struct sct_samples {
    sct_ctx *ctx;
    double *x;
    uint64_t n;
};

namespace {

constexpr int kGmmMax = 16;       // components
constexpr int kGmmThreads = 256;

struct Mixture {
    double pi_c[kGmmMax];         // pi_k (2 pi)^(-1/2)
    double mu[kGmmMax];
    double inv[kGmmMax];          // 1 / var_k
    double det[kGmmMax];          // sqrt(|var_k|)
    int K;
};

__device__ __forceinline__ void densities(const Mixture &m, double x, double (&d)[kGmmMax])
{
#pragma unroll
    for (int k = 0; k < kGmmMax; ++k) {
        if (k < m.K) {
            const double a = x - m.mu[k];
            d[k] = m.pi_c[k] * exp(-0.5 * (a * m.inv[k] * a)) / m.det[k];
        }
    }
}

// gmm.h:145-201, one pass: per workgroup, w_k, s1_k, s2_k and the NLL into
// out[blockIdx.x][3K + 1] (the host adds the workgroups in order, so the
// result does not depend on scheduling).
template <int KC>
__global__ __launch_bounds__(kGmmThreads) void gmm_pass_kernel(const double *__restrict__ x, uint64_t n, Mixture m,
                                                               double *__restrict__ out)
{
    double w[KC], s1[KC], s2[KC], nll = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) w[k] = s1[k] = s2[k] = 0.0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double xi = x[i];
        double d[kGmmMax];
        densities(m, xi, d);
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < KC; ++k)
            if (k < m.K) sum += d[k];
        if (0.0 < sum) {
            nll -= log(sum);
#pragma unroll
            for (int k = 0; k < KC; ++k) d[k] /= sum;
        } else {
            // every density underflowed: the sample is spread uniformly and
            // charged the smallest normal likelihood (gmm.h:170-175)
            nll -= log(DBL_MIN);
#pragma unroll
            for (int k = 0; k < KC; ++k) d[k] = 1.0 / m.K;
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (k < m.K) {
                w[k] += d[k];
                s1[k] += d[k] * xi;
                s2[k] += d[k] * xi * xi;
            }
        }
    }
    __shared__ double red[3 * KC + 1][kGmmThreads / kWave];
    const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
    auto wave_sum = [](double v) {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
        return v;
    };
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const double a = wave_sum(w[k]), b = wave_sum(s1[k]), c = wave_sum(s2[k]);
        if (lane == 0) {
            red[k][wv] = a;
            red[KC + k][wv] = b;
            red[2 * KC + k][wv] = c;
        }
    }
    const double t = wave_sum(nll);
    if (lane == 0) red[3 * KC][wv] = t;
    __syncthreads();
    for (int j = threadIdx.x; j < 3 * KC + 1; j += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < int(blockDim.x / kWave); ++q) s += red[j][q];
        out[uint64_t(blockIdx.x) * (3 * KC + 1) + j] = s;
    }
}

// gmm.h:204-219: the component of largest density (the first on ties).
__global__ __launch_bounds__(kGmmThreads) void gmm_assign_kernel(const double *__restrict__ x, uint64_t n, Mixture m,
                                                                 int32_t *__restrict__ cl)
{
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d[kGmmMax];
    densities(m, x[i], d);
    int best = 0;
    for (int k = 1; k < m.K; ++k)
        if (d[k] > d[best]) best = k;
    cl[i] = best;
}

int mixture(sct_ctx *ctx, int K, const double *pi, const double *mu, const double *var, Mixture &m)
{
    if (K < 1 || K > kGmmMax || !pi || !mu || !var)
        return sct_fail(ctx, SCT_ERR_INVALID, "mixture needs 1..16 components");
    const double pitod = std::pow(1.0 / (2.0 * M_PI), 0.5);
    m = Mixture{};
    m.K = K;
    for (int k = 0; k < K; ++k) {
        m.pi_c[k] = pi[k] * pitod;
        m.mu[k] = mu[k];
        m.inv[k] = 1.0 / var[k];
        m.det[k] = std::sqrt(std::fabs(var[k]));
    }
    return SCT_OK;
}

uint32_t pass_grid(const sct_samples *s)
{
    const uint64_t need = (s->n + kGmmThreads - 1) / kGmmThreads;
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(need, uint64_t(std::max(1u, s->ctx->cu_count)) * 4)));
}

}  // namespace

extern "C" int sct_samples_create(sct_ctx *ctx, const double *x, uint64_t n, sct_samples **out)
{
    if (!ctx) return SCT_ERR_INVALID;
    if (!out || (!x && n)) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    *out = nullptr;
    auto *s = new sct_samples{ctx, nullptr, n};
    int rc = sct_alloc(ctx, (void **)&s->x, sizeof(double) * std::max<uint64_t>(1, n));
    if (!rc && n) {
        hipError_t e = hipMemcpyAsync(s->x, x, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "samples upload");
    }
    if (rc) {
        sct_free(ctx, s->x);
        delete s;
        return rc;
    }
    *out = s;
    return SCT_OK;
}

extern "C" void sct_samples_destroy(sct_samples *s)
{
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    sct_free(s->ctx, s->x);
    delete s;
}

extern "C" int sct_gmm1d_pass(const sct_samples *s, int K, const double *pi, const double *mu, const double *var,
                              double *w, double *s1, double *s2, double *nll)
{
    if (!s) return SCT_ERR_INVALID;
    sct_ctx *ctx = s->ctx;
    if (!w || !s1 || !s2 || !nll) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    Mixture m;
    int rc = mixture(ctx, K, pi, mu, var, m);
    if (rc) return rc;
    const int KC = K <= 4 ? 4 : (K <= 8 ? 8 : 16);
    const uint32_t grid = pass_grid(s);
    const size_t per = size_t(3 * KC + 1);
    double *d_out = nullptr;
    if ((rc = sct_alloc(ctx, (void **)&d_out, sizeof(double) * per * grid))) return rc;
    if (KC == 4) gmm_pass_kernel<4><<<grid, kGmmThreads, 0, ctx->stream>>>(s->x, s->n, m, d_out);
    else if (KC == 8) gmm_pass_kernel<8><<<grid, kGmmThreads, 0, ctx->stream>>>(s->x, s->n, m, d_out);
    else gmm_pass_kernel<16><<<grid, kGmmThreads, 0, ctx->stream>>>(s->x, s->n, m, d_out);
    std::vector<double> h(per * grid);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    sct_free(ctx, d_out);
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "gmm pass");
    for (int k = 0; k < K; ++k) w[k] = s1[k] = s2[k] = 0.0;
    *nll = 0.0;
    for (uint32_t b = 0; b < grid; ++b) {
        const double *p = h.data() + per * b;
        for (int k = 0; k < K; ++k) {
            w[k] += p[k];
            s1[k] += p[KC + k];
            s2[k] += p[2 * KC + k];
        }
        *nll += p[3 * KC];
    }
    return SCT_OK;
}

extern "C" int sct_gmm1d_assign(const sct_samples *s, int K, const double *pi, const double *mu, const double *var,
                                int32_t *clusters)
{
    if (!s) return SCT_ERR_INVALID;
    sct_ctx *ctx = s->ctx;
    if (!clusters && s->n) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    Mixture m;
    int rc = mixture(ctx, K, pi, mu, var, m);
    if (rc || !s->n) return rc;
    int32_t *d_cl = nullptr;
    if ((rc = sct_alloc(ctx, (void **)&d_cl, sizeof(int32_t) * s->n))) return rc;
    gmm_assign_kernel<<<uint32_t((s->n + kGmmThreads - 1) / kGmmThreads), kGmmThreads, 0, ctx->stream>>>(s->x, s->n, m,
                                                                                                          d_cl);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(clusters, d_cl, sizeof(int32_t) * s->n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    sct_free(ctx, d_cl);
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "gmm assign");
    return SCT_OK;
}
