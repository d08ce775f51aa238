// bimm.hip -- the sample likelihood of bimm-peaks' bivariate intensity /
// gradient-magnitude mixture (bimm-peaks/puv.h:42-130, summed as in
// bimm-peaks.cpp:458-472), over (u, v) samples resident on the device.  The
// reference evaluates it on the host with OpenMP and one adaptive GSL
// integral per sample and interface pair; here one thread takes a sample.
//
// Components: K pure phases with density
//   full_p_i(u, v) = 2 sqrt(2) / pi v^2 / (sn^4 (1 - rho)^1.5)
//                    exp(-(u - I_i)^2 / (2 sn^2) - v^2 / a),   a = sn^2 (1 - rho)
// and K(K-1)/2 interfaces (pairs i < j in strict-upper order) with density
//   full_p_ij(u, v) = v / (2 sqrt(2) pi ds sn sqrt(a)) int_{-ds}^{ds} f(t) dt,
//   f(t) = exp(-(u - I(t))^2 / (2 sn^2) - (v - G(t))^2 / a) (1 - e^{-4 v G / a}) / G,
//   I(t) = I_i + (I_j - I_i)(1 + erf(t / sqrt 2)) / 2,
//   G(t) = |I_j - I_i| / (sb sqrt(2 pi)) e^{-t^2 / 2}
// (its v -> G -> 0 limit 4 v / a exp(...) when |t| >= 30 or I_j == I_i).
// The integral is adaptive Gauss-Kronrod (7/15 points, QUADPACK's qk15
// table) on an initial partition of [-ds, ds] with breakpoints where the
// integrand can peak (I(t) = u and G(t) = v), interval halving until the
// Kronrod-Gauss difference is below 1e-10 relative.  Per-workgroup partial
// sums are added on the host in order, so the result is deterministic.
#include "sct_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

struct sct_uv_samples {
    sct_ctx *ctx;
    double *u, *v;
    uint64_t n;
};

namespace {

constexpr int kBimmMaxK = 6;
constexpr int kBimmPairs = kBimmMaxK * (kBimmMaxK - 1) / 2;
constexpr int kBimmThreads = 256;
constexpr int kStack = 48;          // pending intervals per thread
constexpr int kPanels = 8;          // uniform panels of the initial partition
constexpr double kRelTol = 1e-10;
constexpr int kMaxIntervals = 4096;

struct BimmModel {
    int K, npairs;
    double w[kBimmMaxK + kBimmPairs];
    double I[kBimmMaxK];
    int pi[kBimmPairs], pj[kBimmPairs];
    double sn2, a, ds;
    double cpure;   // 2 sqrt(2) / pi / (sn^4 (1 - rho)^1.5)
    double cpair;   // 1 / (2 sqrt(2) pi ds sn sqrt(a))
    double inv_sb;  // 1 / (sb sqrt(2 pi))
};

// QUADPACK qk15: Kronrod abscissae (the Gauss-Legendre 7-point nodes are the
// odd ones), Kronrod weights, Gauss weights.
__constant__ double kXgk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                               0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                               0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                               0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
__constant__ double kWgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                               0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                               0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                               0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
__constant__ double kWg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                              0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

struct Integrand {
    double u, v, Ii, dI, G0, sn2, a;
    __device__ double operator()(double t) const
    {
        const double I = Ii + 0.5 * dI * (1.0 + erf(t * M_SQRT1_2));
        const double du = u - I;
        if (fabs(t) < 30.0 && fabs(dI) > 1e-9) {
            const double G = G0 * exp(-0.5 * t * t);
            const double dv = v - G;
            return exp(-0.5 * du * du / sn2 - dv * dv / a) * (-expm1(-4.0 * v * G / a)) / G;
        }
        return exp(-0.5 * du * du / sn2 - v * v / a) * 4.0 * v / a;
    }
};

// Kronrod 15 and Gauss 7 estimates of f over [x0, x1].
__device__ void gk15(const Integrand &f, double x0, double x1, double &k, double &g)
{
    const double c = 0.5 * (x0 + x1), h = 0.5 * (x1 - x0);
    const double fc = f(c);
    double rk = fc * kWgk[7], rg = fc * kWg[3];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const double d = h * kXgk[j];
        const double s = f(c - d) + f(c + d);
        rk += kWgk[j] * s;
        if (j & 1) rg += kWg[j / 2] * s;
    }
    k = rk * h;
    g = rg * h;
}

// int_{-ds}^{ds} f(t) dt for one sample and one pair.
__device__ double interface_integral(const Integrand &f, double ds)
{
    // initial partition: uniform panels plus the points where the intensity
    // matches u and where the gradient matches v
    double bp[kPanels + 4];
    int nb = 0;
    for (int p = 0; p <= kPanels; ++p) bp[nb++] = -ds + 2.0 * ds * p / kPanels;
    if (f.dI != 0.0) {
        const double s = (f.u - f.Ii) / f.dI;   // Phi(t) at I(t) = u
        if (s > 0.0 && s < 1.0) {
            const double t = M_SQRT2 * erfinv(2.0 * s - 1.0);
            if (t > -ds && t < ds) bp[nb++] = t;
        }
    }
    if (f.G0 > 0.0 && f.v > 0.0 && f.v < f.G0) {
        const double t = sqrt(2.0 * log(f.G0 / f.v));   // G(t) = v
        if (t < ds) {
            bp[nb++] = t;
            bp[nb++] = -t;
        }
    }
    for (int i = 1; i < nb; ++i)   // insertion sort (at most 12 points)
        for (int j = i; j > 0 && bp[j - 1] > bp[j]; --j) {
            const double x = bp[j];
            bp[j] = bp[j - 1];
            bp[j - 1] = x;
        }
    double sa[kStack], sb[kStack];
    int top = 0;
    for (int i = nb - 1; i > 0; --i)
        if (bp[i] > bp[i - 1]) {
            sa[top] = bp[i - 1];
            sb[top] = bp[i];
            ++top;
        }
    double total = 0.0;
    int budget = kMaxIntervals;   // bounds the work of any one integral
    while (top > 0) {
        --top;
        const double x0 = sa[top], x1 = sb[top];
        double k, g;
        gk15(f, x0, x1, k, g);
        const double mid = 0.5 * (x0 + x1);
        if (fabs(k - g) <= kRelTol * fabs(k) || top + 2 > kStack || --budget <= 0 || !(mid > x0 && mid < x1)) {
            total += k;
        } else {
            sa[top] = mid;
            sb[top] = x1;
            ++top;
            sa[top] = x0;
            sb[top] = mid;
            ++top;
        }
    }
    return total;
}

__device__ double full_p(const BimmModel &m, double u, double v)
{
    double p = 0.0;
    const double v2 = v * v;
    for (int i = 0; i < m.K; ++i) {
        const double du = u - m.I[i];
        p += m.w[i] * m.cpure * v2 * exp(-0.5 * du * du / m.sn2 - v2 / m.a);
    }
    for (int k = 0; k < m.npairs; ++k) {
        const double Ii = m.I[m.pi[k]], dI = m.I[m.pj[k]] - Ii;
        const Integrand f{u, v, Ii, dI, fabs(dI) * m.inv_sb, m.sn2, m.a};
        p += m.w[m.K + k] * m.cpair * v * interface_integral(f, m.ds);
    }
    return p;
}

__global__ __launch_bounds__(kBimmThreads) void bimm_nll_kernel(const double *__restrict__ u,
                                                                const double *__restrict__ v, uint64_t first,
                                                                uint64_t count, BimmModel m, double *__restrict__ out)
{
    double s = 0.0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        s -= log(full_p(m, u[first + i], v[first + i]));
    __shared__ double red[kBimmThreads];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kBimmThreads / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

}  // namespace

extern "C" int sct_uv_samples_create(sct_ctx *ctx, const double *u, const double *v, uint64_t n, sct_uv_samples **out)
{
    if (!ctx) return SCT_ERR_INVALID;
    if (!out || (n && (!u || !v))) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    *out = nullptr;
    auto *s = new sct_uv_samples{ctx, nullptr, nullptr, n};
    const size_t bytes = sizeof(double) * std::max<uint64_t>(1, n);
    int rc = sct_alloc(ctx, (void **)&s->u, bytes);
    if (!rc) rc = sct_alloc(ctx, (void **)&s->v, bytes);
    if (!rc && n) {
        hipError_t e = hipMemcpyAsync(s->u, u, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(s->v, v, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "uv samples upload");
    }
    if (rc) {
        sct_free(ctx, s->u);
        sct_free(ctx, s->v);
        delete s;
        return rc;
    }
    *out = s;
    return SCT_OK;
}

extern "C" void sct_uv_samples_destroy(sct_uv_samples *s)
{
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    sct_free(s->ctx, s->u);
    sct_free(s->ctx, s->v);
    delete s;
}

extern "C" int sct_bimm_nll(const sct_uv_samples *s, int K, const double *w, const double *I, double log_sn,
                            double log_sb, double log_ds, double rho, uint64_t first, uint64_t count, double *nll)
{
    if (!s) return SCT_ERR_INVALID;
    sct_ctx *ctx = s->ctx;
    if (!w || !I || !nll) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (K < 1 || K > kBimmMaxK) return sct_fail(ctx, SCT_ERR_INVALID, "K must be in [1, 6]");
    if (first > s->n || count > s->n - first) return sct_fail(ctx, SCT_ERR_INVALID, "sample range out of bounds");
    if (!(rho < 1.0)) return sct_fail(ctx, SCT_ERR_INVALID, "rho must be below 1");
    BimmModel m{};
    m.K = K;
    m.npairs = K * (K - 1) / 2;
    for (int i = 0; i < K + m.npairs; ++i) m.w[i] = w[i];
    for (int i = 0; i < K; ++i) m.I[i] = I[i];
    for (int i = 0, k = 0; i < K - 1; ++i)
        for (int j = i + 1; j < K; ++j, ++k) {
            m.pi[k] = i;
            m.pj[k] = j;
        }
    const double sn = std::exp(log_sn), sb = std::exp(log_sb);
    m.sn2 = std::exp(2.0 * log_sn);
    m.a = m.sn2 * (1.0 - rho);
    m.ds = std::exp(log_ds);
    m.cpure = 2.0 * M_SQRT2 / M_PI / (m.sn2 * m.sn2 * std::pow(1.0 - rho, 1.5));
    m.cpair = 1.0 / (2.0 * M_SQRT2 * M_PI * m.ds * sn * std::sqrt(m.a));
    m.inv_sb = 1.0 / (sb * std::sqrt(2.0 * M_PI));
    *nll = 0.0;
    if (!count) return SCT_OK;
    const uint32_t grid = uint32_t(std::max<uint64_t>(
        1, std::min<uint64_t>((count + kBimmThreads - 1) / kBimmThreads, uint64_t(std::max(1u, ctx->cu_count)) * 8)));
    double *d_out = nullptr;
    int rc = sct_alloc(ctx, (void **)&d_out, sizeof(double) * grid);
    if (rc) return rc;
    bimm_nll_kernel<<<grid, kBimmThreads, 0, ctx->stream>>>(s->u, s->v, first, count, m, d_out);
    std::vector<double> h(grid);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_out, sizeof(double) * grid, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    sct_free(ctx, d_out);
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "bimm nll");
    double t = 0.0;
    for (double x : h) t += x;
    *nll = t;
    return SCT_OK;
}
