// gap.hip -- FP64 primal-dual gap of the Potts saddle problem (no counterpart
// in the reference, which stops after a fixed number of iterations).
//
// The iteration (potts-min-partition.cpp:96-489) is the primal-dual method
// for the saddle function
//     L(V, Xi) = <V, G> + <grad V, Xi>,   G_i = lambda (u - c_i)^2,
// with V in the probability simplex and Xi in K = {|xi_i - xi_j| <= 1} per
// voxel (the K-th components implicit, as in the iteration).  Then
//     primal  E(V)  = <V, G> + sum_x h_K(grad V(x)),
//     dual    D(Xi) = sum_x min_i (G_i - (div Xi)_i),
// E >= D, and E - D -> 0 at the saddle point.  h_K, the support function of
// K, is the cheapest flow on the complete graph on the labels with node
// supplies p_i (conic duality):
//     h_K(p) = min sum_{i<j} |f_ij|  s.t.  p_i = sum_{j>i} f_ij - sum_{j<i} f_ji.
// K = 2: |p_0|.  K = 3: the Fermat-Weber cost of the three points
// {0, p_0, -p_1} in closed form (Torricelli point, or the vertex when an
// angle reaches 120 degrees).  K >= 4: iteratively reweighted least squares
// (conductances = flow magnitudes, label K-1 grounded), 256 steps from unit
// conductances, keeping the cheapest iterate.  Every iterate is a feasible
// flow, so the K >= 4 value bounds h_K from above and the gap stays an upper
// bound of the true gap.
//
// All arithmetic is FP64 on the FP32 (or reduced-storage) state; per-block
// sums are combined in a fixed order, so the result is deterministic on a
// given device.
#include "potts.hpp"

#include <algorithm>
#include <vector>

namespace {

constexpr int kIrlsSteps = 256;
constexpr int kGapThreads = 256;

__device__ __forceinline__ double len3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Minimum total distance from one point to {0, a, -b}.
__device__ double fermat_weber3(const double *a, const double *b)
{
    const double s[3] = {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
    const double la = len3(a), lb = len3(b), ls = len3(s);
    if (-dot3(a, b) <= -0.5 * la * lb) return la + lb;   // angle at 0
    if (dot3(a, s) <= -0.5 * la * ls) return la + ls;    // angle at a
    if (dot3(b, s) <= -0.5 * lb * ls) return lb + ls;    // angle at -b
    const double x[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    return sqrt(0.5 * (la * la + lb * lb + ls * ls) + sqrt(3.0) * len3(x));
}

template <int K>
__device__ double min_flow_irls(const double (&p)[K][3])
{
    constexpr int M = K - 1, NP = K * (K - 1) / 2;
    double scale = 0.0;
    for (int i = 0; i < K; ++i) scale = fmax(scale, len3(p[i]));
    if (scale == 0.0) return 0.0;
    double c[NP], L[M][M], phi[K][3], best = HUGE_VAL;
    for (int e = 0; e < NP; ++e) c[e] = 1.0;
    for (int step = 0; step < kIrlsSteps; ++step) {
        for (int r = 0; r < M; ++r)
            for (int q = 0; q < M; ++q) L[r][q] = 0.0;
        for (int i = 0, e = 0; i < K; ++i)
            for (int j = i + 1; j < K; ++j, ++e) {
                if (i < M) L[i][i] += c[e];
                if (j < M) L[j][j] += c[e];
                if (j < M) {
                    L[i][j] -= c[e];
                    L[j][i] -= c[e];
                }
            }
        // Cholesky in place (lower triangle)
        for (int r = 0; r < M; ++r)
            for (int q = 0; q <= r; ++q) {
                double acc = L[r][q];
                for (int t = 0; t < q; ++t) acc -= L[r][t] * L[q][t];
                L[r][q] = (r == q) ? sqrt(acc) : acc / L[q][q];
            }
        for (int a = 0; a < 3; ++a) {
            double y[M];
            for (int r = 0; r < M; ++r) {
                double acc = p[r][a];
                for (int t = 0; t < r; ++t) acc -= L[r][t] * y[t];
                y[r] = acc / L[r][r];
            }
            for (int r = M - 1; r >= 0; --r) {
                double acc = y[r];
                for (int t = r + 1; t < M; ++t) acc -= L[t][r] * phi[t][a];
                phi[r][a] = acc / L[r][r];
            }
            phi[M][a] = 0.0;
        }
        double cost = 0.0;
        for (int i = 0, e = 0; i < K; ++i)
            for (int j = i + 1; j < K; ++j, ++e) {
                const double f[3] = {c[e] * (phi[i][0] - phi[j][0]), c[e] * (phi[i][1] - phi[j][1]),
                                     c[e] * (phi[i][2] - phi[j][2])};
                const double fl = len3(f);
                cost += fl;
                c[e] = fmax(fl, 1e-12 * scale);
            }
        best = fmin(best, cost);
    }
    return best;
}

template <int K>
__device__ double support_k(const double (&p)[K][3])
{
    if constexpr (K == 2) return len3(p[0]);
    else if constexpr (K == 3) return fermat_weber3(p[0], p[1]);
    else return min_flow_irls<K>(p);
}

template <int K, typename St>
struct GapArgs {
    const LineNbr *nbr;
    uint32_t lines, len, wstride, in_stride, s0;
    uint64_t wspan;
    const float *V;
    const typename St::T *xi;
    const float *u;
    double c[K];
    double lambda;
};

template <int K, typename St>
__global__ __launch_bounds__(kGapThreads) void potts_gap_kernel(GapArgs<K, St> a, double *partials)
{
    constexpr int KM = K - 1;
    double E = 0.0, D = 0.0;
    const uint64_t total = uint64_t(a.lines) * a.len;
    for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += uint64_t(gridDim.x) * blockDim.x) {
        const uint32_t line = uint32_t(t / a.len), n = uint32_t(t % a.len);
        const LineNbr nb = a.nbr[line];
        const uint64_t o = uint64_t(line) * a.wstride + n;
        const bool nx[3] = {n + 1 < a.len, nb.nrow >= 0, nb.ncol >= 0};
        const bool px[3] = {n > 0, nb.prow >= 0, nb.pcol >= 0};
        const uint64_t on[3] = {o + 1, uint64_t(nb.nrow) * a.wstride + n, uint64_t(nb.ncol) * a.wstride + n};
        const uint64_t op[3] = {o - 1, uint64_t(nb.prow) * a.wstride + n, uint64_t(nb.pcol) * a.wstride + n};
        double v[K], p[K][3], dv[K];
        double vk = 1.0;
        for (int i = 0; i < KM; ++i) {
            v[i] = a.V[i * a.wspan + o];
            vk -= v[i];
        }
        v[KM] = vk;
        for (int ax = 0; ax < 3; ++ax) {
            double sk = 0.0;
            for (int i = 0; i < KM; ++i) {
                p[i][ax] = nx[ax] ? double(a.V[i * a.wspan + on[ax]]) - v[i] : 0.0;
                sk -= p[i][ax];
            }
            p[KM][ax] = sk;
        }
        for (int i = 0; i < K; ++i) dv[i] = 0.0;
        for (int ax = 0; ax < 3; ++ax) {
            double own_k = 0.0, prev_k = 0.0;
            for (int i = 0; i < KM; ++i) {
                const double own = St::ld(a.xi[(3 * i + ax) * a.wspan + o]);
                const double prev = px[ax] ? double(St::ld(a.xi[(3 * i + ax) * a.wspan + op[ax]])) : 0.0;
                if (nx[ax]) dv[i] += own;
                if (px[ax]) dv[i] -= prev;
                own_k -= own;
                prev_k -= prev;
            }
            if (nx[ax]) dv[KM] += own_k;
            if (px[ax]) dv[KM] -= prev_k;
        }
        const double uu = a.u[uint64_t(line) * a.in_stride + a.s0 + n];
        double lin = 0.0, mn = HUGE_VAL;
        for (int i = 0; i < K; ++i) {
            const double d = uu - a.c[i];
            const double gi = a.lambda * d * d;
            lin += v[i] * gi;
            mn = fmin(mn, gi - dv[i]);
        }
        E += lin + support_k<K>(p);
        D += mn;
    }
    __shared__ double sE[kGapThreads], sD[kGapThreads];
    sE[threadIdx.x] = E;
    sD[threadIdx.x] = D;
    __syncthreads();
    for (int s = kGapThreads / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            sE[threadIdx.x] += sE[threadIdx.x + s];
            sD[threadIdx.x] += sD[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sE[0];
        partials[2 * blockIdx.x + 1] = sD[0];
    }
}

template <int K, typename St>
int launch_gap(sct_potts_solver *s, double *partials, uint32_t grid)
{
    using T = typename St::T;
    const Geometry &geo = s->g->geo;
    GapArgs<K, St> a;
    a.nbr = geo.d_nbr;
    a.lines = geo.lines;
    a.len = s->s1 - s->s0;
    a.wstride = s->wstride;
    a.in_stride = geo.nstride;
    a.s0 = s->s0;
    a.wspan = s->wspan;
    a.V = s->V;
    a.xi = (const T *)s->xi[s->cur];
    a.u = (const float *)s->g->data;
    for (int i = 0; i < K; ++i) a.c[i] = s->c[i];
    a.lambda = s->lambda;
    potts_gap_kernel<K, St><<<grid, kGapThreads, 0, s->ctx->stream>>>(a, partials);
    SCT_HIP(s->ctx, hipGetLastError());
    return SCT_OK;
}

template <int K>
int launch_gap_k(sct_potts_solver *s, double *partials, uint32_t grid)
{
    switch (s->storage) {
    case SCT_STORAGE_BF16: return launch_gap<K, sct::StorageBF16>(s, partials, grid);
    case SCT_STORAGE_F16: return launch_gap<K, sct::StorageF16>(s, partials, grid);
    default: return launch_gap<K, sct::StorageF32>(s, partials, grid);
    }
}

}  // namespace

int sct::potts_gap(sct_potts_solver *s, double *primal, double *dual)
{
    const uint32_t grid = std::max<uint32_t>(1, s->ctx->cu_count * 4);
    double *d = nullptr;
    int rc = sct_alloc(s->ctx, (void **)&d, sizeof(double) * 2 * grid);
    if (rc) return rc;
    if (s->g->geo.lines && s->s1 > s->s0) {
        switch (s->K) {
        case 2: rc = launch_gap_k<2>(s, d, grid); break;
        case 3: rc = launch_gap_k<3>(s, d, grid); break;
        case 4: rc = launch_gap_k<4>(s, d, grid); break;
        case 5: rc = launch_gap_k<5>(s, d, grid); break;
        default: rc = launch_gap_k<6>(s, d, grid); break;
        }
    } else {
        const hipError_t e = hipMemsetAsync(d, 0, sizeof(double) * 2 * grid, s->ctx->stream);
        if (e != hipSuccess) rc = sct_hip_fail(s->ctx, e, "gap");
    }
    std::vector<double> h(2 * size_t(grid));
    if (!rc) {
        hipError_t e = hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
        if (e != hipSuccess) rc = sct_hip_fail(s->ctx, e, "gap download");
    }
    sct_free(s->ctx, d);
    if (rc) return rc;
    double E = 0.0, D = 0.0;
    for (uint32_t b = 0; b < grid; ++b) {
        E += h[2 * b];
        D += h[2 * b + 1];
    }
    *primal = E;
    *dual = D;
    return SCT_OK;
}

extern "C" int sct_potts_solver_gap(sct_potts_solver *s, double *primal, double *dual)
{
    if (!s || !primal || !dual) return SCT_ERR_INVALID;
    return sct::potts_gap(s, primal, dual);
}
