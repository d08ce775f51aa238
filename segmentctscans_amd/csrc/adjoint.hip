// adjoint.hip -- gradient of a segmentation loss through the unrolled Potts
// iteration (no counterpart in the reference).
//
// Loss after `iters` iterations from the nearest-label start:
//   L = 1/(N K) sum_x sum_{i<K} (V_i(x) - [target(x) = i])^2,
// N in-mask voxels, V_{K-1} = 1 - sum of the stored components.  Gradients
// with respect to the labels c, lambda and u flow through the data term
// G_i = lambda (u - c_i)^2 of every V step; the start (nearest label) is
// piecewise constant, so its gradient is 0 almost everywhere.
//
// Reverse step for one iteration s_t = (V, V-bar, Xi) -> s_{t+1}, two kernels:
//   A  V+ = Pi_S(q), V-bar+ = 2 V+ - V: the simplex projection's Jacobian is
//      diag(a) - a a^T / |a| on its active set a; q = V + tau1 (div Xi+ - G)
//      passes the adjoint to V, to div Xi+ and to G (-> c, lambda, u);
//   B  div's adjoint is -grad; Xi+ = Pi_K(x), x = Xi + tau2 grad V-bar: the
//      Dykstra sweeps are replayed with the same arithmetic (so the same sweep
//      count and the same ball tests) recording each pair's h, and each pair
//      step is back-propagated: the excess map f(h) = h (1 - 1/|h|) for
//      |h| > 1 has Jacobian (1 - 1/|h|) I + h h^T / |h|^3;
//   C  grad's adjoint is -div, into V-bar; evaluated inside the next
//      step's A from B's Xi adjoint at the voxel and its -e_a neighbours,
//      so the V-bar adjoint is never stored.
// The implicit K-th components turn each full-K adjoint g into g_i - g_{K-1}
// on the stored components.
//
// States are checkpointed every `ck` iterations during the forward pass; the
// reverse pass recomputes the ck - 1 states inside each segment from its
// checkpoint (memory ~ (T/ck + ck) states, ck = sqrt(T) by default; ck = 1
// keeps every state and recomputes none) with the forward kernels of the
// solver, so the replayed values are bitwise the forward's.
//
// Every forward step also stores its SweepRecord (per voxel: 0 when the
// Dykstra input was feasible, else the sweeps it took) after the state.  B
// passes the feasible voxels straight through and replays the others in
// buckets of equal sweep count, so that a replay wave does not run as long as
// its slowest of 64 unrelated voxels (about 3x the mean sweeps in queue order).
#include "potts.hpp"
#include "twopass.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace sct;
using namespace sct::tp;

namespace {

constexpr int kAdjThreads = 256;
constexpr int kBuckets = 21;  // replay items by sweep count 1..20 (0: no replay)
constexpr int kShortSweeps = 4;   // replay items whose Dykstra history B2 keeps in registers (replay_short)
// Bucket counters are kept per stripe of consecutive workgroups, so that no
// counter takes the device atomics of more than 1/kStripes of a grid.
constexpr int kStripes = 64;
// [stripe][bucket], one 128-byte line per stripe: device atomics on one line
// are serialised, so a line takes only its own stripe's workgroups.
constexpr int kStripeWords = 32;
static_assert(kBuckets <= kStripeWords, "a stripe's counters fit one line");
constexpr int kCounters = kStripes * kStripeWords;
constexpr int kBucketWords = 2 * kCounters;
static_assert(kStripes == kWave, "one stripe per lane in bucket_starts");

// Work layout of the adjoint: state = V (K-1) | V-bar (K-1) | Xi 3(K-1)
// components of wspan floats each.
struct AdjLayout {
    VolumeLines fwd;    // chunks of 128 (the FP32 solver kernels)
    VolumeLines one;    // chunks of 64, one voxel per lane (adjoint kernels)
    uint64_t wspan;
};

template <int K>
TwoPassArgs<K, StorageF32> state_args(float *st, const float *u, uint64_t wspan)
{
    TwoPassArgs<K, StorageF32> a;
    for (int i = 0; i < K - 1; ++i) {
        a.V[i] = st + uint64_t(i) * wspan;
        a.vbar[i] = st + uint64_t(K - 1 + i) * wspan;
    }
    for (int j = 0; j < 3 * (K - 1); ++j) a.xi[j] = st + uint64_t(2 * (K - 1) + j) * wspan;
    a.u = u;
    a.in_place();
    return a;
}

// One voxel per lane: (chunk, lane) of the 64-voxel chunking.
struct Voxel {
    bool ok;
    Chunk ch;
    uint32_t n, w, uo;
};

__device__ __forceinline__ Voxel voxel_of(const VolumeLines &L)
{
    Voxel v{};
    const uint64_t gw = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    if (gw >= L.count()) return v;
    v.ch = L.chunk<1>(gw);
    const uint32_t lane = threadIdx.x % kWave;
    v.n = v.ch.n0 + lane;
    v.ok = v.n < v.ch.len;
    v.w = v.ch.w + lane;
    v.uo = v.ch.u + lane;
    return v;
}

// Block sum of x[N]; thread k < N adds component k to out[blockIdx.x * N + k]
// (one slot per workgroup, accumulated over launches with the same grid).
template <int N>
__device__ __forceinline__ void block_add(double (&x)[N], double *out)
{
    __shared__ double red[N][kAdjThreads / kWave];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double t = x[i];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, kWave);
        if (threadIdx.x % kWave == 0) red[i][threadIdx.x / kWave] = t;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        double t = 0.0;
        for (int k = 0; k < int(blockDim.x / kWave); ++k) t += red[threadIdx.x][k];
        out[uint64_t(blockIdx.x) * N + threadIdx.x] += t;
    }
}

// Loss and its adjoint with respect to the stored components of V_T.
template <int K>
__global__ __launch_bounds__(kAdjThreads) void adj_loss_kernel(VolumeLines L, uint64_t wspan, const float *st,
                                                               const float *target, float wgt, float *gV,
                                                               double *loss)
{
    const Voxel vx = voxel_of(L);
    double part[1] = {0.0};
    if (vx.ok) {
        const int y = int(target[vx.uo]);
        float e[K], vk = 1.0f;
        for (int i = 0; i < K - 1; ++i) {
            const float v = st[uint64_t(i) * wspan + vx.w];
            vk -= v;
            e[i] = v - (i == y ? 1.0f : 0.0f);
        }
        e[K - 1] = vk - (K - 1 == y ? 1.0f : 0.0f);
        for (int i = 0; i < K; ++i) part[0] += double(wgt) * double(e[i]) * double(e[i]);
        for (int i = 0; i < K - 1; ++i) gV[uint64_t(i) * wspan + vx.w] = 2.0f * wgt * (e[i] - e[K - 1]);
    }
    block_add<1>(part, loss);
}

// Adjoint of V-bar+ from the previous reverse step's B: with g the adjoint of
// Dykstra's input, grad's adjoint is -div of tau2 g on all K components, and
// the implicit K-th component folds into the stored ones as
// -div(tau2 (g_i - g_{K-1})).  g_i - g_{K-1} is gXi, B's output for the Xi
// state, which the next B1 has not yet overwritten when A runs.
template <int K>
__device__ __forceinline__ void vbar_adjoint(const Voxel &vx, uint64_t wspan, const float *gXi, float tau2,
                                             float (&gvb)[K - 1])
{
    const Chunk &ch = vx.ch;
    const uint32_t lane = threadIdx.x % kWave;
    const bool px[3] = {vx.n > 0, ch.pr != kNone, ch.pc != kNone};
    const bool nx[3] = {vx.n + 1 < ch.len, ch.nr != kNone, ch.nc != kNone};
    const uint32_t wp[3] = {vx.w - 1, ch.pr + lane, ch.pc + lane};
    for (int i = 0; i < K - 1; ++i) {
        float dv = 0.0f;
        for (int a = 0; a < 3; ++a) {
            if (nx[a]) dv += gXi[uint64_t(3 * i + a) * wspan + vx.w];
            if (px[a]) dv -= gXi[uint64_t(3 * i + a) * wspan + wp[a]];
        }
        gvb[i] = -tau2 * dv;
    }
}

// Replay items of one reverse step, ordered by sweep count: hist[z][s] counts
// the voxels of stripe z whose Dykstra took s sweeps (from the forward's
// SweepRecord); offset[b][s] is where workgroup b's items of bucket s begin
// within its stripe's (the value of hist[z][s] before b added its count).  Two
// histograms alternate between reverse steps so that A can clear the next one.
struct ReplayBuckets {
    uint32_t *hist_cur, *hist_next, *offset;
};

__device__ __forceinline__ uint32_t stripe_of(uint32_t block, uint32_t blocks)
{
    return uint32_t(uint64_t(block) * kStripes / blocks);
}

// A: through V-bar+ = 2 V+ - V and V+ = Pi_S(q).  gs = gV | gXi.
// Also counts the replay items of the following B1 launch by sweep count
// (rec1: the SweepRecord of the step s -> s1) and places each workgroup's
// items within its stripe.
template <int K>
__global__ __launch_bounds__(kAdjThreads) void adj_v_kernel(VolumeLines L, uint64_t wspan, const float *st,
                                                            const float *st1, const uint8_t *rec1, const float *u,
                                                            SolveParams<K> sp, float tau1, float tau2,
                                                            float *gs, float *GQ, float *gu, double *acc,
                                                            ReplayBuckets rb)
{
    constexpr int KM = K - 1;
    __shared__ uint32_t hb[kBuckets];
    if (threadIdx.x < kBuckets) hb[threadIdx.x] = 0;
    if (blockIdx.x == 0)
        for (uint32_t i = threadIdx.x; i < kCounters; i += blockDim.x) rb.hist_next[i] = 0;
    __syncthreads();
    const Voxel vx = voxel_of(L);
    double part[K + 1] = {};
    if (vx.ok) {
        const uint32_t sw = rec1[vx.w];
        if (sw) atomicAdd(&hb[min(sw, uint32_t(kBuckets - 1))], 1u);
        const Chunk &ch = vx.ch;
        const uint32_t lane = threadIdx.x % kWave;
        float gvb[KM];
        vbar_adjoint<K>(vx, wspan, gs + uint64_t(KM) * wspan, tau2, gvb);
        const float *Xi1 = st1 + uint64_t(2 * KM) * wspan;
        const bool px[3] = {vx.n > 0, ch.pr != kNone, ch.pc != kNone};
        const bool nx[3] = {vx.n + 1 < ch.len, ch.nr != kNone, ch.nc != kNone};
        const uint32_t wp[3] = {vx.w - 1, ch.pr + lane, ch.pc + lane};
        float xo[KM][3], xp[3][KM], v[KM];
        for (int i = 0; i < KM; ++i) {
            v[i] = st[uint64_t(i) * wspan + vx.w];
            for (int a = 0; a < 3; ++a) {
                xo[i][a] = Xi1[uint64_t(3 * i + a) * wspan + vx.w];
                xp[a][i] = px[a] ? Xi1[uint64_t(3 * i + a) * wspan + wp[a]] : 0.0f;
            }
        }
        const float uu = u[vx.uo];
        float vf[K], q[K];
        bool act[K];
        v_input<K>(xo, xp, px, nx, v, uu, sp.c, sp.lambda, tau1, vf, q);
        project_simplex<K>(q, &act);
        float ge[K], mean = 0.0f;
        int na = 0;
        for (int i = 0; i < K; ++i) {
            ge[i] = i < KM ? gs[uint64_t(i) * wspan + vx.w] + 2.0f * gvb[i] : 0.0f;
            if (act[i]) {
                mean += ge[i];
                ++na;
            }
        }
        mean /= float(na);
        float gq[K], gul = 0.0f;
        for (int i = 0; i < K; ++i) {
            gq[i] = act[i] ? ge[i] - mean : 0.0f;
            // gq is the projection of ge onto the active face's tangent space,
            // so it sums to 0 over the labels: GQ keeps K - 1 components and
            // B1 rebuilds the last
            if (i < KM) GQ[uint64_t(i) * wspan + vx.w] = gq[i];
            const float d = uu - sp.c[i];
            const float gG = -tau1 * gq[i];
            gul += gG * 2.0f * sp.lambda * d;
            part[i] += double(gG) * (-2.0 * double(sp.lambda) * double(d));
            part[K] += double(gG) * double(d) * double(d);
        }
        for (int i = 0; i < KM; ++i) gs[uint64_t(i) * wspan + vx.w] = gq[i] - gq[KM] - gvb[i];
        gu[vx.w] += gul;
    }
    block_add<K + 1>(part, acc);   // its barrier also completes hb
    if (threadIdx.x < kBuckets) {
        const uint32_t h = hb[threadIdx.x];
        const uint32_t o =
            h ? atomicAdd(&rb.hist_cur[stripe_of(blockIdx.x, gridDim.x) * kStripeWords + threadIdx.x], h) : 0u;
        rb.offset[uint64_t(blockIdx.x) * kStripeWords + threadIdx.x] = o;
    }
}

// Dykstra's sweeps replayed from x with the solver's arithmetic, recording h.
template <int K>
__device__ __forceinline__ int dykstra_replay(float (&q)[K][3], float (*hist)[K * (K - 1) / 2][3], DykstraLimits lim)
{
    constexpr int NP = K * (K - 1) / 2;
    float r[NP][3];
    for (int p = 0; p < NP; ++p) r[p][0] = r[p][1] = r[p][2] = 0.0f;
    int sw = 0;
    while (true) {
        const float m2 = dykstra_sweep<K>(q, r, lim, hist[sw]);
        ++sw;
        if (m2 < lim.stol || sw == 20) break;
    }
    return sw;
}

// g: adjoint of Dykstra's output on entry, of its input on return.
template <int K>
__device__ __forceinline__ void dykstra_back(int sweeps, const float (*hist)[K * (K - 1) / 2][3], float (&g)[K][3],
                                             DykstraLimits lim)
{
    constexpr int NP = K * (K - 1) / 2;
    float gr[NP][3];
    for (int p = 0; p < NP; ++p) gr[p][0] = gr[p][1] = gr[p][2] = 0.0f;
    for (int s = sweeps - 1; s >= 0; --s)
        for (int i1 = K - 2, p = NP - 1; i1 >= 0; --i1)
            for (int i2 = K - 1; i2 > i1; --i2, --p) {
                const float h0 = hist[s][p][0], h1 = hist[s][p][1], h2 = hist[s][p][2];
                const float ss = __fmaf_rn(h2, h2, __fmaf_rn(h1, h1, h0 * h0));
                float gd[3], ge[3], gh[3] = {0.0f, 0.0f, 0.0f};
                for (int a = 0; a < 3; ++a) {
                    gd[a] = g[i1][a] - g[i2][a];
                    ge[a] = 0.5f * gd[a] + gr[p][a];
                }
                if (ss > lim.t1) {
                    const float l = sqrtf(ss), il = 1.0f / l;
                    const float hg = h0 * ge[0] + h1 * ge[1] + h2 * ge[2];
                    const float k = hg * il * il * il, s1 = 1.0f - il;
                    gh[0] = s1 * ge[0] + k * h0;
                    gh[1] = s1 * ge[1] + k * h1;
                    gh[2] = s1 * ge[2] + k * h2;
                }
                for (int a = 0; a < 3; ++a) {
                    g[i1][a] -= gh[a];
                    g[i2][a] += gh[a];
                    gr[p][a] = -0.5f * gd[a] + gh[a];
                }
            }
}

// One sweep of dykstra_back: the pairs in reverse order, h the sweep's
// recorded pre-projection differences, gr the corrections' adjoints.
template <int K>
__device__ __forceinline__ void back_sweep(const float (&h)[K * (K - 1) / 2][3], float (&g)[K][3],
                                           float (&gr)[K * (K - 1) / 2][3], DykstraLimits lim)
{
#pragma unroll
    for (int i1 = K - 2; i1 >= 0; --i1)
#pragma unroll
        for (int i2 = K - 1; i2 > i1; --i2) {
            const int p = (2 * K - 1 - i1) * i1 / 2 + i2 - i1 - 1;
            const float h0 = h[p][0], h1 = h[p][1], h2 = h[p][2];
            const float ss = __fmaf_rn(h2, h2, __fmaf_rn(h1, h1, h0 * h0));
            float gd[3], ge[3], gh[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                gd[a] = g[i1][a] - g[i2][a];
                ge[a] = 0.5f * gd[a] + gr[p][a];
            }
            if (ss > lim.t1) {
                const float l = sqrtf(ss), il = 1.0f / l;
                const float hg = h0 * ge[0] + h1 * ge[1] + h2 * ge[2];
                const float k = hg * il * il * il, s1 = 1.0f - il;
                gh[0] = s1 * ge[0] + k * h0;
                gh[1] = s1 * ge[1] + k * h1;
                gh[2] = s1 * ge[2] + k * h2;
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                g[i1][a] -= gh[a];
                g[i2][a] += gh[a];
                gr[p][a] = -0.5f * gd[a] + gh[a];
            }
        }
}

// Replay and back-propagation of an item whose Dykstra converges within
// kShortSweeps = 4 sweeps, each sweep's history in its own small array: the
// compiler keeps those in registers, where one 20-sweep array goes to private
// memory.  Returns false, with g unchanged, if it has not converged by then.
template <int K>
__device__ __forceinline__ bool replay_short(const float (&x)[K][3], float (&g)[K][3], DykstraLimits lim)
{
    constexpr int NP = K * (K - 1) / 2;
    float q[K][3], r[NP][3], h0[NP][3], h1[NP][3], h2[NP][3], h3[NP][3];
#pragma unroll
    for (int i = 0; i < K; ++i)
#pragma unroll
        for (int a = 0; a < 3; ++a) q[i][a] = x[i][a];
#pragma unroll
    for (int pp = 0; pp < NP; ++pp) r[pp][0] = r[pp][1] = r[pp][2] = 0.0f;
    int sw = 1;
    if (!(dykstra_sweep<K>(q, r, lim, h0) < lim.stol)) {
        sw = 2;
        if (!(dykstra_sweep<K>(q, r, lim, h1) < lim.stol)) {
            sw = 3;
            if (!(dykstra_sweep<K>(q, r, lim, h2) < lim.stol)) {
                sw = 4;
                if (!(dykstra_sweep<K>(q, r, lim, h3) < lim.stol)) return false;
            }
        }
    }
    static_assert(kShortSweeps == 4, "replay_short records four sweeps");
    float gr[NP][3];
#pragma unroll
    for (int pp = 0; pp < NP; ++pp) gr[pp][0] = gr[pp][1] = gr[pp][2] = 0.0f;
    if (sw >= 4) back_sweep<K>(h3, g, gr, lim);
    if (sw >= 3) back_sweep<K>(h2, g, gr, lim);
    if (sw >= 2) back_sweep<K>(h1, g, gr, lim);
    back_sweep<K>(h0, g, gr, lim);
    return true;
}

// dykstra_back with the sweep loop unrolled (constant history indices) and
// the history of sweep s - 1 loaded while sweep s is back-propagated, so the
// replay waits for one load latency per sweep, not one per pair step.
template <int K, int H = 20>
__device__ __forceinline__ void dykstra_back_unrolled(int sweeps, const float (&hist)[H][K * (K - 1) / 2][3],
                                                      float (&g)[K][3], DykstraLimits lim)
{
    constexpr int NP = K * (K - 1) / 2;
    float gr[NP][3], hn[NP][3];
#pragma unroll
    for (int p = 0; p < NP; ++p) gr[p][0] = gr[p][1] = gr[p][2] = 0.0f;
#pragma unroll
    for (int p = 0; p < NP; ++p) hn[p][0] = hn[p][1] = hn[p][2] = 0.0f;
    if (sweeps >= H) {
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int a = 0; a < 3; ++a) hn[p][a] = hist[H - 1][p][a];
    }
#pragma unroll
    for (int s = H - 1; s >= 0; --s) {
        if (s >= sweeps) {
            if (s - 1 >= 0 && s - 1 < sweeps) {
#pragma unroll
                for (int p = 0; p < NP; ++p)
#pragma unroll
                    for (int a = 0; a < 3; ++a) hn[p][a] = hist[s - 1][p][a];
            }
            continue;
        }
        float hc[NP][3];
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int a = 0; a < 3; ++a) hc[p][a] = hn[p][a];
        if (s - 1 >= 0) {
#pragma unroll
            for (int p = 0; p < NP; ++p)
#pragma unroll
                for (int a = 0; a < 3; ++a) hn[p][a] = hist[s - 1][p][a];
        }
#pragma unroll
        for (int i1 = K - 2; i1 >= 0; --i1)
#pragma unroll
            for (int i2 = K - 1; i2 > i1; --i2) {
                const int p = (2 * K - 1 - i1) * i1 / 2 + i2 - i1 - 1;
                const float h0 = hc[p][0], h1 = hc[p][1], h2 = hc[p][2];
                const float ss = __fmaf_rn(h2, h2, __fmaf_rn(h1, h1, h0 * h0));
                float gd[3], ge[3], gh[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    gd[a] = g[i1][a] - g[i2][a];
                    ge[a] = 0.5f * gd[a] + gr[p][a];
                }
                if (ss > lim.t1) {
                    const float l = sqrtf(ss), il = 1.0f / l;
                    const float hg = h0 * ge[0] + h1 * ge[1] + h2 * ge[2];
                    const float k = hg * il * il * il, s1 = 1.0f - il;
                    gh[0] = s1 * ge[0] + k * h0;
                    gh[1] = s1 * ge[1] + k * h1;
                    gh[2] = s1 * ge[2] + k * h2;
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    g[i1][a] -= gh[a];
                    g[i2][a] += gh[a];
                    gr[p][a] = -0.5f * gd[a] + gh[a];
                }
            }
    }
}

// Writes the Xi-step adjoint of a voxel, g = adjoint of Dykstra's input x, on
// the stored components (the next A also reads it, for V-bar's adjoint).
template <int K>
__device__ __forceinline__ void xi_outputs(uint64_t wspan, uint32_t w, const float (&g)[K][3], float *gXi)
{
    for (int i = 0; i < K - 1; ++i)
        for (int a = 0; a < 3; ++a) gXi[uint64_t(3 * i + a) * wspan + w] = g[i][a] - g[K - 1][a];
}

// Voxels whose Dykstra replay is needed, in buckets of equal sweep count,
// longest first (bucket s starts after the items of buckets s+1..20), each
// bucket stripe by stripe, so that the lanes of a replay wave run the same
// number of sweeps on voxels of nearby workgroups.  An item is x (K x 3), the
// output adjoint g (K - 1 x 3; g_{K-1} is 0) and the voxel's work offset:
// 64 bytes for K = 3.  Items are stored one after the other (a
// workgroup's items of one bucket are contiguous).
struct XiQueue {
    float *items;
    uint32_t cap;      // items
};

// Where stripe z's items of each bucket start in the item list (start[s]),
// and the total (start[kBuckets]), into LDS, in two parts so that the loads
// can be issued early: bucket_loads reads, per lane = stripe, the counts of
// the buckets its wave scans (wave wv: buckets 1 + wv, 1 + wv + nw, ...);
// bucket_starts scans them.  Called by every thread of a kAdjThreads
// workgroup; bucket_starts ends with a barrier.
constexpr uint32_t kScanWaves = kAdjThreads / kWave;
constexpr int kPerWave = int((kBuckets - 1 + kScanWaves - 1) / kScanWaves);

__device__ __forceinline__ void bucket_loads(const uint32_t *hist, uint32_t (&hv)[kPerWave])
{
    const uint32_t lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < kPerWave; ++k) {
        const uint32_t s = 1 + wv + k * kScanWaves;
        hv[k] = s < kBuckets ? hist[lane * kStripeWords + s] : 0u;
    }
}

__device__ void bucket_starts(const uint32_t (&hv)[kPerWave], uint32_t z, uint32_t (&start)[kBuckets + 1])
{
    __shared__ uint32_t tot[kBuckets], before[kBuckets];
    const uint32_t lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < kPerWave; ++k) {
        const uint32_t s = 1 + wv + k * kScanWaves;
        if (s >= kBuckets) break;
        const uint32_t h = hv[k];
        uint32_t incl = h;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, kWave);
            if (lane >= uint32_t(o)) incl += v;
        }
        const uint32_t t = __shfl(incl, kWave - 1, kWave);
        const uint32_t b = __shfl(incl - h, z, kWave);
        if (lane == 0) {
            tot[s] = t;
            before[s] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int s = kBuckets - 1; s >= 1; --s) {
            start[s] = acc + before[s];
            acc += tot[s];
        }
        start[0] = acc;
        start[kBuckets] = acc;
    }
    __syncthreads();
}

template <int K>
constexpr int item_floats() { return 6 * K - 2; }   // x: 3K, g: 3(K - 1) (g_{K-1} = 0), offset

// B1: div's adjoint (-grad of tau1 GQ).  Voxels whose Dykstra input was
// feasible (SweepRecord 0: the projection was the identity) pass the adjoint
// through Pi_K; the others recompute x = Xi + tau2 grad V-bar as the solver
// does and go to the bucket of their sweep count, at the place kernel A
// reserved for the workgroup (same grid).
template <int K>
__global__ __launch_bounds__(kAdjThreads) void adj_xi_kernel(VolumeLines L, uint64_t wspan, const float *st,
                                                             const uint8_t *rec1, float tau1, float tau2,
                                                             const float *GQ, float *gs, XiQueue Q,
                                                             const uint32_t *hist, const uint32_t *offset)
{
    constexpr int KM = K - 1;
    __shared__ uint32_t start[kBuckets + 1], cnt[kBuckets], base[kBuckets];
    const uint32_t z = stripe_of(blockIdx.x, gridDim.x);
    uint32_t hv[kPerWave];
    bucket_loads(hist, hv);   // used after the voxel's work
    if (threadIdx.x < kBuckets) cnt[threadIdx.x] = 0;
    __syncthreads();
    const Voxel vx = voxel_of(L);
    float *gXi = gs + uint64_t(KM) * wspan;
    uint32_t sw = 0, rank = 0;
    float x[K][3], g[K][3];
    if (vx.ok) {
        const Chunk &ch = vx.ch;
        const uint32_t lane = threadIdx.x % kWave;
        const bool nx[3] = {vx.n + 1 < ch.len, ch.nr != kNone, ch.nc != kNone};
        const uint32_t wn[3] = {vx.w + 1, ch.nr + lane, ch.nc + lane};
        // GQ at the voxel and its +e_a neighbours, the last label's as minus
        // the sum of the others (GQ sums to 0 over the labels)
        float own[K], nbq[3][K], full[K][3];
        own[KM] = 0.0f;
        for (int a = 0; a < 3; ++a) nbq[a][KM] = 0.0f;
        for (int i = 0; i < KM; ++i) {
            own[i] = GQ[uint64_t(i) * wspan + vx.w];
            own[KM] -= own[i];
            for (int a = 0; a < 3; ++a) {
                nbq[a][i] = nx[a] ? GQ[uint64_t(i) * wspan + wn[a]] : 0.0f;
                nbq[a][KM] -= nbq[a][i];
            }
        }
        for (int i = 0; i < K; ++i)
            for (int a = 0; a < 3; ++a) full[i][a] = nx[a] ? -tau1 * (nbq[a][i] - own[i]) : 0.0f;
        for (int a = 0; a < 3; ++a) {
            for (int i = 0; i < KM; ++i) g[i][a] = gXi[uint64_t(3 * i + a) * wspan + vx.w] + full[i][a] - full[KM][a];
            g[KM][a] = 0.0f;
        }
        sw = min(uint32_t(rec1[vx.w]), uint32_t(kBuckets - 1));
        if (sw) {
            // recompute x = Xi + tau2 grad V-bar as the solver does
            const float *Vb = st + uint64_t(KM) * wspan, *Xi = st + uint64_t(2 * KM) * wspan;
            float vb[KM], vbn[3][KM], xo[KM][3];
            for (int i = 0; i < KM; ++i) {
                vb[i] = Vb[uint64_t(i) * wspan + vx.w];
                for (int a = 0; a < 3; ++a) {
                    vbn[a][i] = nx[a] ? Vb[uint64_t(i) * wspan + wn[a]] : 0.0f;
                    xo[i][a] = Xi[uint64_t(3 * i + a) * wspan + vx.w];
                }
            }
            xi_input<K>(vb, vbn, nx, xo, tau2, x);
            rank = atomicAdd(&cnt[sw], 1u);
        } else {
            xi_outputs<K>(wspan, vx.w, g, gXi);
        }
    }
    if (threadIdx.x < kBuckets) base[threadIdx.x] = offset[uint64_t(blockIdx.x) * kStripeWords + threadIdx.x];
    bucket_starts(hv, z, start);   // its barriers also publish base
    if (sw) {
        const uint32_t at = start[sw] + base[sw] + rank;
        if (at < Q.cap) {
            float *it = Q.items + uint64_t(at) * item_floats<K>();
            for (int i = 0; i < K; ++i)
                for (int a = 0; a < 3; ++a) {
                    it[3 * i + a] = x[i][a];
                    if (i < KM) it[3 * K + 3 * i + a] = g[i][a];
                }
            it[6 * K - 3] = __uint_as_float(vx.w);
        }
    }
}

// B2: the bucketed voxels, one per lane (longest first): Dykstra replayed from
// x recording each pair's h, then back-propagated.
template <int K>
__global__ __launch_bounds__(kAdjThreads) void adj_xi_replay_kernel(uint64_t wspan, DykstraLimits lim, float *gs,
                                                                    XiQueue Q, const uint32_t *hist)
{
    constexpr int KM = K - 1, NP = K * (K - 1) / 2;
    __shared__ uint32_t start[kBuckets + 1];
    uint32_t hv[kPerWave];
    bucket_loads(hist, hv);
    bucket_starts(hv, 0, start);
    const uint32_t total = min(start[kBuckets], Q.cap);
    float *gXi = gs + uint64_t(KM) * wspan;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const float *it = Q.items + uint64_t(t) * item_floats<K>();
        float x[K][3], g[K][3];
        for (int i = 0; i < K; ++i)
            for (int a = 0; a < 3; ++a) {
                x[i][a] = it[3 * i + a];
                g[i][a] = i < KM ? it[3 * K + 3 * i + a] : 0.0f;
            }
        // the item's bucket: the sweeps the forward's Dykstra took
        uint32_t bk = kBuckets - 1;
        while (bk > 1 && t >= start[bk - 1]) --bk;
        bool replayed = false;
        if constexpr (NP <= 3)
            if (bk <= uint32_t(kShortSweeps)) replayed = replay_short<K>(x, g, lim);
        if (!replayed) {
            if constexpr (NP <= 3) {
                // whole history unrolled (held in private memory)
                float hist[20][NP][3], r[NP][3];
#pragma unroll
                for (int pp = 0; pp < NP; ++pp) r[pp][0] = r[pp][1] = r[pp][2] = 0.0f;
                int sw = 0;
                bool run = true;
#pragma unroll
                for (int s = 0; s < 20; ++s) {
                    if (run) {
                        const float m2 = dykstra_sweep<K>(x, r, lim, hist[s]);
                        sw = s + 1;
                        run = !(m2 < lim.stol);
                    }
                }
                dykstra_back_unrolled<K>(sw, hist, g, lim);
            } else {
                float hist[20][NP][3];
                const int sw = dykstra_replay<K>(x, hist, lim);
                dykstra_back<K>(sw, hist, g, lim);
            }
        }
        xi_outputs<K>(wspan, __float_as_uint(it[6 * K - 3]), g, gXi);
    }
}

struct Buffers {
    // seg: the forward's working state (seg[0], the final state) and the
    // recompute states of even segments; seg2[1..]: those of odd segments
    std::vector<float *> ck, seg, seg2;
    float *gs = nullptr, *GQ = nullptr, *gu = nullptr, *items = nullptr;
    uint32_t *buckets = nullptr;   // ReplayBuckets: the two histograms
    uint32_t *offset = nullptr;    // ReplayBuckets: kStripeWords per workgroup of A and B1
    double *acc = nullptr;   // dc[K], dlambda, loss
};

void free_buffers(sct_ctx *ctx, Buffers &b)
{
    for (float *p : b.ck) sct_free(ctx, p);
    for (float *p : b.seg) sct_free(ctx, p);
    for (float *p : b.seg2) sct_free(ctx, p);
    for (void *p : {(void *)b.gs, (void *)b.GQ, (void *)b.gu, (void *)b.items, (void *)b.buckets, (void *)b.offset,
                    (void *)b.acc})
        sct_free(ctx, p);
    b = Buffers{};
}

// Seed of the reverse pass from dL/dV_i (i < K, fields with the layout of g):
// the stored components receive dV_i - dV_{K-1}.
template <int K>
struct DvFields {
    const float *p[K];
};

template <int K>
__global__ __launch_bounds__(kAdjThreads) void adj_seed_kernel(VolumeLines L, uint64_t wspan, DvFields<K> dv, float *gV)
{
    const Voxel vx = voxel_of(L);
    if (!vx.ok) return;
    const float last = dv.p[K - 1][vx.uo];
    for (int i = 0; i < K - 1; ++i) gV[uint64_t(i) * wspan + vx.w] = dv.p[i][vx.uo] - last;
}

// V_i (i < K) of a state into fields with the layout of g; V_{K-1} = 1 - sum.
template <int K>
struct VFields {
    float *p[K];
};

template <int K>
__global__ __launch_bounds__(kAdjThreads) void adj_export_v_kernel(VolumeLines L, uint64_t wspan, const float *st,
                                                                   VFields<K> out)
{
    const Voxel vx = voxel_of(L);
    if (!vx.ok) return;
    float last = 1.0f;
    for (int i = 0; i < K - 1; ++i) {
        const float v = st[uint64_t(i) * wspan + vx.w];
        last -= v;
        out.p[i][vx.uo] = v;
    }
    out.p[K - 1][vx.uo] = last;
}

struct TapeBase {
    virtual ~TapeBase() = default;
    virtual const sct_field *input() const = 0;
    // seeds: target (MSE loss) or dV[K]
    virtual int backward(const sct_field *target, const sct_field *const *dV, double *loss, double *grad_c,
                         double *grad_lambda, sct_field **grad_u, float *ms) = 0;
};

// The forward pass with its checkpoints, and any number of reverse passes
// from them.
template <int K>
struct Tape : TapeBase {
    sct_ctx *ctx;
    const sct_field *g;
    int iters, ck, nck = 0, seg_len = 0;
    AdjLayout Ly;
    uint64_t S = 1, state = 0, nacc = 0;
    Buffers B;
    XiQueue Q{};
    OneSolve<K> prm;
    DykstraLimits lim;
    float tau1, tau2;
    uint32_t fgrid = 0, xgrid = 0, agrid = 0, rgrid = 0;
    float *final = nullptr;   // the state after the last iteration (seg[0], never recomputed into)
    // The reverse pass recomputes segment k - 1 on `side` while segment k's
    // reverse kernels run on the context's stream; events order the two.
    hipStream_t side = nullptr;
    hipEvent_t recomputed[2] = {}, reversed[2] = {}, started = nullptr;
    int rc = SCT_OK;

    Tape(sct_ctx *c, const sct_field *gf, const float *cv, float lambda, float t1, float t2, int it, int ckp)
        : ctx(c), g(gf), iters(it), ck(ckp), lim(dykstra_limits()), tau1(t1), tau2(t2)
    {
        for (int i = 0; i < K; ++i) prm.p.c[i] = cv[i];
        prm.p.lambda = lambda;
    }
    ~Tape() override
    {
        if (side) (void)hipStreamSynchronize(side);
        free_buffers(ctx, B);
        for (hipEvent_t e : {recomputed[0], recomputed[1], reversed[0], reversed[1], started})
            if (e) (void)hipEventDestroy(e);
        if (side) (void)hipStreamDestroy(side);
    }
    const sct_field *input() const override { return g; }

    void chk(hipError_t e, const char *what)
    {
        if (e != hipSuccess && !rc) rc = sct_hip_fail(ctx, e, what);
    }
    const float *u() const { return (const float *)g->data; }
    // The SweepRecord of the step that produced a state, after its 5 (K - 1)
    // components.
    uint8_t *rec(const float *s) const { return (uint8_t *)(s + 5 * uint64_t(K - 1) * S); }
    void step(float *s)
    {
        hipStream_t st = ctx->stream;
        const auto a = state_args<K>(s, u(), S);
        potts_xi_kernel<K, StorageF32, VolumeLines, kXiBlockWaves, true>
            <<<xgrid, kXiThreads, 0, st>>>(Ly.fwd, a, tau2, lim, SweepRecord{rec(s)});
        potts_v_kernel<K, StorageF32, false><<<fgrid, 256, 0, st>>>(Ly.fwd, a, prm, tau1, nullptr);
    }
    // One iteration from state src into state dst (src unchanged): the Xi
    // step writes dst's Xi, the V step reads it with src's V.
    void step_into(const float *src, float *dst) { step_into(src, dst, ctx->stream); }
    void step_into(const float *src, float *dst, hipStream_t st)
    {
        const auto in = state_args<K>(const_cast<float *>(src), u(), S);
        const auto out = state_args<K>(dst, u(), S);
        auto ax = in;
        for (int j = 0; j < 3 * (K - 1); ++j) ax.xi_out[j] = out.xi[j];
        auto av = in;
        for (int j = 0; j < 3 * (K - 1); ++j) av.xi[j] = out.xi[j];
        for (int i = 0; i < K - 1; ++i) {
            av.V_out[i] = out.V[i];
            av.vbar_out[i] = out.vbar[i];
        }
        potts_xi_kernel<K, StorageF32, VolumeLines, kXiBlockWaves, true>
            <<<xgrid, kXiThreads, 0, st>>>(Ly.fwd, ax, tau2, lim, SweepRecord{rec(dst)});
        potts_v_kernel<K, StorageF32, false><<<fgrid, 256, 0, st>>>(Ly.fwd, av, prm, tau1, nullptr);
    }

    int allocate()
    {
        const Geometry &geo = g->geo;
        const uint32_t n = geo.n_images, ws = sct_get_stride(n);
        Ly.fwd = VolumeLines::make(geo.d_nbr, geo.lines, n, (n + 127) / 128, ws, geo.nstride, 0);
        Ly.one = VolumeLines::make(geo.d_nbr, geo.lines, n, (n + 63) / 64, ws, geo.nstride, 0);
        Ly.wspan = uint64_t(geo.lines) * ws;
        S = std::max<uint64_t>(1, Ly.wspan);
        state = 5 * uint64_t(K - 1) * S + (S + 3) / 4;   // components, then the SweepRecord
        nck = (iters + ck - 1) / ck;
        seg_len = std::min(ck, iters);
        auto alloc = [&](float *&p, uint64_t elems) {
            if (!rc) rc = sct_alloc(ctx, (void **)&p, sizeof(float) * elems);
        };
        B.ck.assign(size_t(nck), nullptr);
        // seg[0]: the forward's working state, which ends as the final state;
        // seg[1 .. seg_len - 1]: the states a segment recomputes between its
        // two checkpoints
        B.seg.assign(size_t(seg_len), nullptr);
        B.seg2.assign(size_t(nck > 1 ? seg_len : 1), nullptr);   // [0] unused
        for (auto &p : B.ck) alloc(p, state);
        for (auto &p : B.seg) alloc(p, state);
        for (size_t j = 1; j < B.seg2.size(); ++j) alloc(B.seg2[j], state);
        if (!rc && hipStreamCreateWithFlags(&side, hipStreamNonBlocking) != hipSuccess) {
            side = nullptr;
            rc = sct_fail(ctx, SCT_ERR_HIP, "hipStreamCreate failed");
        }
        for (hipEvent_t *e : {&recomputed[0], &recomputed[1], &reversed[0], &reversed[1], &started})
            if (!rc && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) {
                *e = nullptr;
                rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventCreate failed");
            }
        alloc(B.gs, 4 * uint64_t(K - 1) * S);   // gV | gXi
        alloc(B.GQ, uint64_t(K - 1) * S);
        alloc(B.gu, S);
        agrid = uint32_t((Ly.one.count() * kWave + kAdjThreads - 1) / kAdjThreads);
        // per-workgroup sums: dc[K], dlambda of every A launch, then the loss
        nacc = uint64_t(agrid) * (K + 2);
        if (!rc) rc = sct_alloc(ctx, (void **)&B.acc, sizeof(double) * std::max<uint64_t>(1, nacc));
        // replay items: at most one per voxel slot of the one-voxel-per-lane grid
        Q = XiQueue{nullptr, agrid * uint32_t(kAdjThreads)};
        alloc(B.items, uint64_t(Q.cap) * item_floats<K>());
        if (!rc) rc = sct_alloc(ctx, (void **)&B.buckets, sizeof(uint32_t) * kBucketWords);
        if (!rc) rc = sct_alloc(ctx, (void **)&B.offset, sizeof(uint32_t) * kStripeWords * uint64_t(agrid));
        Q.items = B.items;
        rgrid = std::max<uint32_t>(1, ctx->cu_count * 4);
        fgrid = wave_grid(Ly.fwd.count());
        xgrid = xi_grid<StorageF32>(Ly.fwd.count());
        return rc;
    }

    // iters iterations from the nearest-label start.  State t = j ck is
    // stepped straight into checkpoint j (no copies); the others are stepped
    // in seg[0], which ends holding the final state.
    int forward(float *ms)
    {
        hipStream_t st = ctx->stream;
        hipEvent_t ev[2] = {};
        for (auto &e : ev)
            if (!rc && hipEventCreate(&e) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventCreate failed");
        float *work = B.seg.empty() ? nullptr : B.seg[0];
        float *cur = B.ck.empty() ? nullptr : B.ck[0];
        if (!rc) {
            chk(hipMemsetAsync(cur, 0, sizeof(float) * state, st), "memset");
            chk(hipMemsetAsync(work, 0, sizeof(float) * state, st), "memset");
            chk(hipEventRecord(ev[0], st), "hipEventRecord");
            potts_init_kernel<K, StorageF32><<<fgrid, 256, 0, st>>>(Ly.fwd, state_args<K>(cur, u(), S), prm);
            for (int it = 0; it < iters && !rc; ++it) {
                const int t = it + 1;
                float *dst = (t % ck == 0 && t / ck < nck) ? B.ck[size_t(t / ck)] : work;
                if (dst == cur)
                    step(cur);
                else
                    step_into(cur, dst);
                cur = dst;
            }
            chk(hipEventRecord(ev[1], st), "hipEventRecord");
            chk(hipGetLastError(), "adjoint forward launch");
            chk(hipStreamSynchronize(st), "adjoint forward");
        }
        if (!rc && ms && hipEventElapsedTime(ms, ev[0], ev[1]) != hipSuccess)
            rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventElapsedTime failed");
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        final = cur;
        return rc;
    }

    int export_v(sct_field **V)
    {
        const Geometry &geo = g->geo;
        VFields<K> out{};
        for (int i = 0; i < K && !rc; ++i) {
            rc = sct_field_like(ctx, g, SCT_F32, &V[i]);
            if (!rc) {
                out.p[i] = (float *)V[i]->data;
                chk(hipMemsetAsync(out.p[i], 0, sizeof(float) * geo.span(), ctx->stream), "memset");
            }
        }
        if (!rc) {
            adj_export_v_kernel<K><<<agrid, kAdjThreads, 0, ctx->stream>>>(Ly.one, S, final, out);
            chk(hipGetLastError(), "adjoint V export");
            chk(hipStreamSynchronize(ctx->stream), "adjoint V export");
        }
        return rc;
    }

    int backward(const sct_field *target, const sct_field *const *dV, double *loss, double *grad_c,
                 double *grad_lambda, sct_field **grad_u, float *ms) override
    {
        const Geometry &geo = g->geo;
        hipStream_t st = ctx->stream;
        const uint32_t n = geo.n_images, ws = sct_get_stride(n);
        hipEvent_t ev[2] = {};
        for (auto &e : ev)
            if (!rc && hipEventCreate(&e) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventCreate failed");
        if (rc) {
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
            return rc;
        }
        chk(hipMemsetAsync(B.gs, 0, sizeof(float) * 4 * (K - 1) * S, st), "memset");
        chk(hipMemsetAsync(B.gu, 0, sizeof(float) * S, st), "memset");
        chk(hipMemsetAsync(B.acc, 0, sizeof(double) * std::max<uint64_t>(1, nacc), st), "memset");
        chk(hipMemsetAsync(B.buckets, 0, sizeof(uint32_t) * kBucketWords, st), "memset");
        if (target) {
            const float wgt = float(1.0 / (double(std::max<uint64_t>(1, geo.data_size)) * K));
            adj_loss_kernel<K><<<agrid, kAdjThreads, 0, st>>>(Ly.one, S, final, (const float *)target->data, wgt,
                                                               B.gs, B.acc + uint64_t(agrid) * (K + 1));
        } else {
            DvFields<K> d{};
            for (int i = 0; i < K; ++i) d.p[i] = (const float *)dV[i]->data;
            adj_seed_kernel<K><<<agrid, kAdjThreads, 0, st>>>(Ly.one, S, d, B.gs);
        }
        // reverse pass, one checkpoint segment at a time; segment k's states
        // t0 .. t1 are checkpoint k, the states between (recomputed from it
        // into the buffer set of k's parity), and t1 = checkpoint k + 1 or
        // the final state.  Segment k - 1 is recomputed on the side stream
        // while segment k's reverse kernels run.
        chk(hipEventRecord(ev[0], st), "hipEventRecord");
        auto set_of = [&](int k) -> const std::vector<float *> & { return (k % 2) ? B.seg2 : B.seg; };
        // states a segment recomputes (none when checkpoints are every iteration)
        auto inner = [&](int k) { return std::min(iters, k * ck + ck) - k * ck - 1; };
        auto recompute = [&](int k) {
            if (inner(k) < 1) return;
            const int t0 = k * ck, m = std::min(iters, t0 + ck) - t0;
            const std::vector<float *> &set = set_of(k);
            const float *prev = B.ck[size_t(k)];
            for (int j = 1; j < m; ++j) {
                step_into(prev, set[size_t(j)], side);
                prev = set[size_t(j)];
            }
            chk(hipEventRecord(recomputed[k % 2], side), "hipEventRecord");
        };
        // the side stream starts after this call's set-up (and any earlier
        // reverse pass that read the recompute buffers)
        if (inner(0) >= 1) {
            chk(hipEventRecord(started, st), "hipEventRecord");
            chk(hipStreamWaitEvent(side, started, 0), "hipStreamWaitEvent");
        }
        if (!rc) recompute(nck - 1);
        int par = 0;   // which bucket histogram the current reverse step uses
        for (int k = nck - 1; k >= 0 && !rc; --k) {
            if (k >= 1 && inner(k - 1) >= 1) {
                // segment k - 1 reuses segment k + 1's buffers: wait for its reverse
                if (k + 1 < nck && inner(k + 1) >= 1)
                    chk(hipStreamWaitEvent(side, reversed[(k + 1) % 2], 0), "hipStreamWaitEvent");
                recompute(k - 1);
            }
            if (inner(k) >= 1) chk(hipStreamWaitEvent(st, recomputed[k % 2], 0), "hipStreamWaitEvent");
            const int t0 = k * ck, t1 = std::min(iters, t0 + ck), m = t1 - t0;
            std::vector<const float *> sg(size_t(m) + 1);
            sg[0] = B.ck[size_t(k)];
            for (int j = 1; j < m; ++j) sg[size_t(j)] = set_of(k)[size_t(j)];
            sg[size_t(m)] = k + 1 < nck ? B.ck[size_t(k) + 1] : final;
            for (int t = t1 - 1; t >= t0; --t) {
                const float *s0 = sg[size_t(t - t0)], *s1 = sg[size_t(t - t0) + 1];
                uint32_t *const bw = B.buckets;
                const ReplayBuckets rb{bw + par * kCounters, bw + (1 - par) * kCounters, B.offset};
                adj_v_kernel<K><<<agrid, kAdjThreads, 0, st>>>(Ly.one, S, s0, s1, rec(s1), u(), prm.p, tau1, tau2,
                                                                B.gs, B.GQ, B.gu, B.acc, rb);
                adj_xi_kernel<K><<<agrid, kAdjThreads, 0, st>>>(Ly.one, S, s0, rec(s1), tau1, tau2, B.GQ, B.gs, Q,
                                                                 rb.hist_cur, rb.offset);
                adj_xi_replay_kernel<K><<<rgrid, kAdjThreads, 0, st>>>(S, lim, B.gs, Q, rb.hist_cur);
                par = 1 - par;
            }
            if (inner(k) >= 1) chk(hipEventRecord(reversed[k % 2], st), "hipEventRecord");
        }
        chk(hipEventRecord(ev[1], st), "hipEventRecord");
        chk(hipGetLastError(), "adjoint launch");
        std::vector<double> acc(nacc);
        (void)hipStreamSynchronize(st);
        if (!rc && nacc) {
            hipError_t e = hipMemcpyAsync(acc.data(), B.acc, sizeof(double) * nacc, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "adjoint");
        }
        if (!rc && grad_u) {
            rc = sct_field_like(ctx, g, SCT_F32, grad_u);
            if (!rc) {
                hipError_t e = hipMemsetAsync((*grad_u)->data, 0, sizeof(float) * geo.span(), st);
                if (e == hipSuccess && geo.lines)
                    e = hipMemcpy2DAsync((*grad_u)->data, sizeof(float) * geo.nstride, B.gu, sizeof(float) * ws,
                                         sizeof(float) * n, geo.lines, hipMemcpyDeviceToDevice, st);
                if (e == hipSuccess) e = hipStreamSynchronize(st);
                if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "adjoint grad_u");
            }
        }
        if (!rc && ms && hipEventElapsedTime(ms, ev[0], ev[1]) != hipSuccess)
            rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventElapsedTime failed");
        for (auto e : ev) (void)hipEventDestroy(e);
        if (rc) return rc;
        double sums[8] = {};
        for (uint32_t b = 0; b < agrid; ++b) {
            for (int i = 0; i <= K; ++i) sums[i] += acc[size_t(b) * (K + 1) + size_t(i)];
            sums[K + 1] += acc[size_t(agrid) * (K + 1) + b];
        }
        for (int i = 0; i < K; ++i) grad_c[i] = sums[i];
        *grad_lambda = sums[K];
        if (loss) *loss = sums[K + 1];
        return SCT_OK;
    }
};

template <int K>
TapeBase *new_tape(sct_ctx *ctx, const sct_field *g, const float *c, float lambda, float tau1, float tau2, int iters,
                   int ck, sct_field **V, float *forward_ms, int &rc)
{
    auto *t = new Tape<K>(ctx, g, c, lambda, tau1, tau2, iters, ck);
    rc = t->allocate();
    if (!rc) rc = t->forward(forward_ms);
    if (!rc && V) rc = t->export_v(V);
    if (rc) {
        delete t;
        return nullptr;
    }
    return t;
}

TapeBase *make_tape(int K, sct_ctx *ctx, const sct_field *g, const float *c, float lambda, float tau1, float tau2,
                    int iters, int ck, sct_field **V, float *forward_ms, int &rc)
{
    switch (K) {
    case 2: return new_tape<2>(ctx, g, c, lambda, tau1, tau2, iters, ck, V, forward_ms, rc);
    case 3: return new_tape<3>(ctx, g, c, lambda, tau1, tau2, iters, ck, V, forward_ms, rc);
    case 4: return new_tape<4>(ctx, g, c, lambda, tau1, tau2, iters, ck, V, forward_ms, rc);
    case 5: return new_tape<5>(ctx, g, c, lambda, tau1, tau2, iters, ck, V, forward_ms, rc);
    default: return new_tape<6>(ctx, g, c, lambda, tau1, tau2, iters, ck, V, forward_ms, rc);
    }
}

bool same_layout(const sct_field *a, const sct_field *b)
{
    const Geometry &x = a->geo, &y = b->geo;
    return x.n_images == y.n_images && x.height == y.height && x.width == y.width && x.nstride == y.nstride &&
           x.col_begin == y.col_begin && x.col_end == y.col_end && x.base == y.base && x.start_row == y.start_row;
}

int check_forward_args(sct_ctx *ctx, const sct_field *g, int K, const float *c, int iters, int checkpoint_every)
{
    if (!g || !c || K < 2 || K > 6 || iters < 1 || checkpoint_every < 0)
        return sct_fail(ctx, SCT_ERR_INVALID, "invalid adjoint arguments");
    if (g->dtype != SCT_F32) return sct_fail(ctx, SCT_ERR_INVALID, "the adjoint needs a float field");
    return SCT_OK;
}

int checkpoint_interval(int iters, int checkpoint_every)
{
    return checkpoint_every > 0 ? checkpoint_every : std::max(1, int(std::lround(std::sqrt(double(iters)))));
}

}  // namespace

struct sct_potts_tape {
    sct_ctx *ctx;
    int K;
    TapeBase *t;
};

extern "C" int sct_potts_tape_forward(sct_ctx *ctx, const sct_field *g, int K, const float *c, float lambda,
                                      float tau1, float tau2, int iters, int checkpoint_every, sct_field **V,
                                      sct_potts_tape **tape, float *forward_ms)
{
    if (!ctx) return SCT_ERR_INVALID;
    if (!tape) return sct_fail(ctx, SCT_ERR_INVALID, "null tape");
    *tape = nullptr;
    if (V)
        for (int i = 0; i < K && i < 6; ++i) V[i] = nullptr;
    int rc = check_forward_args(ctx, g, K, c, iters, checkpoint_every);
    if (rc) return rc;
    TapeBase *t = make_tape(K, ctx, g, c, lambda, tau1, tau2, iters, checkpoint_interval(iters, checkpoint_every), V,
                            forward_ms, rc);
    if (!t) {
        if (V)
            for (int i = 0; i < K; ++i) {
                sct_field_destroy(V[i]);
                V[i] = nullptr;
            }
        return rc;
    }
    *tape = new sct_potts_tape{ctx, K, t};
    return SCT_OK;
}

extern "C" int sct_potts_tape_backward(sct_potts_tape *tape, const sct_field *const *dV, double *grad_c,
                                       double *grad_lambda, sct_field **grad_u, float *backward_ms)
{
    if (!tape) return SCT_ERR_INVALID;
    sct_ctx *ctx = tape->ctx;
    if (grad_u) *grad_u = nullptr;
    if (!dV || !grad_c || !grad_lambda) return sct_fail(ctx, SCT_ERR_INVALID, "invalid adjoint arguments");
    for (int i = 0; i < tape->K; ++i)
        if (!dV[i] || dV[i]->dtype != SCT_F32 || !same_layout(dV[i], tape->t->input()))
            return sct_fail(ctx, SCT_ERR_INVALID, "dV must be K float fields with the layout of g");
    return tape->t->backward(nullptr, dV, nullptr, grad_c, grad_lambda, grad_u, backward_ms);
}

extern "C" void sct_potts_tape_destroy(sct_potts_tape *tape)
{
    if (!tape) return;
    (void)hipStreamSynchronize(tape->ctx->stream);
    delete tape->t;
    delete tape;
}

extern "C" int sct_potts_adjoint(sct_ctx *ctx, const sct_field *g, const sct_field *target, int K, const float *c,
                                 float lambda, float tau1, float tau2, int iters, int checkpoint_every,
                                 double *loss, double *grad_c, double *grad_lambda, sct_field **grad_u,
                                 float *times_ms)
{
    if (!ctx) return SCT_ERR_INVALID;
    if (grad_u) *grad_u = nullptr;
    if (!target || !loss || !grad_c || !grad_lambda) return sct_fail(ctx, SCT_ERR_INVALID, "invalid adjoint arguments");
    int rc = check_forward_args(ctx, g, K, c, iters, checkpoint_every);
    if (rc) return rc;
    if (target->dtype != SCT_F32 || !same_layout(target, g))
        return sct_fail(ctx, SCT_ERR_INVALID, "target must be a float field with the layout of g");
    TapeBase *t = make_tape(K, ctx, g, c, lambda, tau1, tau2, iters, checkpoint_interval(iters, checkpoint_every),
                            nullptr, times_ms, rc);
    if (!t) return rc;
    rc = t->backward(target, nullptr, loss, grad_c, grad_lambda, grad_u, times_ms ? times_ms + 1 : nullptr);
    delete t;
    return rc;
}
