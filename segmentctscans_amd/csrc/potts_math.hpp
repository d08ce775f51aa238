// potts_math.hpp -- per-voxel arithmetic of the Potts primal-dual iteration.
//
// Operation order follows potts-model/potts-min-partition.cpp; fused
// multiply-adds are explicit (__fmaf_rn) and everything else is unfused
// (the library is compiled with -ffp-contract=off), so results are a
// deterministic function of the inputs and match the FP32 host mirror bit
// for bit.  Division and square root are IEEE correctly rounded.
#pragma once

#include <hip/hip_runtime.h>

namespace sct {

// Thresholds that let the kernel skip square roots without changing results:
//   sqrtf(ss) > 1          <=>  ss > t1      (t1: largest float with sqrtf <= 1)
//   (double)sqrtf(m) < 1e-6 <=>  m < stol    (stol: smallest float with
//                                             (double)sqrtf >= 1e-6)
// Both follow from sqrtf being correctly rounded and monotone; the host
// computes them once by bisection over float bit patterns.
// Correctly rounded square root for finite x >= 1.  The hardware estimate
// (within 1 ulp) moves to a neighbouring float when the residual x - s s' says
// that float is closer: the correction the compiler applies to sqrtf, without
// the scaling it adds for denormal and tiny inputs, which cannot occur here.
__device__ __forceinline__ float sqrt_cr_ge1(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sdn = __uint_as_float(__float_as_uint(s) - 1u);
    const float sup = __uint_as_float(__float_as_uint(s) + 1u);
    float r = (__fmaf_rn(-sdn, s, x) <= 0.0f) ? sdn : s;
    r = (__fmaf_rn(-sup, s, x) > 0.0f) ? sup : r;
    return r;
}

// Correctly rounded a / b for normal a, b whose quotient and intermediate
// products stay far from overflow and underflow (here 2^-23 <= a < b): a
// refined reciprocal, a quotient and two residual corrections, the sequence
// the compiler emits for division minus its range scaling and special-case
// fix-up.
__device__ __forceinline__ float div_cr_normal(float a, float b)
{
    const float r0 = __builtin_amdgcn_rcpf(b);
    const float r = __fmaf_rn(__fmaf_rn(-b, r0, 1.0f), r0, r0);
    const float q0 = a * r;
    const float q1 = __fmaf_rn(__fmaf_rn(-b, q0, a), r, q0);
    return __fmaf_rn(__fmaf_rn(-b, q1, a), r, q1);
}

// x / D, correctly rounded, for D in {3, 5, 7}: q0 = x RN(1/D), then one
// residual correction, r = x - D q0 (exact in one FMA), q = q0 + r RN(1/D).
// Equal bit for bit to the IEEE quotient for every finite x except -0 (which
// gives +0); checked exhaustively over all float patterns on the host and by
// sct_selftest_cr_math on the device.  Other D divide (exact for powers of 2).
template <int D>
__host__ __device__ __forceinline__ float div_small(float x)
{
    if constexpr (D == 3 || D == 5 || D == 7) {
        constexpr float c = 1.0f / float(D);
        const float q0 = x * c;
        return __fmaf_rn(__fmaf_rn(-float(D), q0, x), c, q0);
    } else {
        return x / float(D);
    }
}

struct DykstraLimits {
    float t1, stol;
};

template <int K>
__device__ __forceinline__ void sort_desc(float (&u)[K])
{
    // Sorting networks of potts-min-partition.cpp:251-290, insertion sort for
    // K = 6 (:318-323).  ce(a, b) puts the larger value at a.
    auto ce = [&](int a, int b) {
        const float x = u[a], y = u[b];
        u[a] = fmaxf(x, y);
        u[b] = fminf(x, y);
    };
    if constexpr (K == 2) {
        ce(0, 1);
    } else if constexpr (K == 3) {
        ce(0, 2); ce(0, 1); ce(1, 2);
    } else if constexpr (K == 4) {
        ce(0, 2); ce(1, 3); ce(0, 1); ce(2, 3); ce(1, 2);
    } else if constexpr (K == 5) {
        ce(0, 3); ce(1, 4); ce(0, 2); ce(1, 3); ce(0, 1); ce(2, 4); ce(1, 2); ce(3, 4); ce(2, 3);
    } else {
#pragma unroll
        for (int i = 1; i < K; ++i)
#pragma unroll
            for (int jj = 0; jj < i; ++jj) ce(i - 1 - jj, i - jj);
    }
}

// Euclidean projection onto the probability simplex, potts-min-partition.cpp:292-347.
// act (optional) receives the active set {i : q_i > theta}.
template <int K>
__device__ __forceinline__ void project_simplex(float (&q)[K], bool (*act)[K] = nullptr)
{
    float u[K];
#pragma unroll
    for (int i = 0; i < K; ++i) u[i] = q[i];
    sort_desc<K>(u);
    float sum = 0.0f, theta = 0.0f;
    auto candidate = [](float x, int i) {
        switch (i) {   // i is a constant once unrolled; sum - 1 is never -0
        case 0: return div_small<1>(x);
        case 1: return div_small<2>(x);
        case 2: return div_small<3>(x);
        case 3: return div_small<4>(x);
        case 4: return div_small<5>(x);
        default: return div_small<6>(x);
        }
    };
#pragma unroll
    for (int i = 0; i < K; ++i) {
        sum += u[i];
        // (rho_sum - 1) / rho equals the candidate computed at index rho - 1,
        // so the final division of :340 is the one already done here.
        const float th = candidate(sum - 1.0f, i);
        if (u[i] > th) theta = th;
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
        const float a = q[i] - theta;
        q[i] = (a < 0.0f) ? 0.0f : a;    // cuda::std::max(a, 0.0f)
        if (act) (*act)[i] = a > 0.0f;
    }
}

// One sweep of Dykstra's alternating projection onto {|q_i1 - q_i2| <= 1 for
// all pairs}, potts-min-partition.cpp:204-234.  r holds the correction of
// each pair.  Returns the largest squared correction norm of the sweep.
// hist (optional) receives the pre-projection difference h of every pair.
template <int K>
__device__ __forceinline__ float dykstra_sweep(float (&q)[K][3], float (&r)[K * (K - 1) / 2][3], DykstraLimits lim,
                                               float (*hist)[3] = nullptr)
{
    float m2 = 0.0f;
#pragma unroll
    for (int i1 = 0; i1 < K - 1; ++i1) {
#pragma unroll
        for (int i2 = i1 + 1; i2 < K; ++i2) {
            const int p = (2 * K - 1 - i1) * i1 / 2 + i2 - i1 - 1;
            const float h0 = (q[i2][0] - q[i1][0]) + r[p][0];
            const float h1 = (q[i2][1] - q[i1][1]) + r[p][1];
            const float h2 = (q[i2][2] - q[i1][2]) + r[p][2];
            const float ss = __fmaf_rn(h2, h2, __fmaf_rn(h1, h1, h0 * h0));
            if (hist) {
                hist[p][0] = h0;
                hist[p][1] = h1;
                hist[p][2] = h2;
            }
            float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
            if (ss > lim.t1) {
                // ss > t1 gives l > 1, so l - 1 >= 2^-23 and l - 1 < l
                const float l = sqrt_cr_ge1(ss);
                const float s = div_cr_normal(l - 1.0f, l);
                e0 = s * h0;
                e1 = s * h1;
                e2 = s * h2;
            }
            const float d0 = 0.5f * (e0 - r[p][0]);
            const float d1 = 0.5f * (e1 - r[p][1]);
            const float d2 = 0.5f * (e2 - r[p][2]);
            const float dn2 = __fmaf_rn(d2, d2, __fmaf_rn(d1, d1, d0 * d0));
            m2 = (dn2 < m2) ? m2 : dn2;
            q[i1][0] += d0; q[i1][1] += d1; q[i1][2] += d2;
            q[i2][0] -= d0; q[i2][1] -= d1; q[i2][2] -= d2;
            r[p][0] = e0; r[p][1] = e1; r[p][2] = e2;
        }
    }
    return m2;
}

// Dykstra's projection, potts-min-partition.cpp:202-238: at most 20 sweeps,
// stop when the largest correction norm of a sweep is below 1e-6.  Returns
// the sweeps performed.
template <int K>
__device__ __forceinline__ int dykstra(float (&q)[K][3], DykstraLimits lim)
{
    constexpr int NP = K * (K - 1) / 2;
    float r[NP][3];
#pragma unroll
    for (int p = 0; p < NP; ++p) r[p][0] = r[p][1] = r[p][2] = 0.0f;
    int sweep = 0;
    while (sweep < 20) {
        ++sweep;
        if (dykstra_sweep<K>(q, r, lim) < lim.stol) break;
    }
    return sweep;
}

// True when Dykstra stops after its first sweep: with zero corrections every
// pair is already feasible, the sweep changes q only by adding +0 to the
// first index of each pair (turning -0 into +0), and its correction norm is 0.
// In that case q is updated accordingly.  The sweep's differences are
// (q_i2 - q_i1) + 0; only their squares enter the test, which the + 0 (a sign
// of zero) does not change, so it is left out here.
template <int K>
__device__ __forceinline__ bool dykstra_first_sweep_done(float (&q)[K][3], DykstraLimits lim)
{
    bool ok = true;
#pragma unroll
    for (int i1 = 0; i1 < K - 1; ++i1) {
#pragma unroll
        for (int i2 = i1 + 1; i2 < K; ++i2) {
            const float h0 = q[i2][0] - q[i1][0];
            const float h1 = q[i2][1] - q[i1][1];
            const float h2 = q[i2][2] - q[i1][2];
            ok = ok && !(__fmaf_rn(h2, h2, __fmaf_rn(h1, h1, h0 * h0)) > lim.t1);
        }
    }
    if (ok) {
#pragma unroll
        for (int i = 0; i < K - 1; ++i) {
            q[i][0] += 0.0f;
            q[i][1] += 0.0f;
            q[i][2] += 0.0f;
        }
    }
    return ok;
}

// Input of the Xi projection for one voxel, potts-min-partition.cpp:135-200:
// q = Xi + tau2 grad V-bar over all K labels (the K-th components implicit).
//   vb[i]      V-bar at x (explicit components)
//   vbn[a][i]  V-bar at x + e_a, used when nx[a]
//   xi[i][a]   Xi at x
template <int K>
__device__ __forceinline__ void xi_input(const float (&vb)[K - 1], const float (&vbn)[3][K - 1], const bool (&nx)[3],
                                         const float (&xi)[K - 1][3], float tau2, float (&q)[K][3])
{
    float vf[K];
    float vk = 1.0f;
#pragma unroll
    for (int i = 0; i < K - 1; ++i) {
        vf[i] = vb[i];
        vk = vk - vb[i];
    }
    vf[K - 1] = vk;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float pn[K];
        if (nx[a]) {
            float vka = 1.0f;
#pragma unroll
            for (int i = 0; i < K - 1; ++i) {
                pn[i] = vbn[a][i];
                vka = vka - vbn[a][i];
            }
            pn[K - 1] = vka;
#pragma unroll
            for (int i = 0; i < K; ++i) pn[i] = pn[i] - vf[i];
        } else {
#pragma unroll
            for (int i = 0; i < K; ++i) pn[i] = 0.0f;
        }
        float xk = 0.0f;
#pragma unroll
        for (int i = 0; i < K - 1; ++i) {
            xk = xk - xi[i][a];
            q[i][a] = __fmaf_rn(pn[i], tau2, xi[i][a]);
        }
        q[K - 1][a] = __fmaf_rn(pn[K - 1], tau2, xk);
    }
}

// Xi update of one voxel, potts-min-partition.cpp:135-243: xi <- Pi_K(q).
template <int K>
__device__ __forceinline__ int xi_update(const float (&vb)[K - 1], const float (&vbn)[3][K - 1],
                                         const bool (&nx)[3], float (&xi)[K - 1][3], float tau2,
                                         DykstraLimits lim)
{
    float q[K][3];
    xi_input<K>(vb, vbn, nx, xi, tau2, q);
    const int sweeps = dykstra<K>(q, lim);
#pragma unroll
    for (int i = 0; i < K - 1; ++i)
#pragma unroll
        for (int a = 0; a < 3; ++a) xi[i][a] = q[i][a];
    return sweeps;
}

// Input of the simplex projection of the V step, potts-min-partition.cpp:397-467:
// q = V + tau1 (div Xi - lambda (u - c)^2) over all K labels; vf = V with its
// implicit K-th component.
template <int K>
__device__ __forceinline__ void v_input(const float (&xo)[K - 1][3], const float (&xp)[3][K - 1],
                                        const bool (&px)[3], const bool (&nx)[3], const float (&v)[K - 1],
                                        float u, const float (&c)[K], float lambda, float tau1, float (&vf)[K],
                                        float (&q)[K])
{
    float xof[K][3], xpf[3][K];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float xk = 0.0f, xkp = 0.0f;
#pragma unroll
        for (int i = 0; i < K - 1; ++i) {
            xof[i][a] = xo[i][a];
            xk = xk - xo[i][a];
            xpf[a][i] = xp[a][i];
            xkp = xkp - xp[a][i];
        }
        xof[K - 1][a] = xk;
        xpf[a][K - 1] = xkp;
    }
    float vk = 1.0f;
#pragma unroll
    for (int i = 0; i < K - 1; ++i) {
        vf[i] = v[i];
        vk = vk - v[i];
    }
    vf[K - 1] = vk;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        float div = 0.0f;
        if (px[0]) div -= xpf[0][i];
        if (nx[0]) div += xof[i][0];
        if (px[1]) div -= xpf[1][i];
        if (nx[1]) div += xof[i][1];
        if (px[2]) div -= xpf[2][i];
        if (nx[2]) div += xof[i][2];
        const float d = u - c[i];
        q[i] = __fmaf_rn(tau1, __fmaf_rn(-lambda, d * d, div), vf[i]);
    }
}

// V update of one voxel, potts-min-partition.cpp:397-480.
//   xo[i][a]   Xi at x;  xp[a][i] Xi at x - e_a (used when px[a])
//   v[i]       in: V at x, out: projected V;  vbar[i] out: 2 V+ - V
// Returns sum_k (V+_k - V_k)^2 over all K components.
template <int K>
__device__ __forceinline__ float v_update(const float (&xo)[K - 1][3], const float (&xp)[3][K - 1],
                                          const bool (&px)[3], const bool (&nx)[3], float (&v)[K - 1],
                                          float (&vbar)[K - 1], float u, const float (&c)[K],
                                          float lambda, float tau1)
{
    float vf[K], q[K];
    v_input<K>(xo, xp, px, nx, v, u, c, lambda, tau1, vf, q);
    project_simplex<K>(q);
#pragma unroll
    for (int i = 0; i < K - 1; ++i) {
        vbar[i] = 2.0f * q[i] - vf[i];
        v[i] = q[i];
    }
    float dq = 0.0f;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float d = q[k] - vf[k];
        dq = __fmaf_rn(d, d, dq);
    }
    return dq;
}

// Nearest label, strict < so the lowest index wins ties (potts-min-partition.cpp:81-87).
template <int K>
__device__ __forceinline__ int nearest_label(float u, const float (&c)[K])
{
    float best = 3.402823466e+38f;
    int cl = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float d = fabsf(c[k] - u);
        if (d < best) {
            best = d;
            cl = k;
        }
    }
    return cl;
}

// Argmax from 0 with strict >, lowest index wins ties (potts-min-partition.cpp:518-533).
template <int K>
__device__ __forceinline__ int argmax_label(const float (&v)[K - 1])
{
    float best = 0.0f, vk = 1.0f;
    int lab = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        float vi;
        if (i < K - 1) {
            vi = v[i];
            vk = vk - vi;
        } else {
            vi = vk;
        }
        if (vi > best) {
            best = vi;
            lab = i;
        }
    }
    return lab;
}

}  // namespace sct

namespace sct {

// Storage of Xi and V-bar.  Arithmetic is FP32 throughout; loads widen and
// stores round to nearest even.
struct StorageF32 {
    using T = float;
    static __device__ __forceinline__ float ld(T x) { return x; }
    static __device__ __forceinline__ T st(float f) { return f; }
};

struct StorageBF16 {
    using T = uint16_t;   // bfloat16 bit pattern
    static __device__ __forceinline__ float ld(T x) { return __uint_as_float(uint32_t(x) << 16); }
    // v_cvt_pk_bf16_f32: round to nearest even, NaN quietened with its top
    // payload bits kept; two adjacent conversions pack into one instruction
    static __device__ __forceinline__ T st(float f) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f)); }
};

struct StorageF16 {
    using T = _Float16;
    static __device__ __forceinline__ float ld(T x) { return float(x); }
    static __device__ __forceinline__ T st(float f) { return T(f); }   // v_cvt_f16_f32, round to nearest even
};

}  // namespace sct
