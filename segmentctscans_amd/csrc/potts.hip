// potts.hip -- Chambolle-Cremers-Pock Potts solver (potts-model/potts-min-partition.cpp).
//
// State per voxel, component-major in a work layout of `lines` lines of
// `wstride` floats (the input's lines, n re-based to the slab start):
//   Xi    3(K-1) components   (the K-th is -sum, never stored)
//   V     K-1                 (the K-th is 1 - sum)
//   V-bar K-1
// The data term lambda (u - c_i)^2 is recomputed from u every iteration.
#include "sct_internal.hpp"
#include "potts_math.hpp"
#include "potts.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace sct;

namespace {

size_t storage_bytes(int storage) { return storage == SCT_STORAGE_F32 ? 4 : 2; }

float bf16_to_float(uint16_t h)
{
    const uint32_t x = uint32_t(h) << 16;
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

float half_to_float(uint16_t h)
{
    _Float16 v;
    std::memcpy(&v, &h, 2);
    return float(v);
}

struct WorkGeom {
    const LineNbr *nbr;
    uint32_t lines;      // number of lines
    uint32_t len;        // slab length along n
    uint32_t chunks;     // ceil(len / 64)
    uint32_t vchunks;    // ceil(len / (64 VPL)), two-pass kernels
    uint32_t wstride;    // work line stride
    uint64_t wspan;      // lines * wstride (component stride)
    uint32_t in_stride;  // input line stride
    uint32_t s0;         // slab start in input coordinates
};

// One wave = 64 consecutive n of one line; 256-thread blocks in XCD runs of 64
// (about 11 lines of 1426 voxels).
__device__ __forceinline__ bool wave_voxel(const WorkGeom &g, uint32_t &line, uint32_t &n)
{
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x, 64);
    const uint64_t gw = (uint64_t(b) * blockDim.x + threadIdx.x) / kWave;
    line = uint32_t(gw / g.chunks);
    n = uint32_t(gw % g.chunks) * kWave + (threadIdx.x % kWave);
    return line < g.lines && n < g.len;
}

template <int K>
struct Labels {
    float c[K];
};

// init_v_kernel, potts-min-partition.cpp:56-94.
template <int K, typename St>
__global__ __launch_bounds__(256) void potts_init_kernel(WorkGeom g, const float *__restrict__ u,
                                                         float *__restrict__ V, typename St::T *__restrict__ vbar,
                                                         Labels<K> lab)
{
    uint32_t line, n;
    if (!wave_voxel(g, line, n)) return;
    const float x = u[uint64_t(line) * g.in_stride + g.s0 + n];
    const int cl = nearest_label<K>(x, lab.c);
    const uint64_t w = uint64_t(line) * g.wstride + n;
#pragma unroll
    for (int i = 0; i < K - 1; ++i) {
        const float o = (i == cl) ? 1.0f : 0.0f;
        V[i * g.wspan + w] = o;
        vbar[i * g.wspan + w] = St::st(o);
    }
}

// Per-component base pointers of the two-pass kernels; every component is
// addressed as base[comp] + 32-bit element offset (line * wstride + n).
template <int K, typename St>
struct TwoPassArgs {
    typename St::T *xi[3 * (K - 1)];
    typename St::T *vbar[K - 1];
    float *V[K - 1];
    const float *u;
};

// The two-pass kernels give every lane VPL consecutive voxels along n, so that
// each lane moves 8 bytes (FP32, VPL = 2) or 8 bytes (BF16 / FP16, VPL = 4)
// per array access: 16-bit loads and stores one element per lane run the
// memory system at half the rate of 32-bit ones.  A chunk is 64 lanes x VPL
// voxels of one line.  Lanes past the end of a line keep their vector inside
// the line's padding (wstride is a multiple of 32 >= len rounded up to VPL),
// which nothing reads for a valid voxel.
template <typename St>
constexpr int vpl_of() { return sizeof(typename St::T) == 4 ? 2 : 4; }

// VPL consecutive stored elements, one load or store.
template <typename St, int N>
struct Packed {
    using T = typename St::T;
    typedef T V __attribute__((ext_vector_type(N)));
    static __device__ __forceinline__ void ld(const T *p, float (&o)[N])
    {
        const V v = *reinterpret_cast<const V *>(p);
#pragma unroll
        for (int j = 0; j < N; ++j) o[j] = St::ld(v[j]);
    }
    static __device__ __forceinline__ void st(T *p, const float (&x)[N])
    {
        V v;
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = St::st(x[j]);
        *reinterpret_cast<V *>(p) = v;
    }
};

constexpr int kXiVoxelsPerLane = 4;  // per lane and workgroup: kXiVoxelsPerLane / VPL chunks per wave
constexpr int kXiItems = 512;        // work-list capacity per workgroup
constexpr int kXiWorkWaves = 4;      // waves that drain the work list

// grad_kernel, potts-min-partition.cpp:96-247: Xi <- Pi_K(Xi + tau2 grad V-bar),
// in place (each voxel reads its own Xi and its neighbours' V-bar only).
//   A  every lane forms the Dykstra input of its voxels; most are feasible and
//      Dykstra stops after one sweep, the rest join a work list in LDS;
//   B  the work list is drained one sweep per loop trip, each lane taking a new
//      item as soon as its item converges, the result going back to LDS;
//   C  every lane stores the Xi+ of its voxels, VPL at a time.
template <int K, typename St>
__global__ __launch_bounds__(256) void potts_xi_kernel(WorkGeom g, TwoPassArgs<K, St> a, float tau2,
                                                       DykstraLimits lim)
{
    constexpr int KM = K - 1, NP = K * (K - 1) / 2;
    constexpr int VPL = vpl_of<St>(), CPW = kXiVoxelsPerLane / VPL;
    using P = Packed<St, VPL>;
    __shared__ float qb[kXiItems][K][3];   // inputs, then results, of the queued voxels
    __shared__ int nq, next;
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    if (threadIdx.x == 0) {
        nq = 0;
        next = 0;
    }
    __syncthreads();
    const uint64_t total = uint64_t(g.lines) * g.vchunks;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x, 16);
    float res[CPW][VPL][KM][3];
    int pos[CPW][VPL];
    uint32_t wofs[CPW];
    int nv[CPW];   // valid voxels of this lane in chunk k
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
        const uint64_t gw = (uint64_t(blk) * (blockDim.x / kWave) + w) * CPW + k;
        const uint32_t line = uint32_t(gw / g.vchunks);
        const uint32_t n0 = uint32_t(gw % g.vchunks) * (kWave * VPL) + lane * VPL;
        nv[k] = (gw < total && n0 < g.len) ? int(min(uint32_t(VPL), g.len - n0)) : 0;
        wofs[k] = nv[k] ? line * g.wstride + n0 : 0;
#pragma unroll
        for (int j = 0; j < VPL; ++j) pos[k][j] = -1;
        if (gw >= total) continue;   // uniform over the wave
        const LineNbr nb = g.nbr[line];
        const uint32_t w0 = wofs[k];
        float vb[KM][VPL], vr[KM][VPL], vc[KM][VPL], x[KM][3][VPL], vn[KM][VPL];
#pragma unroll
        for (int i = 0; i < KM; ++i) {
#pragma unroll
            for (int j = 0; j < VPL; ++j) vb[i][j] = vr[i][j] = vc[i][j] = 0.0f;
            if (nv[k]) {
                P::ld(a.vbar[i] + w0, vb[i]);
                if (nb.nrow >= 0) P::ld(a.vbar[i] + uint32_t(nb.nrow) * g.wstride + n0, vr[i]);
                if (nb.ncol >= 0) P::ld(a.vbar[i] + uint32_t(nb.ncol) * g.wstride + n0, vc[i]);
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) P::ld(a.xi[3 * i + ax] + w0, x[i][ax]);
            }
            // V-bar at n + 1: the next element of the lane, the next lane's first,
            // or (last lane) the first of the next chunk
#pragma unroll
            for (int j = 0; j + 1 < VPL; ++j) vn[i][j] = vb[i][j + 1];
            vn[i][VPL - 1] = __shfl_down(vb[i][0], 1, kWave);
            if (lane == kWave - 1 && nv[k] == VPL && n0 + VPL < g.len) vn[i][VPL - 1] = St::ld(a.vbar[i][w0 + VPL]);
        }
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const bool valid = j < nv[k];
            const bool nx[3] = {n0 + j + 1 < g.len, nb.nrow >= 0, nb.ncol >= 0};
            float vbj[KM], vbn[3][KM], xj[KM][3];
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                vbj[i] = vb[i][j];
                vbn[0][i] = vn[i][j];
                vbn[1][i] = vr[i][j];
                vbn[2][i] = vc[i][j];
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) xj[i][ax] = x[i][ax][j];
            }
            float q[K][3];
            xi_input<K>(vbj, vbn, nx, xj, tau2, q);
            const bool queued = valid && !dykstra_first_sweep_done<K>(q, lim);
            const uint64_t mask = __ballot(queued);
            int p = 0;
            if (mask) {
                const int leader = __ffsll((unsigned long long)mask) - 1;
                if (lane == leader) p = atomicAdd(&nq, __popcll(mask));
                p = __shfl(p, leader, kWave) + __popcll(mask & ((1ull << lane) - 1ull));
            }
            if (queued && p < kXiItems) {
#pragma unroll
                for (int i = 0; i < K; ++i)
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) qb[p][i][ax] = q[i][ax];
                pos[k][j] = p;
            } else if (queued) {
                dykstra<K>(q, lim);   // work list full: solve in place
            }
#pragma unroll
            for (int i = 0; i < KM; ++i)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) res[k][j][i][ax] = valid ? q[i][ax] : 0.0f;
        }
    }
    __syncthreads();   // work list complete
    const int nwl = min(nq, kXiItems);
    if (w < kXiWorkWaves && w * kWave < nwl) {
        float q[K][3], r[NP][3];
        int sw = 0, t = 0;
        bool have = false, drained = false;
        while (true) {
            const bool need = !have && !drained;
            const uint64_t mask = __ballot(need);
            if (mask) {
                const int leader = __ffsll((unsigned long long)mask) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(&next, __popcll(mask));
                base = __shfl(base, leader, kWave);
                if (need) {
                    t = base + __popcll(mask & ((1ull << lane) - 1ull));
                    if (t < nwl) {
#pragma unroll
                        for (int i = 0; i < K; ++i)
#pragma unroll
                            for (int ax = 0; ax < 3; ++ax) q[i][ax] = qb[t][i][ax];
#pragma unroll
                        for (int pp = 0; pp < NP; ++pp) r[pp][0] = r[pp][1] = r[pp][2] = 0.0f;
                        sw = 0;
                        have = true;
                    } else {
                        drained = true;
                    }
                }
            }
            if (!__any(have)) break;
            if (have) {
                const float m2 = dykstra_sweep<K>(q, r, lim);
                ++sw;
                if (m2 < lim.stol || sw == 20) {
#pragma unroll
                    for (int i = 0; i < KM; ++i)
#pragma unroll
                        for (int ax = 0; ax < 3; ++ax) qb[t][i][ax] = q[i][ax];
                    have = false;
                }
            }
        }
    }
    __syncthreads();   // results of the work list in LDS
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
        if (!nv[k]) continue;
#pragma unroll
        for (int i = 0; i < KM; ++i)
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                float o[VPL];
#pragma unroll
                for (int j = 0; j < VPL; ++j) o[j] = pos[k][j] >= 0 ? qb[pos[k][j]][i][ax] : res[k][j][i][ax];
                P::st(a.xi[3 * i + ax] + wofs[k], o);
            }
    }
}

// div_kernel, potts-min-partition.cpp:349-488: V <- Pi_S(V + tau1 (div Xi - G)),
// V-bar <- 2 V+ - V.  With TRACE, the block maximum of sum_k dV_k^2 goes to *dq.
// One chunk per wave; 256-thread blocks in XCD runs of 64 / VPL (about 11 lines).
template <int K, typename St, bool TRACE>
__global__ __launch_bounds__(256) void potts_v_kernel(WorkGeom g, TwoPassArgs<K, St> a, Labels<K> lab,
                                                      float lambda, float tau1, float *dq)
{
    constexpr int KM = K - 1, VPL = vpl_of<St>();
    using P = Packed<St, VPL>;
    using PF = Packed<StorageF32, VPL>;
    const int lane = threadIdx.x % kWave;
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x, 64 / VPL);
    const uint64_t gw = (uint64_t(b) * blockDim.x + threadIdx.x) / kWave;
    float d = 0.0f;
    if (gw < uint64_t(g.lines) * g.vchunks) {   // uniform over the wave
        const uint32_t line = uint32_t(gw / g.vchunks);
        const uint32_t n0 = uint32_t(gw % g.vchunks) * (kWave * VPL) + lane * VPL;
        const int nv = n0 < g.len ? int(min(uint32_t(VPL), g.len - n0)) : 0;
        const LineNbr nb = g.nbr[line];
        const uint32_t w0 = line * g.wstride + n0;
        float xo[KM][3][VPL], xr[KM][VPL], xc[KM][VPL], xn[KM][VPL], v[KM][VPL], u[VPL];
#pragma unroll
        for (int i = 0; i < KM; ++i) {
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                xr[i][j] = xc[i][j] = v[i][j] = 0.0f;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) xo[i][ax][j] = 0.0f;
            }
            if (nv) {
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) P::ld(a.xi[3 * i + ax] + w0, xo[i][ax]);
                if (nb.prow >= 0) P::ld(a.xi[3 * i + 1] + uint32_t(nb.prow) * g.wstride + n0, xr[i]);
                if (nb.pcol >= 0) P::ld(a.xi[3 * i + 2] + uint32_t(nb.pcol) * g.wstride + n0, xc[i]);
                PF::ld(a.V[i] + w0, v[i]);
            }
            // Xi_n at n - 1: the previous element of the lane, the previous
            // lane's last, or (first lane) the last of the previous chunk
#pragma unroll
            for (int j = 1; j < VPL; ++j) xn[i][j] = xo[i][0][j - 1];
            xn[i][0] = __shfl_up(xo[i][0][VPL - 1], 1, kWave);
            if (lane == 0 && nv && n0 > 0) xn[i][0] = St::ld(a.xi[3 * i][w0 - 1]);
        }
#pragma unroll
        for (int j = 0; j < VPL; ++j) u[j] = j < nv ? a.u[line * g.in_stride + g.s0 + n0 + j] : 0.0f;
        float vb[KM][VPL];
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const bool px[3] = {n0 + j > 0, nb.prow >= 0, nb.pcol >= 0};
            const bool nx[3] = {n0 + j + 1 < g.len, nb.nrow >= 0, nb.ncol >= 0};
            float xoj[KM][3], xpj[3][KM], vj[KM], vbj[KM];
#pragma unroll
            for (int i = 0; i < KM; ++i) {
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) xoj[i][ax] = xo[i][ax][j];
                xpj[0][i] = xn[i][j];
                xpj[1][i] = xr[i][j];
                xpj[2][i] = xc[i][j];
                vj[i] = v[i][j];
            }
            const float dj = v_update<K>(xoj, xpj, px, nx, vj, vbj, u[j], lab.c, lambda, tau1);
            if (j < nv) d = fmaxf(d, dj);
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                v[i][j] = vj[i];
                vb[i][j] = vbj[i];
            }
        }
        if (nv) {
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                PF::st(a.V[i] + w0, v[i]);
                P::st(a.vbar[i] + w0, vb[i]);
            }
        }
    }
    if constexpr (TRACE) {
        __shared__ float red[4];
        d = wave_max(d);
        if ((threadIdx.x % kWave) == 0) red[threadIdx.x / kWave] = d;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = red[0];
            for (int i = 1; i < int(blockDim.x / kWave); ++i) m = fmaxf(m, red[i]);
            atomic_max_nonneg(dq, m);
        }
    }
}

// update_kernel, potts-min-partition.cpp:491-538: res = c[argmax V] on [r0, r1).
template <int K>
__global__ __launch_bounds__(256) void potts_labels_kernel(WorkGeom g, const float *__restrict__ V,
                                                           float *__restrict__ res, Labels<K> lab,
                                                           uint32_t r0, uint32_t r1)
{
    uint32_t line, n;
    if (!wave_voxel(g, line, n)) return;
    const uint32_t na = g.s0 + n;
    if (na < r0 || na >= r1) return;
    const uint64_t w = uint64_t(line) * g.wstride + n;
    float v[K - 1];
#pragma unroll
    for (int i = 0; i < K - 1; ++i) v[i] = V[i * g.wspan + w];
    res[uint64_t(line) * g.in_stride + na] = lab.c[argmax_label<K>(v)];
}

uint32_t grid_for(const WorkGeom &g)
{
    const uint64_t waves = uint64_t(g.lines) * g.chunks;
    return uint32_t((waves + 3) / 4);
}

}  // namespace

namespace {

WorkGeom make_geom(const sct_potts_solver *s)
{
    const Geometry &geo = s->g->geo;
    WorkGeom g;
    g.nbr = geo.d_nbr;
    g.lines = geo.lines;
    g.len = s->s1 - s->s0;
    g.chunks = (g.len + kWave - 1) / kWave;
    g.wstride = s->wstride;
    g.wspan = s->wspan;
    g.in_stride = geo.nstride;
    g.s0 = s->s0;
    return g;
}

template <int K>
int launch_init(sct_potts_solver *s)
{
    const WorkGeom g = make_geom(s);
    Labels<K> lab;
    for (int i = 0; i < K; ++i) lab.c[i] = s->c[i];
    if (g.lines && g.len) {
        const float *u = (const float *)s->g->data;
        if (s->storage == SCT_STORAGE_BF16)
            potts_init_kernel<K, StorageBF16><<<grid_for(g), 256, 0, s->ctx->stream>>>(g, u, s->V, (uint16_t *)s->vbar[0], lab);
        else if (s->storage == SCT_STORAGE_F16)
            potts_init_kernel<K, StorageF16><<<grid_for(g), 256, 0, s->ctx->stream>>>(g, u, s->V, (_Float16 *)s->vbar[0], lab);
        else
            potts_init_kernel<K, StorageF32><<<grid_for(g), 256, 0, s->ctx->stream>>>(g, u, s->V, (float *)s->vbar[0], lab);
        SCT_HIP(s->ctx, hipGetLastError());
    }
    s->cur = 0;
    SCT_HIP(s->ctx, hipMemsetAsync(s->xi[0], 0, storage_bytes(s->storage) * 3 * (K - 1) * s->wspan, s->ctx->stream));
    return SCT_OK;
}

template <int K, typename St>
int twopass_iterate(sct_potts_solver *s, const WorkGeom &g, int iters, float *dq_dev)
{
    using T = typename St::T;
    Labels<K> lab;
    for (int i = 0; i < K; ++i) lab.c[i] = s->c[i];
    const DykstraLimits lim = dykstra_limits();
    TwoPassArgs<K, St> a;
    for (int j = 0; j < 3 * (K - 1); ++j) a.xi[j] = (T *)s->xi[0] + j * s->wspan;
    for (int i = 0; i < K - 1; ++i) {
        a.vbar[i] = (T *)s->vbar[0] + i * s->wspan;
        a.V[i] = s->V + i * s->wspan;
    }
    a.u = (const float *)s->g->data;
    constexpr int VPL = vpl_of<St>();
    WorkGeom gv = g;
    gv.vchunks = (g.len + kWave * VPL - 1) / (kWave * VPL);
    const uint64_t waves = uint64_t(g.lines) * gv.vchunks;
    const uint64_t xi_waves = (waves + kXiVoxelsPerLane / VPL - 1) / (kXiVoxelsPerLane / VPL);
    const uint32_t xi_grid = uint32_t((xi_waves + 3) / 4);
    const uint32_t v_grid = uint32_t((waves + 3) / 4);
    for (int i = 0; i < iters; ++i) {
        potts_mark(s);
        potts_xi_kernel<K, St><<<xi_grid, 256, 0, s->ctx->stream>>>(gv, a, s->tau2, lim);
        potts_mark(s);
        if (dq_dev)
            potts_v_kernel<K, St, true><<<v_grid, 256, 0, s->ctx->stream>>>(gv, a, lab, s->lambda, s->tau1, dq_dev + i);
        else
            potts_v_kernel<K, St, false><<<v_grid, 256, 0, s->ctx->stream>>>(gv, a, lab, s->lambda, s->tau1, nullptr);
    }
    SCT_HIP(s->ctx, hipGetLastError());
    return SCT_OK;
}

template <int K>
int launch_iterate(sct_potts_solver *s, int iters, float *dq_dev)
{
    const WorkGeom g = make_geom(s);
    if (!g.lines || !g.len) return SCT_OK;
    if (s->kernel == SCT_KERNEL_FUSED) return fused_iterate(s, iters, dq_dev);
    switch (s->storage) {
    case SCT_STORAGE_BF16: return twopass_iterate<K, StorageBF16>(s, g, iters, dq_dev);
    case SCT_STORAGE_F16: return twopass_iterate<K, StorageF16>(s, g, iters, dq_dev);
    default: return twopass_iterate<K, StorageF32>(s, g, iters, dq_dev);
    }
}

template <int K>
int launch_labels(sct_potts_solver *s, float *res, uint32_t r0, uint32_t r1)
{
    const WorkGeom g = make_geom(s);
    if (!g.lines || !g.len || r0 >= r1) return SCT_OK;
    Labels<K> lab;
    for (int i = 0; i < K; ++i) lab.c[i] = s->c[i];
    potts_labels_kernel<K><<<grid_for(g), 256, 0, s->ctx->stream>>>(g, s->V, res, lab, r0, r1);
    SCT_HIP(s->ctx, hipGetLastError());
    return SCT_OK;
}

#define SCT_DISPATCH_K(K, fn, ...)              \
    switch (K) {                                \
    case 2: return fn<2>(__VA_ARGS__);          \
    case 3: return fn<3>(__VA_ARGS__);          \
    case 4: return fn<4>(__VA_ARGS__);          \
    case 5: return fn<5>(__VA_ARGS__);          \
    case 6: return fn<6>(__VA_ARGS__);          \
    default: return SCT_ERR_INVALID;            \
    }

int do_init(sct_potts_solver *s) { SCT_DISPATCH_K(s->K, launch_init, s); }
int do_iterate(sct_potts_solver *s, int iters, float *dq) { SCT_DISPATCH_K(s->K, launch_iterate, s, iters, dq); }

// do_iterate with the FP64 gap evaluated after every gap_every-th iteration of
// the slab (count: iterations already done on it) into gap_trace[count / gap_every - 1].
int iterate_with_gap(sct_potts_solver *s, int iters, float *dq, int &count, int capacity)
{
    const int every = s->opts.gap_every;
    if (every <= 0 || !s->opts.gap_trace) {
        count += iters;
        return do_iterate(s, iters, dq);
    }
    int done = 0;
    while (done < iters) {
        const int chunk = std::min(iters - done, every - count % every);
        int rc = do_iterate(s, chunk, dq ? dq + done : nullptr);
        if (rc) return rc;
        done += chunk;
        count += chunk;
        if (count % every == 0 && count / every <= capacity) {
            double E = 0.0, D = 0.0;
            if ((rc = potts_gap(s, &E, &D))) return rc;
            s->opts.gap_trace[count / every - 1] = E != 0.0 ? (E - D) / E : 0.0;
        }
    }
    return SCT_OK;
}
int do_labels(sct_potts_solver *s, float *res, uint32_t r0, uint32_t r1) { SCT_DISPATCH_K(s->K, launch_labels, s, res, r0, r1); }

// Bytes of solver state per voxel of a slab: Xi 3(K-1) and V-bar (K-1) in the
// storage type (doubled for the fused kernel's ping-pong), V (K-1) in FP32.
uint64_t bytes_per_voxel(int K, int kernel, int storage)
{
    const uint64_t copies = kernel == SCT_KERNEL_FUSED ? 2 : 1;
    return (copies * 4 * storage_bytes(storage) + 4) * uint64_t(K - 1);
}

int pick_kernel(const sct_potts_opts &o)
{
    return o.kernel == SCT_KERNEL_FUSED ? SCT_KERNEL_FUSED : SCT_KERNEL_TWOPASS;
}

void release_buffers(sct_potts_solver *s)
{
    for (int b = 0; b < 2; ++b) {
        sct_free(s->ctx, s->xi[b]);
        sct_free(s->ctx, s->vbar[b]);
        s->xi[b] = s->vbar[b] = nullptr;
    }
    sct_free(s->ctx, s->V);
    s->V = nullptr;
    fused_release(s);
}

int alloc_buffers(sct_potts_solver *s, uint32_t slab_len)
{
    const Geometry &geo = s->g->geo;
    s->wstride = sct_get_stride(slab_len);
    s->wspan = uint64_t(geo.lines) * s->wstride;
    const size_t comp = std::max<uint64_t>(1, s->wspan);
    const size_t sb = storage_bytes(s->storage);
    const int nb = s->kernel == SCT_KERNEL_FUSED ? 2 : 1;
    int rc;
    for (int b = 0; b < nb; ++b) {
        if ((rc = sct_alloc(s->ctx, &s->xi[b], sb * comp * 3 * (s->K - 1)))) return rc;
        if ((rc = sct_alloc(s->ctx, &s->vbar[b], sb * comp * (s->K - 1)))) return rc;
    }
    if ((rc = sct_alloc(s->ctx, (void **)&s->V, sizeof(float) * comp * (s->K - 1)))) return rc;
    if (s->kernel == SCT_KERNEL_FUSED && (rc = fused_prepare(s))) return rc;
    return SCT_OK;
}

int validate(sct_ctx *ctx, const sct_field *g, int K, const float *c, int iters, const sct_potts_opts *opts)
{
    if (!ctx || !g || !c) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (g->dtype != SCT_F32) return sct_fail(ctx, SCT_ERR_INVALID, "Potts needs a float field");
    if (K < 2 || K > 6) return sct_fail(ctx, SCT_ERR_INVALID, "K must be in [2, 6]");
    if (opts && (opts->storage < SCT_STORAGE_F32 || opts->storage > SCT_STORAGE_F16))
        return sct_fail(ctx, SCT_ERR_INVALID, "unknown storage type");
    if (iters < 0) return sct_fail(ctx, SCT_ERR_INVALID, "iters must be >= 0");
    if (opts && (opts->gap_every < 0 || opts->gap_capacity < 0))
        return sct_fail(ctx, SCT_ERR_INVALID, "gap_every and gap_capacity must be >= 0");
    return SCT_OK;
}

}  // namespace

// sqrtf-free thresholds (see DykstraLimits).
DykstraLimits sct::dykstra_limits()
{
    static DykstraLimits lim = [] {
        auto bits = [](float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; };
        auto flt = [](uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; };
        // t1: largest x >= 0 with sqrtf(x) <= 1.
        uint32_t lo = 0, hi = bits(4.0f);
        while (hi - lo > 1) {
            uint32_t mid = lo + (hi - lo) / 2;
            if (std::sqrt(flt(mid)) <= 1.0f) lo = mid; else hi = mid;
        }
        DykstraLimits l;
        l.t1 = flt(lo);
        // stol: smallest x >= 0 with (double)sqrtf(x) >= 1e-6.
        lo = 0;
        hi = bits(1.0f);
        while (hi - lo > 1) {
            uint32_t mid = lo + (hi - lo) / 2;
            if (double(std::sqrt(flt(mid))) >= 1e-6) hi = mid; else lo = mid;
        }
        l.stol = flt(hi);
        return l;
    }();
    return lim;
}


extern "C" void sct_potts_opts_init(sct_potts_opts *o)
{
    std::memset(o, 0, sizeof(*o));
    o->size = sizeof(*o);
    o->mem_available = -1;
}

// Slab plan, potts-min-partition.cpp:560-577 and :627-630.
int sct::potts_plan(sct_ctx *ctx, const sct_field *g, int K, int kernel, int iters, const sct_potts_opts *o,
                    PottsPlan *plan)
{
    const Geometry &geo = g->geo;
    const int n = int(geo.n_images);
    int slab_w;
    if (o && o->slab_w > 0) {
        slab_w = o->slab_w;
    } else {
        size_t avail = 0, total = 0;
        if (o && o->mem_available >= 0) {
            avail = size_t(o->mem_available);
        } else {
            int rc = sct_device_memory(ctx, &avail, &total);
            if (rc) return rc;
        }
        const uint64_t per_slice = sizeof(float) * uint64_t(std::max<uint32_t>(1, geo.lines)) * bytes_per_voxel(K, kernel, o ? o->storage : 0) / sizeof(float);
        const uint64_t w = avail / per_slice;
        slab_w = int(sct_prev_stride(uint32_t(std::min<uint64_t>(w, 0x7fffffffu))));
    }
    plan->halo = iters;
    if (n <= slab_w) {
        plan->slab_w = n;
        plan->res_w = n;
        plan->n_slabs = 1;
        return SCT_OK;
    }
    plan->slab_w = slab_w;
    plan->res_w = slab_w - 2 * iters;
    if (plan->res_w <= 0)
        return sct_fail(ctx, SCT_ERR_NOMEM, "Not enough memory. Reduce the mask radius or the number of iterations.");
    plan->n_slabs = 1 + (n - slab_w + plan->res_w - 1) / plan->res_w;
    return SCT_OK;
}

extern "C" int sct_potts_solver_create(sct_ctx *ctx, const sct_field *g, int K, const float *c, float lambda,
                                       float tau1, float tau2, const sct_potts_opts *opts,
                                       sct_potts_solver **out)
{
    if (!out) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    *out = nullptr;
    int rc = validate(ctx, g, K, c, 0, opts);
    if (rc) return rc;
    auto *s = new sct_potts_solver;
    s->ctx = ctx;
    s->g = g;
    s->K = K;
    std::copy(c, c + K, s->c);
    s->lambda = lambda;
    s->tau1 = tau1;
    s->tau2 = tau2;
    if (opts) std::memcpy(&s->opts, opts, std::min<size_t>(sizeof(s->opts), opts->size ? opts->size : sizeof(s->opts)));
    else sct_potts_opts_init(&s->opts);
    s->kernel = pick_kernel(s->opts);
    s->storage = s->opts.storage;
    s->s0 = 0;
    s->s1 = g->geo.n_images;
    rc = alloc_buffers(s, s->s1);
    if (!rc) rc = do_init(s);
    if (rc) {
        release_buffers(s);
        delete s;
        return rc;
    }
    *out = s;
    return SCT_OK;
}

extern "C" int sct_potts_solver_kernels(const sct_potts_solver *s)
{
    if (!s) return SCT_ERR_INVALID;
    return s->kernel == SCT_KERNEL_FUSED ? 1 : 2;
}

extern "C" int sct_potts_solver_iterate(sct_potts_solver *s, int iters)
{
    if (!s || iters < 0) return SCT_ERR_INVALID;
    float *dq = nullptr;
    if (s->opts.dq_trace) {
        const int need = s->it + iters;
        if (need > s->dq_cap) {
            float *nd = nullptr;
            int rc = sct_alloc(s->ctx, (void **)&nd, sizeof(float) * size_t(std::max(need, 1)));
            if (rc) return rc;
            SCT_HIP(s->ctx, hipMemsetAsync(nd, 0, sizeof(float) * size_t(std::max(need, 1)), s->ctx->stream));
            if (s->d_dq && s->it)
                SCT_HIP(s->ctx, hipMemcpyAsync(nd, s->d_dq, sizeof(float) * s->it, hipMemcpyDeviceToDevice, s->ctx->stream));
            SCT_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
            sct_free(s->ctx, s->d_dq);
            s->d_dq = nd;
            s->dq_cap = need;
        }
        dq = s->d_dq + s->it;
    }
    int count = s->it;
    int rc = iterate_with_gap(s, iters, dq, count, s->opts.gap_capacity);
    if (!rc) s->it += iters;
    return rc;
}

extern "C" int sct_potts_solver_iterate_timed(sct_potts_solver *s, int iters, float *kernel_ms)
{
    if (!s || iters < 0 || (iters && !kernel_ms)) return SCT_ERR_INVALID;
    const int nk = sct_potts_solver_kernels(s);
    std::vector<hipEvent_t> ev(size_t(iters) * nk + 1, nullptr);
    int rc = SCT_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            rc = sct_fail(s->ctx, SCT_ERR_HIP, "hipEventCreate failed");
            break;
        }
    if (!rc) {
        s->marks = ev.data();
        s->n_marks = 0;
        rc = sct_potts_solver_iterate(s, iters);
        s->marks = nullptr;
    }
    if (!rc && hipEventRecord(ev.back(), s->ctx->stream) != hipSuccess)
        rc = sct_fail(s->ctx, SCT_ERR_HIP, "hipEventRecord failed");
    if (!rc && hipEventSynchronize(ev.back()) != hipSuccess)
        rc = sct_fail(s->ctx, SCT_ERR_HIP, "hipEventSynchronize failed");
    for (size_t k = 0; !rc && k + 1 < ev.size(); ++k)
        if (hipEventElapsedTime(&kernel_ms[k], ev[k], ev[k + 1]) != hipSuccess)
            rc = sct_fail(s->ctx, SCT_ERR_HIP, "hipEventElapsedTime failed");
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

extern "C" int sct_potts_solver_labels(sct_potts_solver *s, sct_field *res)
{
    if (!s || !res) return SCT_ERR_INVALID;
    if (res->geo.span() != s->g->geo.span() || res->geo.lines != s->g->geo.lines || res->dtype != SCT_F32)
        return sct_fail(s->ctx, SCT_ERR_INVALID, "result field layout differs from the input");
    return do_labels(s, (float *)res->data, s->s0, s->s1);
}

extern "C" int sct_potts_solver_state(sct_potts_solver *s, float *V, float *xi)
{
    if (!s) return SCT_ERR_INVALID;
    const Geometry &geo = s->g->geo;
    const size_t span = geo.span();
    SCT_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    // Copy one component (elem bytes per value, storage type st) to the host
    // layout, widening bfloat16 / half to float.
    std::vector<uint8_t> tmp(s->wspan * sizeof(float));
    auto export_comp = [&](const void *dev, size_t elem, int st, float *host) -> int {
        SCT_HIP(s->ctx, hipMemcpy(tmp.data(), dev, elem * s->wspan, hipMemcpyDeviceToHost));
        std::fill(host, host + span, 0.0f);
        const uint32_t len = s->s1 - s->s0;
        for (uint32_t L = 0; L < geo.lines; ++L) {
            float *dst = host + size_t(L) * geo.nstride + s->s0;
            const uint8_t *src = tmp.data() + size_t(L) * s->wstride * elem;
            for (uint32_t k = 0; k < len; ++k) {
                if (st == SCT_STORAGE_F32) {
                    std::memcpy(dst + k, src + 4 * k, 4);
                } else {
                    uint16_t h;
                    std::memcpy(&h, src + 2 * k, 2);
                    dst[k] = st == SCT_STORAGE_BF16 ? bf16_to_float(h) : half_to_float(h);
                }
            }
        }
        return SCT_OK;
    };
    int rc;
    const size_t sb = storage_bytes(s->storage);
    if (V)
        for (int i = 0; i < s->K - 1; ++i)
            if ((rc = export_comp(s->V + i * s->wspan, 4, SCT_STORAGE_F32, V + i * span))) return rc;
    if (xi)
        for (int j = 0; j < 3 * (s->K - 1); ++j)
            if ((rc = export_comp((const uint8_t *)s->xi[s->cur] + j * s->wspan * sb, sb, s->storage, xi + j * span)))
                return rc;
    return SCT_OK;
}

extern "C" int sct_potts_solver_dq(sct_potts_solver *s, float *dq, int n)
{
    if (!s || !dq || n < 0 || n > s->it || !s->d_dq) return SCT_ERR_INVALID;
    SCT_HIP(s->ctx, hipMemcpyAsync(dq, s->d_dq, sizeof(float) * size_t(n), hipMemcpyDeviceToHost, s->ctx->stream));
    SCT_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return SCT_OK;
}

extern "C" void sct_potts_solver_destroy(sct_potts_solver *s)
{
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    release_buffers(s);
    sct_free(s->ctx, s->d_dq);
    delete s;
}

// potts_min_partition<K>, potts-min-partition.cpp:540-679.
extern "C" int sct_potts_min_partition(sct_ctx *ctx, const sct_field *g, int K, const float *c, float lambda,
                                       float tau1, float tau2, int iters, const sct_potts_opts *opts,
                                       sct_field **out)
{
    if (!out) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    *out = nullptr;
    int rc = validate(ctx, g, K, c, iters, opts);
    if (rc) return rc;
    sct_potts_opts o;
    sct_potts_opts_init(&o);
    if (opts) std::memcpy(&o, opts, std::min<size_t>(sizeof(o), opts->size ? opts->size : sizeof(o)));
    PottsPlan plan;
    if ((rc = potts_plan(ctx, g, K, pick_kernel(o), iters, &o, &plan))) return rc;

    sct_field *res = nullptr;
    sct_layout l{g->geo.n_images, g->geo.height, g->geo.width, g->geo.nstride, g->geo.col_begin, g->geo.col_end,
                 g->geo.start_row.data(), g->geo.base.data()};
    if ((rc = sct_field_create(ctx, &l, SCT_F32, &res))) return rc;
    SCT_HIP(ctx, hipMemsetAsync(res->data, 0, sizeof(float) * res->geo.span(), ctx->stream));

    sct_potts_solver s;
    s.ctx = ctx;
    s.g = g;
    s.K = K;
    std::copy(c, c + K, s.c);
    s.lambda = lambda;
    s.tau1 = tau1;
    s.tau2 = tau2;
    s.opts = o;
    s.kernel = pick_kernel(o);
    s.storage = o.storage;
    const int n = int(g->geo.n_images);
    rc = alloc_buffers(&s, uint32_t(plan.slab_w));
    float *d_dq = nullptr;
    if (!rc && o.dq_trace) {
        rc = sct_alloc(ctx, (void **)&d_dq, sizeof(float) * size_t(std::max(iters, 1)));
    }
    for (int k = 0; !rc && k < plan.n_slabs; ++k) {
        s.s0 = uint32_t(k * plan.res_w);
        s.s1 = uint32_t(std::min(n, k * plan.res_w + plan.slab_w));
        if (d_dq) {
            hipError_t e = hipMemsetAsync(d_dq, 0, sizeof(float) * size_t(std::max(iters, 1)), ctx->stream);
            if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "hipMemsetAsync");
        }
        if (!rc) rc = do_init(&s);
        int count = 0;
        if (!rc) rc = iterate_with_gap(&s, iters, d_dq, count, o.gap_capacity > 0 ? o.gap_capacity
                                                                                 : iters / std::max(1, o.gap_every));
        const uint32_t r0 = s.s0 == 0 ? s.s0 : s.s0 + uint32_t(plan.halo);
        const uint32_t r1 = int(s.s1) == n ? s.s1 : s.s1 - uint32_t(plan.halo);
        if (!rc) rc = do_labels(&s, (float *)res->data, r0, r1);
    }
    if (!rc && d_dq) {
        hipError_t e = hipMemcpyAsync(o.dq_trace, d_dq, sizeof(float) * size_t(iters), hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "dq download");
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "potts");
    }
    release_buffers(&s);
    sct_free(ctx, d_dq);
    if (rc) {
        sct_field_destroy(res);
        return rc;
    }
    *out = res;
    return SCT_OK;
}
