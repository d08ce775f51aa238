// potts_fused.hip -- one HBM pass per Potts iteration.
//
// One launch runs the V step of iteration k (div_kernel of
// potts-min-partition.cpp:349-488) and then the Xi step of iteration k+1
// (grad_kernel, :96-247) on the same voxels:
//
//   V+   = Pi_S(V + tau1 (div Xi - G)),  V-bar+ = 2 V+ - V   (needs Xi at x and
//          its three backward neighbours: all stored by the previous launch)
//   Xi++ = Pi_K(Xi + tau2 grad V-bar+)                        (needs V-bar+ at x
//          and its three forward neighbours: computed in this launch)
//
// so V-bar is never stored: per voxel and iteration the launch reads Xi, V and
// u and writes Xi and V, 68 B at K = 3 in FP32 where the two-kernel form moves
// 108 B.  The forward V-bar+ neighbours outside a tile are recomputed by the
// cheap simplex update (no Dykstra is ever recomputed), and nothing in the
// launch waits for a Dykstra result: a voxel whose projection needs more than
// one sweep goes to its wave's pool, which advances two sweeps per column and
// stores each result when it converges.
//
// Each workgroup owns TN voxels along n (64 lanes x VPL) of the rows
// [r0, r0 + R) and marches along col over a segment [cs, ce):
//   wave w < R owns row r0 + w; wave R computes V-bar+ of the row r0 + R above
//   the tile and, in its lanes, of the voxels n0 + TN of the tile rows;
//   per column cv: V step of column cv (the column ce is the +col halo, not
//   stored), a barrier, Xi step of column cv - 1.
// V-bar+ goes to a three-column LDS ring (+row and +n neighbours); Xi and V
// are double-buffered in HBM.  A call of n iterations runs the two-kernel Xi
// step first, n - 1 launches, and the two-kernel V step last, so that the
// state between calls is the reference's (Xi, V, V-bar) after n iterations.
// Arithmetic per voxel is that of the two-pass kernels (potts_math.hpp),
// V-bar+ rounded to the storage type as there, so results are bitwise the
// same.
#include "potts.hpp"
#include "twopass.hpp"

#include <algorithm>

using namespace sct;

namespace {

constexpr size_t kFusedPtrBytes = 512;   // one pointer table

// Tile shape per K and storage.  K <= 3: 16 waves (15 tile rows), two voxels
// per lane (8-byte accesses of FP32 state); K >= 4 keeps its larger per-voxel
// state in registers with one voxel per lane and 8 waves.
template <int K, typename St>
struct FusedCfg {
    static constexpr int VPL = K <= 3 ? 2 : 1;
    static constexpr int W = K <= 3 ? 16 : 8;   // waves: R tile rows and the row above
    static constexpr int R = W - 1;             // tile rows
    static constexpr int TN = 64 * VPL;         // tile width along n
    static constexpr int NP = K * (K - 1) / 2;
    static constexpr int QC = K <= 3 ? 128 : 64;   // Dykstra items waiting per wave (more: solved in place)
    static constexpr int SWEEPS = 2;            // pool sweeps per column
    static constexpr int SEG = 64;              // columns per segment (workgroup)
};

template <int K, typename St>
struct FusedLds {
    using T = typename St::T;
    using Cfg = FusedCfg<K, St>;
    // V-bar+ of columns by cv % 3: rows r0 .. r0 + R; element [TN] of a tile
    // row is the voxel n0 + TN
    T vb[3][Cfg::W][K - 1][Cfg::TN + Cfg::VPL];
    float qw[Cfg::R][Cfg::QC][K][3];   // Dykstra inputs waiting for a lane, per wave
    uint32_t qo[Cfg::R][Cfg::QC];      // their element offsets
    int nw[Cfg::R];
    // wave R: inputs of the n-halo V step (lane l: row r0 + l at n0 + TN)
    float hx[Cfg::R][3 * (K - 1)];     // Xi
    float hn[Cfg::R][K - 1];           // Xi_n at n0 + TN - 1
    float hr[Cfg::R][K - 1];           // Xi_row at row - 1
    float hv[Cfg::R][K - 1];           // V
    float hu[Cfg::R];
    float red[Cfg::W];
};

// Per-component base pointers, in device memory: every component is
// addressed as base[comp] + 32-bit element offset (lines * wstride < 2^32 is
// checked on the host).  The kernel reads them with scalar loads where it uses
// them rather than holding 3(K-1) + ... 64-bit pointers in scalar registers
// over the whole march.
template <int K, typename St>
struct FusedPtrs {
    const typename St::T *xi_in[3 * (K - 1)];
    typename St::T *xi_out[3 * (K - 1)];
    const float *V_in[K - 1];
    float *V_out[K - 1];
    const float *u;
};

template <int K, typename St>
struct FusedArgs {
    uint32_t len, wstride, in_stride, s0, col_begin, col_end, ntn;
    const FusedPtrs<K, St> *ptrs;
    float c[K];
    float tau1, tau2, lambda;
    DykstraLimits lim;
    float *dq;
};

// Wave-uniform value in a scalar register.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <typename X>
using Const = __attribute__((address_space(4))) X;
template <typename X>
using Glob = __attribute__((address_space(1))) X;

// The pointer table, re-read at each use: the empty asm makes the table's
// address opaque, so its scalar loads are neither hoisted out of the march
// nor kept in registers across it.
template <typename X>
__device__ __forceinline__ const Const<X> *fresh(const X *p)
{
    const Const<X> *q = (const Const<X> *)p;
    asm volatile("" : "+s"(q));
    return q;
}

// Global-memory accesses through address-space-1 pointers: the pointers come
// from the table, so the compiler cannot infer their address space, and flat
// accesses would count against the LDS wait of every barrier.
template <typename St, int N>
__device__ __forceinline__ void gld(const typename St::T *p, float (&o)[N])
{
    typedef typename St::T V __attribute__((ext_vector_type(N)));
    const V v = *(const Glob<V> *)p;
#pragma unroll
    for (int j = 0; j < N; ++j) o[j] = St::ld(v[j]);
}

template <typename St>
__device__ __forceinline__ float gld1(const typename St::T *p)
{
    return St::ld(*(const Glob<typename St::T> *)p);
}

template <typename St, int N>
__device__ __forceinline__ void gst_nt(typename St::T *p, const float (&x)[N])
{
    typedef typename St::T V __attribute__((ext_vector_type(N)));
    V v;
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = St::st(x[j]);
    __builtin_nontemporal_store(v, (Glob<V> *)p);
}

template <typename St>
__device__ __forceinline__ void gst1(typename St::T *p, float x)
{
    *(Glob<typename St::T> *)p = St::st(x);
}

// LDS accesses of one wave that other lanes of the same wave read next: keep
// the compiler from moving them across (the LDS executes a wave's accesses in
// order).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Col {
    int line0, rb, re;
    __device__ __forceinline__ bool live(int r) const { return r >= rb && r < re; }
    __device__ __forceinline__ uint32_t line(int r) const { return uint32_t(line0 + (r - rb)); }
};

__device__ __forceinline__ Col load_col(const ColInfo *cols, int c, int cb, int ce)
{
    if (c < cb || c >= ce) return Col{0, 0, 0};
    const ColInfo *p = cols + c;
    return Col{uniform(int(p->line0)), uniform(int(p->rb)), uniform(int(p->re))};
}

template <int K, typename St, bool TRACE>
__global__ __launch_bounds__((FusedCfg<K, St>::W * 64), 1) void potts_fused_kernel(FusedArgs<K, St> a,
                                                                                   const ColInfo *__restrict__ cols,
                                                                                   const FusedTile *__restrict__ tiles)
{
    using Cfg = FusedCfg<K, St>;
    using T = typename St::T;
    constexpr int KM = K - 1, NC = 3 * KM, NP = Cfg::NP;
    constexpr int VPL = Cfg::VPL, W = Cfg::W, R = Cfg::R, TN = Cfg::TN;
    typedef T TV __attribute__((ext_vector_type(VPL)));
    extern __shared__ __align__(16) unsigned char smem_raw[];
    FusedLds<K, St> &L = *reinterpret_cast<FusedLds<K, St> *>(smem_raw);

    // n-tiles of one (segment, band) are consecutive and go to one XCD, so the
    // n-halo reads of a tile hit the lines its neighbours fetch.
    const uint32_t ti = xcd_block(blockIdx.x, gridDim.x, a.ntn);
    const FusedTile tile = tiles[ti];
    const int lane = threadIdx.x % kWave, w = uniform(threadIdx.x / kWave);
    const int len = int(a.len);
    const int n0 = int(tile.n0), r0 = int(tile.r0);
    const int cs = int(tile.cs), ce = int(tile.ce);
    const int cb = int(a.col_begin), cend = int(a.col_end);
    const int nl0 = lane * VPL;                            // lane's first voxel in the tile
    const int nv = max(0, min(VPL, len - (n0 + nl0)));     // valid voxels of the lane
    const int rr = r0 + w;                                 // own row (w = R: the row above the tile)
    const bool tile_row = w < R;
    const uint32_t lofs = uint32_t(n0 + nl0);
    const int hr = r0 + lane;                              // wave R: n-halo lane's row
    const bool h_ok = lane < R && n0 + TN < len;           // wave R: n-halo lane in use

    // ---- registers ----
    float H[NC][VPL];      // Xi(rr, cv - 1): own Xi of the Xi step, -col of the V step
    float pxi[NC][VPL];    // Xi(rr, cv)
    float pxr[KM][VPL];    // Xi_row(rr - 1, cv)
    float pxn[KM];         // lane 0: Xi_n(n0 - 1, rr, cv)
    float pv[KM][VPL], pu[VPL];
    float hcol[KM];        // wave R, n-halo lane: Xi_col(n0 + TN, row, cv - 1)
    auto zero = [](float(&x)[VPL]) {
#pragma unroll
        for (int j = 0; j < VPL; ++j) x[j] = 0.0f;
    };
    // Xi, Xi_row of the row below, Xi_n at n0 - 1, V and u of the own row at column cv
    auto prefetch = [&](int cv) {
        const Col cc = load_col(cols, cv, cb, cend);
        const Const<FusedPtrs<K, St>> *pt = fresh(a.ptrs);
        const bool own = nv && cc.live(rr), below = nv && cc.live(rr - 1);
        const uint32_t off = own ? cc.line(rr) * a.wstride + lofs : 0;
        const uint32_t offb = below ? cc.line(rr - 1) * a.wstride + lofs : 0;
        // Every lane loads (an absent voxel from offset 0, inside every array)
        // and keeps zero where the voxel is absent: no branches between the loads.
        auto keep = [&](bool ok, float(&x)[VPL]) {
#pragma unroll
            for (int j = 0; j < VPL; ++j) x[j] = ok ? x[j] : 0.0f;
        };
#pragma unroll
        for (int k = 0; k < NC; ++k) gld<St, VPL>(pt->xi_in[k] + off, pxi[k]);
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            gld<St, VPL>(pt->xi_in[3 * i + 1] + offb, pxr[i]);
            gld<StorageF32, VPL>(pt->V_in[i] + off, pv[i]);
            pxn[i] = (own && lane == 0 && n0 > 0) ? gld1<St>(pt->xi_in[3 * i] + off - 1) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) keep(own, pxi[k]);
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            keep(below, pxr[i]);
            keep(own, pv[i]);
        }
        const uint32_t uo = own ? cc.line(rr) * a.in_stride + a.s0 + lofs : 0;
#pragma unroll
        for (int j = 0; j < VPL; ++j) pu[j] = (own && j < nv) ? gld1<StorageF32>(pt->u + uo + j) : 0.0f;
    };
    // wave R: the n-halo V-step inputs of column cv, through LDS
    auto halo_inputs = [&](int cv) {
        const Col cc = load_col(cols, cv, cb, cend);
        const Const<FusedPtrs<K, St>> *pt = fresh(a.ptrs);
        const bool own = h_ok && cc.live(hr), below = h_ok && cc.live(hr - 1);
        const uint32_t off = own ? cc.line(hr) * a.wstride + uint32_t(n0 + TN) : 0;
        const uint32_t offb = below ? cc.line(hr - 1) * a.wstride + uint32_t(n0 + TN) : 0;
        float x[NC], xn[KM], xr[KM], v[KM], u;
#pragma unroll
        for (int k = 0; k < NC; ++k) x[k] = own ? gld1<St>(pt->xi_in[k] + off) : 0.0f;
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            xn[i] = own ? gld1<St>(pt->xi_in[3 * i] + off - 1) : 0.0f;
            xr[i] = below ? gld1<St>(pt->xi_in[3 * i + 1] + offb) : 0.0f;
            v[i] = own ? gld1<StorageF32>(pt->V_in[i] + off) : 0.0f;
        }
        u = own ? gld1<StorageF32>(pt->u + cc.line(hr) * a.in_stride + a.s0 + uint32_t(n0 + TN)) : 0.0f;
        if (lane < R) {
#pragma unroll
            for (int k = 0; k < NC; ++k) L.hx[lane][k] = x[k];
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                L.hn[lane][i] = xn[i];
                L.hr[lane][i] = xr[i];
                L.hv[lane][i] = v[i];
            }
            L.hu[lane] = u;
        }
    };
    // the wave's Dykstra pool: one item per lane
    float dq_[K][3], dr[NP][3];
    int dsw = 0;
    uint32_t dofs = 0;
    bool dhave = false;
    // store Xi++ of one voxel from its Dykstra state
    auto store_item = [&](uint32_t ofs, const float (&q)[K][3]) {
        const Const<FusedPtrs<K, St>> *pt = fresh(a.ptrs);
#pragma unroll
        for (int i = 0; i < KM; ++i)
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) gst1<St>(pt->xi_out[3 * i + ax] + ofs, q[i][ax]);
    };
    // idle lanes take waiting items
    auto refill = [&]() {
        const int waiting = L.nw[w];
        const uint64_t idle = __ballot(!dhave);
        const int take = min(waiting, __popcll(idle));
        const int rank = __popcll(idle & ((1ull << lane) - 1ull));
        if (!dhave && rank < take) {
            const int t = waiting - 1 - rank;
#pragma unroll
            for (int i = 0; i < K; ++i)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) dq_[i][ax] = L.qw[w][t][i][ax];
            dofs = L.qo[w][t];
#pragma unroll
            for (int pp = 0; pp < NP; ++pp) dr[pp][0] = dr[pp][1] = dr[pp][2] = 0.0f;
            dsw = 0;
            dhave = true;
        }
        wave_sync();
        if (lane == 0) L.nw[w] = waiting - take;
        wave_sync();
    };
    // refill, then `sweeps` sweeps of the pool (the two-sweep cadence per
    // column keeps the waves' work per column even; a pool that swept only
    // when nearly full was slower)
    auto pool_run = [&](int sweeps) {
        refill();
        for (int s = 0; s < sweeps && __any(dhave); ++s) {
            if (dhave) {
                const float m2 = dykstra_sweep<K>(dq_, dr, a.lim);
                ++dsw;
                if (m2 < a.lim.stol || dsw == 20) {
                    store_item(dofs, dq_);
                    dhave = false;
                }
            }
        }
    };

    if (tile_row && lane == 0) L.nw[w] = 0;
    // prologue: H = Xi(cs - 1), the inputs of column cs
    {
        const Col cx = load_col(cols, cs - 1, cb, cend);
        const Const<FusedPtrs<K, St>> *pt = fresh(a.ptrs);
        const bool own = nv && cx.live(rr);
        const uint32_t off = own ? cx.line(rr) * a.wstride + lofs : 0;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if (own) gld<St, VPL>(pt->xi_in[k] + off, H[k]);
            else zero(H[k]);
        }
        if (w == R) {
            const bool hown = h_ok && cx.live(hr);
            const uint32_t hoff = hown ? cx.line(hr) * a.wstride + uint32_t(n0 + TN) : 0;
#pragma unroll
            for (int i = 0; i < KM; ++i) hcol[i] = hown ? gld1<St>(pt->xi_in[3 * i + 2] + hoff) : 0.0f;
            halo_inputs(cs);
        }
        prefetch(cs);
    }
    float dqmax = 0.0f;

    for (int cv = cs; cv <= ce; ++cv) {
        const Col cp = load_col(cols, cv - 1, cb, cend);
        const Col cc = load_col(cols, cv, cb, cend);
        const Col cn = load_col(cols, cv + 1, cb, cend);
        const int bv = cv % 3;
        // ---- V step of column cv (own row) ----
        {
            float vo[KM][VPL];
            TV vbt[KM];
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                const int n = n0 + nl0 + j;
                const bool px[3] = {n > 0, cc.live(rr - 1), cp.live(rr)};
                const bool nx[3] = {n + 1 < len, cc.live(rr + 1), cn.live(rr)};
                float xo[KM][3], xp[3][KM], vv[KM], vbj[KM];
#pragma unroll
                for (int i = 0; i < KM; ++i) {
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) xo[i][ax] = pxi[3 * i + ax][j];
                    float xn = j > 0 ? pxi[3 * i][j - 1] : __shfl_up(pxi[3 * i][VPL - 1], 1, kWave);
                    if (j == 0 && lane == 0) xn = pxn[i];
                    xp[0][i] = xn;
                    xp[1][i] = pxr[i][j];
                    xp[2][i] = H[3 * i + 2][j];
                    vv[i] = pv[i][j];
                }
                const float d = v_update<K>(xo, xp, px, nx, vv, vbj, pu[j], a.c, a.lambda, a.tau1);
                if (TRACE && tile_row && cv < ce && j < nv && cc.live(rr)) dqmax = fmaxf(dqmax, d);
#pragma unroll
                for (int i = 0; i < KM; ++i) {
                    vo[i][j] = vv[i];
                    vbt[i][j] = St::st(vbj[i]);
                }
            }
#pragma unroll
            for (int i = 0; i < KM; ++i) *reinterpret_cast<TV *>(&L.vb[bv][w][i][nl0]) = vbt[i];
            if (tile_row && cv < ce && nv && cc.live(rr)) {
                const Const<FusedPtrs<K, St>> *pt = fresh(a.ptrs);
                const uint32_t off = cc.line(rr) * a.wstride + lofs;
#pragma unroll
                for (int i = 0; i < KM; ++i) gst_nt<StorageF32, VPL>(pt->V_out[i] + off, vo[i]);
            }
        }
        // ---- wave R: the V step at n0 + TN of the tile rows, one per lane ----
        if (w == R) {
            const bool ok = h_ok && cc.live(hr);
            const bool px[3] = {true, cc.live(hr - 1), cp.live(hr)};
            const bool nx[3] = {n0 + TN + 1 < len, cc.live(hr + 1), cn.live(hr)};
            float xo[KM][3], xp[3][KM], vv[KM], vbj[KM];
            const int l = lane < R ? lane : 0;
#pragma unroll
            for (int i = 0; i < KM; ++i) {
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) xo[i][ax] = L.hx[l][3 * i + ax];
                xp[0][i] = L.hn[l][i];
                xp[1][i] = L.hr[l][i];
                xp[2][i] = hcol[i];
                vv[i] = L.hv[l][i];
            }
            v_update<K>(xo, xp, px, nx, vv, vbj, L.hu[l], a.c, a.lambda, a.tau1);
            if (lane < R) {
#pragma unroll
                for (int i = 0; i < KM; ++i) L.vb[bv][lane][i][TN] = St::st(ok ? vbj[i] : 0.0f);
            }
#pragma unroll
            for (int i = 0; i < KM; ++i) hcol[i] = xo[i][2];
        }
        // Xi(cv) becomes H after the Xi step of column cv - 1
        float Hn[NC][VPL];
#pragma unroll
        for (int k = 0; k < NC; ++k)
#pragma unroll
            for (int j = 0; j < VPL; ++j) Hn[k][j] = pxi[k][j];
        if (cv + 1 <= ce) prefetch(cv + 1);
        __syncthreads();   // V-bar+ of column cv in LDS

        if (w == R) {
            if (cv + 1 <= ce) halo_inputs(cv + 1);
        } else if (cv > cs && cp.live(rr)) {
            // ---- Xi step of column c = cv - 1 ----
            const int c = cv - 1, b = c % 3;
            const Col &cx = cp;                  // column c
            const uint32_t off = cx.line(rr) * a.wstride + lofs;
            float res[NC][VPL];
            bool feas[VPL];
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                const int nn = nl0 + j;
                const bool valid = j < nv;
                const bool nx[3] = {n0 + nn + 1 < len, cx.live(rr + 1), cc.live(rr)};
                float vbj[KM], vbn[3][KM], xj[KM][3];
#pragma unroll
                for (int i = 0; i < KM; ++i) {
                    vbj[i] = St::ld(L.vb[b][w][i][nn]);
                    vbn[0][i] = St::ld(L.vb[b][w][i][nn + 1]);
                    vbn[1][i] = St::ld(L.vb[b][w + 1][i][nn]);
                    vbn[2][i] = St::ld(L.vb[bv][w][i][nn]);
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) xj[i][ax] = H[3 * i + ax][j];
                }
                float q[K][3];
                xi_input<K>(vbj, vbn, nx, xj, a.tau2, q);
                const bool queued = valid && !dykstra_first_sweep_done<K>(q, a.lim);
                // queue it in the wave's pool (wave-aggregated append)
                const uint64_t mask = __ballot(queued);
                const int base = L.nw[w];
                const int p = base + __popcll(mask & ((1ull << lane) - 1ull));
                if (queued) {
                    if (p < Cfg::QC) {
#pragma unroll
                        for (int i = 0; i < K; ++i)
#pragma unroll
                            for (int ax = 0; ax < 3; ++ax) L.qw[w][p][i][ax] = q[i][ax];
                        L.qo[w][p] = off + uint32_t(j);
                    } else {
                        dykstra<K>(q, a.lim);   // pool full: solve in place
                    }
                }
                wave_sync();
                if (lane == 0) L.nw[w] = min(Cfg::QC, base + __popcll(mask));
                wave_sync();
                feas[j] = !(queued && p < Cfg::QC);
#pragma unroll
                for (int i = 0; i < KM; ++i)
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) res[3 * i + ax][j] = q[i][ax];
            }
            // store the finished voxels (the pool stores the others when they
            // converge): whole vectors where every voxel is finished
            {
                const Const<FusedPtrs<K, St>> *pt = fresh(a.ptrs);
                bool all = nv > 0;
#pragma unroll
                for (int j = 0; j < VPL; ++j) all = all && feas[j];
                if (all) {
#pragma unroll
                    for (int k = 0; k < NC; ++k) gst_nt<St, VPL>(pt->xi_out[k] + off, res[k]);
                } else {
#pragma unroll
                    for (int j = 0; j < VPL; ++j) {
                        if (j < nv && feas[j]) {
#pragma unroll
                            for (int k = 0; k < NC; ++k) gst1<St>(pt->xi_out[k] + off + j, res[k][j]);
                        }
                    }
                }
            }
            pool_run(Cfg::SWEEPS);
        }
#pragma unroll
        for (int k = 0; k < NC; ++k)
#pragma unroll
            for (int j = 0; j < VPL; ++j) H[k][j] = Hn[k][j];
    }
    // flush the pool
    if (tile_row) {
        while (__any(dhave) || L.nw[w] > 0) pool_run(4);
    }

    if constexpr (TRACE) {
        dqmax = wave_max(dqmax);
        if (lane == 0) L.red[w] = dqmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = L.red[0];
            for (int i = 1; i < W; ++i) m = fmaxf(m, L.red[i]);
            atomic_max_nonneg(a.dq, m);
        }
    }
}

// The lines of the solver's slab for the two-pass kernels (storage St).
template <typename St>
tp::VolumeLines slab_lines(const sct_potts_solver *s)
{
    const Geometry &geo = s->g->geo;
    constexpr uint32_t cw = kWave * tp::vpl_of<St>();
    const uint32_t len = s->s1 - s->s0;
    return tp::VolumeLines::make(geo.d_nbr, geo.lines, len, (len + cw - 1) / cw, s->wstride, geo.nstride, s->s0);
}

// The solver's current state for the two-pass kernels (in place).
template <int K, typename St>
tp::TwoPassArgs<K, St> slab_args(const sct_potts_solver *s)
{
    using T = typename St::T;
    tp::TwoPassArgs<K, St> a;
    for (int j = 0; j < 3 * (K - 1); ++j) a.xi[j] = (T *)s->xi[s->cur] + j * s->wspan;
    for (int i = 0; i < K - 1; ++i) {
        a.vbar[i] = (T *)s->vbar[0] + i * s->wspan;
        a.V[i] = s->V + i * s->wspan;
    }
    a.u = (const float *)s->g->data;
    a.in_place();
    return a;
}

int fused_tiles(sct_potts_solver *s);

template <int K, typename St>
int launch_fused(sct_potts_solver *s, int iters, float *dq_dev)
{
    using T = typename St::T;
    using Cfg = FusedCfg<K, St>;
    const Geometry &geo = s->g->geo;
    if (iters <= 0 || s->s1 <= s->s0) return SCT_OK;
    // the Xi step of the first iteration: the two-kernel form's
    const tp::VolumeLines VL = slab_lines<St>(s);
    tp::OneSolve<K> prm;
    for (int i = 0; i < K; ++i) prm.p.c[i] = s->c[i];
    prm.p.lambda = s->lambda;
    const DykstraLimits lim = dykstra_limits();
    tp::potts_xi_kernel<K, St><<<tp::xi_grid<St>(VL.count()), tp::kXiThreads, 0, s->ctx->stream>>>(VL, slab_args<K, St>(s),
                                                                                                 s->tau2, lim);
    if (iters > 1) {
        FusedArgs<K, St> a;
        a.len = s->s1 - s->s0;
        a.wstride = s->wstride;
        a.in_stride = geo.nstride;
        a.s0 = s->s0;
        a.col_begin = geo.col_begin;
        a.col_end = geo.col_end;
        a.ntn = (a.len + Cfg::TN - 1) / Cfg::TN;
        for (int i = 0; i < K; ++i) a.c[i] = s->c[i];
        a.tau1 = s->tau1;
        a.tau2 = s->tau2;
        a.lambda = s->lambda;
        a.lim = lim;
        if (int rc = fused_tiles(s)) return rc;
        const size_t lds = sizeof(FusedLds<K, St>);
        static bool attr_set[2] = {false, false};
        for (int t = 0; t < 2; ++t) {
            if (attr_set[t]) continue;
            const void *fn = t ? (const void *)potts_fused_kernel<K, St, true> : (const void *)potts_fused_kernel<K, St, false>;
            SCT_HIP(s->ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            attr_set[t] = true;
        }
        // the two pointer tables (read buffer set b of Xi and V)
        static_assert(sizeof(FusedPtrs<K, St>) <= kFusedPtrBytes, "pointer table size");
        FusedPtrs<K, St> pt[2];
        for (int b = 0; b < 2; ++b) {
            for (int j = 0; j < 3 * (K - 1); ++j) {
                pt[b].xi_in[j] = (const T *)s->xi[b] + j * s->wspan;
                pt[b].xi_out[j] = (T *)s->xi[b ^ 1] + j * s->wspan;
            }
            for (int i = 0; i < K - 1; ++i) {
                pt[b].V_in[i] = s->Vb[b] + i * s->wspan;
                pt[b].V_out[i] = s->Vb[b ^ 1] + i * s->wspan;
            }
            pt[b].u = (const float *)s->g->data;
            SCT_HIP(s->ctx, hipMemcpyAsync((char *)s->d_fptrs + b * kFusedPtrBytes, &pt[b], sizeof(pt[b]),
                                           hipMemcpyHostToDevice, s->ctx->stream));
        }
        SCT_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
        for (int it = 0; it + 1 < iters; ++it) {
            a.ptrs = reinterpret_cast<const FusedPtrs<K, St> *>((const char *)s->d_fptrs + s->cur * kFusedPtrBytes);
            potts_mark(s);
            if (s->n_tiles) {
                if (dq_dev) {
                    a.dq = dq_dev + it;
                    potts_fused_kernel<K, St, true><<<s->n_tiles, Cfg::W * 64, lds, s->ctx->stream>>>(a, s->d_cols, s->d_tiles);
                } else {
                    a.dq = nullptr;
                    potts_fused_kernel<K, St, false><<<s->n_tiles, Cfg::W * 64, lds, s->ctx->stream>>>(a, s->d_cols, s->d_tiles);
                }
            }
            s->cur ^= 1;
            s->V = s->Vb[s->cur];
        }
    }
    // the V step of the last iteration: the two-kernel form's, which also
    // stores V-bar
    potts_mark(s);
    const uint32_t vg = tp::wave_grid(VL.count());
    if (dq_dev)
        tp::potts_v_kernel<K, St, true><<<vg, 256, 0, s->ctx->stream>>>(VL, slab_args<K, St>(s), prm, s->tau1, dq_dev + iters - 1);
    else
        tp::potts_v_kernel<K, St, false><<<vg, 256, 0, s->ctx->stream>>>(VL, slab_args<K, St>(s), prm, s->tau1, nullptr);
    SCT_HIP(s->ctx, hipGetLastError());
    return SCT_OK;
}

// Tile geometry of the solver's K and storage.
template <int K, typename St>
void fused_shape(int *rows, int *tn, int *seg)
{
    using Cfg = FusedCfg<K, St>;
    *rows = Cfg::R;
    *tn = Cfg::TN;
    *seg = Cfg::SEG;
}

template <typename St>
void fused_shape_k(int K, int *rows, int *tn, int *seg)
{
    switch (K) {
    case 2: fused_shape<2, St>(rows, tn, seg); break;
    case 3: fused_shape<3, St>(rows, tn, seg); break;
    case 4: fused_shape<4, St>(rows, tn, seg); break;
    case 5: fused_shape<5, St>(rows, tn, seg); break;
    default: fused_shape<6, St>(rows, tn, seg); break;
    }
}

}  // namespace

int sct::fused_prepare(sct_potts_solver *s)
{
    fused_release(s);
    const Geometry &geo = s->g->geo;
    std::vector<ColInfo> cols(std::max<uint32_t>(geo.width, 1), ColInfo{0, 0, 0, 0});
    for (uint32_t c = geo.col_begin; c < geo.col_end; ++c) {
        const uint32_t rows = (geo.base[c + 1] - geo.base[c]) / geo.nstride;
        if (!rows) continue;
        cols[c] = ColInfo{geo.base[c] / geo.nstride, geo.start_row[c], geo.start_row[c] + rows, 0};
    }
    int rc;
    if ((rc = sct_alloc(s->ctx, (void **)&s->d_cols, sizeof(ColInfo) * cols.size()))) return rc;
    if ((rc = sct_alloc(s->ctx, &s->d_fptrs, 2 * kFusedPtrBytes))) return rc;
    SCT_HIP(s->ctx, hipMemcpy(s->d_cols, cols.data(), sizeof(ColInfo) * cols.size(), hipMemcpyHostToDevice));
    return SCT_OK;
}

namespace {

// The tile table of the current slab (its length s1 - s0 sets the n-tiles),
// rebuilt when the slab length changes.
int fused_tiles(sct_potts_solver *s)
{
    const uint32_t len = s->s1 - s->s0;
    if (s->d_tiles && s->tiles_len == len) return SCT_OK;
    const Geometry &geo = s->g->geo;
    int R = 0, TN = 0, SEG = 0;
    switch (s->storage) {
    case SCT_STORAGE_BF16: fused_shape_k<StorageBF16>(s->K, &R, &TN, &SEG); break;
    case SCT_STORAGE_F16: fused_shape_k<StorageF16>(s->K, &R, &TN, &SEG); break;
    default: fused_shape_k<StorageF32>(s->K, &R, &TN, &SEG); break;
    }
    const uint32_t ntn = (len + TN - 1) / TN;
    // bands of R rows, each over the columns where one of its rows is live,
    // cut into segments of SEG columns; tiles ordered (segment, band, n-tile)
    struct Band {
        uint32_t r0, cs, ce;
    };
    std::vector<Band> bands;
    for (uint32_t r0 = 0; r0 < geo.height; r0 += uint32_t(R)) {
        uint32_t cs = ~0u, ce = 0;
        for (uint32_t c = geo.col_begin; c < geo.col_end; ++c) {
            const uint32_t rb = geo.start_row[c], re = geo.row_end(c);
            if (re > rb && rb < r0 + uint32_t(R) && re > r0) {
                cs = std::min(cs, c);
                ce = c + 1;
            }
        }
        if (ce > 0) bands.push_back(Band{r0, cs, ce});
    }
    std::vector<FusedTile> tiles;
    for (uint32_t k = 0;; ++k) {
        bool any = false;
        for (const Band &b : bands) {
            const uint32_t cs = b.cs + k * uint32_t(SEG);
            if (cs >= b.ce) continue;
            any = true;
            const uint32_t ce = std::min(b.ce, cs + uint32_t(SEG));
            for (uint32_t j = 0; j < ntn; ++j) tiles.push_back(FusedTile{j * uint32_t(TN), b.r0, cs, ce});
        }
        if (!any) break;
    }
    sct_free(s->ctx, s->d_tiles);
    s->d_tiles = nullptr;
    s->n_tiles = 0;
    int rc;
    if ((rc = sct_alloc(s->ctx, (void **)&s->d_tiles, sizeof(FusedTile) * std::max<size_t>(1, tiles.size())))) return rc;
    if (!tiles.empty())
        SCT_HIP(s->ctx, hipMemcpy(s->d_tiles, tiles.data(), sizeof(FusedTile) * tiles.size(), hipMemcpyHostToDevice));
    s->n_tiles = uint32_t(tiles.size());
    s->tiles_len = len;
    return SCT_OK;
}

}  // namespace

void sct::fused_release(sct_potts_solver *s)
{
    sct_free(s->ctx, s->d_cols);
    sct_free(s->ctx, s->d_tiles);
    sct_free(s->ctx, s->d_fptrs);
    s->d_cols = nullptr;
    s->d_tiles = nullptr;
    s->d_fptrs = nullptr;
    s->n_tiles = 0;
    s->tiles_len = 0;
}

template <typename St>
static int fused_iterate_k(sct_potts_solver *s, int iters, float *dq_dev)
{
    switch (s->K) {
    case 2: return launch_fused<2, St>(s, iters, dq_dev);
    case 3: return launch_fused<3, St>(s, iters, dq_dev);
    case 4: return launch_fused<4, St>(s, iters, dq_dev);
    case 5: return launch_fused<5, St>(s, iters, dq_dev);
    case 6: return launch_fused<6, St>(s, iters, dq_dev);
    default: return SCT_ERR_INVALID;
    }
}

int sct::fused_iterate(sct_potts_solver *s, int iters, float *dq_dev)
{
    switch (s->storage) {
    case SCT_STORAGE_BF16: return fused_iterate_k<StorageBF16>(s, iters, dq_dev);
    case SCT_STORAGE_F16: return fused_iterate_k<StorageF16>(s, iters, dq_dev);
    default: return fused_iterate_k<StorageF32>(s, iters, dq_dev);
    }
}
