// potts_fused.hip -- one HBM pass per Potts iteration.
//
// Fuses grad_kernel and div_kernel of potts-min-partition.cpp (:96-247 and
// :349-488) into one launch.  V+(x) needs Xi+ at x and at its three backward
// neighbours, and Xi+(y) needs V-bar at y and its three forward neighbours,
// so each workgroup owns a tile of TN (n) x TR (rows) and marches along col:
//
//   P = tile plus a one-voxel low-side halo in n and in rows, (TN+1) x (TR+1):
//       Xi+ is computed here (halo points are recomputed, not exchanged);
//   T = the tile itself: V+ and V-bar+ are computed and stored here.
//
// Lanes run along n (lane 0 is the n-halo, lanes 1..63 tile voxels); each of
// the 8 waves owns two P rows.  Per column:
//   A  every owner lane forms the Dykstra input q = Xi + tau2 grad V-bar of its
//      points.  Most points are already feasible, so Dykstra stops after its
//      first sweep; that case is finished in place.  The others are appended
//      to a work list in LDS.
//   B  the work list is drained one Dykstra sweep per loop trip, each lane
//      taking a new item as soon as its item converges, so a wave's time
//      follows the total number of sweeps rather than the slowest of its 64
//      points; only as many waves join as there are items per 64 lanes.
//   C  owner lanes store Xi+ and compute V+, V-bar+ of their tile points, with
//      Xi+ of the -n / -row neighbours from LDS and of the -col neighbour from
//      their registers (kept from the previous column).
// The next column's HBM loads are issued before B and consumed after it.
// Xi and V-bar are double-buffered in HBM (read the previous iteration, write
// the next), so no tile ever sees a neighbour's half-updated state; V and u
// are only touched at a tile's own voxels.  Xi+ is rounded to the storage type
// as soon as it is formed, so the V step reads exactly what is stored.
// Arithmetic per voxel is that of the two-pass kernels (potts_math.hpp), so
// results are bitwise identical.
#include "potts.hpp"

#include <algorithm>
#include <numeric>

using namespace sct;

namespace {

constexpr int kTNP = 64;             // P width along n (lane 0 = halo)
constexpr int kTN = kTNP - 1;        // tile width along n
constexpr int kTRP = 16;             // P rows (row 0 = halo)
constexpr int kTR = kTRP - 1;        // tile rows
constexpr int kBoxR = kTRP + 1;      // V-bar box rows (P rows + the +row neighbour)
constexpr int kBoxN = kTNP + 1;      // V-bar box width (P + the +n neighbour)
constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;
constexpr int kRowsPerWave = kTRP / kWaves;
constexpr int kXcds = 8;             // MI355X: workgroups are dealt round-robin to 8 XCDs

// Per-component base pointers: the kernel addresses every component as
// base[comp] + 32-bit element offset (line * wstride + n < 2^32 is checked on
// the host), so addresses stay in scalar registers.
template <int K, typename St>
struct FusedArgs {
    uint32_t nbands, ntn, len, wstride, in_stride, s0, col_begin, col_end;
    const typename St::T *xi_in[3 * (K - 1)];
    typename St::T *xi_out[3 * (K - 1)];
    const typename St::T *vb_in[K - 1];
    typename St::T *vb_out[K - 1];
    float *V[K - 1];
    const float *u;
    float c[K];
    float tau1, tau2, lambda;
    DykstraLimits lim;
    float *dq;
};

// Wave-uniform value in a scalar register.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ ColInfo load_col(const ColInfo *cols, int c)
{
    const ColInfo *p = cols + uniform(c);
    return ColInfo{uint32_t(uniform(int(p->line0))), uint32_t(uniform(int(p->rb))), uint32_t(uniform(int(p->re))), 0};
}

constexpr int kMaxItems = 512;        // work-list capacity per column (overflow is solved in place)
constexpr int kItemsPerLane = 1;      // waves join the work list while each lane gets this many items

template <int K>
struct FusedLds {
    float box[2][K - 1][kBoxR][kBoxN];   // V-bar of columns c and c + 1 (slot = col & 1)
    float xl[kTRP][3 * (K - 1)][kTNP];    // Xi+ of the P points
    float qb[kMaxItems][K][3];            // Dykstra inputs of the queued points
    uint16_t wl[kMaxItems];               // queued point = row * 64 + lane
    int wl_n, wl_next;
    float red[kWaves];
};

__device__ __forceinline__ bool row_live(const ColInfo &ci, int r)
{
    return r >= int(ci.rb) && r < int(ci.re);
}

template <int K, typename St, bool TRACE>
__global__ __launch_bounds__(kThreads, (K <= 3 ? 4 : 2)) void potts_fused_kernel(FusedArgs<K, St> a,
                                                                                  const ColInfo *__restrict__ cols,
                                                                                  const BandInfo *__restrict__ bands)
{
    constexpr int KM = K - 1;
    constexpr int NP = K * (K - 1) / 2;
    constexpr int BOXP = (KM * kBoxR + kWaves - 1) / kWaves;   // box rows loaded per wave
    extern __shared__ __align__(16) unsigned char smem_raw[];
    FusedLds<K> &L = *reinterpret_cast<FusedLds<K> *>(smem_raw);

    // Workgroup -> (band, n-tile): all n-tiles of a band on one XCD, so the
    // n-halo reads of one tile hit the L2 lines its neighbours just fetched.
    const uint32_t xcd = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
    const uint32_t b = (k / a.ntn) * kXcds + xcd, j = k % a.ntn;
    if (b >= a.nbands) return;   // uniform per workgroup
    const BandInfo band = bands[b];
    const int r0 = int(band.r0);
    const int n0 = int(j) * kTN;
    const int lane = threadIdx.x % kWave, w = uniform(threadIdx.x / kWave);
    const int len = int(a.len);
    const int n = n0 - 1 + lane;
    const bool n_ok = n >= 0 && n < len;
    const bool t_lane = lane >= 1;
    const int n2 = n0 - 1 + kTNP;        // the box's extra +n column (lane 0 loads it)
    const ColInfo dead{0, 0, 0, 0};

    // Registers holding one column's HBM data (filled one column ahead).
    struct Col {
        float xin[kRowsPerWave][KM][3];
        float vv[kRowsPerWave][KM];
        float uu[kRowsPerWave];
        float bx[BOXP], bx2[BOXP];       // V-bar of the next column for the box
    };

    auto fetch = [&](int c, Col &f) {
        const ColInfo cc = load_col(cols, c);
        const ColInfo cn = (c + 1 < int(a.col_end)) ? load_col(cols, c + 1) : dead;
#pragma unroll
        for (int p = 0; p < kRowsPerWave; ++p) {
            const int pr = w * kRowsPerWave + p;
            const int r = r0 - 1 + pr;
            f.uu[p] = 0.0f;
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                f.vv[p][i] = 0.0f;
                f.xin[p][i][0] = f.xin[p][i][1] = f.xin[p][i][2] = 0.0f;
            }
            if (n_ok && row_live(cc, r)) {
                const uint32_t line = cc.line0 + uint32_t(r - int(cc.rb));
                const uint32_t off = line * a.wstride + uint32_t(n);
#pragma unroll
                for (int i = 0; i < KM; ++i)
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) f.xin[p][i][ax] = St::ld(a.xi_in[3 * i + ax][off]);
                if (pr >= 1 && t_lane) {
#pragma unroll
                    for (int i = 0; i < KM; ++i) f.vv[p][i] = a.V[i][off];
                    f.uu[p] = a.u[line * a.in_stride + a.s0 + uint32_t(n)];
                }
            }
        }
#pragma unroll
        for (int t = 0; t < BOXP; ++t) {
            const int pair = w + t * kWaves;
            f.bx[t] = f.bx2[t] = 0.0f;
            if (pair < KM * kBoxR) {
                const int i = pair / kBoxR, q = pair % kBoxR;
                const int r = r0 - 1 + q;
                if (row_live(cn, r)) {
                    const uint32_t off = (cn.line0 + uint32_t(r - int(cn.rb))) * a.wstride;
                    if (n_ok) f.bx[t] = St::ld(a.vb_in[i][off + uint32_t(n)]);
                    if (lane == 0 && n2 < len) f.bx2[t] = St::ld(a.vb_in[i][off + uint32_t(n2)]);
                }
            }
        }
    };

    auto commit_box = [&](const Col &f, int slot) {
#pragma unroll
        for (int t = 0; t < BOXP; ++t) {
            const int pair = w + t * kWaves;
            if (pair < KM * kBoxR) {
                const int i = pair / kBoxR, q = pair % kBoxR;
                L.box[slot][i][q][lane] = f.bx[t];
                if (lane == 0) L.box[slot][i][q][kTNP] = f.bx2[t];
            }
        }
    };

    float prev[kRowsPerWave][KM][3];
#pragma unroll
    for (int p = 0; p < kRowsPerWave; ++p)
#pragma unroll
        for (int i = 0; i < KM; ++i) prev[p][i][0] = prev[p][i][1] = prev[p][i][2] = 0.0f;
    float dqmax = 0.0f;

    // Prologue: V-bar of the first column into its box, HBM data of the first column.
    Col cur, nxt;
    fetch(int(band.cs) - 1 >= int(a.col_begin) ? int(band.cs) - 1 : int(band.cs), nxt);
    if (int(band.cs) - 1 >= int(a.col_begin)) {
        commit_box(nxt, int(band.cs) & 1);
    } else {
        // no previous column to fetch from: load the first column's V-bar box directly
        const ColInfo ci = load_col(cols, int(band.cs));
#pragma unroll
        for (int t = 0; t < BOXP; ++t) {
            const int pair = w + t * kWaves;
            if (pair < KM * kBoxR) {
                const int i = pair / kBoxR, q = pair % kBoxR;
                const int r = r0 - 1 + q;
                float v0 = 0.0f, v1 = 0.0f;
                if (row_live(ci, r)) {
                    const uint32_t off = (ci.line0 + uint32_t(r - int(ci.rb))) * a.wstride;
                    if (n_ok) v0 = St::ld(a.vb_in[i][off + uint32_t(n)]);
                    if (lane == 0 && n2 < len) v1 = St::ld(a.vb_in[i][off + uint32_t(n2)]);
                }
                L.box[int(band.cs) & 1][i][q][lane] = v0;
                if (lane == 0) L.box[int(band.cs) & 1][i][q][kTNP] = v1;
            }
        }
    }
    fetch(int(band.cs), cur);

    auto step = [&](const int c, const Col &cur, Col &nxt) {
        const ColInfo cc = load_col(cols, c);
        const ColInfo cn = (c + 1 < int(a.col_end)) ? load_col(cols, c + 1) : dead;
        const ColInfo cp = (c > int(a.col_begin)) ? load_col(cols, c - 1) : dead;
        commit_box(cur, (c + 1) & 1);
        if (threadIdx.x == 0) {
            L.wl_n = 0;
            L.wl_next = 0;
        }
        __syncthreads();   // S1: boxes of c and c+1 in LDS; counters reset
        if (c + 1 < int(band.ce)) fetch(c + 1, nxt);

        // ---- A: Dykstra inputs; finish the points feasible after one sweep ----
        const int sc = c & 1, sn = (c + 1) & 1;
        bool pval[kRowsPerWave];
#pragma unroll
        for (int p = 0; p < kRowsPerWave; ++p) {
            const int pr = w * kRowsPerWave + p;
            const int r = r0 - 1 + pr;
            pval[p] = n_ok && row_live(cc, r);
            bool queued = false;
            if (pval[p]) {
                const bool nx[3] = {n + 1 < len, r + 1 < int(cc.re), row_live(cn, r)};
                float vb[KM], vbn[3][KM];
#pragma unroll
                for (int i = 0; i < KM; ++i) {
                    vb[i] = L.box[sc][i][pr][lane];
                    vbn[0][i] = L.box[sc][i][pr][lane + 1];
                    vbn[1][i] = L.box[sc][i][pr + 1][lane];
                    vbn[2][i] = L.box[sn][i][pr][lane];
                }
                float q[K][3];
                xi_input<K>(vb, vbn, nx, cur.xin[p], a.tau2, q);
                queued = !dykstra_first_sweep_done<K>(q, a.lim);
                // wave-aggregated append to the work list
                const uint64_t mask = __ballot(queued);
                int pos = 0;
                if (mask) {
                    const int leader = __ffsll((unsigned long long)mask) - 1;
                    if (lane == leader) pos = atomicAdd(&L.wl_n, __popcll(mask));
                    pos = __shfl(pos, leader, kWave) + __popcll(mask & ((1ull << lane) - 1ull));
                }
                if (queued && pos < kMaxItems) {
#pragma unroll
                    for (int i = 0; i < K; ++i)
#pragma unroll
                        for (int ax = 0; ax < 3; ++ax) L.qb[pos][i][ax] = q[i][ax];
                    L.wl[pos] = uint16_t(pr * kTNP + lane);
                } else {
                    if (queued) dykstra<K>(q, a.lim);   // work list full: solve in place
#pragma unroll
                    for (int i = 0; i < KM; ++i)
#pragma unroll
                        for (int ax = 0; ax < 3; ++ax) L.xl[pr][3 * i + ax][lane] = St::ld(St::st(q[i][ax]));
                }
            }
        }
        __syncthreads();   // S2: work list complete

        // ---- B: drain the work list, one sweep per trip ----
        // Only as many waves as keep about kItemsPerLane items per lane busy
        // join; the others go straight to the barrier.
        {
            const int nwl = min(L.wl_n, kMaxItems);
            const int waves_needed = (nwl + kItemsPerLane * kWave - 1) / (kItemsPerLane * kWave);
            if (w < waves_needed) {
                float q[K][3], r[NP][3];
                int sw = 0, t = 0;
                bool have = false, drained = false;
                while (true) {
                    const bool need = !have && !drained;
                    const uint64_t mask = __ballot(need);
                    if (mask) {
                        const int leader = __ffsll((unsigned long long)mask) - 1;
                        int base = 0;
                        if (lane == leader) base = atomicAdd(&L.wl_next, __popcll(mask));
                        base = __shfl(base, leader, kWave);
                        if (need) {
                            t = base + __popcll(mask & ((1ull << lane) - 1ull));
                            if (t < nwl) {
#pragma unroll
                                for (int i = 0; i < K; ++i)
#pragma unroll
                                    for (int ax = 0; ax < 3; ++ax) q[i][ax] = L.qb[t][i][ax];
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
                        const float m2 = dykstra_sweep<K>(q, r, a.lim);
                        ++sw;
                        if (m2 < a.lim.stol || sw == 20) {
                            const int pt = L.wl[t];
                            const int pr = pt / kTNP, ln = pt % kTNP;
#pragma unroll
                            for (int i = 0; i < KM; ++i)
#pragma unroll
                                for (int ax = 0; ax < 3; ++ax) L.xl[pr][3 * i + ax][ln] = St::ld(St::st(q[i][ax]));
                            have = false;
                        }
                    }
                }
            }
        }
        __syncthreads();   // S3: Xi+ of every P point in LDS

        // ---- C: store Xi+, V step on the tile points ----
#pragma unroll
        for (int p = 0; p < kRowsPerWave; ++p) {
            const int pr = w * kRowsPerWave + p;
            const int r = r0 - 1 + pr;
            float own[KM][3];
#pragma unroll
            for (int i = 0; i < KM; ++i)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) own[i][ax] = pval[p] ? L.xl[pr][3 * i + ax][lane] : 0.0f;
            if (pval[p] && pr >= 1 && t_lane) {
                const uint32_t off = (cc.line0 + uint32_t(r - int(cc.rb))) * a.wstride + uint32_t(n);
#pragma unroll
                for (int i = 0; i < KM; ++i)
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) a.xi_out[3 * i + ax][off] = St::st(own[i][ax]);
                const bool px[3] = {n > 0, r > int(cc.rb), row_live(cp, r)};
                const bool nx[3] = {n + 1 < len, r + 1 < int(cc.re), row_live(cn, r)};
                float xp[3][KM];
#pragma unroll
                for (int i = 0; i < KM; ++i) {
                    xp[0][i] = L.xl[pr][3 * i + 0][lane - 1];
                    xp[1][i] = L.xl[pr - 1][3 * i + 1][lane];
                    xp[2][i] = prev[p][i][2];
                }
                float vbo[KM];
                float vv[KM];
#pragma unroll
                for (int i = 0; i < KM; ++i) vv[i] = cur.vv[p][i];
                const float d = v_update<K>(own, xp, px, nx, vv, vbo, cur.uu[p], a.c, a.lambda, a.tau1);
                if constexpr (TRACE) dqmax = fmaxf(dqmax, d);
#pragma unroll
                for (int i = 0; i < KM; ++i) {
                    a.V[i][off] = vv[i];
                    a.vb_out[i][off] = St::st(vbo[i]);
                }
            }
#pragma unroll
            for (int i = 0; i < KM; ++i)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) prev[p][i][ax] = own[i][ax];
        }
    };
    for (int c = int(band.cs); c < int(band.ce); c += 2) {
        step(c, cur, nxt);
        if (c + 1 < int(band.ce)) step(c + 1, nxt, cur);
    }

    if constexpr (TRACE) {
        dqmax = wave_max(dqmax);
        if (lane == 0) L.red[w] = dqmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = L.red[0];
            for (int i = 1; i < kWaves; ++i) m = fmaxf(m, L.red[i]);
            atomic_max_nonneg(a.dq, m);
        }
    }
}

template <int K, typename St>
int launch_fused(sct_potts_solver *s, int iters, float *dq_dev)
{
    using T = typename St::T;
    const Geometry &geo = s->g->geo;
    FusedArgs<K, St> a;
    a.nbands = s->n_bands;
    a.len = s->s1 - s->s0;
    a.ntn = (a.len + kTN - 1) / kTN;
    a.wstride = s->wstride;
    a.in_stride = geo.nstride;
    a.s0 = s->s0;
    a.col_begin = geo.col_begin;
    a.col_end = geo.col_end;
    for (int i = 0; i < K - 1; ++i) a.V[i] = s->V + i * s->wspan;
    a.u = (const float *)s->g->data;
    for (int i = 0; i < K; ++i) a.c[i] = s->c[i];
    a.tau1 = s->tau1;
    a.tau2 = s->tau2;
    a.lambda = s->lambda;
    a.lim = dykstra_limits();
    if (!a.nbands || !a.len) return SCT_OK;
    const size_t lds = sizeof(FusedLds<K>);
    static bool attr_set[2] = {false, false};
    if (lds > 65536) {
        for (int t = 0; t < 2; ++t) {
            if (attr_set[t]) continue;
            const void *fn = t ? (const void *)potts_fused_kernel<K, St, true> : (const void *)potts_fused_kernel<K, St, false>;
            SCT_HIP(s->ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            attr_set[t] = true;
        }
    }
    const uint32_t grid = ((a.nbands + kXcds - 1) / kXcds) * kXcds * a.ntn;
    for (int it = 0; it < iters; ++it) {
        for (int j = 0; j < 3 * (K - 1); ++j) {
            a.xi_in[j] = (const T *)s->xi[s->cur] + j * s->wspan;
            a.xi_out[j] = (T *)s->xi[s->cur ^ 1] + j * s->wspan;
        }
        for (int i = 0; i < K - 1; ++i) {
            a.vb_in[i] = (const T *)s->vbar[s->cur] + i * s->wspan;
            a.vb_out[i] = (T *)s->vbar[s->cur ^ 1] + i * s->wspan;
        }
        potts_mark(s);
        if (dq_dev) {
            a.dq = dq_dev + it;
            potts_fused_kernel<K, St, true><<<grid, kThreads, lds, s->ctx->stream>>>(a, s->d_cols, s->d_bands);
        } else {
            a.dq = nullptr;
            potts_fused_kernel<K, St, false><<<grid, kThreads, lds, s->ctx->stream>>>(a, s->d_cols, s->d_bands);
        }
        s->cur ^= 1;
    }
    SCT_HIP(s->ctx, hipGetLastError());
    return SCT_OK;
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
    std::vector<BandInfo> bands;
    for (uint32_t r0 = 0; r0 < geo.height; r0 += kTR) {
        uint32_t cs = ~0u, ce = 0;
        for (uint32_t c = geo.col_begin; c < geo.col_end; ++c) {
            const ColInfo &ci = cols[c];
            if (ci.re > ci.rb && ci.rb < r0 + kTR && ci.re > r0) {
                cs = std::min(cs, c);
                ce = c + 1;
            }
        }
        if (ce > 0) bands.push_back(BandInfo{r0, cs, ce, 0});
    }
    // Longest marches first, so the short edge bands fill the tail.
    std::stable_sort(bands.begin(), bands.end(),
                     [](const BandInfo &x, const BandInfo &y) { return x.ce - x.cs > y.ce - y.cs; });
    s->n_bands = uint32_t(bands.size());
    int rc;
    if ((rc = sct_alloc(s->ctx, (void **)&s->d_cols, sizeof(ColInfo) * cols.size()))) return rc;
    if ((rc = sct_alloc(s->ctx, (void **)&s->d_bands, sizeof(BandInfo) * std::max<size_t>(1, bands.size())))) return rc;
    SCT_HIP(s->ctx, hipMemcpy(s->d_cols, cols.data(), sizeof(ColInfo) * cols.size(), hipMemcpyHostToDevice));
    if (!bands.empty())
        SCT_HIP(s->ctx, hipMemcpy(s->d_bands, bands.data(), sizeof(BandInfo) * bands.size(), hipMemcpyHostToDevice));
    return SCT_OK;
}

void sct::fused_release(sct_potts_solver *s)
{
    sct_free(s->ctx, s->d_cols);
    sct_free(s->ctx, s->d_bands);
    s->d_cols = nullptr;
    s->d_bands = nullptr;
    s->n_bands = 0;
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
