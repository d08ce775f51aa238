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
// Wave w owns P rows 4w..4w+3; lane 0 is the n-halo, lanes 1..63 are tile
// voxels, so the halo costs no extra pass.  Xi+ of the current column goes to
// LDS for the -n / -row neighbours and stays in registers as the -col
// neighbour of the next column.  V-bar of columns c and c+1 sit in two LDS
// boxes that alternate.  Xi and V-bar are double-buffered in HBM (read the
// previous iteration, write the next), so no tile ever sees a neighbour's
// half-updated state; V and u are only touched at a tile's own voxels.
//
// Per voxel and iteration the kernel moves the algorithmic 4(10(K-1)+1) bytes
// plus halo re-reads, which mostly hit L2 (neighbour tiles of one band are
// placed on one XCD).  Arithmetic per voxel is exactly that of the two-pass
// kernels (potts_math.hpp), so results are bitwise identical.
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
constexpr int kWaves = 4;
constexpr int kRowsPerWave = kTRP / kWaves;
constexpr int kXcds = 8;             // MI355X: workgroups are dealt round-robin to 8 XCDs

template <int K>
struct FusedArgs {
    const ColInfo *cols;
    const BandInfo *bands;
    uint32_t nbands, ntn, len, wstride, in_stride, s0, col_begin, col_end;
    uint64_t wspan;
    const float *xi_in;
    float *xi_out;
    const float *vb_in;
    float *vb_out;
    float *V;
    const float *u;
    float c[K];
    float tau1, tau2, lambda;
    DykstraLimits lim;
    float *dq;
};

template <int K>
constexpr size_t fused_lds_bytes()
{
    return sizeof(float) * (size_t(2) * (K - 1) * kBoxR * kBoxN + size_t(kTRP) * 3 * (K - 1) * kTNP);
}

__device__ __forceinline__ bool row_live(const ColInfo &ci, int r)
{
    return r >= int(ci.rb) && r < int(ci.re);
}

template <int K, bool TRACE>
__global__ __launch_bounds__(256) void potts_fused_kernel(FusedArgs<K> a)
{
    constexpr int KM = K - 1;
    constexpr int NX = 3 * KM;
    extern __shared__ float smem[];
    float(*box)[KM][kBoxR][kBoxN] = reinterpret_cast<float(*)[KM][kBoxR][kBoxN]>(smem);
    float(*xl)[NX][kTNP] = reinterpret_cast<float(*)[NX][kTNP]>(smem + 2 * KM * kBoxR * kBoxN);

    // Workgroup -> (band, n-tile): all n-tiles of a band on one XCD, so the
    // n-halo reads of one tile hit the L2 lines its neighbours just fetched.
    const uint32_t xcd = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
    const uint32_t b = (k / a.ntn) * kXcds + xcd, j = k % a.ntn;
    if (b >= a.nbands) return;   // uniform per workgroup
    const BandInfo band = a.bands[b];
    const int r0 = int(band.r0);
    const int n0 = int(j) * kTN;
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    const int len = int(a.len);
    const int n = n0 - 1 + lane;
    const bool n_ok = n >= 0 && n < len;
    const bool t_lane = lane >= 1;
    const ColInfo dead{0, 0, 0, 0};

    float cur[kRowsPerWave][KM][3];
    float prev[kRowsPerWave][KM][3];
#pragma unroll
    for (int p = 0; p < kRowsPerWave; ++p)
#pragma unroll
        for (int i = 0; i < KM; ++i) prev[p][i][0] = prev[p][i][1] = prev[p][i][2] = 0.0f;
    float dqmax = 0.0f;

    auto load_box = [&](int col, int slot) {
        const ColInfo ci = (col < int(a.col_end)) ? a.cols[col] : dead;
        for (int pair = w; pair < KM * kBoxR; pair += kWaves) {
            const int i = pair / kBoxR, q = pair % kBoxR;
            const int r = r0 - 1 + q;
            float v0 = 0.0f, v1 = 0.0f;
            if (row_live(ci, r)) {
                const float *src = a.vb_in + i * a.wspan + uint64_t(ci.line0 + uint32_t(r - int(ci.rb))) * a.wstride;
                if (n_ok) v0 = src[n];
                const int n2 = n0 - 1 + kTNP;
                if (lane == 0 && n2 < len) v1 = src[n2];
            }
            box[slot][i][q][lane] = v0;
            if (lane == 0) box[slot][i][q][kTNP] = v1;
        }
    };

    for (int c = int(band.cs); c < int(band.ce); ++c) {
        const ColInfo cc = a.cols[c];
        const ColInfo cn = (c + 1 < int(a.col_end)) ? a.cols[c + 1] : dead;
        const ColInfo cp = (c > int(a.col_begin)) ? a.cols[c - 1] : dead;
        if (c == int(band.cs)) load_box(c, c & 1);
        load_box(c + 1, (c + 1) & 1);

        bool pval[kRowsPerWave];
        float vv[kRowsPerWave][KM], uu[kRowsPerWave];
#pragma unroll
        for (int p = 0; p < kRowsPerWave; ++p) {
            const int pr = w * kRowsPerWave + p;
            const int r = r0 - 1 + pr;
            pval[p] = n_ok && row_live(cc, r);
            uu[p] = 0.0f;
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                vv[p][i] = 0.0f;
                cur[p][i][0] = cur[p][i][1] = cur[p][i][2] = 0.0f;
            }
            if (pval[p]) {
                const uint64_t line = uint64_t(cc.line0 + uint32_t(r - int(cc.rb)));
                const uint64_t off = line * a.wstride + uint64_t(n);
#pragma unroll
                for (int i = 0; i < KM; ++i)
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) cur[p][i][ax] = a.xi_in[(3 * i + ax) * a.wspan + off];
                if (pr >= 1 && t_lane) {
#pragma unroll
                    for (int i = 0; i < KM; ++i) vv[p][i] = a.V[i * a.wspan + off];
                    uu[p] = a.u[line * a.in_stride + a.s0 + uint32_t(n)];
                }
            }
        }
        __syncthreads();   // boxes visible; the previous column's xl reads are done

        // ---- Xi+ on P ----
        const int sc = c & 1, sn = (c + 1) & 1;
#pragma unroll
        for (int p = 0; p < kRowsPerWave; ++p) {
            const int pr = w * kRowsPerWave + p;
            const int r = r0 - 1 + pr;
            if (pval[p]) {
                const bool nx[3] = {n + 1 < len, r + 1 < int(cc.re), row_live(cn, r)};
                float vb[KM], vbn[3][KM];
#pragma unroll
                for (int i = 0; i < KM; ++i) {
                    vb[i] = box[sc][i][pr][lane];
                    vbn[0][i] = box[sc][i][pr][lane + 1];
                    vbn[1][i] = box[sc][i][pr + 1][lane];
                    vbn[2][i] = box[sn][i][pr][lane];
                }
                xi_update<K>(vb, vbn, nx, cur[p], a.tau2, a.lim);
                if (pr >= 1 && t_lane) {
                    const uint64_t off = uint64_t(cc.line0 + uint32_t(r - int(cc.rb))) * a.wstride + uint64_t(n);
#pragma unroll
                    for (int i = 0; i < KM; ++i)
#pragma unroll
                        for (int ax = 0; ax < 3; ++ax) a.xi_out[(3 * i + ax) * a.wspan + off] = cur[p][i][ax];
                }
            }
#pragma unroll
            for (int i = 0; i < KM; ++i)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) xl[pr][3 * i + ax][lane] = cur[p][i][ax];
        }
        __syncthreads();   // Xi+ of column c visible

        // ---- V+ on T ----
#pragma unroll
        for (int p = 0; p < kRowsPerWave; ++p) {
            const int pr = w * kRowsPerWave + p;
            const int r = r0 - 1 + pr;
            if (pval[p] && pr >= 1 && t_lane) {
                const bool px[3] = {n > 0, r > int(cc.rb), row_live(cp, r)};
                const bool nx[3] = {n + 1 < len, r + 1 < int(cc.re), row_live(cn, r)};
                float xp[3][KM];
#pragma unroll
                for (int i = 0; i < KM; ++i) {
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) {
                        // backward neighbour along axis ax, component ax of label i
                        const float fn = xl[pr][3 * i + ax][lane - 1];
                        const float fr = xl[pr - 1][3 * i + ax][lane];
                        const float fc = prev[p][i][ax];
                        if (ax == 0) xp[0][i] = fn;
                        if (ax == 1) xp[1][i] = fr;
                        if (ax == 2) xp[2][i] = fc;
                    }
                }
                float vbo[KM];
                const float d = v_update<K>(cur[p], xp, px, nx, vv[p], vbo, uu[p], a.c, a.lambda, a.tau1);
                if constexpr (TRACE) dqmax = fmaxf(dqmax, d);
                const uint64_t off = uint64_t(cc.line0 + uint32_t(r - int(cc.rb))) * a.wstride + uint64_t(n);
#pragma unroll
                for (int i = 0; i < KM; ++i) {
                    a.V[i * a.wspan + off] = vv[p][i];
                    a.vb_out[i * a.wspan + off] = vbo[i];
                }
            }
        }
#pragma unroll
        for (int p = 0; p < kRowsPerWave; ++p)
#pragma unroll
            for (int i = 0; i < KM; ++i)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) prev[p][i][ax] = cur[p][i][ax];
    }

    if constexpr (TRACE) {
        __shared__ float red[kWaves];
        dqmax = wave_max(dqmax);
        if (lane == 0) red[w] = dqmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = red[0];
            for (int i = 1; i < kWaves; ++i) m = fmaxf(m, red[i]);
            atomic_max_nonneg(a.dq, m);
        }
    }
}

template <int K>
int launch_fused(sct_potts_solver *s, int iters, float *dq_dev)
{
    const Geometry &geo = s->g->geo;
    FusedArgs<K> a;
    a.cols = s->d_cols;
    a.bands = s->d_bands;
    a.nbands = s->n_bands;
    a.len = s->s1 - s->s0;
    a.ntn = (a.len + kTN - 1) / kTN;
    a.wstride = s->wstride;
    a.in_stride = geo.nstride;
    a.s0 = s->s0;
    a.col_begin = geo.col_begin;
    a.col_end = geo.col_end;
    a.wspan = s->wspan;
    a.V = s->V;
    a.u = (const float *)s->g->data;
    for (int i = 0; i < K; ++i) a.c[i] = s->c[i];
    a.tau1 = s->tau1;
    a.tau2 = s->tau2;
    a.lambda = s->lambda;
    a.lim = dykstra_limits();
    if (!a.nbands || !a.len) return SCT_OK;
    const size_t lds = fused_lds_bytes<K>();
    static bool attr_set[2] = {false, false};
    if (lds > 65536) {
        for (int t = 0; t < 2; ++t) {
            if (attr_set[t]) continue;
            const void *fn = t ? (const void *)potts_fused_kernel<K, true> : (const void *)potts_fused_kernel<K, false>;
            SCT_HIP(s->ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            attr_set[t] = true;
        }
    }
    const uint32_t grid = ((a.nbands + kXcds - 1) / kXcds) * kXcds * a.ntn;
    for (int it = 0; it < iters; ++it) {
        a.xi_in = s->xi[s->cur];
        a.xi_out = s->xi[s->cur ^ 1];
        a.vb_in = s->vbar[s->cur];
        a.vb_out = s->vbar[s->cur ^ 1];
        if (dq_dev) {
            a.dq = dq_dev + it;
            potts_fused_kernel<K, true><<<grid, kWaves * kWave, lds, s->ctx->stream>>>(a);
        } else {
            a.dq = nullptr;
            potts_fused_kernel<K, false><<<grid, kWaves * kWave, lds, s->ctx->stream>>>(a);
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

int sct::fused_iterate(sct_potts_solver *s, int iters, float *dq_dev)
{
    switch (s->K) {
    case 2: return launch_fused<2>(s, iters, dq_dev);
    case 3: return launch_fused<3>(s, iters, dq_dev);
    case 4: return launch_fused<4>(s, iters, dq_dev);
    case 5: return launch_fused<5>(s, iters, dq_dev);
    case 6: return launch_fused<6>(s, iters, dq_dev);
    default: return SCT_ERR_INVALID;
    }
}
