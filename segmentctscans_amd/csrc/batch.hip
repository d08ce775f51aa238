// batch.hip -- many independent Potts solves in one pass per half-step (no
// counterpart in the reference, whose driver solves one volume per call,
// potts-model/potts-model.cpp:164).
//
// Problems are grouped by K.  A group lays the lines of all its solves side by
// side in one work layout (each solve keeps its own line stride, neighbour
// links never cross solves), so one Xi launch and one V launch per iteration
// cover every solve of the group; the V step reads each line's labels and
// lambda from a per-solve table.  Problems that share an input field share
// one gathered copy of it.  With SCT_BATCH_GRAPH the launches of all
// iterations of all groups are captured once into a hipGraph and run as a
// single graph launch.  Every solve computes exactly what
// sct_potts_min_partition computes for it (same kernels, same arithmetic).
#include "potts.hpp"
#include "twopass.hpp"

#include <algorithm>
#include <map>
#include <memory>
#include <vector>

using namespace sct;
using namespace sct::tp;

namespace {

struct Group {
    int K = 0;
    int bw = kXiBlockWaves;            // Xi workgroup tile: waves per workgroup
    std::vector<int> probs;            // problem indices
    std::vector<uint64_t> r_off;       // result offset of each solve in R
    uint64_t wspan = 0, n_chunks = 0, r_total = 0, u_total = 0;
    void *xi = nullptr, *vbar = nullptr;
    float *V = nullptr, *U = nullptr, *R = nullptr;
    uint2 *chunks = nullptr;
    BatchLine *lines = nullptr;
    void *params = nullptr;
    // host copies kept alive until the uploads complete
    std::vector<uint2> h_chunks;
    std::vector<BatchLine> h_lines;
    std::vector<char> h_params;
};

void free_group(sct_ctx *ctx, Group &g)
{
    for (void *p : {g.xi, g.vbar, (void *)g.V, (void *)g.U, (void *)g.R, (void *)g.chunks, (void *)g.lines, g.params})
        sct_free(ctx, p);
    g.xi = g.vbar = g.params = nullptr;
    g.V = g.U = g.R = nullptr;
    g.chunks = nullptr;
    g.lines = nullptr;
}

template <int K, typename St>
int build_group(sct_ctx *ctx, const sct_potts_problem *pr, Group &G)
{
    constexpr uint32_t cw = kWave * vpl_of<St>();
    // inputs: one gathered copy per distinct field
    std::map<const sct_field *, uint64_t> u_of;
    for (int p : G.probs) {
        const sct_field *f = pr[p].g;
        if (u_of.count(f)) continue;
        u_of[f] = G.u_total;
        G.u_total += f->geo.span();
    }
    std::vector<SolveParams<K>> params(G.probs.size());
    uint64_t wb = 0, rb = 0;
    for (size_t s = 0; s < G.probs.size(); ++s) {
        const sct_potts_problem &P = pr[G.probs[s]];
        const Geometry &geo = P.g->geo;
        const uint32_t n = geo.n_images, ws = sct_get_stride(n);
        const uint64_t ub = u_of[P.g];
        for (int i = 0; i < K; ++i) params[s].c[i] = P.c[i];
        params[s].lambda = P.lambda;
        G.r_off.push_back(rb);
        const uint32_t line0 = uint32_t(G.h_lines.size());
        for (uint32_t l = 0; l < geo.lines; ++l) {
            const LineNbr nb = geo.tables->h_nbr[l];
            const auto at = [&](int32_t x) { return x >= 0 ? uint32_t(wb + uint64_t(x) * ws) : kNone; };
            BatchLine L{};
            L.w = uint32_t(wb + uint64_t(l) * ws);
            L.u = uint32_t(ub + uint64_t(l) * geo.nstride);
            L.r = uint32_t(rb + uint64_t(l) * geo.nstride);
            L.len = n;
            L.nr = at(nb.nrow);
            L.nc = at(nb.ncol);
            L.pr = at(nb.prow);
            L.pc = at(nb.pcol);
            L.solve = uint32_t(s);
            G.h_lines.push_back(L);
            for (uint32_t n0 = 0; n0 < n; n0 += cw) G.h_chunks.push_back(uint2{line0 + l, n0});
        }
        wb += uint64_t(geo.lines) * ws;
        rb += geo.span();
    }
    G.wspan = wb;
    G.r_total = rb;
    G.n_chunks = G.h_chunks.size();
    if (G.wspan >= kNone || G.r_total >= kNone || G.u_total >= kNone || G.h_lines.size() >= kNone)
        return sct_fail(ctx, SCT_ERR_SIZE, "batch group exceeds 32-bit offsets; split the batch");
    const size_t sb = sizeof(typename St::T);
    const size_t comp = std::max<uint64_t>(1, G.wspan);
    int rc;
    if ((rc = sct_alloc(ctx, &G.xi, sb * comp * 3 * (K - 1)))) return rc;
    if ((rc = sct_alloc(ctx, &G.vbar, sb * comp * (K - 1)))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&G.V, sizeof(float) * comp * (K - 1)))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&G.U, sizeof(float) * std::max<uint64_t>(1, G.u_total)))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&G.R, sizeof(float) * std::max<uint64_t>(1, G.r_total)))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&G.chunks, sizeof(uint2) * std::max<size_t>(1, G.h_chunks.size())))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&G.lines, sizeof(BatchLine) * std::max<size_t>(1, G.h_lines.size())))) return rc;
    if ((rc = sct_alloc(ctx, &G.params, sizeof(SolveParams<K>) * params.size()))) return rc;
    G.h_params.assign((const char *)params.data(), (const char *)(params.data() + params.size()));
    hipStream_t st = ctx->stream;
    SCT_HIP(ctx, hipMemcpyAsync(G.chunks, G.h_chunks.data(), sizeof(uint2) * G.h_chunks.size(), hipMemcpyHostToDevice, st));
    SCT_HIP(ctx, hipMemcpyAsync(G.lines, G.h_lines.data(), sizeof(BatchLine) * G.h_lines.size(), hipMemcpyHostToDevice, st));
    SCT_HIP(ctx, hipMemcpyAsync(G.params, G.h_params.data(), G.h_params.size(), hipMemcpyHostToDevice, st));
    for (const auto &kv : u_of)
        SCT_HIP(ctx, hipMemcpyAsync(G.U + kv.second, kv.first->data, sizeof(float) * kv.first->geo.span(),
                                    hipMemcpyDeviceToDevice, st));
    SCT_HIP(ctx, hipMemsetAsync(G.R, 0, sizeof(float) * std::max<uint64_t>(1, G.r_total), st));
    return SCT_OK;
}

template <int K, typename St>
TwoPassArgs<K, St> group_args(const Group &G)
{
    using T = typename St::T;
    TwoPassArgs<K, St> a;
    for (int j = 0; j < 3 * (K - 1); ++j) a.xi[j] = (T *)G.xi + j * G.wspan;
    for (int i = 0; i < K - 1; ++i) {
        a.vbar[i] = (T *)G.vbar + i * G.wspan;
        a.V[i] = G.V + i * G.wspan;
    }
    a.u = G.U;
    a.in_place();
    return a;
}

// Phases of a group: 0 build, 1 init, 2 one iteration, 3 labels.
template <int K, typename St>
int group_phase(sct_ctx *ctx, const sct_potts_problem *pr, Group &G, int phase, float tau1, float tau2)
{
    if (phase == 0) return build_group<K, St>(ctx, pr, G);
    if (!G.n_chunks) return SCT_OK;
    const BatchLines L{G.chunks, G.lines, G.n_chunks};
    const TwoPassArgs<K, St> a = group_args<K, St>(G);
    const SolveTable<K> prm{(const SolveParams<K> *)G.params};
    hipStream_t st = ctx->stream;
    if (phase == 1) {
        potts_init_kernel<K, St><<<wave_grid(L.count()), 256, 0, st>>>(L, a, prm);
        potts_xi_zero_kernel<K, St><<<wave_grid(L.count()), 256, 0, st>>>(L, a);
    } else if (phase == 2) {
        const DykstraLimits lim = dykstra_limits();
        if (G.bw == 1)
            potts_xi_kernel<K, St, BatchLines, 1><<<xi_grid<St, 1>(L.count()), 64, 0, st>>>(L, a, tau2, lim);
        else if (G.bw == 2)
            potts_xi_kernel<K, St, BatchLines, 2><<<xi_grid<St, 2>(L.count()), 128, 0, st>>>(L, a, tau2, lim);
        else
            potts_xi_kernel<K, St, BatchLines, 4><<<xi_grid<St, 4>(L.count()), 256, 0, st>>>(L, a, tau2, lim);
        potts_v_kernel<K, St, false><<<wave_grid(L.count()), 256, 0, st>>>(L, a, prm, tau1, nullptr);
    } else {
        potts_labels_kernel<K, St><<<wave_grid(L.count()), 256, 0, st>>>(L, a, prm, G.R, 0u, 0xffffffffu);
    }
    SCT_HIP(ctx, hipGetLastError());
    return SCT_OK;
}

template <int K>
int group_phase_k(sct_ctx *ctx, const sct_potts_problem *pr, Group &G, int storage, int phase, float t1, float t2)
{
    switch (storage) {
    case SCT_STORAGE_BF16: return group_phase<K, StorageBF16>(ctx, pr, G, phase, t1, t2);
    case SCT_STORAGE_F16: return group_phase<K, StorageF16>(ctx, pr, G, phase, t1, t2);
    default: return group_phase<K, StorageF32>(ctx, pr, G, phase, t1, t2);
    }
}

int run_phase(sct_ctx *ctx, const sct_potts_problem *pr, Group &G, int storage, int phase, float t1, float t2)
{
    switch (G.K) {
    case 2: return group_phase_k<2>(ctx, pr, G, storage, phase, t1, t2);
    case 3: return group_phase_k<3>(ctx, pr, G, storage, phase, t1, t2);
    case 4: return group_phase_k<4>(ctx, pr, G, storage, phase, t1, t2);
    case 5: return group_phase_k<5>(ctx, pr, G, storage, phase, t1, t2);
    default: return group_phase_k<6>(ctx, pr, G, storage, phase, t1, t2);
    }
}

// The autotuner: for each group, times kTuneIters iterations with each Xi
// workgroup tile (BW = 4, 2, 1 waves: 1024, 512, 256 voxels and work lists
// of 512, 256, 128 entries), keeps the fastest, and starts the group again
// from its initial state, so the result is the untuned one.  Choices are
// cached in the context per (K, storage, size class of the group).
constexpr int kTuneIters = 3;
constexpr int kTiles[3] = {4, 2, 1};

uint64_t tile_key(const Group &G, int storage)
{
    uint64_t cls = 0;
    for (uint64_t c = G.n_chunks; c > 1; c >>= 1) ++cls;   // log2 of the chunk count
    return (uint64_t(G.K) << 16) | (uint64_t(storage) << 8) | cls;
}

int autotune(sct_ctx *ctx, const sct_potts_problem *pr, std::vector<Group> &groups, int storage, float t1, float t2)
{
    hipStream_t st = ctx->stream;
    hipEvent_t ev[4] = {};
    int rc = SCT_OK;
    for (auto &e : ev)
        if (!rc && hipEventCreate(&e) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventCreate failed");
    for (Group &G : groups) {
        if (rc || !G.n_chunks || ctx->forced_tiles[G.K - 2]) continue;
        const uint64_t key = tile_key(G, storage);
        auto hit = ctx->tiles.find(key);
        if (hit != ctx->tiles.end()) {
            G.bw = hit->second;
            continue;
        }
        float best = 0;
        int best_bw = kXiBlockWaves;
        for (int i = 0; i < 3 && !rc; ++i) {
            G.bw = kTiles[i];
            // one untimed iteration, then kTuneIters timed ones
            rc = run_phase(ctx, pr, G, storage, 2, t1, t2);
            if (!rc && hipEventRecord(ev[0], st) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventRecord");
            for (int it = 0; it < kTuneIters && !rc; ++it) rc = run_phase(ctx, pr, G, storage, 2, t1, t2);
            if (!rc && hipEventRecord(ev[1], st) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventRecord");
            float ms = 0;
            if (!rc && (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
                rc = sct_fail(ctx, SCT_ERR_HIP, "autotune timing failed");
            if (!rc && (i == 0 || ms < best)) {
                best = ms;
                best_bw = kTiles[i];
            }
        }
        G.bw = best_bw;
        if (!rc) ctx->tiles[key] = best_bw;
        if (!rc) rc = run_phase(ctx, pr, G, storage, 1, t1, t2);   // back to the initial state
    }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

// Runs the iterations of all groups, directly or as one captured graph, and
// times them with events on the context's stream.
int run_iterations(sct_ctx *ctx, const sct_potts_problem *pr, std::vector<Group> &groups, int storage, int iters,
                   float t1, float t2, bool graph, float *elapsed_ms)
{
    hipStream_t st = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    int rc = SCT_OK;
    auto launch_all = [&]() {
        for (int it = 0; it < iters && !rc; ++it)
            for (Group &G : groups)
                if (!rc) rc = run_phase(ctx, pr, G, storage, 2, t1, t2);
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventCreate failed");
    if (!rc && graph && iters > 0 && !groups.empty()) {
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "hipStreamBeginCapture");
        if (!rc) launch_all();
        e = hipStreamEndCapture(st, &g);
        if (!rc && e != hipSuccess) rc = sct_hip_fail(ctx, e, "hipStreamEndCapture");
        if (!rc && (e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0)) != hipSuccess)
            rc = sct_hip_fail(ctx, e, "hipGraphInstantiate");
    }
    if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventRecord failed");
    if (!rc) {
        if (ge) {
            const hipError_t e = hipGraphLaunch(ge, st);
            if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "hipGraphLaunch");
        } else {
            launch_all();
        }
    }
    if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventRecord failed");
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "batch iterations failed");
    if (!rc && elapsed_ms && hipEventElapsedTime(elapsed_ms, e0, e1) != hipSuccess)
        rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventElapsedTime failed");
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

}  // namespace

extern "C" int sct_potts_batch_set_tiles(sct_ctx *ctx, const int *waves)
{
    if (!ctx || !waves) return SCT_ERR_INVALID;
    for (int k = 0; k < 5; ++k)
        if (waves[k] != 0 && waves[k] != 1 && waves[k] != 2 && waves[k] != 4)
            return sct_fail(ctx, SCT_ERR_INVALID, "tiles are 1, 2 or 4 waves (0: not set)");
    for (int k = 0; k < 5; ++k) ctx->forced_tiles[k] = waves[k];
    return SCT_OK;
}

extern "C" int sct_potts_batch_tiles(sct_ctx *ctx, int *waves)
{
    if (!ctx || !waves) return SCT_ERR_INVALID;
    for (int k = 0; k < 5; ++k) waves[k] = ctx->last_tiles[k];
    return SCT_OK;
}

extern "C" int sct_potts_batch(sct_ctx *ctx, int n, const sct_potts_problem *problems, float tau1, float tau2,
                               int iters, int storage, int flags, sct_field **out, float *elapsed_ms)
{
    if (!ctx) return SCT_ERR_INVALID;
    if (n < 0 || (n && (!problems || !out)) || iters < 0 || storage < SCT_STORAGE_F32 || storage > SCT_STORAGE_F16)
        return sct_fail(ctx, SCT_ERR_INVALID, "invalid batch arguments");
    for (int p = 0; p < n; ++p) {
        out[p] = nullptr;
        if (!problems[p].g || problems[p].g->dtype != SCT_F32 || problems[p].K < 2 || problems[p].K > 6)
            return sct_fail(ctx, SCT_ERR_INVALID, "batch problem " + std::to_string(p) + ": bad field or K");
        if (problems[p].g->ctx != ctx)
            return sct_fail(ctx, SCT_ERR_INVALID, "batch problem " + std::to_string(p) + ": field of another context");
    }
    std::vector<Group> groups;
    for (int K = 2; K <= 6; ++K) {
        Group G;
        G.K = K;
        for (int p = 0; p < n; ++p)
            if (problems[p].K == K) G.probs.push_back(p);
        if (!G.probs.empty()) groups.push_back(std::move(G));
    }
    int rc = SCT_OK;
    for (Group &G : groups)
        if (!rc) rc = run_phase(ctx, problems, G, storage, 0, tau1, tau2);
    for (Group &G : groups)
        if (!rc) rc = run_phase(ctx, problems, G, storage, 1, tau1, tau2);
    for (Group &G : groups)
        if (ctx->forced_tiles[G.K - 2]) G.bw = ctx->forced_tiles[G.K - 2];
    if (!rc && (flags & SCT_BATCH_AUTOTUNE) && iters > 0) rc = autotune(ctx, problems, groups, storage, tau1, tau2);
    for (int K = 2; K <= 6; ++K) ctx->last_tiles[K - 2] = 0;
    for (const Group &G : groups) ctx->last_tiles[G.K - 2] = G.n_chunks ? G.bw : 0;
    if (!rc) rc = run_iterations(ctx, problems, groups, storage, iters, tau1, tau2, (flags & SCT_BATCH_GRAPH) != 0,
                                 elapsed_ms);
    for (Group &G : groups)
        if (!rc) rc = run_phase(ctx, problems, G, storage, 3, tau1, tau2);
    for (Group &G : groups) {
        for (size_t s = 0; s < G.probs.size() && !rc; ++s) {
            const int p = G.probs[s];
            const Geometry &geo = problems[p].g->geo;
            if ((rc = sct_field_like(ctx, problems[p].g, SCT_F32, &out[p]))) break;
            const hipError_t e = hipMemcpyAsync(out[p]->data, G.R + G.r_off[s], sizeof(float) * geo.span(),
                                                hipMemcpyDeviceToDevice, ctx->stream);
            if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "batch result copy");
        }
    }
    if (!rc) {
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = sct_hip_fail(ctx, e, "batch");
    } else {
        (void)hipStreamSynchronize(ctx->stream);
    }
    for (Group &G : groups) free_group(ctx, G);
    if (rc)
        for (int p = 0; p < n; ++p) {
            sct_field_destroy(out[p]);
            out[p] = nullptr;
        }
    return rc;
}
