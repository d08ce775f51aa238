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
// single graph launch.  With SCT_BATCH_PERSISTENT each group runs all its
// iterations in one persistent launch: its workgroups take Xi tiles, then V
// tiles, from one ordered sequence of tickets, and start a half-step's tiles
// once every ticket of the half-step before is complete.  Every solve computes exactly what sct_potts_min_partition
// computes for it (same per-tile code, same arithmetic).
#include "potts.hpp"
#include "twopass.hpp"

#include <algorithm>
#include <cstddef>
#include <map>
#include <memory>
#include <vector>

using namespace sct;
using namespace sct::tp;

namespace {

struct Group {
    int K = 0;
    int bw = kXiBlockWaves;            // Xi workgroup tile: waves per workgroup
    int occ = 0;                       // persistent launch: workgroups per CU (0: as many as fit, -1: one fewer)
    int per_cu = 0;                    // persistent launch: workgroups per CU of the last launch
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

// Tile tickets and completion count of a persistent launch (zeroed by
// sync_reset_kernel before it).
constexpr size_t kPersistArgBytes = 1024;
struct GridSync {
    unsigned long long ticket;   // next ticket to hand out
    unsigned long long done;     // tickets completed and published
    unsigned err, pad[11];
    alignas(16) unsigned char args[kPersistArgBytes];   // the launch's PersistArgs
};

// Arguments of a persistent launch, kept in device memory: for each ticket its
// workgroups read the pointers of that half-step afresh (scalar loads), so
// neither the other half-step's pointers nor the last ticket's are held in
// registers across the ticket loop.
template <int K, typename St>
struct PersistArgs {
    BatchLines L;
    TwoPassArgs<K, St> a;
    SolveTable<K> prm;
};
static_assert(sizeof(PersistArgs<6, StorageF32>) <= kPersistArgBytes, "persistent argument table size");

template <typename X>
using ConstAS = __attribute__((address_space(4))) X;
template <typename X>
using GlobalAS = __attribute__((address_space(1))) X;

// p, which points to global memory: the cast there and back lets the compiler
// use global rather than flat accesses through pointers read from the table.
template <typename X>
__device__ __forceinline__ X *global_ptr(X *p)
{
    return (X *)(GlobalAS<X> *)p;
}

// The argument table, read afresh: the empty asm makes its address opaque, so
// its loads are neither hoisted out of the ticket loop nor kept across it.
template <int K, typename St>
__device__ __forceinline__ const ConstAS<PersistArgs<K, St>> *fresh_args(const PersistArgs<K, St> *p)
{
    const ConstAS<PersistArgs<K, St>> *q = (const ConstAS<PersistArgs<K, St>> *)p;
    asm volatile("" : "+s"(q));
    return q;
}

template <int K, typename St>
__device__ __forceinline__ BatchLines lines_of(const ConstAS<PersistArgs<K, St>> *p)
{
    return BatchLines{global_ptr(p->L.chunks), global_ptr(p->L.lines), p->L.n_chunks};
}

template <int K, typename St>
__device__ __forceinline__ TwoPassArgs<K, St> args_of(const ConstAS<PersistArgs<K, St>> *p)
{
    TwoPassArgs<K, St> a;
#pragma unroll
    for (int j = 0; j < 3 * (K - 1); ++j) {
        a.xi[j] = global_ptr(p->a.xi[j]);
        a.xi_out[j] = global_ptr(p->a.xi_out[j]);
    }
#pragma unroll
    for (int i = 0; i < K - 1; ++i) {
        a.vbar[i] = global_ptr(p->a.vbar[i]);
        a.V[i] = global_ptr(p->a.V[i]);
        a.vbar_out[i] = global_ptr(p->a.vbar_out[i]);
        a.V_out[i] = global_ptr(p->a.V_out[i]);
    }
    a.u = global_ptr(p->a.u);
    return a;
}

// A waiting workgroup gives up after this many polls (of the order of a
// second; a half-step of the batched workloads takes milliseconds), flags the
// launch as failed and leaves the kernel, as does every workgroup that then
// finds the flag set, so a launch that stops making progress ends with an
// error instead of waiting for ever.
constexpr uint32_t kSpinLimit = 1u << 21;

// Zeroes the counters and stores the launch's argument table.
template <int K, typename St>
__global__ void sync_reset_kernel(GridSync *g, PersistArgs<K, St> args)
{
    if (threadIdx.x == 0) {
        g->ticket = 0;
        g->done = 0;
        g->err = 0;
        *reinterpret_cast<PersistArgs<K, St> *>(g->args) = args;
    }
}

// Adds the n tickets this workgroup completed since its last call to the
// completion count, its writes visible to every workgroup that then reads the
// count: each wave's writes are complete within the workgroup (workgroup-scope
// release), then one thread releases them at agent scope (the workgroups of
// one launch span all XCDs; one L2 write-back per workgroup and half-step).
__device__ __forceinline__ void publish_done(GridSync *g, unsigned n)
{
    if (!n) return;   // uniform over the workgroup
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&g->done, (unsigned long long)n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Waits until the completion count reaches target, then acquires what was
// published for the workgroup.  One thread polls with relaxed loads and
// acquires once, after the wait, so the polling invalidates no cache.
// Returns false (in every thread of the workgroup) when the launch has failed.
__device__ __forceinline__ bool wait_done(GridSync *g, unsigned long long target)
{
    __shared__ unsigned ok;
    if (threadIdx.x == 0) {
        uint32_t polls = 0;
        unsigned good = 1;
        while (__hip_atomic_load(&g->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (__hip_atomic_load(&g->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                good = 0;
                break;
            }
            if (++polls == kSpinLimit) {
                __hip_atomic_store(&g->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                good = 0;
                break;
            }
        }
        ok = good;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return ok != 0u;
}

// x, the same in every lane of the wave, as a scalar.
__device__ __forceinline__ unsigned long long uniform64(unsigned long long x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x)), hi = __builtin_amdgcn_readfirstlane(uint32_t(x >> 32));
    return (unsigned long long)hi << 32 | lo;
}

constexpr unsigned kTilesPerTicket = 8;
constexpr int kMaxForcedPerCu = 64;   // sct_potts_batch_set_occupancy's limit

// Waves per SIMD the persistent kernel is compiled for (launch-bounds hint).
template <int K>
constexpr int persist_waves_per_simd() { return K == 2 ? 8 : K == 3 ? 6 : 1; }

// All iterations of one K group in one launch.  The work is one sequence of
// tickets, each worth kTilesPerTicket tiles: per iteration the Xi half-step's
// tickets (a tile is the two-pass Xi kernel's workgroup of chunks), then the
// V half-step's (a tile is BW chunks, one per wave).  A workgroup draws
// tickets in order; before its first tile of a half-step it publishes the
// tickets it has completed and waits until every ticket of the half-steps
// before is complete.  Those tickets were drawn earlier, by workgroups that
// are running, so the wait always ends, whether or not the whole grid is
// resident at once.
template <int K, typename St, int BW>
__global__ __launch_bounds__(64 * BW, (persist_waves_per_simd<K>())) void potts_batch_persistent_kernel(
    const PersistArgs<K, St> *pa, float tau1, float tau2, DykstraLimits lim, int iters, GridSync *gs)
{
    constexpr uint64_t cpt = uint64_t(xi_cpw<St>()) * BW;   // chunks per Xi tile
    const uint64_t count = fresh_args(pa)->L.n_chunks;
    const unsigned long long nxi = (count + cpt - 1) / cpt, nvt = (count + BW - 1) / BW;
    const unsigned long long kxi = (nxi + kTilesPerTicket - 1) / kTilesPerTicket,
                             kvt = (nvt + kTilesPerTicket - 1) / kTilesPerTicket;
    const unsigned long long per_it = kxi + kvt, total = per_it * (unsigned long long)iters;
    __shared__ unsigned long long tk;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    unsigned long long reached = 0;   // completion count this workgroup has waited for
    unsigned mine = 0;                // tickets completed here, not yet published
    while (true) {
        if (threadIdx.x == 0) tk = atomicAdd(&gs->ticket, 1ull);
        __syncthreads();
        // the same in every lane: kept in scalar registers, so that the tile
        // and chunk indices derived from it stay scalar too
        const unsigned long long t = uniform64(tk);
        __syncthreads();
        if (t >= total) break;
        const unsigned long long it = t / per_it, r = t - it * per_it;
        const bool xi = r < kxi;
        const unsigned long long need = it * per_it + (xi ? 0 : kxi);   // tickets of the half-steps before
        if (need > reached) {
            publish_done(gs, mine);
            mine = 0;
            if (!wait_done(gs, need)) return;
            reached = need;
        }
        if (xi) {
            const auto *p = fresh_args(pa);
            const BatchLines L = lines_of(p);
            const TwoPassArgs<K, St> a = args_of(p);
            const unsigned long long t1 = min(nxi, (r + 1) * kTilesPerTicket);
#pragma unroll 1
            for (unsigned long long u = r * kTilesPerTicket; u < t1; ++u)
                xi_tile<K, St, BatchLines, BW, false, true>(L, a, tau2, lim, SweepRecord{}, uint32_t(u));
        } else {
            const auto *p = fresh_args(pa);
            const BatchLines L = lines_of(p);
            const TwoPassArgs<K, St> a = args_of(p);
            const SolveTable<K> prm{global_ptr(p->prm.p)};
            const unsigned long long k = r - kxi, t1 = min(nvt, (k + 1) * kTilesPerTicket);
#pragma unroll 1
            for (unsigned long long u = k * kTilesPerTicket; u < t1; ++u)
                v_chunk<K, St, BatchLines, SolveTable<K>, true>(L, a, prm, tau1, u * BW + w);
        }
        ++mine;
    }
    publish_done(gs, mine);
}

template <int K, typename St, int BW>
int launch_persistent(sct_ctx *ctx, Group &G, const BatchLines &L, const TwoPassArgs<K, St> &a,
                      const SolveTable<K> &prm, float tau1, float tau2, int iters, GridSync *gs)
{
    const void *fn = (const void *)potts_batch_persistent_kernel<K, St, BW>;
    int fit = 0;
    SCT_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, fn, 64 * BW, 0));
    if (fit < 1) return sct_fail(ctx, SCT_ERR_HIP, "persistent batch kernel does not fit a CU");
    // The query counts more workgroups than a CU holds at once for these
    // kernels (24 of 4 waves for K = 3, whose 80 VGPRs allow 6 waves per
    // SIMD); bound it by the kernel's VGPRs (512 per SIMD lane, in granules of
    // 8, at most 8 waves per SIMD, 4 SIMDs per CU) and its LDS.
    hipFuncAttributes fa{};
    SCT_HIP(ctx, hipFuncGetAttributes(&fa, fn));
    int lds_cu = 0;
    SCT_HIP(ctx, hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, ctx->device));
    if (fa.numRegs > 0) fit = std::min(fit, std::min(8, 512 / ((fa.numRegs + 7) / 8 * 8)) * 4 / BW);
    if (fa.sharedSizeBytes > 0 && lds_cu > 0) fit = std::min(fit, int(size_t(lds_cu) / fa.sharedSizeBytes));
    fit = std::max(fit, 1);
    // As many workgroups per CU as fit, or one fewer (the autotuner's
    // choice): no wait depends on a workgroup that has not started, so the
    // grid need not be resident all at once.
    const int forced = ctx->forced_occ[K - 2];
    const int per_cu = forced > 0 ? forced
                     : G.occ > 0 ? std::min(G.occ, fit) : std::max(1, fit + std::max(G.occ, -1));
    G.per_cu = per_cu;
    const uint32_t grid = uint32_t(per_cu) * std::max<uint32_t>(1, ctx->cu_count);
    sync_reset_kernel<K, St><<<1, 64, 0, ctx->stream>>>(gs, PersistArgs<K, St>{L, a, prm});
    const auto *pa = reinterpret_cast<const PersistArgs<K, St> *>(reinterpret_cast<const char *>(gs) +
                                                                   offsetof(GridSync, args));
    potts_batch_persistent_kernel<K, St, BW><<<grid, 64 * BW, 0, ctx->stream>>>(pa, tau1, tau2, dykstra_limits(),
                                                                                iters, gs);
    SCT_HIP(ctx, hipGetLastError());
    return SCT_OK;
}

// Phases of a group: 0 build, 1 init, 2 one iteration, 3 labels; 4 `iters`
// iterations in one persistent launch.
template <int K, typename St>
int group_phase(sct_ctx *ctx, const sct_potts_problem *pr, Group &G, int phase, float tau1, float tau2,
                int iters = 1, GridSync *gs = nullptr)
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
    } else if (phase == 4) {
        if (iters <= 0) return SCT_OK;
        if (G.bw == 1) return launch_persistent<K, St, 1>(ctx, G, L, a, prm, tau1, tau2, iters, gs);
        if (G.bw == 2) return launch_persistent<K, St, 2>(ctx, G, L, a, prm, tau1, tau2, iters, gs);
        return launch_persistent<K, St, 4>(ctx, G, L, a, prm, tau1, tau2, iters, gs);
    } else {
        potts_labels_kernel<K, St><<<wave_grid(L.count()), 256, 0, st>>>(L, a, prm, G.R, 0u, 0xffffffffu);
    }
    SCT_HIP(ctx, hipGetLastError());
    return SCT_OK;
}

template <int K>
int group_phase_k(sct_ctx *ctx, const sct_potts_problem *pr, Group &G, int storage, int phase, float t1, float t2,
                  int iters, GridSync *gs)
{
    switch (storage) {
    case SCT_STORAGE_BF16: return group_phase<K, StorageBF16>(ctx, pr, G, phase, t1, t2, iters, gs);
    case SCT_STORAGE_F16: return group_phase<K, StorageF16>(ctx, pr, G, phase, t1, t2, iters, gs);
    default: return group_phase<K, StorageF32>(ctx, pr, G, phase, t1, t2, iters, gs);
    }
}

int run_phase(sct_ctx *ctx, const sct_potts_problem *pr, Group &G, int storage, int phase, float t1, float t2,
              int iters = 1, GridSync *gs = nullptr)
{
    switch (G.K) {
    case 2: return group_phase_k<2>(ctx, pr, G, storage, phase, t1, t2, iters, gs);
    case 3: return group_phase_k<3>(ctx, pr, G, storage, phase, t1, t2, iters, gs);
    case 4: return group_phase_k<4>(ctx, pr, G, storage, phase, t1, t2, iters, gs);
    case 5: return group_phase_k<5>(ctx, pr, G, storage, phase, t1, t2, iters, gs);
    default: return group_phase_k<6>(ctx, pr, G, storage, phase, t1, t2, iters, gs);
    }
}

// The completion count of each group's persistent launches: a launch that
// gave up waiting flags its GridSync.
int check_sync(sct_ctx *ctx, const GridSync *gs, size_t n)
{
    std::vector<GridSync> h(n);
    SCT_HIP(ctx, hipMemcpyAsync(h.data(), gs, sizeof(GridSync) * n, hipMemcpyDeviceToHost, ctx->stream));
    SCT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (const GridSync &g : h)
        if (g.err) return sct_fail(ctx, SCT_ERR_HIP, "persistent batch launch: a wait for the previous half-step timed out");
    return SCT_OK;
}

// The autotuner: for each group, times kTuneIters iterations of each
// configuration, keeps the fastest, and starts the group again from its
// initial state, so the result is the untuned one.  The configurations are
// the Xi workgroup tile (BW = 4, 2, 1 waves: 1024, 512, 256 voxels and work
// lists of 512, 256, 128 entries) and, for the persistent launch, the
// workgroups per CU (as many as fit, or one fewer).  Choices are cached in the
// context per (K, storage, launch form, size class of the group).
constexpr int kTuneIters = 3;
constexpr int kTiles[3] = {4, 2, 1};
constexpr int kOcc[2] = {0, -1};

uint64_t tile_key(const Group &G, int storage, bool persistent)
{
    uint64_t cls = 0;
    for (uint64_t c = G.n_chunks; c > 1; c >>= 1) ++cls;   // log2 of the chunk count
    return (uint64_t(persistent) << 24) | (uint64_t(G.K) << 16) | (uint64_t(storage) << 8) | cls;
}

int autotune(sct_ctx *ctx, const sct_potts_problem *pr, std::vector<Group> &groups, int storage, float t1, float t2,
             GridSync *gs)
{
    const bool persistent = gs != nullptr;
    hipStream_t st = ctx->stream;
    hipEvent_t ev[2] = {};
    int rc = SCT_OK;
    for (auto &e : ev)
        if (!rc && hipEventCreate(&e) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventCreate failed");
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        Group &G = groups[gi];
        if (rc || !G.n_chunks) continue;
        const int forced = ctx->forced_tiles[G.K - 2];
        const uint64_t key = tile_key(G, storage, persistent);
        auto hit = ctx->tiles.find(key);
        if (hit != ctx->tiles.end()) {
            G.bw = forced ? forced : (hit->second & 0xff);
            G.occ = (hit->second >> 8) - 1;   // stored + 1 (occ >= -1)
            continue;
        }
        float best = 0;
        int best_bw = G.bw, best_occ = 0;
        bool first = true;
        for (int i = 0; i < 3 && !rc; ++i) {
            if (forced && kTiles[i] != forced) continue;
            for (int j = 0; j < (persistent && !ctx->forced_occ[G.K - 2] ? 2 : 1) && !rc; ++j) {
                G.bw = kTiles[i];
                G.occ = persistent ? kOcc[j] : 0;
                // one untimed iteration, then kTuneIters timed ones
                auto run = [&](int n) {
                    if (persistent) return run_phase(ctx, pr, G, storage, 4, t1, t2, n, gs + gi);
                    int r = SCT_OK;
                    for (int it = 0; it < n && !r; ++it) r = run_phase(ctx, pr, G, storage, 2, t1, t2);
                    return r;
                };
                rc = run(1);
                if (!rc && hipEventRecord(ev[0], st) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventRecord");
                if (!rc) rc = run(kTuneIters);
                if (!rc && hipEventRecord(ev[1], st) != hipSuccess) rc = sct_fail(ctx, SCT_ERR_HIP, "hipEventRecord");
                float ms = 0;
                if (!rc && (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess))
                    rc = sct_fail(ctx, SCT_ERR_HIP, "autotune timing failed");
                if (!rc && persistent) rc = check_sync(ctx, gs + gi, 1);
                if (!rc && (first || ms < best)) {
                    best = ms;
                    best_bw = G.bw;
                    best_occ = G.occ;
                    first = false;
                }
            }
        }
        G.bw = best_bw;
        G.occ = best_occ;
        if (!rc && !forced) ctx->tiles[key] = best_bw | ((best_occ + 1) << 8);
        if (!rc) rc = run_phase(ctx, pr, G, storage, 1, t1, t2);   // back to the initial state
    }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

// Runs the iterations of all groups, directly or as one captured graph, and
// times them with events on the context's stream.
int run_iterations(sct_ctx *ctx, const sct_potts_problem *pr, std::vector<Group> &groups, int storage, int iters,
                   float t1, float t2, bool graph, GridSync *gs, float *elapsed_ms)
{
    hipStream_t st = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    int rc = SCT_OK;
    auto launch_all = [&]() {
        if (gs) {
            for (size_t gi = 0; gi < groups.size() && !rc; ++gi)
                rc = run_phase(ctx, pr, groups[gi], storage, 4, t1, t2, iters, gs + gi);
            return;
        }
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
    if (!rc && gs) rc = check_sync(ctx, gs, groups.size());
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

extern "C" int sct_potts_batch_set_occupancy(sct_ctx *ctx, const int *per_cu)
{
    if (!ctx || !per_cu) return SCT_ERR_INVALID;
    for (int k = 0; k < 5; ++k)
        if (per_cu[k] < 0 || per_cu[k] > kMaxForcedPerCu)
            return sct_fail(ctx, SCT_ERR_INVALID, "workgroups per CU are 1 to 64 (0: not set)");
    for (int k = 0; k < 5; ++k) ctx->forced_occ[k] = per_cu[k];
    return SCT_OK;
}

extern "C" int sct_potts_batch_tiles(sct_ctx *ctx, int *waves)
{
    if (!ctx || !waves) return SCT_ERR_INVALID;
    for (int k = 0; k < 5; ++k) waves[k] = ctx->last_tiles[k];
    return SCT_OK;
}

extern "C" int sct_potts_batch_occupancy(sct_ctx *ctx, int *per_cu)
{
    if (!ctx || !per_cu) return SCT_ERR_INVALID;
    for (int k = 0; k < 5; ++k) per_cu[k] = ctx->last_occ[k];
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
    const bool persistent = (flags & SCT_BATCH_PERSISTENT) != 0;
    GridSync *gs = nullptr;
    if (!rc && persistent) rc = sct_alloc(ctx, (void **)&gs, sizeof(GridSync) * std::max<size_t>(1, groups.size()));
    if (!rc && (flags & SCT_BATCH_AUTOTUNE) && iters > 0)
        rc = autotune(ctx, problems, groups, storage, tau1, tau2, gs);
    if (!rc) rc = run_iterations(ctx, problems, groups, storage, iters, tau1, tau2, (flags & SCT_BATCH_GRAPH) != 0,
                                 gs, elapsed_ms);
    for (int K = 2; K <= 6; ++K) ctx->last_tiles[K - 2] = ctx->last_occ[K - 2] = 0;
    for (const Group &G : groups) {
        ctx->last_tiles[G.K - 2] = G.n_chunks ? G.bw : 0;
        if (persistent && G.n_chunks) ctx->last_occ[G.K - 2] = G.per_cu;
    }
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
    sct_free(ctx, gs);
    if (rc)
        for (int p = 0; p < n; ++p) {
            sct_field_destroy(out[p]);
            out[p] = nullptr;
        }
    return rc;
}
