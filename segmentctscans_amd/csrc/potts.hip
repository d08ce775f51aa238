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
#include "twopass.hpp"

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>

using namespace sct;
using namespace sct::tp;

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

}  // namespace

namespace {

// The lines of the solver's slab, for the two-pass kernels with storage St.
template <typename St>
VolumeLines volume_lines(const sct_potts_solver *s)
{
    const Geometry &geo = s->g->geo;
    constexpr uint32_t cw = kWave * vpl_of<St>();
    const uint32_t len = s->s1 - s->s0;
    return VolumeLines::make(geo.d_nbr, geo.lines, len, (len + cw - 1) / cw, s->wstride, geo.nstride, s->s0);
}

// Base pointers of the state of Xi / V-bar buffer set b.
template <int K, typename St>
TwoPassArgs<K, St> twopass_args(const sct_potts_solver *s, int b)
{
    using T = typename St::T;
    TwoPassArgs<K, St> a;
    for (int j = 0; j < 3 * (K - 1); ++j) a.xi[j] = (T *)s->xi[b] + j * s->wspan;
    for (int i = 0; i < K - 1; ++i) {
        a.vbar[i] = (T *)s->vbar[b] + i * s->wspan;
        a.V[i] = s->V + i * s->wspan;
    }
    a.u = (const float *)s->g->data;
    a.in_place();
    return a;
}

template <int K>
OneSolve<K> one_solve(const sct_potts_solver *s)
{
    OneSolve<K> o;
    for (int i = 0; i < K; ++i) o.p.c[i] = s->c[i];
    o.p.lambda = s->lambda;
    return o;
}

bool slab_empty(const sct_potts_solver *s) { return !s->g->geo.lines || s->s1 <= s->s0; }

template <int K, typename St>
int init_st(sct_potts_solver *s)
{
    const VolumeLines L = volume_lines<St>(s);
    potts_init_kernel<K, St><<<wave_grid(L.count()), 256, 0, s->ctx->stream>>>(L, twopass_args<K, St>(s, 0), one_solve<K>(s));
    SCT_HIP(s->ctx, hipGetLastError());
    return SCT_OK;
}

template <int K>
int launch_init(sct_potts_solver *s)
{
    s->cur = 0;
    if (s->Vb[0]) s->V = s->Vb[0];
    if (!slab_empty(s)) {
        int rc;
        switch (s->storage) {
        case SCT_STORAGE_BF16: rc = init_st<K, StorageBF16>(s); break;
        case SCT_STORAGE_F16: rc = init_st<K, StorageF16>(s); break;
        default: rc = init_st<K, StorageF32>(s); break;
        }
        if (rc) return rc;
    }
    s->cur = 0;
    SCT_HIP(s->ctx, hipMemsetAsync(s->xi[0], 0, storage_bytes(s->storage) * 3 * (K - 1) * s->wspan, s->ctx->stream));
    return SCT_OK;
}

template <int K, typename St>
int twopass_iterate(sct_potts_solver *s, int iters, float *dq_dev)
{
    const DykstraLimits lim = dykstra_limits();
    const TwoPassArgs<K, St> a = twopass_args<K, St>(s, 0);
    const VolumeLines L = volume_lines<St>(s);
    const OneSolve<K> prm = one_solve<K>(s);
    const uint32_t xg = xi_grid<St>(L.count()), vg = wave_grid(L.count());
    for (int i = 0; i < iters; ++i) {
        potts_mark(s);
        potts_xi_kernel<K, St><<<xg, kXiThreads, 0, s->ctx->stream>>>(L, a, s->tau2, lim);
        potts_mark(s);
        if (dq_dev)
            potts_v_kernel<K, St, true><<<vg, 256, 0, s->ctx->stream>>>(L, a, prm, s->tau1, dq_dev + i);
        else
            potts_v_kernel<K, St, false><<<vg, 256, 0, s->ctx->stream>>>(L, a, prm, s->tau1, nullptr);
    }
    SCT_HIP(s->ctx, hipGetLastError());
    return SCT_OK;
}

template <int K>
int launch_iterate(sct_potts_solver *s, int iters, float *dq_dev)
{
    if (slab_empty(s)) return SCT_OK;
    if (s->kernel == SCT_KERNEL_FUSED) return fused_iterate(s, iters, dq_dev);
    switch (s->storage) {
    case SCT_STORAGE_BF16: return twopass_iterate<K, StorageBF16>(s, iters, dq_dev);
    case SCT_STORAGE_F16: return twopass_iterate<K, StorageF16>(s, iters, dq_dev);
    default: return twopass_iterate<K, StorageF32>(s, iters, dq_dev);
    }
}

template <int K>
int launch_labels(sct_potts_solver *s, float *res, uint32_t r0, uint32_t r1)
{
    if (slab_empty(s) || r0 >= r1) return SCT_OK;
    // V is FP32 in every storage mode; the FP32 chunking reads it.
    const VolumeLines L = volume_lines<StorageF32>(s);
    potts_labels_kernel<K, StorageF32><<<wave_grid(L.count()), 256, 0, s->ctx->stream>>>(
        L, twopass_args<K, StorageF32>(s, 0), one_solve<K>(s), res, r0, r1);
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
// With stop, the iterations end at the first evaluation whose relative gap
// (E - D) / E is below opts.gap_tol (the stopping test; count then holds the
// iterations performed).
int iterate_with_gap(sct_potts_solver *s, int iters, float *dq, int &count, int capacity, bool stop = false)
{
    const int every = s->opts.gap_every;
    stop = stop && s->opts.gap_tol > 0.0;
    if (every <= 0 || (!s->opts.gap_trace && !stop)) {
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
        const bool record = s->opts.gap_trace && count / every <= capacity;
        if (count % every == 0 && (record || stop)) {
            double E = 0.0, D = 0.0;
            if ((rc = potts_gap(s, &E, &D))) return rc;
            const double rel = E != 0.0 ? (E - D) / E : 0.0;
            if (record) s->opts.gap_trace[count / every - 1] = rel;
            if (stop && rel < s->opts.gap_tol) break;
        }
    }
    return SCT_OK;
}
int do_labels(sct_potts_solver *s, float *res, uint32_t r0, uint32_t r1) { SCT_DISPATCH_K(s->K, launch_labels, s, res, r0, r1); }

// Bytes of solver state per voxel of a slab: Xi 3(K-1) and V-bar (K-1) in the
// storage type, V (K-1) in FP32; the fused kernel double-buffers Xi and V.
uint64_t bytes_per_voxel(int K, int kernel, int storage)
{
    const uint64_t copies = kernel == SCT_KERNEL_FUSED ? 2 : 1;
    return ((copies * 3 + 1) * storage_bytes(storage) + copies * 4) * uint64_t(K - 1);
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
        sct_free(s->ctx, s->Vb[b]);
        s->xi[b] = s->vbar[b] = nullptr;
        s->Vb[b] = nullptr;
    }
    s->V = nullptr;
    fused_release(s);
}

int alloc_buffers(sct_potts_solver *s, uint32_t slab_len)
{
    const Geometry &geo = s->g->geo;
    sct_pool_trim(s->ctx);   // the plan counted the pool's idle memory as available
    s->wstride = sct_get_stride(slab_len);
    s->wspan = uint64_t(geo.lines) * s->wstride;
    const size_t comp = std::max<uint64_t>(1, s->wspan);
    const size_t sb = storage_bytes(s->storage);
    const int nb = s->kernel == SCT_KERNEL_FUSED ? 2 : 1;
    int rc;
    for (int b = 0; b < nb; ++b) {
        if ((rc = sct_alloc(s->ctx, &s->xi[b], sb * comp * 3 * (s->K - 1)))) return rc;
        if ((rc = sct_alloc(s->ctx, (void **)&s->Vb[b], sizeof(float) * comp * (s->K - 1)))) return rc;
    }
    if ((rc = sct_alloc(s->ctx, &s->vbar[0], sb * comp * (s->K - 1)))) return rc;
    s->V = s->Vb[0];
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
    if (opts && opts->size >= offsetof(sct_potts_opts, iters_done) && !(opts->gap_tol >= 0.0))
        return sct_fail(ctx, SCT_ERR_INVALID, "gap_tol must be >= 0");
    if (opts && opts->size >= offsetof(sct_potts_opts, iters_done) && opts->gap_tol > 0.0 && opts->gap_every <= 0)
        return sct_fail(ctx, SCT_ERR_INVALID, "gap_tol needs gap_every > 0");
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
    if ((rc = sct_field_like(ctx, g, SCT_F32, &res))) return rc;
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
    int done_max = 0;
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
                                                                                 : iters / std::max(1, o.gap_every),
                                       true);
        done_max = std::max(done_max, count);
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
    if (o.iters_done) *o.iters_done = done_max;
    *out = res;
    return SCT_OK;
}
