// context.hip -- HIP context, device fields and layout helpers.
//
// Replaces the reference's CUDA context (include/cuda-context.h:7-37,
// include/cuda-details/cuda-context-impl.h:35-66, cuda-context/cuda-context.cpp):
// one device, one stream per context, fields that own their voxels plus device
// copies of the mapping tables.  No CUDA compatibility layer.
#include "sct_internal.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <set>

// ---- errors ------------------------------------------------------------------
int sct_fail(sct_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

int sct_hip_fail(sct_ctx *ctx, hipError_t e, const char *what)
{
    return sct_fail(ctx, SCT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int sct_alloc(sct_ctx *ctx, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return sct_fail(ctx, SCT_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + " bytes): " +
                                              hipGetErrorString(e));
    }
    return SCT_OK;
}

void sct_free(sct_ctx *, void *p)
{
    if (p) (void)hipFree(p);
}

// ---- field voxels ----------------------------------------------------------------
// A field's voxels come from the device's stream-ordered memory pool, in the
// order of its context's stream: a call that makes and drops an output field
// (grad_magn, convert_to_fp) then costs no hipMalloc / hipFree, and the pool
// keeps up to 8 GiB of freed memory for the next field (counted as available
// by sct_device_memory).  A field outliving its context is freed after the
// device has drained.  The live contexts are tracked for that.
namespace {
std::mutex g_live_mu;
std::set<const sct_ctx *> g_live;

int field_alloc(sct_ctx *ctx, sct_field *f, size_t bytes)
{
    f->device = ctx->device;
    if (bytes == 0) bytes = 256;
    if (hipMallocAsync(&f->data, bytes, ctx->stream) == hipSuccess) {
        f->pooled = true;
        return SCT_OK;
    }
    (void)hipGetLastError();
    f->data = nullptr;
    f->pooled = false;
    return sct_alloc(ctx, &f->data, bytes);
}

void field_free(sct_field *f)
{
    if (!f->data) return;
    if (!f->pooled) {
        (void)hipFree(f->data);   // waits for the work that uses it
        return;
    }
    std::lock_guard<std::mutex> lk(g_live_mu);
    if (g_live.count(f->ctx)) {
        (void)hipFreeAsync(f->data, f->ctx->stream);
    } else {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(f->device);
        (void)hipDeviceSynchronize();
        (void)hipFreeAsync(f->data, nullptr);
        (void)hipSetDevice(cur);
    }
    f->data = nullptr;
}
}  // namespace

// Returns idle pool memory to the device before a large plain allocation.
void sct_pool_trim(sct_ctx *ctx)
{
    if (!ctx || !ctx->pool) return;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipMemPoolTrimTo(ctx->pool, 0);
}

// Sizes round up to 1 MiB (device) and 64 KiB (host), so that calls of
// slightly different sizes do not reallocate.
int sct_ctx_scratch(sct_ctx *ctx, size_t bytes, void **p)
{
    *p = nullptr;
    if (ctx->scratch_bytes < bytes) {
        sct_free(ctx, ctx->scratch);   // hipFree waits for the work that uses it
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
        const size_t want = (bytes + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
        const int rc = sct_alloc(ctx, &ctx->scratch, want);
        if (rc) return rc;
        ctx->scratch_bytes = want;
    }
    *p = ctx->scratch;
    return SCT_OK;
}

int sct_ctx_host_scratch(sct_ctx *ctx, size_t bytes, void **p)
{
    *p = nullptr;
    if (ctx->host_scratch_bytes < bytes) {
        if (ctx->host_scratch) (void)hipHostFree(ctx->host_scratch);
        ctx->host_scratch = nullptr;
        ctx->host_scratch_bytes = 0;
        const size_t want = (bytes + (size_t(1) << 16) - 1) & ~((size_t(1) << 16) - 1);
        const hipError_t e = hipHostMalloc(&ctx->host_scratch, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            ctx->host_scratch = nullptr;
            (void)hipGetLastError();
            return sct_hip_fail(ctx, e, "hipHostMalloc");
        }
        ctx->host_scratch_bytes = want;
    }
    *p = ctx->host_scratch;
    return SCT_OK;
}

// ---- layout helpers ---------------------------------------------------------------
// layout_cylinder::mapping_storage(extents, xc, yc, r, nstride), layout_cylinder.h:152-195:
// a column is live when r^2 - (col - xc)^2 > 0; its rows are the integers strictly
// inside (yc - d, yc + d), clipped to [0, height - 1].
extern "C" int sct_cylinder_layout(uint32_t n_images, uint32_t height, uint32_t width, double xc,
                                   double yc, double r, uint32_t nstride, uint32_t *start_row,
                                   uint32_t *base, uint32_t *col_begin, uint32_t *col_end,
                                   uint64_t *data_size)
{
    if (!start_row || !base || !col_begin || !col_end || !data_size || nstride == 0 || height == 0)
        return SCT_ERR_INVALID;
    const double r2 = r * r;
    uint64_t off = 0;
    uint32_t cb = ~0u, ce = 0;
    for (uint32_t col = 0; col < width; ++col) {
        const double dx = double(col) - xc;
        const double d2 = r2 - dx * dx;
        base[col] = uint32_t(off);
        double lo = 0.0, hi = 0.0;
        if (d2 > 0) {
            const double d = std::sqrt(d2);
            double c = std::ceil(yc - d);
            if (yc - d == c) c += 1.0;          // strictly inside
            double f = std::floor(yc + d);
            if (yc + d == f) f -= 1.0;
            lo = std::max(0.0, c);
            hi = std::min(double(height - 1), f);
        }
        if (d2 > 0 && lo <= hi) {
            const uint32_t r0 = uint32_t(lo), r1 = uint32_t(hi);
            if (cb == ~0u) {
                cb = col;
                ce = col + 1;
            } else {
                ++ce;
            }
            start_row[col] = r0;
            off += uint64_t(r1 - r0 + 1) * nstride;
            if (off > 0xffffffffull) return SCT_ERR_INVALID;   // 32-bit index type
        } else {
            start_row[col] = ~0u;
        }
    }
    base[width] = uint32_t(off);
    uint64_t ds = 0;
    for (uint32_t col = cb; col < ce; ++col) ds += (base[col + 1] - base[col]) / nstride;
    *col_begin = cb;
    *col_end = ce;
    *data_size = ds * n_images;
    return SCT_OK;
}

extern "C" uint32_t sct_get_stride(uint32_t n) { return ((n + 31u) / 32u) * 32u; }
extern "C" uint32_t sct_prev_stride(uint32_t n) { return (n / 32u) * 32u; }

// ---- geometry -------------------------------------------------------------------
int sct_geometry_build(sct_ctx *ctx, const sct_layout *l, Geometry &g)
{
    if (!l || !l->start_row || !l->base || l->nstride == 0 || l->nstride < l->n_images ||
        l->width == 0 || l->col_end > l->width || (l->col_begin > l->col_end && l->col_end != 0))
        return sct_fail(ctx, SCT_ERR_INVALID, "invalid layout");
    g.n_images = l->n_images;
    g.height = l->height;
    g.width = l->width;
    g.nstride = l->nstride;
    g.col_begin = l->col_begin;
    g.col_end = l->col_end;
    g.start_row.assign(l->start_row, l->start_row + l->width);
    g.base.assign(l->base, l->base + l->width + 1);
    if (g.col_begin == ~0u) g.col_begin = g.col_end = 0;   // empty mask
    for (uint32_t c = 0; c < g.width; ++c) {
        if (g.base[c + 1] < g.base[c] || g.base[c] % g.nstride != 0)
            return sct_fail(ctx, SCT_ERR_INVALID, "layout base[] must be nondecreasing multiples of nstride");
        const bool live = c >= g.col_begin && c < g.col_end;
        if (!live && g.base[c + 1] != g.base[c])
            return sct_fail(ctx, SCT_ERR_INVALID, "layout has rows outside [col_begin, col_end)");
        if (live && g.start_row[c] + (g.base[c + 1] - g.base[c]) / g.nstride > g.height)
            return sct_fail(ctx, SCT_ERR_INVALID, "layout rows exceed height");
    }
    if (g.base[g.width] % g.nstride != 0) return sct_fail(ctx, SCT_ERR_INVALID, "span not a multiple of nstride");
    g.lines = g.base[g.width] / g.nstride;
    auto t = std::make_shared<GeoTables>();
    t->h_nbr.assign(g.lines, LineNbr{-1, -1, -1, -1});
    t->line_col.assign(g.lines, 0);
    t->line_row.assign(g.lines, 0);
    uint64_t rows = 0;
    auto line_of = [&](uint32_t row, uint32_t col) { return int32_t(g.base[col] / g.nstride + (row - g.start_row[col])); };
    auto in_col = [&](uint32_t row, uint32_t col) {
        return g.base[col + 1] > g.base[col] && g.start_row[col] <= row && row < g.row_end(col);
    };
    for (uint32_t col = g.col_begin; col < g.col_end; ++col) {
        const uint32_t rb = g.start_row[col], re = g.row_end(col);
        for (uint32_t row = rb; row < re; ++row) {
            const int32_t L = line_of(row, col);
            LineNbr nb{-1, -1, -1, -1};
            if (row > rb) nb.prow = L - 1;
            if (row + 1 < re) nb.nrow = L + 1;
            if (col > g.col_begin && in_col(row, col - 1)) nb.pcol = line_of(row, col - 1);
            if (col + 1 < g.col_end && in_col(row, col + 1)) nb.ncol = line_of(row, col + 1);
            t->h_nbr[size_t(L)] = nb;
            t->line_col[size_t(L)] = col;
            t->line_row[size_t(L)] = row;
            ++rows;
        }
    }
    g.data_size = rows * g.n_images;
    std::vector<uint32_t> order;
    order.reserve(g.lines);
    const uint32_t bands = kBandRows ? (g.height + kBandRows - 1) / kBandRows : 1;
    for (uint32_t band = 0; band < bands; ++band) {
        const uint32_t r0 = kBandRows ? band * kBandRows : 0, r1 = kBandRows ? r0 + kBandRows : g.height;
        for (uint32_t col = g.col_begin; col < g.col_end; ++col) {
            const uint32_t rb = std::max(r0, g.start_row[col]), re = std::min(r1, g.row_end(col));
            for (uint32_t row = rb; row < re; ++row) order.push_back(uint32_t(line_of(row, col)));
        }
    }
    if (order.size() != g.lines) return sct_fail(ctx, SCT_ERR_INVALID, "layout lines do not tile the mask");
    std::vector<OrderedLine> oline(g.lines);
    for (uint32_t k = 0; k < g.lines; ++k) {
        const uint32_t L = order[k];
        const LineNbr &nb = t->h_nbr[L];
        OrderedLine &o = oline[k];
        o = OrderedLine{int32_t(L), nb.prow, nb.nrow, nb.pcol, nb.ncol, -1, -1, 0u};
        if (nb.nrow >= 0) {
            const LineNbr &r = t->h_nbr[size_t(nb.nrow)];
            o.nrow_pcol = r.pcol;
            o.flags |= (r.nrow >= 0 ? kNrowHasNrow : 0u) | (r.ncol >= 0 ? kNrowHasNcol : 0u);
        }
        if (nb.ncol >= 0) {
            const LineNbr &c = t->h_nbr[size_t(nb.ncol)];
            o.ncol_prow = c.prow;
            o.flags |= (c.nrow >= 0 ? kNcolHasNrow : 0u) | (c.ncol >= 0 ? kNcolHasNcol : 0u);
        }
    }
    std::vector<uint2> runs;
    for (uint32_t k = 0; k < g.lines; ++k) {
        const uint32_t L = order[k];
        if (k && L == order[k - 1] + 1 && t->line_col[L] == t->line_col[L - 1] && runs.back().y < kMarchRows)
            ++runs.back().y;
        else
            runs.push_back(make_uint2(k, 1u));
    }
    g.n_runs = uint32_t(runs.size());
    int rc;
    // t frees what it holds if anything below fails
    if ((rc = sct_alloc(ctx, (void **)&t->d_runs, sizeof(uint2) * std::max<size_t>(1, runs.size())))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&t->d_nbr, sizeof(LineNbr) * std::max<size_t>(1, g.lines)))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&t->d_oline, sizeof(OrderedLine) * std::max<size_t>(1, g.lines)))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&t->d_start_row, sizeof(uint32_t) * g.width))) return rc;
    if ((rc = sct_alloc(ctx, (void **)&t->d_base, sizeof(uint32_t) * (g.width + 1)))) return rc;
    g.d_runs = t->d_runs;
    g.d_nbr = t->d_nbr;
    g.d_oline = t->d_oline;
    g.d_start_row = t->d_start_row;
    g.d_base = t->d_base;
    if (g.lines) {
        SCT_HIP(ctx, hipMemcpyAsync(g.d_nbr, t->h_nbr.data(), sizeof(LineNbr) * g.lines, hipMemcpyHostToDevice, ctx->stream));
        SCT_HIP(ctx, hipMemcpyAsync(g.d_oline, oline.data(), sizeof(OrderedLine) * g.lines, hipMemcpyHostToDevice,
                                    ctx->stream));
        SCT_HIP(ctx, hipMemcpyAsync(g.d_runs, runs.data(), sizeof(uint2) * runs.size(), hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    SCT_HIP(ctx, hipMemcpyAsync(g.d_start_row, g.start_row.data(), sizeof(uint32_t) * g.width, hipMemcpyHostToDevice, ctx->stream));
    SCT_HIP(ctx, hipMemcpyAsync(g.d_base, g.base.data(), sizeof(uint32_t) * (g.width + 1), hipMemcpyHostToDevice, ctx->stream));
    SCT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    g.tables = std::move(t);
    return SCT_OK;
}

GeoTables::~GeoTables()
{
    // hipFree waits for the work that uses the tables; no context needed
    for (void *p : {(void *)d_nbr, (void *)d_oline, (void *)d_runs, (void *)d_start_row, (void *)d_base})
        sct_free(nullptr, p);
}

void sct_geometry_release(sct_ctx *, Geometry &g)
{
    g.tables.reset();
    g.d_nbr = nullptr;
    g.d_oline = nullptr;
    g.d_runs = nullptr;
    g.d_start_row = g.d_base = nullptr;
}

// A copy of a built geometry: the host mapping tables by value, the line
// tables shared (they never change).
int sct_geometry_copy(sct_ctx *, const Geometry &src, Geometry &dst)
{
    dst = src;
    return SCT_OK;
}

// ---- context ---------------------------------------------------------------------
extern "C" int sct_ctx_create(int device, sct_ctx **out)
{
    if (!out) return SCT_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SCT_ERR_HIP;
    if (device < 0 || device >= n) return SCT_ERR_INVALID;
    auto *c = new sct_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SCT_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = uint32_t(prop.multiProcessorCount);
    // The default pool keeps up to 8 GiB of freed stream-ordered memory
    // (per-call temporaries) instead of releasing it at every synchronisation.
    if (hipDeviceGetDefaultMemPool(&c->pool, device) == hipSuccess && c->pool) {
        uint64_t keep = uint64_t(8) << 30;
        (void)hipMemPoolSetAttribute(c->pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    if (sct_noalloc_scratch_setup(c) != SCT_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return SCT_ERR_HIP;
    }
    {
        std::lock_guard<std::mutex> lk(g_live_mu);
        g_live.insert(c);
    }
    *out = c;
    return SCT_OK;
}

extern "C" void sct_ctx_destroy(sct_ctx *ctx)
{
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> lk(g_live_mu);
        g_live.erase(ctx);
    }
    (void)hipStreamSynchronize(ctx->stream);
    sct_free(ctx, ctx->scratch);
    if (ctx->host_scratch) (void)hipHostFree(ctx->host_scratch);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *sct_ctx_last_error(const sct_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }
extern "C" void *sct_ctx_stream(sct_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

extern "C" int sct_ctx_synchronize(sct_ctx *ctx)
{
    if (!ctx) return SCT_ERR_INVALID;
    SCT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SCT_OK;
}

// device_memory(), cuda-context.cpp:22-41: free memory plus what the pool
// holds but does not use.
extern "C" int sct_device_memory(sct_ctx *ctx, size_t *available, size_t *total)
{
    if (!ctx || !available || !total) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    SCT_HIP(ctx, hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    SCT_HIP(ctx, hipMemGetInfo(&f, &t));
    uint64_t reserved = 0, used = 0;
    if (ctx->pool) {
        (void)hipMemPoolGetAttribute(ctx->pool, hipMemPoolAttrReservedMemCurrent, &reserved);
        (void)hipMemPoolGetAttribute(ctx->pool, hipMemPoolAttrUsedMemCurrent, &used);
    }
    *available = f + size_t(reserved >= used ? reserved - used : 0);
    *total = t;
    return SCT_OK;
}

// ---- fields ----------------------------------------------------------------------
static size_t elem_size(int dtype) { return dtype == SCT_U16 ? 2 : 4; }

extern "C" int sct_field_create(sct_ctx *ctx, const sct_layout *layout, int dtype, sct_field **out)
{
    if (!ctx || !out) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (dtype != SCT_F32 && dtype != SCT_U16) return sct_fail(ctx, SCT_ERR_INVALID, "bad dtype");
    auto *f = new sct_field;
    f->ctx = ctx;
    f->dtype = dtype;
    int rc = sct_geometry_build(ctx, layout, f->geo);
    if (!rc) rc = field_alloc(ctx, f, f->geo.span() * elem_size(dtype));
    if (rc) {
        sct_geometry_release(ctx, f->geo);
        delete f;
        return rc;
    }
    *out = f;
    return SCT_OK;
}

// A new field with the layout of src (its geometry copied on the device).
int sct_field_like(sct_ctx *ctx, const sct_field *src, int dtype, sct_field **out)
{
    if (!ctx || !src || !out) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (dtype != SCT_F32 && dtype != SCT_U16) return sct_fail(ctx, SCT_ERR_INVALID, "bad dtype");
    auto *f = new sct_field;
    f->ctx = ctx;
    f->dtype = dtype;
    int rc = sct_geometry_copy(ctx, src->geo, f->geo);
    if (!rc) rc = field_alloc(ctx, f, std::max<size_t>(1, f->geo.span()) * elem_size(dtype));
    if (rc) {
        sct_geometry_release(ctx, f->geo);
        delete f;
        return rc;
    }
    *out = f;
    return SCT_OK;
}

extern "C" int sct_upload(sct_ctx *ctx, const sct_layout *layout, int dtype, const void *host, sct_field **out)
{
    if (!host) return sct_fail(ctx, SCT_ERR_INVALID, "null host pointer");
    int rc = sct_field_create(ctx, layout, dtype, out);
    if (rc) return rc;
    sct_field *f = *out;
    hipError_t e = hipMemcpyAsync(f->data, host, f->geo.span() * elem_size(dtype), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        sct_field_destroy(f);
        *out = nullptr;
        return sct_hip_fail(ctx, e, "upload");
    }
    return SCT_OK;
}

extern "C" int sct_download(sct_ctx *ctx, const sct_field *f, void *host, size_t span_elems)
{
    if (!ctx || !f || !host) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (span_elems != f->geo.span())
        return sct_fail(ctx, SCT_ERR_SIZE, "Host and device images have different sizes.");
    SCT_HIP(ctx, hipMemcpyAsync(host, f->data, f->geo.span() * elem_size(f->dtype), hipMemcpyDeviceToHost, ctx->stream));
    SCT_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SCT_OK;
}

extern "C" void sct_field_destroy(sct_field *f)
{
    if (!f) return;
    field_free(f);
    sct_geometry_release(nullptr, f->geo);
    delete f;
}

extern "C" void *sct_field_data(sct_field *f) { return f ? f->data : nullptr; }
extern "C" size_t sct_field_span(const sct_field *f) { return f ? f->geo.span() : 0; }
extern "C" uint64_t sct_field_data_size(const sct_field *f) { return f ? f->geo.data_size : 0; }
