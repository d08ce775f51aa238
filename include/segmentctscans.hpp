// segmentctscans.hpp -- C++ host API with the reference's signatures, over the
// C ABI of sct.h (header-only; link with segmentctscans_amd/_lib/libsct.so).
//
// Reference declarations mirrored (paths relative to the reference repository):
//   include/cuda-context.h:7-37            make_context, upload, download, device_memory,
//                                          DeviceField<T>, device_field_ptr<T>, max_phases
//   include/field.h:23-153                 Field<T> (host volume, lo/hi)
//   include/layout_cylinder.h:29-195       the cylinder mapping
//   potts-model/potts-min-partition.h:8-16 potts_min_partition<K>
//   totalvarmin/chambolle.h:5-6            chambolle
//   include/calc_hist.h:8-12               calc_hist (nbins widened to uint32_t)
//   include/min_max.h:7                    min_max
//   include/grad_magn.h:5-6                grad_magn
//   uv-hist/uv-hist.h:11-44                uv_hist, UVHistReturn
//   transpose/convert_to_fp.h:8-10         convert_to_fp<uint16_t | float>
//   transpose/find_intensity_range.h:8-9   find_intensity_range
//   include/field.h:92-153                 Field<T>::save / load as save_bin / load_bin
// and, with no counterpart in the reference, PottsFunction<K> (forward /
// backward through the unrolled iteration).
//
// Differences a caller sees: the context type is HipContext; Field<T>::view()
// returns a CylinderView (element access view(n, row, col)) in place of a
// std::mdspan with layout_cylinder, because the toolchain's libstdc++ has no
// <mdspan>.  Errors throw std::runtime_error (reference: cuda_error /
// runtime_error); potts_min_partition returns an empty pointer when a slab
// cannot hold the halo, as the reference does.
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "sct.h"

static constexpr auto max_phases = 6;

// ---- layout ------------------------------------------------------------------
class CylinderMapping {
public:
    CylinderMapping() = default;
    // mapping_storage(extents{n, height, width}, xc, yc, r, nstride), layout_cylinder.h:152-195
    CylinderMapping(uint32_t n_images, uint32_t height, uint32_t width, double xc, double yc, double r,
                    uint32_t nstride)
        : n_(n_images), h_(height), w_(width), nstride_(nstride), start_row_(width), base_(width + 1)
    {
        uint64_t ds = 0;
        if (sct_cylinder_layout(n_images, height, width, xc, yc, r, nstride, start_row_.data(), base_.data(),
                                &col_begin_, &col_end_, &ds) != SCT_OK)
            throw std::invalid_argument("invalid cylinder layout");
        if (col_begin_ == ~0u) col_begin_ = col_end_ = 0;
        data_size_ = ds;
    }
    // A mapping from its tables (a deserialised layout_cylinder::mapping_storage,
    // layout_cylinder.h:233-318): start_row[width], base[width + 1].
    CylinderMapping(uint32_t n_images, uint32_t height, uint32_t width, uint32_t nstride, uint32_t col_begin,
                    uint32_t col_end, std::vector<uint32_t> start_row, std::vector<uint32_t> base, uint64_t data_size)
        : n_(n_images), h_(height), w_(width), nstride_(nstride), col_begin_(col_begin), col_end_(col_end),
          data_size_(data_size), start_row_(std::move(start_row)), base_(std::move(base))
    {
        if (start_row_.size() != w_ || base_.size() != size_t(w_) + 1)
            throw std::invalid_argument("mapping tables do not match the width");
    }

    uint32_t operator()(uint32_t n, uint32_t row, uint32_t col) const
    {
        return base_[col] + (row - start_row_[col]) * nstride_ + n;
    }
    uint32_t n_images() const { return n_; }
    uint32_t height() const { return h_; }
    uint32_t width() const { return w_; }
    uint32_t nstride() const { return nstride_; }
    uint32_t col_begin() const { return col_begin_; }
    uint32_t col_end() const { return col_end_; }
    uint32_t row_begin(uint32_t col) const { return start_row_[col]; }
    uint32_t row_end(uint32_t col) const { return start_row_[col] + (base_[col + 1] - base_[col]) / nstride_; }
    uint32_t required_span_size() const { return base_.empty() ? 0 : base_.back(); }
    uint64_t data_size() const { return data_size_; }

    sct_layout c_layout() const
    {
        return sct_layout{n_, h_, w_, nstride_, col_begin_, col_end_, start_row_.data(), base_.data()};
    }

private:
    uint32_t n_ = 0, h_ = 0, w_ = 0, nstride_ = 0, col_begin_ = 0, col_end_ = 0;
    uint64_t data_size_ = 0;
    std::vector<uint32_t> start_row_, base_;
};

template <typename T>
struct CylinderView {
    T *data;
    const CylinderMapping *m;
    T &operator()(uint32_t n, uint32_t row, uint32_t col) const { return data[(*m)(n, row, col)]; }
    const CylinderMapping &mapping() const { return *m; }
    T *data_handle() const { return data; }
};

// ---- host field (field.h:23-153) -----------------------------------------------------
template <typename T>
class Field {
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, uint16_t>, "float or uint16_t voxels");

public:
    using mapping_type = CylinderMapping;
    static constexpr uint32_t get_stride(uint32_t n) { return ((n + 31u) / 32u) * 32u; }
    static constexpr uint32_t prev_stride(uint32_t n) { return (n / 32u) * 32u; }

    Field() = default;
    // Field(num_images, width, height, xc, yc, r), field.h:47-58
    Field(uint32_t num_images, uint32_t width, uint32_t height, double xc, double yc, double r)
        : map_(num_images, height, width, xc, yc, r, get_stride(num_images)), storage_(map_.required_span_size())
    {}
    explicit Field(const CylinderMapping &m) : map_(m), storage_(m.required_span_size()) {}

    CylinderView<T> view() { return {storage_.data(), &map_}; }
    CylinderView<const T> view() const { return {storage_.data(), &map_}; }
    const CylinderMapping &mapping() const { return map_; }
    T *data() { return storage_.data(); }
    const T *data() const { return storage_.data(); }

    std::vector<T> &storage() { return storage_; }

    void lo(float x) { lo_ = x; }
    void hi(float x) { hi_ = x; }
    float lo() const { return lo_; }
    float hi() const { return hi_; }

private:
    CylinderMapping map_;
    std::vector<T> storage_;
    float lo_ = 0, hi_ = 0;
};

// ---- context and device fields (cuda-context.h) ------------------------------------
namespace sct_detail {
[[noreturn]] inline void fail(sct_ctx *ctx, int rc, const char *what)
{
    throw std::runtime_error(std::string(what) + ": " + (ctx ? sct_ctx_last_error(ctx) : "") +
                             " (status " + std::to_string(rc) + ")");
}
inline void check(sct_ctx *ctx, int rc, const char *what)
{
    if (rc != SCT_OK) fail(ctx, rc, what);
}
template <typename T>
constexpr int dtype() { return std::is_same_v<T, float> ? SCT_F32 : SCT_U16; }
}  // namespace sct_detail

class HipContext {
public:
    explicit HipContext(int device = 0)
    {
        if (sct_ctx_create(device, &ctx_) != SCT_OK) throw std::runtime_error("no HIP device");
    }
    ~HipContext() { sct_ctx_destroy(ctx_); }
    HipContext(const HipContext &) = delete;
    HipContext &operator=(const HipContext &) = delete;
    sct_ctx *get() const { return ctx_; }
    void *stream() const { return sct_ctx_stream(ctx_); }

private:
    sct_ctx *ctx_ = nullptr;
};

struct HipContextDeleter {
    void operator()(HipContext *c) const noexcept { delete c; }
};
using hip_context_ptr = std::unique_ptr<HipContext, HipContextDeleter>;
inline hip_context_ptr make_context() { return hip_context_ptr{new HipContext{}}; }

struct device_memory_info {
    std::size_t available;
    std::size_t total;
};
inline device_memory_info device_memory(const HipContext &ctx)
{
    device_memory_info m{};
    sct_detail::check(ctx.get(), sct_device_memory(ctx.get(), &m.available, &m.total), "device_memory");
    return m;
}

template <typename T>
class DeviceField {
public:
    DeviceField(sct_field *f, const CylinderMapping &m) : f_(f), map_(m) {}
    ~DeviceField() { sct_field_destroy(f_); }
    DeviceField(const DeviceField &) = delete;
    DeviceField &operator=(const DeviceField &) = delete;
    sct_field *handle() const { return f_; }
    const CylinderMapping &mapping() const { return map_; }
    T *data() const { return static_cast<T *>(sct_field_data(f_)); }

private:
    sct_field *f_;
    CylinderMapping map_;
};

template <typename T>
struct DeviceFieldDeleter {
    void operator()(DeviceField<T> *f) const noexcept { delete f; }
};
template <typename T>
using device_field_ptr = std::unique_ptr<DeviceField<T>, DeviceFieldDeleter<T>>;

template <typename T>
device_field_ptr<T> upload(HipContext &ctx, const Field<T> &a)
{
    sct_layout l = a.mapping().c_layout();
    sct_field *f = nullptr;
    sct_detail::check(ctx.get(), sct_upload(ctx.get(), &l, sct_detail::dtype<T>(), a.data(), &f), "upload");
    return device_field_ptr<T>{new DeviceField<T>{f, a.mapping()}};
}

// download(ctx, field, out), cuda-context.cpp:59-77: throws on span mismatch.
template <typename T>
void download(HipContext &ctx, const DeviceField<T> &in, Field<T> &out)
{
    const int rc = sct_download(ctx.get(), in.handle(), out.data(), out.mapping().required_span_size());
    if (rc == SCT_ERR_SIZE) throw std::runtime_error("Host and device images have different sizes.");
    sct_detail::check(ctx.get(), rc, "download");
}

// ---- solvers ---------------------------------------------------------------------
// potts_min_partition<K>, potts-min-partition.h:8-16.  cpu_map is accepted for
// signature compatibility; the device field carries its own mapping tables.
template <size_t K>
device_field_ptr<float> potts_min_partition(HipContext &ctx, DeviceField<float> &g, const CylinderMapping &cpu_map,
                                            std::array<float, K> ci, float lambda, float tau1, float tau2, int iters,
                                            const sct_potts_opts *opts = nullptr)
{
    static_assert(K >= 2 && K <= max_phases, "K must be in [2, 6]");
    (void)cpu_map;
    sct_field *out = nullptr;
    const int rc = sct_potts_min_partition(ctx.get(), g.handle(), int(K), ci.data(), lambda, tau1, tau2, iters, opts, &out);
    if (rc == SCT_ERR_NOMEM) return {};
    sct_detail::check(ctx.get(), rc, "potts_min_partition");
    return device_field_ptr<float>{new DeviceField<float>{out, g.mapping()}};
}

// chambolle, chambolle.h:5-6: overwrites g.
inline void chambolle(HipContext &ctx, DeviceField<float> &g, float lambda, float tau, int iters)
{
    sct_detail::check(ctx.get(), sct_chambolle(ctx.get(), g.handle(), lambda, tau, iters, nullptr), "chambolle");
}

// ---- reductions --------------------------------------------------------------------
inline std::tuple<float, float> min_max(HipContext &ctx, const DeviceField<float> &in)
{
    float a = 0, b = 0;
    sct_detail::check(ctx.get(), sct_min_max(ctx.get(), in.handle(), &a, &b), "min_max");
    return {a, b};
}

inline std::vector<double> calc_hist(HipContext &ctx, DeviceField<float> &in, float range_start, float range_end,
                                     uint32_t nbins, float delta, uint64_t seed = 0)
{
    std::vector<uint32_t> counts(nbins);
    std::vector<double> density(nbins);
    sct_detail::check(ctx.get(),
                      sct_calc_hist(ctx.get(), in.handle(), range_start, range_end, nbins, delta, seed, counts.data(),
                                    density.data()),
                      "calc_hist");
    return density;
}

inline std::tuple<std::vector<double>, float, float> calc_hist(HipContext &ctx, DeviceField<float> &in, uint32_t nbins,
                                                               float delta, uint64_t seed = 0)
{
    auto [mn, mx] = min_max(ctx, in);
    return {calc_hist(ctx, in, mn, mx, nbins, delta, seed), mn, mx};
}

inline device_field_ptr<float> grad_magn(HipContext &ctx, DeviceField<float> &u)
{
    sct_field *v = nullptr;
    sct_detail::check(ctx.get(), sct_grad_magn(ctx.get(), u.handle(), &v), "grad_magn");
    return device_field_ptr<float>{new DeviceField<float>{v, u.mapping()}};
}

// convert_to_fp.h:8-10 (ElementType uint16_t or float).
template <typename ElementType>
device_field_ptr<float> convert_to_fp(HipContext &ctx, const DeviceField<ElementType> &in, float lo, float hi)
{
    static_assert(std::is_same_v<ElementType, uint16_t> || std::is_same_v<ElementType, float>,
                  "convert_to_fp takes uint16_t or float fields");
    sct_field *out = nullptr;
    sct_detail::check(ctx.get(), sct_convert_to_fp(ctx.get(), in.handle(), lo, hi, &out), "convert_to_fp");
    return device_field_ptr<float>{new DeviceField<float>{out, in.mapping()}};
}

// find_intensity_range.h:8-9.
inline std::tuple<uint16_t, uint16_t> find_intensity_range(HipContext &ctx, DeviceField<uint16_t> &field,
                                                           double tail_thr)
{
    uint16_t lo = 0, hi = 0;
    sct_detail::check(ctx.get(), sct_find_intensity_range(ctx.get(), field.handle(), tail_thr, &lo, &hi, nullptr),
                      "find_intensity_range");
    return {lo, hi};
}

// uv-hist.h:11-16: histogram rows of `stride` elements.
struct HistMapping {
    uint32_t rows, cols, stride;
    uint32_t operator()(uint32_t y, uint32_t x) const { return y * stride + x; }
    uint32_t required_span_size() const { return rows ? (rows - 1) * stride + cols : 0; }
};

struct UVHistReturn {
    std::vector<uint32_t> hist;
    HistMapping mapping;
    float u_start, u_end, v_start, v_end;
};

inline std::tuple<std::vector<uint32_t>, HistMapping> uv_hist(HipContext &ctx, DeviceField<float> &u,
                                                              DeviceField<float> &v, float u_start, float u_end,
                                                              float v_start, float v_end, uint32_t nbins)
{
    HistMapping hm{nbins, nbins, ((nbins + 31u) / 32u) * 32u};
    std::vector<uint32_t> hist(hm.required_span_size());
    uint32_t stride = 0;
    sct_detail::check(ctx.get(),
                      sct_uv_hist(ctx.get(), u.handle(), v.handle(), u_start, u_end, v_start, v_end, nbins, hist.data(),
                                  &stride),
                      "uv_hist");
    hm.stride = stride;
    return {hist, hm};
}

inline UVHistReturn uv_hist(HipContext &ctx, DeviceField<float> &u, DeviceField<float> &v, uint32_t nbins)
{
    float v0 = 0, v1 = 0;
    sct_detail::check(ctx.get(), sct_grad_min_max(ctx.get(), v.handle(), &v0, &v1), "uv_hist");
    auto [u0, u1] = min_max(ctx, u);
    auto [hist, hm] = uv_hist(ctx, u, v, u0, u1, v0, v1, nbins);
    return {hist, hm, u0, u1, v0, v1};
}

// ---- differentiable Potts (no counterpart in the reference) -------------------------
// A forward / backward Function object over the unrolled primal-dual iteration:
// forward(g, c, lambda) runs iters iterations from the nearest-label start and
// returns V_0..V_{K-1} (the soft segmentation), keeping checkpoints every
// checkpoint_every iterations (0: sqrt(iters)); backward(dV) takes dL/dV_i and
// returns dL/dc, dL/dlambda and dL/du, and may be called repeatedly.  g must
// outlive the forward pass's tape.
template <size_t K>
class PottsFunction {
    static_assert(K >= 2 && K <= max_phases, "K must be in [2, 6]");

public:
    struct Gradients {
        std::array<double, K> c{};
        double lambda = 0;
        device_field_ptr<float> u;
    };

    PottsFunction(HipContext &ctx, float tau1, float tau2, int iters, int checkpoint_every = 0)
        : ctx_(ctx), tau1_(tau1), tau2_(tau2), iters_(iters), ck_(checkpoint_every)
    {}
    ~PottsFunction() { sct_potts_tape_destroy(tape_); }
    PottsFunction(const PottsFunction &) = delete;
    PottsFunction &operator=(const PottsFunction &) = delete;

    std::array<device_field_ptr<float>, K> forward(DeviceField<float> &g, std::array<float, K> c, float lambda)
    {
        sct_potts_tape_destroy(tape_);
        tape_ = nullptr;
        std::array<sct_field *, K> v{};
        sct_detail::check(ctx_.get(),
                          sct_potts_tape_forward(ctx_.get(), g.handle(), int(K), c.data(), lambda, tau1_, tau2_, iters_,
                                                 ck_, v.data(), &tape_, &forward_ms_),
                          "PottsFunction::forward");
        map_ = g.mapping();
        std::array<device_field_ptr<float>, K> out;
        for (size_t i = 0; i < K; ++i) out[i] = device_field_ptr<float>{new DeviceField<float>{v[i], map_}};
        return out;
    }

    Gradients backward(const std::array<const DeviceField<float> *, K> &dV)
    {
        if (!tape_) throw std::logic_error("PottsFunction::backward before forward");
        std::array<const sct_field *, K> d{};
        for (size_t i = 0; i < K; ++i) d[i] = dV[i]->handle();
        Gradients gr;
        sct_field *gu = nullptr;
        sct_detail::check(ctx_.get(),
                          sct_potts_tape_backward(tape_, d.data(), gr.c.data(), &gr.lambda, &gu, &backward_ms_),
                          "PottsFunction::backward");
        gr.u = device_field_ptr<float>{new DeviceField<float>{gu, map_}};
        return gr;
    }

    float forward_ms() const { return forward_ms_; }
    float backward_ms() const { return backward_ms_; }

private:
    HipContext &ctx_;
    float tau1_, tau2_;
    int iters_, ck_;
    sct_potts_tape *tape_ = nullptr;
    CylinderMapping map_;
    float forward_ms_ = 0, backward_ms_ = 0;
};

// ---- the .bin archive: Field<T>::save / load, field.h:92-153 ------------------------
// The reference serialises with Boost archives over its bit streams; the same
// stream is produced and read here through the C ABI, on host memory.
template <typename T>
std::vector<unsigned char> save_bin(const Field<T> &f)
{
    const sct_layout l = f.mapping().c_layout();
    std::vector<unsigned char> buf(sct_bin_bound(&l, sct_detail::dtype<T>()));
    size_t written = 0;
    sct_detail::check(nullptr,
                      sct_bin_write(&l, sct_detail::dtype<T>(), f.lo(), f.hi(), f.data(), buf.data(), buf.size(),
                                    &written),
                      "save");
    buf.resize(written);
    return buf;
}

template <typename T>
void save_bin(const Field<T> &f, const std::string &path)
{
    const auto buf = save_bin(f);
    std::FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error("cannot open " + path);
    const size_t n = std::fwrite(buf.data(), 1, buf.size(), fp);
    const bool ok = n == buf.size() && std::fclose(fp) == 0;
    if (!ok) throw std::runtime_error("cannot write " + path);
}

template <typename T>
Field<T> load_bin(const void *buf, size_t len)
{
    sct_bin_info info{};
    sct_detail::check(nullptr, sct_bin_parse(buf, len, sct_detail::dtype<T>(), &info, nullptr, nullptr), "load");
    std::vector<uint32_t> start_row(info.width), base(size_t(info.width) + 1);
    sct_detail::check(nullptr, sct_bin_parse(buf, len, sct_detail::dtype<T>(), &info, start_row.data(), base.data()),
                      "load");
    Field<T> f(CylinderMapping(info.n_images, info.height, info.width, info.nstride, info.col_begin, info.col_end,
                               std::move(start_row), std::move(base), info.data_size));
    sct_detail::check(nullptr, sct_bin_read(buf, len, sct_detail::dtype<T>(), &info, f.data()), "load");
    f.lo(info.lo);
    f.hi(info.hi);
    return f;
}

template <typename T>
Field<T> load_bin(const std::string &path)
{
    std::FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error("cannot open " + path);
    std::vector<unsigned char> buf;
    unsigned char chunk[1 << 16];
    size_t n;
    while ((n = std::fread(chunk, 1, sizeof chunk, fp)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    std::fclose(fp);
    return load_bin<T>(buf.data(), buf.size());
}
