// api_example.cpp -- the reference's driver flow (potts-model.cpp:126-182,
// totalvarmin.cpp:113-156, uv-hist.cpp:86-94) written against segmentctscans.hpp.
// Exits 0 and prints "api_example ok" when every check passes.
#include <segmentctscans.hpp>

#include <cmath>
#include <cstdio>

static int fails = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            ++fails;                                                       \
        }                                                                  \
    } while (0)

int main()
{
    // 48 images of a 40 x 40 cylinder: left half intensity 0.2, right half 0.8
    auto img = Field<float>(48, 40, 40, 19.5, 19.5, 18.0);
    img.lo(0.0f);
    img.hi(1.0f);
    auto v = img.view();
    const auto &m = img.mapping();
    for (auto col = m.col_begin(); col < m.col_end(); ++col)
        for (auto row = m.row_begin(col); row < m.row_end(col); ++row)
            for (uint32_t n = 0; n < m.n_images(); ++n) v(n, row, col) = col < 20 ? 0.2f : 0.8f;

    auto ctx = make_context();
    auto [avail, total] = device_memory(*ctx);
    CHECK(avail > 0 && total >= avail);

    // potts-model.cpp:141-170
    const float tau = 0.99f * std::sqrt(1.0f / 12.0f);
    auto d_img = upload(*ctx, img);
    auto ci = std::array<float, 2>{0.2f, 0.8f};
    auto d_res = potts_min_partition(*ctx, *d_img, img.mapping(), ci, 30.0f, tau, tau, 100);
    CHECK(d_res != nullptr);
    auto res = Field<float>(img.mapping());
    download(*ctx, *d_res, res);
    auto rv = res.view();
    for (auto col = m.col_begin(); col < m.col_end(); ++col)
        for (auto row = m.row_begin(col); row < m.row_end(col); ++row)
            for (uint32_t n = 0; n < m.n_images(); ++n) CHECK(rv(n, row, col) == (col < 20 ? 0.2f : 0.8f));

    // a slab that cannot hold 2 * iters planes: the reference's null return
    sct_potts_opts o;
    sct_potts_opts_init(&o);
    o.slab_w = 32;
    CHECK(potts_min_partition(*ctx, *d_img, img.mapping(), ci, 30.0f, tau, tau, 100, &o) == nullptr);

    // totalvarmin.cpp:123-156
    auto [hist, mn, mx] = calc_hist(*ctx, *d_img, 1024u, 0.0f);
    CHECK(mn == 0.2f && mx == 0.8f);
    double mass = 0;
    for (double h : hist) mass += h * (double(mx) - double(mn)) / 1024;
    CHECK(std::fabs(mass - 1.0) < 1e-9);
    chambolle(*ctx, *d_img, 0.06f, 0.99f / 6.0f, 150);
    auto [mn2, mx2] = min_max(*ctx, *d_img);
    CHECK(mn2 > 0.2f && mx2 < 0.8f);   // TV shrinks the step
    auto hist2 = calc_hist(*ctx, *d_img, mn, mx, 1024u, 0.0f);
    CHECK(hist2.size() == 1024);

    // uv-hist.cpp:86-94
    auto d_v = grad_magn(*ctx, *d_img);
    auto uv = uv_hist(*ctx, *d_img, *d_v, 64u);
    CHECK(uv.mapping.stride == 64 && uv.v_start > 0);

    // forward / backward through the unrolled iteration: L = sum of V_0 over
    // the left half (dL/dV_0 = 1 there); twice the upstream gradient gives
    // exactly twice the gradients
    {
        auto fn = PottsFunction<2>(*ctx, tau, tau, 20, 5);
        auto d_in = upload(*ctx, img);
        auto V = fn.forward(*d_in, ci, 30.0f);
        auto hv0 = Field<float>(img.mapping()), hv1 = Field<float>(img.mapping());
        download(*ctx, *V[0], hv0);
        download(*ctx, *V[1], hv1);
        auto w = Field<float>(img.mapping()), w2 = Field<float>(img.mapping()), z = Field<float>(img.mapping());
        auto wv = w.view(), w2v = w2.view();
        for (auto col = m.col_begin(); col < m.col_end(); ++col)
            for (auto row = m.row_begin(col); row < m.row_end(col); ++row)
                for (uint32_t n = 0; n < m.n_images(); ++n) {
                    CHECK(hv0.view()(n, row, col) + hv1.view()(n, row, col) == 1.0f);
                    wv(n, row, col) = col < 20 ? 1.0f : 0.0f;
                    w2v(n, row, col) = col < 20 ? 2.0f : 0.0f;
                }
        auto d_w = upload(*ctx, w), d_w2 = upload(*ctx, w2), d_z = upload(*ctx, z);
        auto g1 = fn.backward({d_w.get(), d_z.get()});
        auto g2 = fn.backward({d_w2.get(), d_z.get()});
        CHECK(std::isfinite(g1.c[0]) && std::isfinite(g1.lambda));
        CHECK(g2.c[0] == 2 * g1.c[0] && g2.c[1] == 2 * g1.c[1] && g2.lambda == 2 * g1.lambda);
        auto hu1 = Field<float>(img.mapping()), hu2 = Field<float>(img.mapping());
        download(*ctx, *g1.u, hu1);
        download(*ctx, *g2.u, hu2);
        for (uint32_t i = 0; i < m.required_span_size(); ++i) CHECK(hu2.data()[i] == 2.0f * hu1.data()[i]);
    }

    // download into a field of another size throws (cuda-context.cpp:69-71)
    bool threw = false;
    try {
        auto small = Field<float>(8, 8, 8, 3.5, 3.5, 3.0);
        download(*ctx, *d_res, small);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);

    if (fails) {
        std::fprintf(stderr, "api_example: %d failures\n", fails);
        return 1;
    }
    std::printf("api_example ok\n");
    return 0;
}
