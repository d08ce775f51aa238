// archive_example.cpp -- Field<T>::save / load (field.h:92-153) through
// save_bin / load_bin of segmentctscans.hpp, host only.  Writes
// <dir>/f32.bin and <dir>/u16.bin for a 37 x 29 x 23 cylinder (radius 11.2)
// filled with a fixed pattern, reads them back and checks every field.
// Prints "archive_example ok" when every check passes.
#include <segmentctscans.hpp>

#include <cstdio>
#include <string>

static int fails = 0;
#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__);   \
            ++fails;                                                                 \
        }                                                                            \
    } while (0)

template <typename T>
static T pattern(uint32_t n, uint32_t row, uint32_t col)
{
    if constexpr (std::is_same_v<T, float>)
        return float(n + 3 * row + 7 * col) / 1000.0f;
    else
        return uint16_t((n * 31u + row * 17u + col * 13u) & 0xffffu);
}

template <typename T>
static void round_trip(const std::string &path)
{
    Field<T> f(37, 29, 23, 14.0, 11.0, 11.2);
    if constexpr (std::is_same_v<T, float>) {
        f.lo(-0.25f);
        f.hi(1.5f);
    }
    auto v = f.view();
    const auto &m = f.mapping();
    for (auto col = m.col_begin(); col < m.col_end(); ++col)
        for (auto row = m.row_begin(col); row < m.row_end(col); ++row)
            for (uint32_t n = 0; n < m.n_images(); ++n) v(n, row, col) = pattern<T>(n, row, col);
    save_bin(f, path);
    const Field<T> g = load_bin<T>(path);
    const auto &gm = g.mapping();
    CHECK(gm.n_images() == m.n_images() && gm.height() == m.height() && gm.width() == m.width());
    CHECK(gm.nstride() == m.nstride() && gm.col_begin() == m.col_begin() && gm.col_end() == m.col_end());
    CHECK(gm.required_span_size() == m.required_span_size() && gm.data_size() == m.data_size());
    CHECK(g.lo() == f.lo() && g.hi() == f.hi());
    for (uint32_t i = 0; i < m.required_span_size(); ++i) CHECK(g.data()[i] == f.data()[i]);
    const auto bytes = save_bin(g);
    const auto again = save_bin(f);
    CHECK(bytes == again);
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    round_trip<float>(dir + "/f32.bin");
    round_trip<uint16_t>(dir + "/u16.bin");
    bool threw = false;
    try {
        const unsigned char junk[3] = {0, 0, 0};
        (void)load_bin<float>(junk, sizeof junk);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    if (fails) {
        std::fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    std::printf("archive_example ok\n");
    return 0;
}
