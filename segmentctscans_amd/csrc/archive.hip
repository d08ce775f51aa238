// archive.hip -- the reference's .bin volume archive (host code).
//
// A Field<T> archive is one MSB-first bit stream (io/outbitstream.h:10-55):
//   mapping header (layout_cylinder.h:233-318): ue(n_images), ue(height),
//     ue(width), ue(nstride), ue(col_begin), ue(col_end), ue(#start_row),
//     ue(start_row[0]), se(start_row[i] - start_row[i-1])..., ue(#base),
//     ue(base[0]), se(base[i] - base[i-1])...
//   T = float only: lo, hi as u<32> bit patterns (field.h:95-101)
//   the in-mask voxels, (col, row, n) order, as u<16> / u<32> (field.h:102-115)
// ue is order-0 Exponential-Golomb, se maps k > 0 to 2k - 1 and k <= 0 to -2k
// (io/writeprimitives.h:55-66, io/readprimitives.h:53-75).  The file is the
// stream's bytes, the last one zero-padded (transpose.cpp:303-308).
//
// Every voxel has the same width, so voxel k starts at bit H + k * bits (H =
// header bits): lines are split among threads, each packs its voxels with a
// 64-bit accumulator, and the (at most two) bytes a chunk shares with its
// neighbours are merged afterwards.
#include "sct_internal.hpp"

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct BitWriter {
    uint8_t *buf;
    size_t cap;
    uint64_t bit = 0;
    bool overflow = false;

    void put(uint64_t v, unsigned n)   // the low n bits of v, most significant first
    {
        for (unsigned i = n; i-- > 0;) {
            const size_t byte = bit >> 3;
            if (byte >= cap) {
                overflow = true;
                return;
            }
            const unsigned sh = 7 - unsigned(bit & 7);
            if (sh == 7) buf[byte] = 0;
            buf[byte] |= uint8_t(((v >> i) & 1u) << sh);
            ++bit;
        }
    }
    void ue(uint64_t v)
    {
        const unsigned nb = 64 - unsigned(__builtin_clzll(v + 1));
        put(0, nb - 1);
        put(v + 1, nb);
    }
    void se(int64_t v)
    {
        uint64_t a = 2 * uint64_t(v < 0 ? -v : v);
        if (v > 0) --a;
        ue(a);
    }
};

struct BitReader {
    const uint8_t *buf;
    size_t len;
    uint64_t bit = 0;
    bool fail = false;

    uint64_t get(unsigned n)
    {
        uint64_t v = 0;
        for (unsigned i = 0; i < n; ++i) {
            const size_t byte = bit >> 3;
            if (byte >= len) {
                fail = true;
                return 0;
            }
            v = (v << 1) | ((buf[byte] >> (7 - (bit & 7))) & 1u);
            ++bit;
        }
        return v;
    }
    uint64_t ue()
    {
        unsigned z = 0;
        while (!fail && get(1) == 0) {
            if (++z > 63) {
                fail = true;
                return 0;
            }
        }
        return z ? (uint64_t(1) << z) + get(z) - 1 : 0;
    }
    int64_t se()
    {
        const uint64_t u = ue();
        return (u & 1) ? int64_t((u + 1) >> 1) : -int64_t(u >> 1);
    }
};

unsigned elem_bits(int dtype) { return dtype == SCT_U16 ? 16 : 32; }

uint32_t lines_of(const uint32_t *base, uint32_t width, uint32_t nstride)
{
    return nstride ? base[width] / nstride : 0;
}

// Header of a layout: bits and (if buf) the bits themselves.
void write_header(BitWriter &w, const sct_layout &l, int dtype, float lo, float hi)
{
    w.ue(l.n_images);
    w.ue(l.height);
    w.ue(l.width);
    w.ue(l.nstride);
    w.ue(l.col_begin);
    w.ue(l.col_end);
    w.ue(l.width);
    if (l.width) {
        w.ue(l.start_row[0]);
        for (uint32_t i = 1; i < l.width; ++i) w.se(int64_t(l.start_row[i]) - int64_t(l.start_row[i - 1]));
    }
    w.ue(uint64_t(l.width) + 1);
    w.ue(l.base[0]);
    for (uint32_t i = 1; i <= l.width; ++i) w.se(int64_t(l.base[i]) - int64_t(l.base[i - 1]));
    if (dtype == SCT_F32) {
        uint32_t b;
        std::memcpy(&b, &lo, 4);
        w.put(b, 32);
        std::memcpy(&b, &hi, 4);
        w.put(b, 32);
    }
}

// Splits [0, lines) into chunks for the available threads and runs f(l0, l1, chunk).
template <typename F>
void parallel_lines(uint32_t lines, F f)
{
    const unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), std::max(1u, lines / 64)));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        const uint32_t l0 = uint32_t(uint64_t(lines) * t / nt), l1 = uint32_t(uint64_t(lines) * (t + 1) / nt);
        if (nt == 1) f(l0, l1, t);
        else th.emplace_back(f, l0, l1, t);
    }
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" size_t sct_bin_bound(const sct_layout *l, int dtype)
{
    if (!l || !l->base || (l->width && !l->start_row) || l->nstride == 0) return 0;
    // ue/se of a 64-bit value takes at most 129 bits
    const uint64_t header = 129ull * (2ull * l->width + 9) + 64;
    const uint64_t vox = uint64_t(lines_of(l->base, l->width, l->nstride)) * l->n_images * elem_bits(dtype);
    return size_t((header + vox + 7) / 8);
}

static int sct_bin_write_impl(const sct_layout *l, int dtype, float lo, float hi, const void *data, void *out,
                             size_t cap, size_t *written)
{
    if (!l || !data || !out || !written || !l->base || (l->width && !l->start_row) || l->nstride == 0 ||
        l->nstride < l->n_images || (dtype != SCT_F32 && dtype != SCT_U16))
        return SCT_ERR_INVALID;
    uint8_t *buf = static_cast<uint8_t *>(out);
    BitWriter w{buf, cap};
    write_header(w, *l, dtype, lo, hi);
    if (w.overflow) return SCT_ERR_SIZE;
    const uint64_t H = w.bit;
    const unsigned nb = elem_bits(dtype);
    const uint32_t lines = lines_of(l->base, l->width, l->nstride), n = l->n_images;
    const uint64_t total_bits = H + uint64_t(lines) * n * nb;
    if ((total_bits + 7) / 8 > cap) return SCT_ERR_SIZE;
    if (H & 7) buf[H >> 3] &= uint8_t(0xFF << (8 - (H & 7)));   // keep the header's bits of the shared byte
    struct Edge {
        size_t pos[2];
        uint8_t v[2];
        int n;
    };
    std::vector<Edge> edges(std::max(1u, std::thread::hardware_concurrency()));
    parallel_lines(lines, [&](uint32_t l0, uint32_t l1, unsigned t) {
        Edge &e = edges[t];
        e.n = 0;
        if (l0 >= l1 || n == 0) return;
        const uint64_t s = H + uint64_t(l0) * n * nb;
        size_t pos = size_t(s >> 3);
        const unsigned phase = unsigned(s & 7);
        uint64_t acc = 0;
        unsigned bits = phase;   // leading bits belong to the previous chunk (zeros here)
        bool first = phase != 0;
        auto emit = [&](uint8_t b) {
            if (first) {
                e.pos[e.n] = pos;
                e.v[e.n++] = b;
                first = false;
            } else {
                buf[pos] = b;
            }
            ++pos;
        };
        for (uint32_t line = l0; line < l1; ++line) {
            const size_t o = size_t(line) * l->nstride;
            for (uint32_t k = 0; k < n; ++k) {
                const uint64_t v = nb == 16 ? static_cast<const uint16_t *>(data)[o + k]
                                            : static_cast<const uint32_t *>(data)[o + k];
                acc = (acc << nb) | v;
                bits += nb;
                while (bits >= 8) {
                    emit(uint8_t(acc >> (bits - 8)));
                    bits -= 8;
                }
                acc &= (uint64_t(1) << bits) - 1;
            }
        }
        if (bits) {   // last byte, shared with the next chunk (or the stream's padding)
            e.pos[e.n] = pos;
            e.v[e.n++] = uint8_t(acc << (8 - bits));
        }
    });
    // merge the shared bytes: OR of the header's bits (kept above) and both chunks' parts
    std::vector<std::pair<size_t, uint8_t>> parts;
    for (const Edge &e : edges)
        for (int i = 0; i < e.n; ++i) parts.emplace_back(e.pos[i], e.v[i]);
    std::sort(parts.begin(), parts.end());
    for (size_t i = 0; i < parts.size(); ++i) {
        const size_t p = parts[i].first;
        const bool header_byte = (H & 7) && p == (H >> 3);
        if (!header_byte && (i == 0 || parts[i - 1].first != p)) buf[p] = 0;
        buf[p] |= parts[i].second;
    }
    *written = size_t((total_bits + 7) / 8);
    return SCT_OK;
}

static int sct_bin_parse_impl(const void *in, size_t len, int dtype, sct_bin_info *info, uint32_t *start_row,
                             uint32_t *base)
{
    if (!in || !info || (dtype != SCT_F32 && dtype != SCT_U16)) return SCT_ERR_INVALID;
    BitReader r{static_cast<const uint8_t *>(in), len};
    sct_bin_info I{};
    const uint64_t e0 = r.ue(), e1 = r.ue(), e2 = r.ue(), ns = r.ue(), cb = r.ue(), ce = r.ue();
    const uint64_t nsr = r.ue();
    if (r.fail || e0 > 0xffffffffu || e1 > 0xffffffffu || e2 > 0xffffffffu || ns > 0xffffffffu ||
        cb > 0xffffffffu || ce > 0xffffffffu || nsr != e2)
        return SCT_ERR_INVALID;
    I.n_images = uint32_t(e0);
    I.height = uint32_t(e1);
    I.width = uint32_t(e2);
    I.nstride = uint32_t(ns);
    I.col_begin = uint32_t(cb);
    I.col_end = uint32_t(ce);
    // start_row[width] and base[width + 1] follow, one Exp-Golomb code (at
    // least one bit) each, plus the base count: a width the remaining bits
    // cannot hold is rejected before anything is sized by it
    if (2 * uint64_t(I.width) + 2 > uint64_t(len) * 8 - std::min<uint64_t>(r.bit, uint64_t(len) * 8))
        return SCT_ERR_SIZE;
    std::vector<uint32_t> sr(I.width), bs(size_t(I.width) + 1);
    if (I.width) {
        sr[0] = uint32_t(r.ue());
        for (uint32_t i = 1; i < I.width; ++i) sr[i] = uint32_t(sr[i - 1] + r.se());
    }
    if (r.ue() != uint64_t(I.width) + 1 || r.fail) return SCT_ERR_INVALID;
    bs[0] = uint32_t(r.ue());
    for (uint32_t i = 1; i <= I.width; ++i) bs[i] = uint32_t(bs[i - 1] + r.se());
    if (r.fail) return SCT_ERR_SIZE;
    if (dtype == SCT_F32) {
        const uint32_t a = uint32_t(r.get(32)), b = uint32_t(r.get(32));
        std::memcpy(&I.lo, &a, 4);
        std::memcpy(&I.hi, &b, 4);
        if (r.fail) return SCT_ERR_SIZE;
    }
    // the mapping must be a cylinder layout: whole lines, monotone bases
    if (I.nstride == 0 || I.nstride < I.n_images || bs[0] != 0) return SCT_ERR_INVALID;
    uint64_t ds = 0;
    for (uint32_t c = 0; c < I.width; ++c) {
        if (bs[c + 1] < bs[c] || (bs[c + 1] - bs[c]) % I.nstride) return SCT_ERR_INVALID;
        if (c >= I.col_begin && c < I.col_end) ds += (bs[c + 1] - bs[c]) / I.nstride;
    }
    I.span = bs[I.width];
    I.data_size = ds * I.n_images;
    I.header_bits = r.bit;
    const uint64_t total_bits = r.bit + uint64_t(I.span / I.nstride) * I.n_images * elem_bits(dtype);
    if ((total_bits + 7) / 8 > len) return SCT_ERR_SIZE;
    *info = I;
    if (start_row) std::copy(sr.begin(), sr.end(), start_row);
    if (base) std::copy(bs.begin(), bs.end(), base);
    return SCT_OK;
}

static int sct_bin_read_impl(const void *in, size_t len, int dtype, const sct_bin_info *info, void *data)
{
    if (!in || !info || !data || (dtype != SCT_F32 && dtype != SCT_U16) || info->nstride == 0) return SCT_ERR_INVALID;
    const unsigned nb = elem_bits(dtype);
    const uint32_t lines = uint32_t(info->span / info->nstride), n = info->n_images;
    const uint64_t H = info->header_bits;
    if ((H + uint64_t(lines) * n * nb + 7) / 8 > len) return SCT_ERR_SIZE;
    const uint8_t *buf = static_cast<const uint8_t *>(in);
    std::memset(data, 0, size_t(info->span) * (nb / 8));
    parallel_lines(lines, [&](uint32_t l0, uint32_t l1, unsigned) {
        if (l0 >= l1 || n == 0) return;
        const uint64_t s = H + uint64_t(l0) * n * nb;
        size_t pos = size_t(s >> 3);
        uint64_t acc = buf[pos++] & (0xFFu >> (s & 7));
        unsigned bits = 8 - unsigned(s & 7);
        for (uint32_t line = l0; line < l1; ++line) {
            const size_t o = size_t(line) * info->nstride;
            for (uint32_t k = 0; k < n; ++k) {
                while (bits < nb) {
                    acc = (acc << 8) | buf[pos++];
                    bits += 8;
                }
                const uint64_t v = (acc >> (bits - nb)) & ((uint64_t(1) << nb) - 1);
                bits -= nb;
                acc &= (uint64_t(1) << bits) - 1;
                if (nb == 16) static_cast<uint16_t *>(data)[o + k] = uint16_t(v);
                else static_cast<uint32_t *>(data)[o + k] = uint32_t(v);
            }
        }
    });
    return SCT_OK;
}

// The C ABI lets no exception through: a failed allocation (or any other
// exception) of the host code above is an error status.
extern "C" int sct_bin_parse(const void *in, size_t len, int dtype, sct_bin_info *info, uint32_t *start_row,
                             uint32_t *base)
{
    try {
        return sct_bin_parse_impl(in, len, dtype, info, start_row, base);
    } catch (...) {
        return SCT_ERR_INVALID;
    }
}

extern "C" int sct_bin_read(const void *in, size_t len, int dtype, const sct_bin_info *info, void *data)
{
    try {
        return sct_bin_read_impl(in, len, dtype, info, data);
    } catch (...) {
        return SCT_ERR_INVALID;
    }
}

extern "C" int sct_bin_write(const sct_layout *layout, int dtype, float lo, float hi, const void *data, void *buf,
                             size_t cap, size_t *written)
{
    try {
        return sct_bin_write_impl(layout, dtype, lo, hi, data, buf, cap, written);
    } catch (...) {
        return SCT_ERR_INVALID;
    }
}
