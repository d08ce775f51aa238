// phantom.cpp -- seeded synthetic micro-CT phantom (see phantom.h).
#include "phantom.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

// Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
// SC'11): 10 rounds of two 32x32->64 multiplies with Weyl key bumps.
struct Philox {
    uint32_t k0, k1;
    explicit Philox(uint64_t seed) : k0(uint32_t(seed)), k1(uint32_t(seed >> 32)) {}
    void operator()(const uint32_t ctr[4], uint32_t out[4]) const
    {
        uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
        uint32_t a = k0, b = k1;
        for (int r = 0; r < 10; ++r) {
            uint64_t p0 = uint64_t(0xD2511F53u) * c0;
            uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
            uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ a;
            uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ b;
            c0 = n0;
            c1 = uint32_t(p1);
            c2 = n2;
            c3 = uint32_t(p0);
            a += 0x9E3779B9u;
            b += 0xBB67AE85u;
        }
        out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
    }
};

inline double unit(uint32_t x) { return double(x) * (1.0 / 4294967296.0); }   // [0, 1)

struct Sphere { double z, y, x, R; uint8_t phase; };

}  // namespace

extern "C" void sct_phantom_defaults(sct_phantom_params *p, uint32_t n_images, uint32_t side,
                                     double rmin, double rmax, uint64_t seed)
{
    std::memset(p, 0, sizeof(*p));
    p->n_images = n_images;
    p->height = side;
    p->width = side;
    p->xc = p->yc = (double(side) - 1.0) / 2.0;
    p->r = 0.45 * double(side);
    p->rmin = rmin;
    p->rmax = rmax;
    p->porosity = 0.35;
    p->c_matrix = 0.0339;
    p->c_phase1 = 0.0409;
    p->c_phase2 = 0.0527;
    p->p_phase2 = 0.5;
    p->blur_sigma = 1.0;
    p->blur_radius = 4;
    p->noise_sd = 0.008;
    p->lo = 0.019182531;
    p->hi = 0.075767376;
    p->seed = seed;
    p->max_attempts = 2000000;
}

extern "C" int sct_phantom_generate(const sct_phantom_params *p, const uint32_t *start_row,
                                    const uint32_t *base, uint32_t nstride, uint32_t col_begin,
                                    uint32_t col_end, float *u, uint8_t *labels,
                                    sct_phantom_stats *stats, int nthreads)
{
    if (!p || !start_row || !base || !u || nstride < p->n_images || p->hi == p->lo ||
        p->rmin <= 0 || p->rmax < p->rmin || p->blur_radius < 0)
        return 1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const size_t N = p->n_images, H = p->height, W = p->width;
    const size_t dense = N * H * W;
    auto row_end = [&](uint32_t col) { return start_row[col] + (base[col + 1] - base[col]) / nstride; };
    auto in_mask = [&](size_t row, size_t col) {
        return col >= col_begin && col < col_end && row >= start_row[col] && row < row_end(uint32_t(col));
    };

    std::vector<float> a, b;
    std::vector<uint8_t> lab;
    try {
        a.assign(dense, 0.0f);
        b.assign(dense, 0.0f);
        lab.assign(dense, 0);
    } catch (...) {
        return 2;
    }
    auto D = [&](size_t n, size_t row, size_t col) { return (col * H + row) * N + n; };

    // 1. matrix inside the mask
#pragma omp parallel for schedule(static)
    for (long col = 0; col < long(W); ++col)
        for (size_t row = 0; row < H; ++row)
            if (in_mask(row, size_t(col)))
                for (size_t n = 0; n < N; ++n) a[D(n, row, size_t(col))] = float(p->c_matrix);

    // 2. random sequential addition of non-overlapping spheres
    const Philox rng(p->seed);
    const double target = p->porosity * M_PI * p->r * p->r * double(N);
    const double cell = 2.0 * p->rmax;
    std::vector<Sphere> spheres;
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    auto key = [](int64_t i, int64_t j, int64_t k) {
        return (uint64_t(i + 1048576) << 42) | (uint64_t(j + 1048576) << 21) | uint64_t(k + 1048576);
    };
    double vol = 0.0;
    for (int64_t att = 0; att < p->max_attempts && vol < target; ++att) {
        uint32_t c1[4] = {uint32_t(att), uint32_t(uint64_t(att) >> 32), 0u, 0u}, r1[4];
        uint32_t c2[4] = {uint32_t(att), uint32_t(uint64_t(att) >> 32), 1u, 0u}, r2[4];
        rng(c1, r1);
        rng(c2, r2);
        Sphere s;
        s.R = p->rmin + (p->rmax - p->rmin) * unit(r1[0]);
        double reach = p->r - s.R;
        if (reach <= 0) continue;
        double rho = reach * std::sqrt(unit(r1[1]));
        double phi = 2.0 * M_PI * unit(r1[2]);
        s.x = p->xc + rho * std::cos(phi);
        s.y = p->yc + rho * std::sin(phi);
        s.z = double(N) * unit(r1[3]);
        s.phase = unit(r2[0]) < p->p_phase2 ? 2 : 1;
        int64_t gi = int64_t(std::floor(s.z / cell)), gj = int64_t(std::floor(s.y / cell)),
                gk = int64_t(std::floor(s.x / cell));
        bool ok = true;
        for (int64_t di = -1; di <= 1 && ok; ++di)
            for (int64_t dj = -1; dj <= 1 && ok; ++dj)
                for (int64_t dk = -1; dk <= 1 && ok; ++dk) {
                    auto it = grid.find(key(gi + di, gj + dj, gk + dk));
                    if (it == grid.end()) continue;
                    for (uint32_t id : it->second) {
                        const Sphere &o = spheres[id];
                        double dz = o.z - s.z, dy = o.y - s.y, dx = o.x - s.x;
                        if (dz * dz + dy * dy + dx * dx < (o.R + s.R) * (o.R + s.R)) { ok = false; break; }
                    }
                }
        if (!ok) continue;
        grid[key(gi, gj, gk)].push_back(uint32_t(spheres.size()));
        spheres.push_back(s);
        vol += 4.0 / 3.0 * M_PI * s.R * s.R * s.R;
    }
    // Spheres are disjoint, so rasterising them in parallel writes disjoint voxels.
#pragma omp parallel for schedule(dynamic, 4)
    for (long si = 0; si < long(spheres.size()); ++si) {
        const Sphere &s = spheres[size_t(si)];
        float val = float(s.phase == 2 ? p->c_phase2 : p->c_phase1);
        long c0 = std::max(0L, long(std::ceil(s.x - s.R))), c1 = std::min(long(W) - 1, long(std::floor(s.x + s.R)));
        long r0 = std::max(0L, long(std::ceil(s.y - s.R))), r1 = std::min(long(H) - 1, long(std::floor(s.y + s.R)));
        long n0 = std::max(0L, long(std::ceil(s.z - s.R))), n1 = std::min(long(N) - 1, long(std::floor(s.z + s.R)));
        for (long col = c0; col <= c1; ++col)
            for (long row = r0; row <= r1; ++row) {
                if (!in_mask(size_t(row), size_t(col))) continue;
                double dxy = (double(col) - s.x) * (double(col) - s.x) + (double(row) - s.y) * (double(row) - s.y);
                for (long n = n0; n <= n1; ++n) {
                    double d2 = dxy + (double(n) - s.z) * (double(n) - s.z);
                    if (d2 <= s.R * s.R) {
                        size_t o = D(size_t(n), size_t(row), size_t(col));
                        a[o] = val;
                        lab[o] = s.phase;
                    }
                }
            }
    }

    // 3. separable Gaussian blur, clamped edges, fixed summation order
    const int R = p->blur_radius;
    std::vector<double> w(size_t(2 * R + 1));
    {
        double sum = 0;
        for (int k = -R; k <= R; ++k) {
            w[size_t(k + R)] = std::exp(-double(k) * double(k) / (2.0 * p->blur_sigma * p->blur_sigma));
            sum += w[size_t(k + R)];
        }
        for (auto &x : w) x /= sum;
    }
    auto clampi = [](long i, long n) { return i < 0 ? 0L : (i >= n ? n - 1 : i); };
    if (R > 0) {
        // along n: a -> b
#pragma omp parallel for schedule(static)
        for (long line = 0; line < long(W * H); ++line) {
            const float *src = &a[size_t(line) * N];
            float *dst = &b[size_t(line) * N];
            for (long n = 0; n < long(N); ++n) {
                double s = 0;
                for (int k = -R; k <= R; ++k) s += w[size_t(k + R)] * src[clampi(n + k, long(N))];
                dst[n] = float(s);
            }
        }
        // along rows: b -> a
#pragma omp parallel for schedule(static)
        for (long col = 0; col < long(W); ++col) {
            std::vector<double> acc(N);
            for (long row = 0; row < long(H); ++row) {
                std::fill(acc.begin(), acc.end(), 0.0);
                for (int k = -R; k <= R; ++k) {
                    const float *src = &b[D(0, size_t(clampi(row + k, long(H))), size_t(col))];
                    double wk = w[size_t(k + R)];
                    for (size_t n = 0; n < N; ++n) acc[n] += wk * src[n];
                }
                float *dst = &a[D(0, size_t(row), size_t(col))];
                for (size_t n = 0; n < N; ++n) dst[n] = float(acc[n]);
            }
        }
        // along cols: a -> b
#pragma omp parallel for schedule(static)
        for (long col = 0; col < long(W); ++col) {
            std::vector<double> acc(N);
            for (long row = 0; row < long(H); ++row) {
                std::fill(acc.begin(), acc.end(), 0.0);
                for (int k = -R; k <= R; ++k) {
                    const float *src = &a[D(0, size_t(row), size_t(clampi(col + k, long(W))))];
                    double wk = w[size_t(k + R)];
                    for (size_t n = 0; n < N; ++n) acc[n] += wk * src[n];
                }
                float *dst = &b[D(0, size_t(row), size_t(col))];
                for (size_t n = 0; n < N; ++n) dst[n] = float(acc[n]);
            }
        }
    } else {
        b = a;
    }

    // 4-5. noise on in-mask voxels, normalisation, store in the cylinder layout
    uint64_t in_mask_count = 0, pore_count = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : in_mask_count, pore_count)
    for (long col = long(col_begin); col < long(col_end); ++col) {
        for (uint32_t row = start_row[col]; row < row_end(uint32_t(col)); ++row) {
            size_t line = base[col] / nstride + (row - start_row[col]);
            for (size_t n = 0; n < N; ++n) {
                size_t d = D(n, row, size_t(col));
                uint32_t ctr[4] = {uint32_t(d), uint32_t(uint64_t(d) >> 32), 2u, 0u}, r[4];
                rng(ctr, r);
                double u1 = (double(r[0]) + 1.0) * (1.0 / 4294967296.0);   // (0, 1]
                double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * unit(r[1]));
                double x = double(b[d]) + p->noise_sd * z;
                u[line * nstride + n] = float((x - p->lo) / (p->hi - p->lo));
                if (labels) labels[line * nstride + n] = lab[d];
                ++in_mask_count;
                pore_count += lab[d] != 0;
            }
            for (size_t n = N; n < nstride; ++n) {
                u[line * nstride + n] = 0.0f;
                if (labels) labels[line * nstride + n] = 0;
            }
        }
    }
    if (stats) {
        stats->n_spheres = uint32_t(spheres.size());
        stats->porosity = in_mask_count ? double(pore_count) / double(in_mask_count) : 0.0;
    }
    return 0;
}
