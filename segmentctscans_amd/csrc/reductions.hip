// reductions.hip -- min_max, calc_hist, grad_magn, grad_min_max, uv_hist
// (min_max/min_max.cpp, calc_hist/calc_hist.cpp, grad-magn/grad_magn.cpp,
// uv-hist/uv-hist-kernel.cpp).
//
// All counts are integer adds and all extrema are order-independent, so the
// results do not depend on scheduling.  Waves stream 64 consecutive n of one
// line; work is distributed over (line, 64-chunk) pairs.
#include "sct_internal.hpp"
#include "potts_math.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

namespace {

struct VoxGeom {
    const LineNbr *nbr;
    uint32_t lines, len, chunks, stride;
};

VoxGeom vox_geom(const sct_field *f)
{
    VoxGeom g;
    g.nbr = f->geo.d_nbr;
    g.lines = f->geo.lines;
    g.len = f->geo.n_images;
    g.chunks = (g.len + kWave - 1) / kWave;
    g.stride = f->geo.nstride;
    return g;
}

// True when the voxels of the lines cover the whole span (no padding after
// n_images, no gaps between columns), so that a kernel writing every voxel
// of every line leaves nothing to clear.
bool lines_cover_span(const Geometry &geo)
{
    return geo.nstride == geo.n_images && geo.span() == uint64_t(geo.lines) * geo.nstride;
}

// Grid of a wave_lines kernel whose workgroups end in device atomics on a
// few words: four workgroups per CU (a wave per line, two 16-byte loads in
// flight per lane keep enough bytes moving), lines permitting.
uint32_t line_grid(const sct_ctx *ctx, const VoxGeom &g)
{
    const uint32_t cap = std::max<uint32_t>(ctx->cu_count, 1) * 4;
    return std::max<uint32_t>(1, std::min<uint32_t>((g.lines + 3) / 4, cap));
}

uint32_t stream_grid(const sct_ctx *ctx, const VoxGeom &g)
{
    const uint64_t waves = uint64_t(g.lines) * g.chunks;
    const uint64_t cap = uint64_t(std::max<uint32_t>(ctx->cu_count, 1)) * 16;
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((waves + 3) / 4, cap)));
}

// Calls f(x, line, n) for every voxel of the lines first, first + step, ...
// below end, a line per wave at a time, x[k] the voxel's value in array k.
// Each lane takes 4 consecutive voxels per trip, as one 16-byte load per array
// when the line stride is a multiple of 4 (field data is 256-byte aligned),
// with two trips' loads in flight; other strides load one voxel per lane.
// Called by all lanes of the wave.
template <int NA, typename F>
__device__ __forceinline__ void wave_lines_n(const VoxGeom &g, const float *const (&arr)[NA], uint32_t first,
                                             uint32_t end, uint32_t step, F &&f)
{
    const uint32_t lane = threadIdx.x % kWave;
    if (g.stride % 4 == 0) {
        for (uint32_t line = first; line < end; line += step) {
            const uint64_t lo = uint64_t(line) * g.stride;
            for (uint32_t n0 = lane * 4; n0 < g.len; n0 += 2 * kWave * 4) {
                float v[2][NA][4];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const uint32_t n = n0 + uint32_t(k) * kWave * 4;
#pragma unroll
                    for (int c = 0; c < NA; ++c) {
                        const float *row = arr[c] + lo;
                        if (n + 4 <= g.len) {
                            const float4 q = *reinterpret_cast<const float4 *>(row + n);
                            v[k][c][0] = q.x;
                            v[k][c][1] = q.y;
                            v[k][c][2] = q.z;
                            v[k][c][3] = q.w;
                        } else {
#pragma unroll
                            for (int j = 0; j < 4; ++j) v[k][c][j] = n + j < g.len ? row[n + j] : 0.0f;
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < 2; ++k)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t n = n0 + uint32_t(k) * kWave * 4 + uint32_t(j);
                        float x[NA];
#pragma unroll
                        for (int c = 0; c < NA; ++c) x[c] = v[k][c][j];
                        if (n < g.len) f(x, line, n);
                    }
            }
        }
    } else {
        for (uint32_t line = first; line < end; line += step) {
            const uint64_t lo = uint64_t(line) * g.stride;
            for (uint32_t n = lane; n < g.len; n += kWave) {
                float x[NA];
#pragma unroll
                for (int c = 0; c < NA; ++c) x[c] = arr[c][lo + n];
                f(x, line, n);
            }
        }
    }
}

template <typename F>
__device__ __forceinline__ void wave_lines(const VoxGeom &g, const float *u, uint32_t first, uint32_t end,
                                           uint32_t step, F &&f)
{
    const float *const arr[1] = {u};
    wave_lines_n<1>(g, arr, first, end, step, [&](const float (&x)[1], uint32_t line, uint32_t n) { f(x[0], line, n); });
}

// The min and max keys sit on separate 128-byte lines: device atomics on one
// line are serialised.
constexpr uint32_t kKeyMax = 32;

// Block-wide reduction of (mn, mx) followed by one atomic per block on
// order-preserving keys (min_max.cpp:45-53 uses one device atomic per block).
__device__ void block_minmax_atomic(float mn, float mx, uint32_t *keys)
{
    __shared__ float smn[16], smx[16];
    mn = wave_min(mn);
    mx = wave_max(mx);
    const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
    if ((threadIdx.x % kWave) == 0) {
        smn[w] = mn;
        smx[w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < nw; ++i) {
            mn = fminf(mn, smn[i]);
            mx = fmaxf(mx, smx[i]);
        }
        atomicMin(&keys[0], float_key(mn));
        atomicMax(&keys[kKeyMax], float_key(mx));
    }
}

// min_max_kernel (min_max.cpp:18-56); POSITIVE: grad_min_max_kernel
// (uv-hist-kernel.cpp:31-71), which only considers v > 0.
template <bool POSITIVE>
__global__ __launch_bounds__(256) void minmax_kernel(VoxGeom g, const float *__restrict__ u, uint32_t *keys)
{
    float mn = FLT_MAX, mx = -FLT_MAX;
    const uint32_t nwaves = gridDim.x * (blockDim.x / kWave);
    wave_lines(g, u, blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave, g.lines, nwaves,
               [&](float x, uint32_t, uint32_t) {
                   if (!POSITIVE || 0.0f < x) {
                       mn = (x < mn) ? x : mn;
                       mx = (mx < x) ? x : mx;
                   }
               });
    block_minmax_atomic(mn, mx, keys);
}

// Philox4x32-10 (Salmon et al., SC'11) for the optional histogram dither.
__device__ __forceinline__ uint32_t philox_first(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1)
{
    uint32_t c2 = 0, c3 = 0;
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c0;
}

struct HistArgs {
    float rs, re, delta;
    uint32_t nbins;
    uint32_t k0, k1;        // dither key
    uint32_t lines_per_block;
    uint32_t fast;          // hist_frac may take the short division (re - rs in range)
};

// re - rs in [2^-32, 2^32] (as the device computes it): the range of
// denominators for which hist_frac's short division is exact.
inline uint32_t short_div_ok(float rs, float re)
{
    const float d = re - rs;
    return d >= 0x1p-32f && d <= 0x1p32f;
}

// t = a / b (a = x - range start, b = range length) for hist_bin.  With b in
// [2^-32, 2^32] (fast) and |a| >= b 2^-40 the compiler's IEEE division neither
// scales its operands (v_div_scale) nor fixes up its result (v_div_fixup), so
// its core -- reciprocal, a Newton step, the quotient and two residual
// corrections, div_cr_normal -- is the correctly rounded quotient on its own;
// for smaller |a| both give |t| < 2^-39 (or a t of a's sign), which bins to 0
// whatever its last bits.  x is in the range, so a <= b.
__device__ __forceinline__ float hist_frac(float a, float b, bool fast)
{
    return fast ? sct::div_cr_normal(a, b) : a / b;
}

// bin = clamp(int(round(nbins * t - 0.5f)), 0, nbins - 1), t = (x - rs) / (re - rs)
// (calc_hist.cpp:61-63); nbins * t - 0.5f is one fused multiply-add.  For x
// in the range (callers test it first) t <= 1, so y = nbins t - 0.5 < nbins:
// the clamp at 0 goes first, y' = max(y, 0) (y in [-0.5, 0) rounds to -1 or
// 0, both clamped to 0; NaN, from rs = re, gives 0 as its conversion would),
// and round half away from zero of y' >= 0 is trunc(y') + (fract(y') >= 0.5),
// fract(y') = y' - floor(y') being exact.
__device__ __forceinline__ int32_t hist_bin(float x, float rs, float re, uint32_t nbins, bool fast)
{
    const float t = hist_frac(x - rs, re - rs, fast);
    const float y = fmaxf(__fmaf_rn(float(nbins), t, -0.5f), 0.0f);
    const int32_t b = int32_t(y) + int32_t(__builtin_amdgcn_fractf(y) >= 0.5f);
    return min(int32_t(nbins) - 1, b);
}

// calc_hist_kernel (calc_hist.cpp:23-80).  MODE 0: u32 counters in LDS
// (up to kSmallBins bins); MODE 2: device atomics only (bin counts too large
// for LDS).  Larger histograms up to kPartBins * kMaxParts bins use
// hist_part_kernel below.
template <int MODE, bool DITHER>
__global__ __launch_bounds__(256) void hist_kernel(VoxGeom g, const float *__restrict__ u, HistArgs a,
                                                   uint32_t *__restrict__ counts)
{
    extern __shared__ uint32_t sh[];
    const uint32_t nwords = MODE == 0 ? a.nbins : 0;
    if constexpr (MODE != 2) {
        for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) sh[i] = 0;
        __syncthreads();
    }
    const uint32_t l0 = blockIdx.x * a.lines_per_block;
    const uint32_t l1 = min(g.lines, l0 + a.lines_per_block);
    const uint32_t nw = blockDim.x / kWave;
    const float half = a.delta / 2;
    wave_lines(g, u, l0 + threadIdx.x / kWave, l1, nw, [&](float x, uint32_t line, uint32_t n) {
        if constexpr (DITHER) {
            const uint32_t r = philox_first(line, n, a.k0, a.k1);
            x = x + (-half + a.delta * (float(r >> 8) * (1.0f / 16777216.0f)));
        }
        if (a.rs - half <= x && x + half <= a.re) {
            const int32_t b = hist_bin(x, a.rs, a.re, a.nbins, a.fast);
            if constexpr (MODE == 0) atomicAdd(&sh[b], 1u);
            else atomicAdd(&counts[b], 1u);
        }
    });
    if constexpr (MODE != 2) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) {
            const uint32_t c = sh[i];
            if (c == 0) continue;
            atomicAdd(&counts[i], c);
        }
    }
}

// ---- wide histograms: 16-bit counters in LDS ---------------------------------
// Histograms of more than kSmallBins bins cut the bins into partitions of up
// to kWideBins counters and the lines into slices.  Workgroup (slice,
// partition) counts the voxels of its slice that fall in its partition with
// 16-bit counters in LDS, two per 32-bit word (65536 bins in 128 KB, so a
// histogram of up to 65536 bins reads every voxel once), and stores them as
// the (slice, partition) block of `rows`; wide_sum_kernel adds the blocks.
// 1024-thread workgroups, one per CU (the LDS): 16 waves.
//
// Counters stay within 16 bits by settling: the adds (LDS atomics whose
// result is not read) run in periods of at most 32 voxels per thread, 32768
// adds per workgroup; between two barriers after each period, settle() takes
// kSpill = 2^15 off every counter at kSpill or more and appends its bin to the
// workgroup's spill list, an entry per kSpill counts.  Every counter starts a
// period below 2^15 and gains at most 32768 in it, so it never passes 65535
// and never carries into its neighbour.  An entry stands for kSpill adds, so
// a slice of at most kWideSliceVoxels voxel slots fills at most kSpillCap.
constexpr uint32_t kSmallBins = 8192;
constexpr uint32_t kWideBins = 65536;                  // counters per partition
constexpr uint32_t kWideWords = kWideBins / 2;         // their LDS words
constexpr uint32_t kSpill = 32768;
constexpr uint32_t kSpillCap = 256;
constexpr uint64_t kWideSliceVoxels = uint64_t(kSpillCap) * kSpill;   // 2^23
constexpr uint32_t kPeriodVoxels = 32;                 // per thread between settles
// A (slice, partition) block of rows: the counter words, the spill count and
// the spill bins (a multiple of 256 bytes, so every block is 16-byte aligned)
constexpr uint32_t kWideRowWords = kWideWords + 512;
constexpr uint32_t kMaxParts = 8;
constexpr int kPartThreads = 1024;
static_assert(kWideWords + 1 + kSpillCap <= kWideRowWords, "spill list fits the block");
static_assert(uint64_t(kPartThreads) * kPeriodVoxels <= 65536 - kSpill, "a period keeps counters in 16 bits");

// Counter words of nb counters, a multiple of 4 (settle reads 16 bytes).
__host__ __device__ inline uint32_t wide_words(uint32_t nb) { return ((nb + 1) / 2 + 3) & ~3u; }

// LDS words of a wide kernel: the counter words, the spill list, its count
// (132 KB of the CU's 160 KB).
constexpr uint32_t kWideLdsWords = kWideWords + kSpillCap + 1;

struct WideHist {
    uint32_t *w, *list, *n;
    uint32_t nw;
    __device__ WideHist(uint32_t *sh, uint32_t nb)
        : w(sh), list(sh + wide_words(nb)), n(sh + wide_words(nb) + kSpillCap), nw(wide_words(nb)) {}
    // Followed by a barrier before the first add.
    __device__ void clear() const
    {
        for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) w[i] = 0;
        if (threadIdx.x == 0) *n = 0;
    }
    __device__ __forceinline__ void add(uint32_t b) const { atomicAdd(&w[b >> 1], 1u << ((b & 1u) << 4)); }
    // Called by every thread of the workgroup after each period.
    __device__ void settle() const
    {
        __syncthreads();
        for (uint32_t i = 4 * threadIdx.x; i < nw; i += 4 * blockDim.x) {
            const uint4 q = *reinterpret_cast<const uint4 *>(w + i);
            if (!((q.x | q.y | q.z | q.w) & 0x80008000u)) continue;
            const uint32_t qw[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t x = qw[k];
#pragma unroll
                for (uint32_t half = 0; half < 2; ++half)
                    if ((x >> (16 * half)) & kSpill) {
                        x -= kSpill << (16 * half);
                        list[atomicAdd(n, 1u)] = 2 * (i + k) + half;
                    }
                w[i + k] = x;
            }
        }
        __syncthreads();
    }
    // After a barrier that follows the last add: the block of rows.
    __device__ void store(uint32_t *blk) const
    {
        for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) blk[i] = w[i];
        const uint32_t ns = *n;
        if (threadIdx.x == 0) blk[kWideWords] = ns;
        for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) blk[kWideWords + 1 + i] = list[i];
    }
};

// Workgroups of a wide kernel: (slice, partition) pairs, the partitions of a
// slice on one XCD (workgroups go to the 8 XCDs round-robin) so that its
// voxels come from that XCD's L2 after the first partition reads them.
inline uint32_t wide_grid(uint32_t used, uint32_t parts)
{
    return parts == 1 ? used : ((used + 7) / 8) * 8 * parts;
}

__device__ __forceinline__ bool wide_block(uint32_t parts, uint32_t used, uint32_t &slice, uint32_t &part)
{
    const uint32_t b = blockIdx.x;
    if (parts == 1) {
        slice = b;
        part = 0;
    } else {
        part = (b / 8) % parts;
        slice = b % 8 + 8 * (b / (8 * parts));
    }
    return slice < used;
}

constexpr uint32_t kNoBin = ~0u;

// Counts the voxels of lines [l0, l1) into h: bin_of(x, line, n) is a voxel's
// counter in h (kNoBin for none), x[k] its value in array k.  The lines are cut
// into chunks of 64 x VEC voxels (VEC consecutive elements per lane: one
// 16-byte load per array when VEC = 16 / sizeof(T), else one element); every
// round each wave takes CPR chunks (four, two for two arrays), issues the
// next round's loads and then counts this round's voxels; settle() follows every kPeriodVoxels voxels per
// lane.  All waves run the same rounds, those past the last chunk on empty
// ones (a load clamped to the slice, no voxel counted).  Lanes past a line's
// end load its first elements.  Called by all kPartThreads threads.
template <uint32_t VEC, typename T, int NA, typename F>
__device__ __forceinline__ void count_lines(const WideHist &h, const VoxGeom &g, const T *const (&arr)[NA],
                                            uint32_t l0, uint32_t l1, F &&bin_of)
{
    constexpr uint32_t CH = kWave * VEC, NW = kPartThreads / kWave;
    constexpr uint32_t CPR = NA == 1 ? 4 : 2;   // chunks per wave and round (registers of two arrays)
    constexpr uint32_t kRoundsPerSettle = kPeriodVoxels / (CPR * VEC);
    static_assert(kRoundsPerSettle >= 1 && kPeriodVoxels % (CPR * VEC) == 0, "whole rounds per period");
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t cpl = (g.len + CH - 1) / CH;   // chunks per line
    const uint32_t total = l0 < l1 ? (l1 - l0) * cpl : 0;
    const uint32_t rounds = (total + CPR * NW - 1) / (CPR * NW);
    if (!rounds) return;   // uniform over the workgroup
    struct Batch {
        T v[CPR][NA][VEC];
        uint32_t line[CPR], n[CPR];   // n <= len; n = len: nothing to count
    };
    auto load = [&](uint32_t r, Batch &bt) {
#pragma unroll
        for (uint32_t k = 0; k < CPR; ++k) {
            const uint32_t c = (CPR * r + k) * NW + w;
            const uint32_t ce = min(c, total - 1);
            const uint32_t li = ce / cpl;
            const uint32_t nn = (ce - li * cpl) * CH + lane * VEC;
            bt.line[k] = l0 + li;
            bt.n[k] = c < total && nn < g.len ? nn : g.len;
            const uint64_t o = uint64_t(bt.line[k]) * g.stride + (nn < g.len ? nn : 0);
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                if constexpr (VEC * sizeof(T) == 16) {
                    const uint4 q = *reinterpret_cast<const uint4 *>(arr[a] + o);
                    __builtin_memcpy(bt.v[k][a], &q, 16);
                } else {
#pragma unroll
                    for (uint32_t j = 0; j < VEC; ++j) bt.v[k][a][j] = nn + j < g.len ? arr[a][o + j] : T(0);
                }
            }
        }
    };
    auto count = [&](const Batch &bt, auto full) {
#pragma unroll
        for (uint32_t k = 0; k < CPR; ++k)
#pragma unroll
            for (uint32_t j = 0; j < VEC; ++j) {
                if (!full && j >= g.len - bt.n[k]) continue;   // n <= len
                T x[NA];
#pragma unroll
                for (int a = 0; a < NA; ++a) x[a] = bt.v[k][a][j];
                const uint32_t b = bin_of(x, bt.line[k], bt.n[k] + j);
                if (b != kNoBin) h.add(b);
            }
    };
    Batch cur;
    load(0, cur);
    for (uint32_t r = 0; r < rounds; ++r) {
        Batch nxt;
        load(r + 1, nxt);   // past the last round: clamped, never counted
        bool full = true;
#pragma unroll
        for (uint32_t k = 0; k < CPR; ++k) full = full && cur.n[k] + VEC <= g.len;
        if (__all(full))   // every voxel of the wave's chunks is in its line
            count(cur, std::true_type());
        else
            count(cur, std::false_type());
        if ((r + 1) % kRoundsPerSettle == 0) h.settle();
        cur = nxt;
    }
}

// count_lines with 16-byte loads when the line stride allows them.
template <typename T, int NA, typename F>
__device__ __forceinline__ void count_slice(const WideHist &h, const VoxGeom &g, const T *const (&arr)[NA],
                                            uint32_t l0, uint32_t l1, F &&bin_of)
{
    constexpr uint32_t VEC = 16 / sizeof(T);
    if (g.stride % VEC == 0)
        count_lines<VEC>(h, g, arr, l0, l1, bin_of);
    else
        count_lines<kPeriodVoxels / 4>(h, g, arr, l0, l1, bin_of);
}

// ONE: a single partition (nbins <= kWideBins), whose bins need no test.
template <bool DITHER, bool ONE>
__global__ __launch_bounds__(kPartThreads) void hist_wide_kernel(VoxGeom g, const float *__restrict__ u, HistArgs a,
                                                                 uint32_t parts, uint32_t used,
                                                                 uint32_t *__restrict__ rows)
{
    uint32_t slice, part;
    if (!wide_block(parts, used, slice, part)) return;
    __shared__ uint32_t sh[kWideLdsWords];
    const uint32_t b0 = part * kWideBins, nb = min(kWideBins, a.nbins - b0);
    const WideHist h(sh, nb);
    h.clear();
    __syncthreads();
    const uint32_t l0 = slice * a.lines_per_block;
    const uint32_t l1 = min(g.lines, l0 + a.lines_per_block);
    const float half = a.delta / 2;
    const float *const arr[1] = {u};
    count_slice<float, 1>(h, g, arr, l0, l1, [&](const float (&x)[1], uint32_t line, uint32_t n) -> uint32_t {
        float v = x[0];
        if constexpr (DITHER) {
            const uint32_t r = philox_first(line, n, a.k0, a.k1);
            v = v + (-half + a.delta * (float(r >> 8) * (1.0f / 16777216.0f)));
            if (!(a.rs - half <= v && v + half <= a.re)) return kNoBin;
        } else {
            if (!(a.rs <= v && v <= a.re)) return kNoBin;   // delta = 0: v + 0 compares as v
        }
        const uint32_t b = uint32_t(hist_bin(v, a.rs, a.re, a.nbins, a.fast)) - b0;
        if constexpr (ONE) return b;
        return b < nb ? b : kNoBin;
    });
    __syncthreads();
    h.store(rows + (uint64_t(slice) * parts + part) * kWideRowWords);
}

// Sums the blocks of `rows` of `slices` slices into out[nbins]: workgroup b
// takes bins [512 b, 512 b + 512) (one partition), lane l of wave w their
// 8 bins from 8 l on (one 16-byte load per block) of slices w, w + 4, ...;
// then the spill entries of every slice that fall in the range.
template <typename T>
__global__ __launch_bounds__(256) void wide_sum_kernel(const uint32_t *__restrict__ rows, uint32_t slices,
                                                       uint32_t parts, uint32_t nbins, T *__restrict__ out)
{
    __shared__ T acc[4][512];
    __shared__ T tot[512];
    const uint32_t part = blockIdx.x / (kWideBins / 512), c0 = (blockIdx.x % (kWideBins / 512)) * 512;
    const uint32_t lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    T s[8] = {};
#pragma unroll 4
    for (uint32_t k = w; k < slices; k += 4) {
        const uint4 q = *reinterpret_cast<const uint4 *>(rows + (uint64_t(k) * parts + part) * kWideRowWords + c0 / 2 +
                                                         4 * lane);
        const uint32_t qw[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            s[2 * j] += T(qw[j] & 0xffffu);
            s[2 * j + 1] += T(qw[j] >> 16);
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[w][8 * lane + j] = s[j];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 512; i += blockDim.x) tot[i] = acc[0][i] + acc[1][i] + acc[2][i] + acc[3][i];
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < slices; k += blockDim.x) {
        const uint32_t *blk = rows + (uint64_t(k) * parts + part) * kWideRowWords;
        const uint32_t ns = blk[kWideWords];
        for (uint32_t e = 0; e < ns; ++e) {
            const uint32_t b = blk[kWideWords + 1 + e] - c0;
            if (b < 512) atomicAdd(&tot[b], T(kSpill));
        }
    }
    __syncthreads();
    const uint32_t gb = part * kWideBins + c0;
    for (uint32_t i = threadIdx.x; i < 512 && gb + i < nbins; i += blockDim.x) out[gb + i] = tot[i];
}

// Slices of lines for the wide kernels: a workgroup per CU over all
// partitions where the lines allow, at most kWideSliceVoxels voxel slots per
// slice, and the blocks of rows within max_words words.  ok is false when
// these cannot all hold (a line longer than kWideSliceVoxels, more than
// kMaxParts partitions, or too little room for the rows).
struct WidePlan {
    bool ok = false;
    uint32_t parts = 0, lps = 0, used = 0;   // partitions, lines per slice, slices
    uint64_t words = 0;                      // of rows
};

WidePlan wide_plan(const sct_ctx *ctx, const VoxGeom &g, uint64_t nbins, uint64_t max_words)
{
    WidePlan p;
    p.parts = uint32_t((nbins + kWideBins - 1) / kWideBins);
    if (!g.lines || !g.len || g.len > kWideSliceVoxels || p.parts > kMaxParts) return p;
    const uint32_t lps_cap = uint32_t(kWideSliceVoxels / g.len);
    uint64_t slices = std::max<uint32_t>(1, std::min<uint32_t>(g.lines, std::max<uint32_t>(ctx->cu_count, 1) / p.parts));
    slices = std::min<uint64_t>(slices, max_words / (uint64_t(p.parts) * kWideRowWords));
    if (!slices) return p;
    p.lps = std::min<uint32_t>(uint32_t((g.lines + slices - 1) / slices), lps_cap);
    p.used = (g.lines + p.lps - 1) / p.lps;
    p.words = uint64_t(p.used) * p.parts * kWideRowWords;
    p.ok = p.words <= max_words;
    return p;
}

// grad_kernel of grad_magn.cpp:17-58 for one voxel: x its value's
// neighbours along n (xm, xp), along rows (rm, rp) and columns (cm, cp).
__device__ __forceinline__ float grad_value(bool inner, float xm, float xp, float rm, float rp, float cm, float cp)
{
    if (!inner) return -1.0f;
    const float q1 = xp - xm, q2 = rp - rm, q3 = cp - cm;
    return 0.5f * sqrtf(__fmaf_rn(q3, q3, __fmaf_rn(q2, q2, q1 * q1)));
}

// A line per wave; each lane takes 4 consecutive voxels per trip, as 16-byte
// loads of the line and its four neighbour lines (stride a multiple of 4);
// the values at n - 1 and n + 4 come from the neighbouring lanes.  Voxels
// without all six neighbours get -1.
__global__ __launch_bounds__(256) void grad_magn_kernel(VoxGeom g, const float *__restrict__ u, float *__restrict__ v)
{
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t nwaves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t line = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; line < g.lines; line += nwaves) {
        const LineNbr nb = g.nbr[line];
        const bool inner_line = nb.nrow >= 0 && nb.ncol >= 0 && nb.prow >= 0 && nb.pcol >= 0;
        const uint64_t o = uint64_t(line) * g.stride;
        const float *r = u + o;
        if (g.stride % 4 != 0) {
            for (uint32_t n = lane; n < g.len; n += kWave) {
                const bool inner = inner_line && n > 0 && n + 1 < g.len;
                v[o + n] = inner ? grad_value(true, r[n - 1], r[n + 1], u[uint64_t(nb.prow) * g.stride + n],
                                              u[uint64_t(nb.nrow) * g.stride + n], u[uint64_t(nb.pcol) * g.stride + n],
                                              u[uint64_t(nb.ncol) * g.stride + n])
                                 : -1.0f;
            }
            continue;
        }
        const float *nl[4] = {u + uint64_t(max(nb.prow, 0)) * g.stride, u + uint64_t(max(nb.nrow, 0)) * g.stride,
                              u + uint64_t(max(nb.pcol, 0)) * g.stride, u + uint64_t(max(nb.ncol, 0)) * g.stride};
        // trips are uniform over the wave (the shuffles need every lane)
        for (uint32_t base = 0; base < g.len; base += kWave * 4) {
            const uint32_t n0 = base + lane * 4;
            const bool full = n0 + 4 <= g.len;
            auto load4 = [&](const float *row, float (&x)[4]) {
                if (full) {
                    const float4 q = *reinterpret_cast<const float4 *>(row + n0);
                    x[0] = q.x;
                    x[1] = q.y;
                    x[2] = q.z;
                    x[3] = q.w;
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j) x[j] = n0 + j < g.len ? row[n0 + j] : 0.0f;
                }
            };
            float c[4], nbv[4][4] = {};
            load4(r, c);
            if (inner_line)
#pragma unroll
                for (int k = 0; k < 4; ++k) load4(nl[k], nbv[k]);
            float left = __shfl_up(c[3], 1, kWave), right = __shfl_down(c[0], 1, kWave);
            if (lane == 0) left = n0 > 0 && n0 - 1 < g.len ? r[n0 - 1] : 0.0f;
            if (lane == kWave - 1) right = n0 + 4 < g.len ? r[n0 + 4] : 0.0f;
            float out[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t n = n0 + uint32_t(j);
                const bool inner = inner_line && n > 0 && n + 1 < g.len;
                const float xm = j == 0 ? left : c[j - 1], xp = j == 3 ? right : c[j + 1];
                out[j] = grad_value(inner, xm, xp, nbv[0][j], nbv[1][j], nbv[2][j], nbv[3][j]);
            }
            if (full) {
                *reinterpret_cast<float4 *>(v + o + n0) = make_float4(out[0], out[1], out[2], out[3]);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (n0 + j < g.len) v[o + n0 + j] = out[j];
            }
        }
    }
}

struct UvArgs {
    float u0, u1, v0, v1;
    uint32_t nbins, hstride;
    uint32_t fast_u, fast_v;   // short_div_ok of the two ranges
};

// counts[bin] += 1 for every lane of the wave with bin >= 0.  The lanes that
// share the leader's bin are added with one device atomic, for up to
// kAggRounds distinct bins; the rest add one each.  Neighbouring voxels
// mostly fall in the same few bins (intensity phases, small gradients), so
// this removes most of the contention on the hot counters.  Called by all
// lanes of the wave.
constexpr int kAggRounds = 8;

__device__ __forceinline__ void wave_count(uint32_t *counts, long long bin)
{
    const uint32_t lane = threadIdx.x % kWave;
    bool active = bin >= 0;
    for (int r = 0; r < kAggRounds; ++r) {
        const uint64_t act = __ballot(active);
        if (!act) return;
        const int leader = __ffsll((unsigned long long)act) - 1;
        const long long key = __shfl(bin, leader, kWave);
        const bool mine = active && bin == key;
        const uint64_t same = __ballot(mine);
        if (int(lane) == leader) atomicAdd(&counts[key], uint32_t(__popcll(same)));
        if (mine) active = false;
    }
    if (active) atomicAdd(&counts[bin], 1u);
}

// The bin-id functor and histogram of uv-hist-kernel.cpp:121-185.
__global__ __launch_bounds__(256) void uv_hist_kernel(VoxGeom g, const float *__restrict__ u,
                                                      const float *__restrict__ v, UvArgs a, uint32_t *counts)
{
    const uint64_t total = uint64_t(g.lines) * g.chunks;
    const uint64_t nwaves = uint64_t(gridDim.x) * (blockDim.x / kWave);
    const uint32_t lane = threadIdx.x % kWave;
    for (uint64_t w = uint64_t(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; w < total; w += nwaves) {
        const uint32_t line = uint32_t(w / g.chunks);
        const uint32_t n = uint32_t(w % g.chunks) * kWave + lane;
        long long bin = -1;
        if (n < g.len) {
            const uint64_t o = uint64_t(line) * g.stride + n;
            const float x = u[o], y = v[o];
            if (a.u0 <= x && x <= a.u1 && a.v0 <= y && y <= a.v1) {
                const int32_t xb = hist_bin(x, a.u0, a.u1, a.nbins, a.fast_u);
                const int32_t yb = hist_bin(y, a.v0, a.v1, a.nbins, a.fast_v);
                bin = (long long)yb * a.hstride + xb;
            }
        }
        wave_count(counts, bin);
    }
}

// uv_hist with up to kWideBins * kMaxParts counters: workgroup (slice,
// partition) counts the slice's voxels whose (v, u) bin falls in its 65536
// 16-bit counters of LDS and stores them as its (slice, partition) block of
// rows; wide_sum_kernel adds the blocks.  Hot bins (the phases' intensities
// at small gradients) then cost LDS atomics within a workgroup instead of
// device atomics from every CU on one address.
// The joint bin of every voxel (~0u outside the ranges), for the partitioned
// count when more than kBinPassParts partitions would each recompute it.
constexpr uint32_t kBinPassParts = 2;

__global__ __launch_bounds__(256) void uv_bin_kernel(VoxGeom g, const float *__restrict__ u,
                                                     const float *__restrict__ v, UvArgs a, uint32_t *__restrict__ bins)
{
    const float *const arr[2] = {u, v};
    wave_lines_n<2>(g, arr, blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave, g.lines,
                    gridDim.x * (blockDim.x / kWave), [&](const float (&xy)[2], uint32_t line, uint32_t n) {
                        const float x = xy[0], y = xy[1];
                        uint32_t b = ~0u;
                        if (a.u0 <= x && x <= a.u1 && a.v0 <= y && y <= a.v1)
                            b = uint32_t(hist_bin(y, a.v0, a.v1, a.nbins, a.fast_v)) * a.hstride +
                                uint32_t(hist_bin(x, a.u0, a.u1, a.nbins, a.fast_u));
                        bins[uint64_t(line) * g.stride + n] = b;
                    });
}

// BINS: bins holds uv_bin_kernel's bins and u, v are unused; otherwise the
// workgroup computes the bins from u and v.
template <bool BINS>
__global__ __launch_bounds__(kPartThreads) void uv_wide_kernel(VoxGeom g, const float *__restrict__ u,
                                                               const float *__restrict__ v,
                                                               const uint32_t *__restrict__ bins, UvArgs a,
                                                               uint32_t parts, uint32_t used, uint32_t lines_per_slice,
                                                               uint32_t total, uint32_t *__restrict__ rows)
{
    uint32_t slice, part;
    if (!wide_block(parts, used, slice, part)) return;
    __shared__ uint32_t sh[kWideLdsWords];
    const uint32_t b0 = part * kWideBins, nb = min(kWideBins, total - b0);
    const WideHist h(sh, nb);
    h.clear();
    __syncthreads();
    const uint32_t l0 = slice * lines_per_slice;
    const uint32_t l1 = min(g.lines, l0 + lines_per_slice);
    if constexpr (BINS) {
        const uint32_t *const arr[1] = {bins};
        count_slice<uint32_t, 1>(h, g, arr, l0, l1, [&](const uint32_t (&x)[1], uint32_t, uint32_t) -> uint32_t {
            const uint32_t b = x[0] - b0;   // ~0u - b0 >= nb
            return b < nb ? b : kNoBin;
        });
    } else {
        const float *const arr[2] = {u, v};
        count_slice<float, 2>(h, g, arr, l0, l1, [&](const float (&xy)[2], uint32_t, uint32_t) -> uint32_t {
            const float x = xy[0], y = xy[1];
            if (!(a.u0 <= x && x <= a.u1 && a.v0 <= y && y <= a.v1)) return kNoBin;
            const uint32_t xb = uint32_t(hist_bin(x, a.u0, a.u1, a.nbins, a.fast_u));
            const uint32_t yb = uint32_t(hist_bin(y, a.v0, a.v1, a.nbins, a.fast_v));
            const uint32_t b = yb * a.hstride + xb - b0;
            return b < nb ? b : kNoBin;
        });
    }
    __syncthreads();
    h.store(rows + (uint64_t(slice) * parts + part) * kWideRowWords);
}

int run_minmax(sct_ctx *ctx, const sct_field *f, bool positive, float *mn, float *mx)
{
    if (!ctx || !f || !mn || !mx) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (f->dtype != SCT_F32) return sct_fail(ctx, SCT_ERR_INVALID, "float field required");
    const VoxGeom g = vox_geom(f);
    uint32_t *keys = nullptr, *h = nullptr;
    constexpr size_t kKeyBytes = sizeof(uint32_t) * (kKeyMax + 1);
    int rc = sct_ctx_scratch(ctx, kKeyBytes, (void **)&keys);
    if (!rc) rc = sct_ctx_host_scratch(ctx, kKeyBytes, (void **)&h);
    if (rc) return rc;
    {
        // keys start at (FLT_MAX, lowest), the reference's initial values (min_max.cpp:61-65)
        const float init[2] = {FLT_MAX, -FLT_MAX};
        for (int i = 0; i < 2; ++i) {
            uint32_t b;
            std::memcpy(&b, &init[i], 4);
            h[i * kKeyMax] = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
        }
    }
    hipError_t e = hipMemcpyAsync(keys, h, kKeyBytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && g.lines && g.len) {
        if (positive)
            minmax_kernel<true><<<line_grid(ctx, g), 256, 0, ctx->stream>>>(g, (const float *)f->data, keys);
        else
            minmax_kernel<false><<<line_grid(ctx, g), 256, 0, ctx->stream>>>(g, (const float *)f->data, keys);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, keys, kKeyBytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) (void)hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "min_max");
    *mn = key_float(h[0]);
    *mx = key_float(h[kKeyMax]);
    return SCT_OK;
}

}  // namespace

extern "C" int sct_min_max(sct_ctx *ctx, const sct_field *u, float *mn, float *mx)
{
    return run_minmax(ctx, u, false, mn, mx);
}

extern "C" int sct_grad_min_max(sct_ctx *ctx, const sct_field *v, float *mn, float *mx)
{
    return run_minmax(ctx, v, true, mn, mx);
}

namespace {

__global__ __launch_bounds__(256) void fill_zero_kernel(uint32_t *__restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// How calc_hist counts nbins bins of g: mode 0 (LDS counters, device atomic
// flush), 3 (16-bit LDS counters by partitions of kWideBins bins, a block of
// rows per slice and partition, then wide_sum_kernel) or 2 (device atomics,
// when mode 3's plan does not hold).  The rows of mode 3 take at most
// max_partial words.
struct HistPlan {
    int mode = 0;
    uint32_t parts = 1, used = 0;   // partitions, slices
    uint64_t partial_words = 0;
};

HistPlan hist_plan(const sct_ctx *ctx, const VoxGeom &g, HistArgs &a, uint64_t max_partial)
{
    HistPlan p;
    p.mode = a.nbins <= kSmallBins ? 0 : 3;
    if (!g.lines || !g.len) return p;
    if (p.mode == 3) {
        const WidePlan w = wide_plan(ctx, g, a.nbins, max_partial);
        if (w.ok) {
            a.lines_per_block = w.lps;
            p.parts = w.parts;
            p.used = w.used;
            p.partial_words = w.words;
            return p;
        }
        p.mode = 2;
    }
    {
        const uint32_t blocks = std::max<uint32_t>(1, ctx->cu_count * 4);
        a.lines_per_block = std::max<uint32_t>(1, (g.lines + blocks - 1) / blocks);
    }
    return p;
}

// The kernels of calc_hist on the context's stream: counts d[nbins] (cleared
// here by a kernel), partial rows in partial (mode 3).
hipError_t hist_passes(sct_ctx *ctx, const sct_field *f, const VoxGeom &g, const HistArgs &a, const HistPlan &p,
                       uint32_t *d, uint32_t *partial)
{
    const bool dither = a.delta != 0.0f;
    const uint32_t nbins = a.nbins;
    if (p.mode != 3 || !g.lines || !g.len) {
        fill_zero_kernel<<<(nbins + 255) / 256, 256, 0, ctx->stream>>>(d, nbins);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || !g.lines || !g.len) return e;
    }
    if (p.mode == 3) {
        auto launch = [&](auto kernel) -> hipError_t {
            kernel<<<wide_grid(p.used, p.parts), kPartThreads, 0, ctx->stream>>>(g, (const float *)f->data, a,
                                                                                 p.parts, p.used, partial);
            return hipGetLastError();
        };
        hipError_t e = p.parts == 1 ? (dither ? launch(hist_wide_kernel<true, true>) : launch(hist_wide_kernel<false, true>))
                                    : (dither ? launch(hist_wide_kernel<true, false>)
                                              : launch(hist_wide_kernel<false, false>));
        if (e != hipSuccess) return e;
        wide_sum_kernel<uint32_t><<<(nbins + 511) / 512, 256, 0, ctx->stream>>>(partial, p.used, p.parts, nbins, d);
        return hipGetLastError();
    }
    const size_t lds = p.mode == 0 ? sizeof(uint32_t) * nbins : 0;
    const uint32_t grid = std::max<uint32_t>(1, (g.lines + a.lines_per_block - 1) / a.lines_per_block);
    if (p.mode == 0) {
        if (dither) hist_kernel<0, true><<<grid, 256, lds, ctx->stream>>>(g, (const float *)f->data, a, d);
        else hist_kernel<0, false><<<grid, 256, lds, ctx->stream>>>(g, (const float *)f->data, a, d);
    } else {
        if (dither) hist_kernel<2, true><<<grid, 256, 0, ctx->stream>>>(g, (const float *)f->data, a, d);
        else hist_kernel<2, false><<<grid, 256, 0, ctx->stream>>>(g, (const float *)f->data, a, d);
    }
    return hipGetLastError();
}

int hist_check(sct_ctx *ctx, const sct_field *f, uint32_t nbins, const uint32_t *counts)
{
    if (!ctx || !f || !counts) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (f->dtype != SCT_F32) return sct_fail(ctx, SCT_ERR_INVALID, "float field required");
    if (nbins == 0) return sct_fail(ctx, SCT_ERR_INVALID, "nbins must be positive");
    return SCT_OK;
}

HistArgs hist_args(float rs, float re, uint32_t nbins, float delta, uint64_t seed)
{
    HistArgs a;
    a.rs = rs;
    a.re = re;
    a.delta = delta;
    a.nbins = nbins;
    a.k0 = uint32_t(seed);
    a.k1 = uint32_t(seed >> 32);
    a.lines_per_block = 1;
    a.fast = short_div_ok(rs, re);
    return a;
}

// density = count / (data_size * bin_width), in FP64 (calc_hist.cpp:131-137).
void hist_density(const sct_field *f, float rs, float re, uint32_t nbins, const uint32_t *counts, double *density)
{
    if (!density) return;
    const double bin_width = (double(re) - double(rs)) / nbins;
    const double scale = 1.0 / (double(f->geo.data_size) * bin_width);
    for (uint32_t i = 0; i < nbins; ++i) density[i] = counts[i] * scale;
}

}  // namespace

// calc_hist(ctx, in, range_start, range_end, nbins, delta), calc_hist.cpp:94-138.
extern "C" int sct_calc_hist(sct_ctx *ctx, const sct_field *f, float rs, float re, uint32_t nbins, float delta,
                             uint64_t seed, uint32_t *counts, double *density)
{
    int rc = hist_check(ctx, f, nbins, counts);
    if (rc) return rc;
    const VoxGeom g = vox_geom(f);
    HistArgs a = hist_args(rs, re, nbins, delta, seed);
    const HistPlan p = hist_plan(ctx, g, a, ~uint64_t(0));
    // counts d, then (partitioned mode) a row of partial counts per slice of
    // lines, in the context's scratch
    const size_t dbytes = (sizeof(uint32_t) * size_t(nbins) + 255) & ~size_t(255);
    void *mem = nullptr, *host = nullptr;
    rc = sct_ctx_scratch(ctx, dbytes + sizeof(uint32_t) * p.partial_words, &mem);
    if (!rc) rc = sct_ctx_host_scratch(ctx, sizeof(uint32_t) * size_t(nbins), &host);
    if (rc) return rc;
    uint32_t *d = (uint32_t *)mem, *partial = p.used ? (uint32_t *)((char *)mem + dbytes) : nullptr;
    hipError_t e = hist_passes(ctx, f, g, a, p, d, partial);
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, sizeof(uint32_t) * nbins, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) (void)hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "calc_hist");
    std::memcpy(counts, host, sizeof(uint32_t) * nbins);
    hist_density(f, rs, re, nbins, counts, density);
    return SCT_OK;
}

// ---- calc_hist with no allocation ------------------------------------------
// The device temporaries of sct_calc_hist_noalloc live in one module-scope
// array per GPU, `scratch`: the counts (SCT_NOALLOC_MAX_BINS words) and then
// the blocks of rows of the partitioned mode (kNoAllocPartial words).  Each
// GPU holds its own copy at its own address, found once per device with
// hipGetSymbolAddress when a context is created on it; host code uses that
// address, never the array's name.
constexpr uint64_t kNoAllocPartial = uint64_t(4) << 20;   // 16 MiB of rows
static_assert(uint64_t(SCT_NOALLOC_MAX_BINS + kWideBins - 1) / kWideBins * kWideRowWords <= kNoAllocPartial,
              "the blocks of one slice must fit");
static __device__ __align__(256) uint32_t scratch[SCT_NOALLOC_MAX_BINS + kNoAllocPartial];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];
}  // namespace

int sct_noalloc_scratch_setup(sct_ctx *ctx)
{
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return sct_fail(ctx, SCT_ERR_INVALID, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void *ptr = nullptr;
        // the context's device is current (sct_ctx_create set it)
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return sct_hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                         "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[ctx->device] = ptr;
    }
    ctx->noalloc_scratch = g_scratch_ptr[ctx->device];
    return SCT_OK;
}

// sct_calc_hist with its device temporaries in the module's fixed scratch:
// no hipMalloc, hipFree or pinned staging (the counts are copied straight
// into the caller's array).
extern "C" int sct_calc_hist_noalloc(sct_ctx *ctx, const sct_field *f, float rs, float re, uint32_t nbins,
                                     float delta, uint64_t seed, uint32_t *counts, double *density)
{
    int rc = hist_check(ctx, f, nbins, counts);
    if (rc) return rc;
    if (nbins > SCT_NOALLOC_MAX_BINS)
        return sct_fail(ctx, SCT_ERR_INVALID, "sct_calc_hist_noalloc: nbins " + std::to_string(nbins) +
                                                  " exceeds SCT_NOALLOC_MAX_BINS (" +
                                                  std::to_string(SCT_NOALLOC_MAX_BINS) + ")");
    if (!ctx->noalloc_scratch) return sct_fail(ctx, SCT_ERR_HIP, "no-allocation scratch not set up");
    const VoxGeom g = vox_geom(f);
    HistArgs a = hist_args(rs, re, nbins, delta, seed);
    const HistPlan p = hist_plan(ctx, g, a, kNoAllocPartial);
    if (p.partial_words > kNoAllocPartial) return sct_fail(ctx, SCT_ERR_INVALID, "partial rows exceed the scratch");
    uint32_t *d = (uint32_t *)ctx->noalloc_scratch;
    uint32_t *partial = p.used ? d + SCT_NOALLOC_MAX_BINS : nullptr;
    hipError_t e = hist_passes(ctx, f, g, a, p, d, partial);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, d, sizeof(uint32_t) * nbins, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) (void)hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "calc_hist_noalloc");
    hist_density(f, rs, re, nbins, counts, density);
    return SCT_OK;
}

// calc_hist(ctx, in, nbins, delta), calc_hist.cpp:83-92.
extern "C" int sct_calc_hist_auto(sct_ctx *ctx, const sct_field *f, uint32_t nbins, float delta, uint64_t seed,
                                  uint32_t *counts, double *density, float *range_start, float *range_end)
{
    float mn = 0, mx = 0;
    int rc = sct_min_max(ctx, f, &mn, &mx);
    if (rc) return rc;
    if (range_start) *range_start = mn;
    if (range_end) *range_end = mx;
    return sct_calc_hist(ctx, f, mn, mx, nbins, delta, seed, counts, density);
}

// grad_magn(), grad_magn.cpp:60-79 (asynchronous, like the reference).
extern "C" int sct_grad_magn(sct_ctx *ctx, const sct_field *u, sct_field **v)
{
    if (!ctx || !u || !v) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    *v = nullptr;
    if (u->dtype != SCT_F32) return sct_fail(ctx, SCT_ERR_INVALID, "float field required");
    sct_field *out = nullptr;
    int rc = sct_field_like(ctx, u, SCT_F32, &out);
    if (rc) return rc;
    const VoxGeom g = vox_geom(u);
    // the kernel writes every voxel of every line
    hipError_t e = hipSuccess;
    if (!lines_cover_span(out->geo)) e = hipMemsetAsync(out->data, 0, sizeof(float) * out->geo.span(), ctx->stream);
    if (e == hipSuccess && g.lines && g.len) {
        grad_magn_kernel<<<stream_grid(ctx, g), 256, 0, ctx->stream>>>(g, (const float *)u->data,
                                                                     (float *)out->data);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        sct_field_destroy(out);
        return sct_hip_fail(ctx, e, "grad_magn");
    }
    *v = out;
    return SCT_OK;
}

// uv_hist(ctx, u, v, u0, u1, v0, v1, nbins), uv-hist-kernel.cpp:101-198.
extern "C" int sct_uv_hist(sct_ctx *ctx, const sct_field *u, const sct_field *v, float u_start, float u_end,
                           float v_start, float v_end, uint32_t nbins, uint32_t *counts, uint32_t *stride)
{
    if (!ctx || !u || !v || !counts) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (u->dtype != SCT_F32 || v->dtype != SCT_F32) return sct_fail(ctx, SCT_ERR_INVALID, "float fields required");
    if (u->geo.span() != v->geo.span() || u->geo.lines != v->geo.lines)
        return sct_fail(ctx, SCT_ERR_INVALID, "u and v layouts differ");
    if (nbins == 0) return sct_fail(ctx, SCT_ERR_INVALID, "nbins must be positive");
    UvArgs a;
    a.u0 = u_start;
    a.u1 = u_end;
    a.v0 = v_start;
    a.v1 = v_end;
    a.nbins = nbins;
    a.hstride = ((nbins + 31u) / 32u) * 32u;   // 128-byte rows, uv-hist-kernel.cpp:115-121
    a.fast_u = short_div_ok(u_start, u_end);
    a.fast_v = short_div_ok(v_start, v_end);
    const size_t total = size_t(nbins - 1) * a.hstride + nbins;
    if (stride) *stride = a.hstride;
    const VoxGeom g = vox_geom(u);
    // counts d, then (partitioned mode) the blocks of rows, in the context's
    // scratch
    const size_t dbytes = (sizeof(uint32_t) * total + 255) & ~size_t(255);
    const WidePlan wp = wide_plan(ctx, g, total, ~uint64_t(0));
    const bool part_mode = wp.ok;
    const uint32_t parts = wp.parts, lps = wp.lps, used = part_mode ? wp.used : 0;
    const size_t pbytes = part_mode ? sizeof(uint32_t) * wp.words : 0;
    void *mem = nullptr, *host = nullptr;
    int rc = sct_ctx_scratch(ctx, dbytes + pbytes, &mem);
    if (!rc) rc = sct_ctx_host_scratch(ctx, sizeof(uint32_t) * total, &host);
    if (rc) return rc;
    uint32_t *d = (uint32_t *)mem, *partial = used ? (uint32_t *)((char *)mem + dbytes) : nullptr;
    // With many partitions the joint bins are computed once, into 4 bytes per
    // voxel slot taken from the device's memory pool for this call only (the
    // pool keeps it for the next call and the memory query counts it as
    // available); without that memory the partitions compute the bins
    // themselves.
    uint32_t *bins = nullptr;
    if (part_mode && parts > kBinPassParts) {
        if (hipMallocAsync((void **)&bins, sizeof(uint32_t) * u->geo.span(), ctx->stream) != hipSuccess) {
            (void)hipGetLastError();
            bins = nullptr;
        }
    }
    const bool bin_pass = bins != nullptr;
    hipError_t e = hipSuccess;
    if (part_mode) {
        // 16-bit counters in LDS, by partitions of the bins (see uv_wide_kernel)
        const uint32_t grid = wide_grid(used, parts);
        if (bin_pass) {
            uv_bin_kernel<<<stream_grid(ctx, g), 256, 0, ctx->stream>>>(g, (const float *)u->data,
                                                                         (const float *)v->data, a, bins);
            e = hipGetLastError();
            if (e == hipSuccess) {
                uv_wide_kernel<true><<<grid, kPartThreads, 0, ctx->stream>>>(
                    g, nullptr, nullptr, bins, a, parts, used, lps, uint32_t(total), partial);
                e = hipGetLastError();
            }
        } else {
            uv_wide_kernel<false><<<grid, kPartThreads, 0, ctx->stream>>>(
                g, (const float *)u->data, (const float *)v->data, nullptr, a, parts, used, lps, uint32_t(total),
                partial);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            wide_sum_kernel<uint32_t><<<uint32_t((total + 511) / 512), 256, 0, ctx->stream>>>(partial, used, parts,
                                                                                               uint32_t(total), d);
            e = hipGetLastError();
        }
    } else {
        e = hipMemsetAsync(d, 0, sizeof(uint32_t) * total, ctx->stream);
        if (e == hipSuccess && g.lines && g.len) {
            uv_hist_kernel<<<stream_grid(ctx, g), 256, 0, ctx->stream>>>(g, (const float *)u->data,
                                                                          (const float *)v->data, a, d);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(host, d, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, ctx->stream);
    if (bins) {
        const hipError_t ef = hipFreeAsync(bins, ctx->stream);
        if (e == hipSuccess) e = ef;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) (void)hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "uv_hist");
    std::memcpy(counts, host, sizeof(uint32_t) * total);
    return SCT_OK;
}

// uv_hist(ctx, u, v, nbins), uv-hist-kernel.cpp:73-99.
extern "C" int sct_uv_hist_auto(sct_ctx *ctx, const sct_field *u, const sct_field *v, uint32_t nbins,
                                uint32_t *counts, uint32_t *stride, float *ranges)
{
    float v0 = 0, v1 = 0, u0 = 0, u1 = 0;
    int rc = sct_grad_min_max(ctx, v, &v0, &v1);
    if (!rc) rc = sct_min_max(ctx, u, &u0, &u1);
    if (rc) return rc;
    if (ranges) {
        ranges[0] = u0;
        ranges[1] = u1;
        ranges[2] = v0;
        ranges[3] = v1;
    }
    return sct_uv_hist(ctx, u, v, u0, u1, v0, v1, nbins, counts, stride);
}

// ---------------------------------------------------------------------------
// Ingest: the steps before the path on scanner data (transpose/).
// ---------------------------------------------------------------------------
namespace {

// convert_to_fp_kernel, convert_to_fp.cpp:11-46: v <= lo -> 0, v >= hi -> 1,
// else (v - lo) / (hi - lo) in float.
// A line per wave; each lane converts 4 consecutive voxels per trip, with one
// vector load and one 16-byte store when the stride is a multiple of 4.
__device__ __forceinline__ float convert_one(float v, float lo, float hi)
{
    return v <= lo ? 0.0f : (v >= hi ? 1.0f : (v - lo) / (hi - lo));
}

template <typename T>
__global__ __launch_bounds__(256) void convert_kernel(VoxGeom g, const T *__restrict__ in, float *__restrict__ out,
                                                      float lo, float hi)
{
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t nwaves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t line = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; line < g.lines; line += nwaves) {
        const uint64_t o = uint64_t(line) * g.stride;
        const T *r = in + o;
        float *w = out + o;
        if (g.stride % 4 != 0) {
            for (uint32_t n = lane; n < g.len; n += kWave) w[n] = convert_one(float(r[n]), lo, hi);
            continue;
        }
        for (uint32_t n0 = lane * 4; n0 < g.len; n0 += kWave * 4) {
            if (n0 + 4 <= g.len) {
                float v[4];
                if constexpr (sizeof(T) == 2) {
                    const uint2 q = *reinterpret_cast<const uint2 *>(r + n0);
                    v[0] = float(q.x & 0xffffu);
                    v[1] = float(q.x >> 16);
                    v[2] = float(q.y & 0xffffu);
                    v[3] = float(q.y >> 16);
                } else {
                    const float4 q = *reinterpret_cast<const float4 *>(r + n0);
                    v[0] = float(q.x);
                    v[1] = float(q.y);
                    v[2] = float(q.z);
                    v[3] = float(q.w);
                }
                *reinterpret_cast<float4 *>(w + n0) = make_float4(convert_one(v[0], lo, hi), convert_one(v[1], lo, hi),
                                                                  convert_one(v[2], lo, hi), convert_one(v[3], lo, hi));
            } else {
                for (uint32_t n = n0; n < g.len; ++n) w[n] = convert_one(float(r[n]), lo, hi);
            }
        }
    }
}

// histogram_kernel of find_intensity_range.cpp:22-46 (65536 bins, one per u16
// value) with the wide histogram: workgroup (slice of lines) counts its
// slice's values in 65536 16-bit LDS counters and stores them as its block of
// rows; wide_sum_kernel adds the blocks into 64-bit counts.  Each lane takes 8
// values per chunk (one 16-byte load when the stride is a multiple of 8).
__global__ __launch_bounds__(kPartThreads) void u16_wide_kernel(VoxGeom g, const uint16_t *__restrict__ in,
                                                                uint32_t used, uint32_t lines_per_slice,
                                                                uint32_t *__restrict__ rows)
{
    uint32_t slice, part;
    if (!wide_block(1, used, slice, part)) return;
    __shared__ uint32_t sh[kWideLdsWords];
    const WideHist h(sh, kWideBins);
    h.clear();
    __syncthreads();
    const uint32_t l0 = slice * lines_per_slice, l1 = min(g.lines, l0 + lines_per_slice);
    const uint16_t *const arr[1] = {in};
    count_slice<uint16_t, 1>(h, g, arr, l0, l1,
                             [&](const uint16_t (&x)[1], uint32_t, uint32_t) -> uint32_t { return x[0]; });
    __syncthreads();
    h.store(rows + uint64_t(slice) * kWideRowWords);
}

// The same histogram with device atomics, for a device without 128 KB of LDS
// per workgroup.
__global__ __launch_bounds__(256) void u16_hist_kernel(VoxGeom g, const uint16_t *__restrict__ in,
                                                       uint32_t lines_per_block, unsigned long long *__restrict__ counts)
{
    const uint32_t l0 = blockIdx.x * lines_per_block, l1 = min(g.lines, l0 + lines_per_block);
    const uint32_t lane = threadIdx.x % kWave, nw = blockDim.x / kWave;
    const uint32_t pairs = l0 < l1 ? (l1 - l0) * g.chunks : 0;
    for (uint32_t p = threadIdx.x / kWave; p < pairs; p += nw) {
        const uint32_t line = l0 + p / g.chunks, n = (p % g.chunks) * kWave + lane;
        if (n >= g.len) continue;
        atomicAdd(&counts[in[uint64_t(line) * g.stride + n]], 1ull);
    }
}

}  // namespace

// convert_to_fp<ElementType>(ctx, in, lo, hi), convert_to_fp.cpp:48-71
// (ElementType uint16_t or float).  *out: new float field, padding zero.
extern "C" int sct_convert_to_fp(sct_ctx *ctx, const sct_field *in, float lo, float hi, sct_field **out)
{
    if (!ctx || !in || !out) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    *out = nullptr;
    const Geometry &geo = in->geo;
    sct_field *f = nullptr;
    int rc = sct_field_like(ctx, in, SCT_F32, &f);
    if (rc) return rc;
    const VoxGeom g = vox_geom(in);
    // the kernel writes every voxel of every line
    hipError_t e = hipSuccess;
    if (!lines_cover_span(geo)) e = hipMemsetAsync(f->data, 0, sizeof(float) * geo.span(), ctx->stream);
    if (e == hipSuccess && g.lines && g.len) {
        if (in->dtype == SCT_U16)
            convert_kernel<uint16_t><<<stream_grid(ctx, g), 256, 0, ctx->stream>>>(g, (const uint16_t *)in->data,
                                                                                (float *)f->data, lo, hi);
        else
            convert_kernel<float><<<stream_grid(ctx, g), 256, 0, ctx->stream>>>(g, (const float *)in->data,
                                                                             (float *)f->data, lo, hi);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        sct_field_destroy(f);
        return sct_hip_fail(ctx, e, "convert_to_fp");
    }
    *out = f;
    return SCT_OK;
}

// find_intensity_range(ctx, field, tail_thr), find_intensity_range.cpp:63-110:
// the 65536-bin histogram of a u16 field and the tail cut.  hist (optional)
// receives the 65536 counts.  Counts are 64-bit, as the reference's
// (find_intensity_range.cpp:48-60): the blocks of a slice hold 16-bit
// counters and spill entries of at most 2^23 voxels, the sums 64-bit ones.
// Synchronises.
extern "C" int sct_find_intensity_range(sct_ctx *ctx, const sct_field *f, double tail_thr, uint16_t *lo,
                                        uint16_t *hi, uint64_t *hist)
{
    if (!ctx || !f || !lo || !hi) return sct_fail(ctx, SCT_ERR_INVALID, "null argument");
    if (f->dtype != SCT_U16) return sct_fail(ctx, SCT_ERR_INVALID, "u16 field required");
    constexpr uint32_t nbins = 65536;
    const VoxGeom g = vox_geom(f);
    // counts d, then the blocks of rows of the slices of lines, in the
    // context's scratch
    const WidePlan wp = wide_plan(ctx, g, nbins, ~uint64_t(0));
    const size_t dbytes = sizeof(uint64_t) * nbins;
    uint64_t *h = nullptr;
    void *mem = nullptr;
    int rc = sct_ctx_scratch(ctx, dbytes + (wp.ok ? sizeof(uint32_t) * wp.words : 0), &mem);
    if (!rc) rc = sct_ctx_host_scratch(ctx, sizeof(uint64_t) * nbins, (void **)&h);
    if (rc) return rc;
    uint64_t *d = (uint64_t *)mem;
    uint32_t *partial = (uint32_t *)((char *)mem + dbytes);
    hipError_t e = hipSuccess;
    if (wp.ok) {
        u16_wide_kernel<<<wp.used, kPartThreads, 0, ctx->stream>>>(g, (const uint16_t *)f->data, wp.used, wp.lps,
                                                                   partial);
        e = hipGetLastError();
        if (e == hipSuccess) {
            wide_sum_kernel<unsigned long long><<<nbins / 512, 256, 0, ctx->stream>>>(partial, wp.used, 1, nbins,
                                                                                       (unsigned long long *)d);
            e = hipGetLastError();
        }
    } else {
        // no voxels, or lines too long for a slice: device atomics
        e = hipMemsetAsync(d, 0, sizeof(uint64_t) * nbins, ctx->stream);
        if (e == hipSuccess && g.lines && g.len) {
            const uint32_t blocks = std::max<uint32_t>(1, ctx->cu_count * 4);
            const uint32_t lpb = std::max<uint32_t>(1, (g.lines + blocks - 1) / blocks);
            u16_hist_kernel<<<(g.lines + lpb - 1) / lpb, 256, 0, ctx->stream>>>(
                g, (const uint16_t *)f->data, lpb, (unsigned long long *)d);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(uint64_t) * nbins, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) (void)hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return sct_hip_fail(ctx, e, "find_intensity_range");
    uint64_t total = 0;
    for (uint32_t i = 0; i < nbins; ++i) total += h[i];
    const double pct_lo = tail_thr * double(total), pct_hi = (1.0 - tail_thr) * double(total);
    size_t l = 0, u = nbins;
    uint64_t cum = 0;
    for (size_t i = 0; i < nbins; ++i) {
        cum += h[i];
        if (pct_lo >= double(cum)) l = i;
        if (pct_hi <= double(cum)) {
            u = i;
            break;
        }
    }
    *lo = uint16_t(l);
    *hi = uint16_t(u);   // 65536 wraps to 0 as in the reference's uint16_t(hi)
    if (hist)
        for (uint32_t i = 0; i < nbins; ++i) hist[i] = h[i];
    return SCT_OK;
}
