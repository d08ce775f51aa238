// sct_internal.hpp -- shared host/device plumbing of the HIP library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sct.h"

// ---------------------------------------------------------------------------
// Line geometry.  The cylinder layout stores each live (row, col) pair as one
// contiguous line of nstride elements, lines ordered by (col, row); since
// base[col] is a multiple of nstride, offset(n, row, col) = line * nstride + n
// (layout_cylinder.h:81-84).  Every stencil needs, per line, the line indices
// of its four in-plane neighbours, or -1 where the reference's predicates say
// the neighbour is absent (potts-min-partition.cpp:128,131,390,393).  The
// table is built once per layout on the host and read once per wave.
// ---------------------------------------------------------------------------
struct LineNbr {
    int32_t prow, nrow, pcol, ncol;
};

// Band order of the lines: the (row, col) plane is cut into bands of
// kBandRows rows; band by band, column by column, the band's rows of that
// column.  A stencil that walks the lines in this order finds its +col / -col
// neighbour kBandRows lines away instead of a whole column (about 2r lines
// for a cylinder of radius r), so the neighbour lines are still in the L2
// that fetched them.  Only the lines at band edges have a row neighbour in
// another band.
constexpr uint32_t kBandRows = 32;
constexpr uint32_t kMarchRows = 8;   // lines a marching wave walks (Chambolle)

// One line of the band order with everything a stencil reaching one line in
// each direction, and the diagonal lines a divergence at the +row / +col line
// needs, reads about it: one 32-byte scalar load per wave.
struct OrderedLine {
    int32_t line, prow, nrow, pcol, ncol;
    int32_t nrow_pcol;   // -col line of the +row line, or -1
    int32_t ncol_prow;   // -row line of the +col line, or -1
    uint32_t flags;      // kNrowHasNrow | kNrowHasNcol | kNcolHasNrow | kNcolHasNcol
};
constexpr uint32_t kNrowHasNrow = 1, kNrowHasNcol = 2, kNcolHasNrow = 4, kNcolHasNcol = 8;

// The line tables of a built geometry.  They never change, so every field
// with that geometry shares them (a field made like another copies only the
// pointer); the last owner frees the device tables.
struct GeoTables {
    std::vector<LineNbr> h_nbr;
    std::vector<uint32_t> line_col, line_row;
    LineNbr *d_nbr = nullptr;
    OrderedLine *d_oline = nullptr;
    uint2 *d_runs = nullptr;
    uint32_t *d_start_row = nullptr, *d_base = nullptr;
    GeoTables() = default;
    GeoTables(const GeoTables &) = delete;
    GeoTables &operator=(const GeoTables &) = delete;
    ~GeoTables();
};

struct Geometry {
    uint32_t n_images = 0, height = 0, width = 0, nstride = 0, col_begin = 0, col_end = 0;
    std::vector<uint32_t> start_row, base;   // host copies of the mapping tables
    uint32_t lines = 0;
    uint64_t data_size = 0;
    std::shared_ptr<const GeoTables> tables;
    // the device tables of `tables`, for kernel arguments
    LineNbr *d_nbr = nullptr;                // device [lines]
    OrderedLine *d_oline = nullptr;          // device [lines], in band order
    // Runs of up to kMarchRows consecutive rows of one column in the band
    // order: (first position in the band order, number of lines).
    uint32_t n_runs = 0;
    uint2 *d_runs = nullptr;                 // device [n_runs]
    uint32_t *d_start_row = nullptr, *d_base = nullptr;   // device mapping tables

    uint32_t row_end(uint32_t col) const { return start_row[col] + (base[col + 1] - base[col]) / nstride; }
    size_t span() const { return base[width]; }
};

struct sct_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipMemPool_t pool = nullptr;
    std::string err;
    uint32_t cu_count = 0;
    std::map<uint64_t, int> tiles;   // batch autotuner: chosen Xi workgroup tiles
    int last_tiles[5] = {};          // tiles of the last sct_potts_batch call, K = 2..6
    int last_occ[5] = {};            // its persistent launches' workgroups per CU (0: no persistent launch)
    int forced_tiles[5] = {};        // tiles set by sct_potts_batch_set_tiles (0: not set)
    int forced_occ[5] = {};          // workgroups per CU set by sct_potts_batch_set_occupancy (0: not set)
    void *scratch = nullptr;         // device scratch of the reductions (sct_ctx_scratch)
    size_t scratch_bytes = 0;
    void *host_scratch = nullptr;    // pinned host staging of their results
    size_t host_scratch_bytes = 0;
    void *noalloc_scratch = nullptr; // device address of this GPU's fixed scratch (sct_calc_hist_noalloc)
};

struct sct_field {
    sct_ctx *ctx = nullptr;
    int dtype = SCT_F32;
    void *data = nullptr;
    Geometry geo;
    int device = 0;
    bool pooled = false;   // data from the device's stream-ordered pool (else hipMalloc)
};

// ---- error handling --------------------------------------------------------
int sct_fail(sct_ctx *ctx, int code, const std::string &msg);
int sct_hip_fail(sct_ctx *ctx, hipError_t e, const char *what);

#define SCT_HIP(ctx, expr)                                        \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return sct_hip_fail((ctx), e_, #expr); \
    } while (0)

// ---- memory (hipMalloc / hipFree) -------------------------------------------
int sct_alloc(sct_ctx *ctx, void **p, size_t bytes);
void sct_free(sct_ctx *ctx, void *p);
// Waits for the context's stream and returns the pool's idle memory (kept for
// field allocations) to the device, before a large sct_alloc.
void sct_pool_trim(sct_ctx *ctx);
// Device / pinned host scratch of at least `bytes`, owned by the context and
// grown on demand (kept until the context is destroyed).  For temporaries
// within one call: the next call on the context's stream reuses the same
// memory (stream order keeps the uses apart; the host side must have
// synchronised before it reuses host scratch).
int sct_ctx_scratch(sct_ctx *ctx, size_t bytes, void **p);
int sct_ctx_host_scratch(sct_ctx *ctx, size_t bytes, void **p);

// Finds the device address of the fixed scratch of sct_calc_hist_noalloc on
// the context's (current) device, once per device, into ctx->noalloc_scratch.
int sct_noalloc_scratch_setup(sct_ctx *ctx);

// ---- geometry construction ------------------------------------------------
int sct_geometry_build(sct_ctx *ctx, const sct_layout *l, Geometry &g);
void sct_geometry_release(sct_ctx *ctx, Geometry &g);
int sct_geometry_copy(sct_ctx *ctx, const Geometry &src, Geometry &dst);
int sct_field_like(sct_ctx *ctx, const sct_field *src, int dtype, sct_field **out);

// ---- device helpers --------------------------------------------------------
constexpr int kWave = 64;
constexpr uint32_t kXcdCount = 8;   // MI355X: 8 XCDs, each with its own L2

// Workgroup b of a grid of nb is dispatched to XCD b % 8.  xcd_block(b, nb, G)
// is the logical block it works on: within every run of 8G logical blocks,
// XCD x takes the G consecutive blocks [xG, xG + G), so neighbouring work
// shares an L2 while the whole grid still advances as one front.  Blocks past
// the last full run keep their index.  A bijection on [0, nb).
__host__ __device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb, uint32_t G)
{
    const uint32_t run = kXcdCount * G;
    if (b >= nb / run * run) return b;
    const uint32_t x = b % kXcdCount, s = b / kXcdCount;
    return (s / G) * run + x * G + s % G;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}

// Atomic max of a non-negative float through its ordered bit pattern.
__device__ __forceinline__ void atomic_max_nonneg(float *addr, float v)
{
    atomicMax(reinterpret_cast<unsigned int *>(addr), __float_as_uint(v));
}

// Order-preserving float <-> uint key, for exact atomic min/max of any sign.
__device__ __forceinline__ uint32_t float_key(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
inline float key_float(uint32_t k)
{
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
