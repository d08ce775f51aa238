// potts.hpp -- host-side pieces of the Potts solver shared between translation units.
#pragma once

#include "sct_internal.hpp"
#include "potts_math.hpp"

namespace sct {

// Slab decomposition along n, potts-min-partition.cpp:560-577 and :627-630.
struct PottsPlan {
    int slab_w = 0;   // n-width of one slab
    int res_w = 0;    // slab_w - 2 * halo (planes kept per slab)
    int halo = 0;     // = iters
    int n_slabs = 0;
};

// Per-column row range and first line, for kernels that march along col.
struct ColInfo {
    uint32_t line0;   // line index of (start_row, col)
    uint32_t rb, re;  // live rows [rb, re); rb == re for dead columns
    uint32_t pad;
};

// One workgroup of the fused kernel: the n-range [n0, n0 + TN) of the rows
// [r0, r0 + R), marched over the columns [cs, ce) (V is updated there; Xi+ is
// also recomputed at the row r0 - 1, at n0 - 1 and at the column cs - 1).
struct FusedTile {
    uint32_t n0, r0;
    uint32_t cs, ce;
};

DykstraLimits dykstra_limits();

}  // namespace sct

struct sct_potts_solver {
    sct_ctx *ctx = nullptr;
    const sct_field *g = nullptr;
    int K = 0;
    float c[6] = {};
    float lambda = 0, tau1 = 0, tau2 = 0;
    sct_potts_opts opts{};
    int kernel = SCT_KERNEL_FUSED;
    uint32_t s0 = 0, s1 = 0, wstride = 0;
    uint64_t wspan = 0;
    // Xi and V are double-buffered for the fused kernel (read [cur], write
    // [cur ^ 1]; V points at Vb[cur]); the two-pass kernels update [0] in
    // place.  V-bar has one buffer (vbar[1] is unused).
    // Element type of Xi and V-bar follows opts.storage (float, bfloat16 or half).
    void *xi[2] = {nullptr, nullptr};
    void *vbar[2] = {nullptr, nullptr};
    int storage = SCT_STORAGE_F32;
    float *V = nullptr;
    float *Vb[2] = {nullptr, nullptr};
    int cur = 0;
    float *d_dq = nullptr;
    int dq_cap = 0;
    int it = 0;   // iterations done on this slab
    // kernel timing: when marks is set, an event is recorded before every
    // kernel launch of an iteration (sct_potts_solver_iterate_timed)
    hipEvent_t *marks = nullptr;
    int n_marks = 0;
    // fused-kernel tables (device)
    sct::ColInfo *d_cols = nullptr;
    sct::FusedTile *d_tiles = nullptr;
    uint32_t n_tiles = 0, tiles_len = 0;   // tiles of a slab of length tiles_len
    void *d_fptrs = nullptr;   // the fused kernel's two component pointer tables
};

namespace sct {

// FP64 primal and dual energies of the current state (gap.hip).  Synchronises.
int potts_gap(sct_potts_solver *s, double *primal, double *dual);

inline void potts_mark(sct_potts_solver *s)
{
    if (s->marks) (void)hipEventRecord(s->marks[s->n_marks++], s->ctx->stream);
}

int potts_plan(sct_ctx *ctx, const sct_field *g, int K, int kernel, int iters, const sct_potts_opts *o,
               PottsPlan *plan);

// Fused one-pass iteration (potts_fused.hip).
int fused_prepare(sct_potts_solver *s);
int fused_iterate(sct_potts_solver *s, int iters, float *dq_dev);
void fused_release(sct_potts_solver *s);

}  // namespace sct
