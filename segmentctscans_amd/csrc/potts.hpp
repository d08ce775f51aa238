// potts.hpp -- host-side pieces of the Potts solver shared between translation units.
#pragma once

#include "sct_internal.hpp"

namespace sct {

// Slab decomposition along n, potts-min-partition.cpp:560-577 and :627-630.
struct PottsPlan {
    int slab_w = 0;   // n-width of one slab
    int res_w = 0;    // slab_w - 2 * halo (planes kept per slab)
    int halo = 0;     // = iters
    int n_slabs = 0;
};

int potts_plan(sct_ctx *ctx, const sct_field *g, int K, int iters, const sct_potts_opts *o, PottsPlan *plan);

}  // namespace sct
