// Pathwise adjoint Greeks: launch interface shared by the host side
// (ofg_api.cpp) and the kernels (ofg_adjoint.hip).
//
// One fused kernel per sample: a forward sweep that checkpoints the path state
// every adj_segment(model) steps into per-wave scratch, then a backward sweep that
// walks the segments in reverse, recomputing each from its checkpoint with the
// segment's draws regenerated by stepping the RNG stream backwards, and sweeps
// the adjoints of the state and of the model parameters through it.
#pragma once

#include "../../include/openferric_amd.h"
#include "ofg_internal.hpp"

#include <hip/hip_runtime.h>

namespace ofg {

// Steps between checkpoints: 8 for GBM, 4 for Heston (its segment keeps
// S, v, both normals and the growth factor of every step in registers).
__host__ __device__ constexpr int adj_segment(int model) { return model == 1 ? 4 : 8; }
constexpr int ADJ_NG = 8;        // gradient slots

// Gradient slots: 0 S0, 1 sigma (GBM) | v0 (Heston), 2 r, 3 T, 4 kappa,
// 5 theta, 6 xi, 7 rho.
struct AdjPartial {
    double spv, spv2, g[ADJ_NG], n;
};

struct AdjOut {
    double price, stderr_, g[ADJ_NG], n;
};

struct AdjArgs {
    const OptDev *opts;
    const WorkItem *items;
    const int32_t *order;
    int32_t order_begin, order_end;
    uint32_t *queue_head;
    AdjPartial *partials;
    const double *obs_w;
    const double *normals1, *normals2;
    double *scratch;          // checkpoint scratch, one slot per resident wave
    int64_t wave_stride;      // doubles per wave slot
    int32_t rng_mode, model, antithetic, block, grid;
};

hipError_t launch_adjoint(const AdjArgs &a, hipStream_t stream);
hipError_t launch_adjoint_reduce(const OptDev *opts, const int64_t *item_first,
                                 const int32_t *item_count, const AdjPartial *partials,
                                 AdjOut *out, int32_t n_opts, hipStream_t stream);
int adjoint_grid(int block, int rng_mode, int model, int antithetic, int device);

}  // namespace ofg
