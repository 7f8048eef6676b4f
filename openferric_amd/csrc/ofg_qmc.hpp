// Device layouts and launch interface of the quasi-Monte Carlo (Sobol)
// European pricer, ofg_qmc.hip; shared with the host side (ofg_qmc_api.cpp).
//
// Reference: mc_european_qmc_with_seed (src/engines/monte_carlo/mc_qmc.rs:
// 40-144) on SobolSequence (src/math/sobol.rs). Point k (k = 1, 2, ..., the
// reference's index after k calls of next_into) of dimension d is
//   x = XOR of v[d][j] over the set bits j of g = k ^ (k >> 1),  then ^ scr[d],
// v the direction numbers and scr the scrambler. A wave covers 64
// consecutive k aligned to 64, so the bits of g above 5 are the same on every
// lane: their XOR is formed once per wave (per dimension), and the low 6 bits
// index `lows[d][g & 63]`, the XOR of v[d][0..5] precomputed for all 64
// patterns.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ofg {

constexpr int QMC_BLOCK = 256;
constexpr int QMC_SEG = 256;          // dimensions per LDS segment
constexpr int QMC_MAX_TILES = 1024;

struct alignas(16) QmcDev {
    int32_t put;
    int32_t steps;          // Sobol dimensions
    int32_t tile_first;
    int32_t n_tiles;
    int32_t ppt;            // point indices per thread: a tile covers 256 * ppt indices
    int32_t pad;
    int64_t paths;          // points k = 1 .. paths
    int64_t dim_offset;     // first dimension of this option in the tables
    double s0, strike, dt_drift, dt_vol, discount;
};

struct QmcPartial {
    double sum, sum_sq;
};

struct QmcOut {
    double price, stderr_;
};

struct QmcArgs {
    const QmcDev *opts;
    const int32_t *tile_opt;
    const uint64_t *dirs;   // [dim][64] direction numbers
    const uint64_t *lows;   // [dim][64] XOR of v[dim][0..5] by 6-bit pattern
    const uint64_t *scr;    // [dim] scramblers
    QmcPartial *partials;   // one per tile
    QmcOut *out;            // one per option
    int32_t n_tiles;
    int32_t n_opts;
};

hipError_t launch_qmc(const QmcArgs &a, hipStream_t stream);

}  // namespace ofg
