// Device layouts and launch interface of the least-squares Monte Carlo
// (Longstaff-Schwartz) kernels, ofg_lsm.hip. Shared with the host side
// (ofg_lsm_api.cpp).
//
// One option's exercise dates are the path indices 1..n-1 at which the
// backward induction regresses (src/engines/lsm/longstaff_schwartz.rs:
// can_exercise :567-584, step_schedule :314-320); the terminal date is always
// exercised and is not among them. HBM holds, per option,
//   spots  [date][path]  the spot of every path at every exercise date,
//   values [path]        the running discounted cash flow of each path,
// and the backward induction is one launch per exercise date: apply the
// exercise decision of the date regressed by the previous launch, discount to
// the next earlier date, accumulate the regression sums there. Each launch is
// followed by a per-option close kernel (a fixed-order reduction of the tile
// partials) that solves the 3x3 normal equations for the next launch.
#pragma once

#include "ofg_internal.hpp"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ofg {

enum LsmDynamics : int32_t { LSM_GBM = 0, LSM_LOCAL_VOL = 1, LSM_HESTON = 2 };

constexpr int LSM_BLOCK = 256;       // threads per tile
constexpr int LSM_MAX_TILES = 1024;  // tiles per option (bounds the final reduce)

struct alignas(16) LsmDev {
    OptDev m;              // path constants: s0, dt, drift, diffusion, Heston, fast scalings, seed
    int32_t dynamics;      // LsmDynamics
    int32_t put;
    int32_t n_ex;          // exercise dates strictly before expiry
    int32_t barrier;       // 1: barrier instrument priced by plain Monte Carlo
    int32_t barrier_down;
    int32_t knock_in;
    int32_t tile_first;    // first tile (and partial slot) of this option
    int32_t n_tiles;
    int32_t date_offset;   // into the LsmDate / LsmDateOut arrays
    int32_t ppt;           // paths per thread: a tile covers LSM_BLOCK * ppt paths
    int32_t term_gap;      // discount steps from expiry to the latest date (or to 0)
    int32_t pad;
    int64_t paths;
    int64_t s_offset;      // spots, [date][path]
    int64_t v_offset;      // values
    int64_t seed_offset;   // reference-stream start states, [path][word]
    int64_t normals_offset;  // external normals, [step][path]
    double disc;           // exp(-r dt)
    double terminal_strike;
    double barrier_level, rebate, df_total;
};

struct LsmDate {
    int32_t step;          // path index of the date
    int32_t gap;           // discount steps back to the previous date (first date: to 0)
    double strike;
};

// Partial sums of one tile in one launch.
struct alignas(16) LsmPartial {
    double reg[8];         // sums of 1, S, S^2, S^3, S^4, y, S y, S^2 y over ITM paths
    double ex_n, ex_ext;   // exercised paths and their extreme spot (max put / min call)
    double term_itm;       // paths in the money at expiry
    double fsum, fsq;      // discounted values at time 0: sum, sum of squares
    double pad[3];
};
constexpr int LSM_NACC = 13;

// Per exercise date: regression result (written by the launch that
// regresses it) and exercise statistics (written by the next launch).
struct LsmDateOut {
    double beta[3];
    double itm;
    double exercised;
    double extreme;
    int32_t valid;         // itm >= 3: the date's exercise decision is applied
    int32_t pad;
};

struct LsmOptOut {
    double price, stderr_, term_itm, pad;
};

struct LsmArgs {
    const LsmDev *opts;
    const int32_t *tile_opt;     // tile -> option
    const LsmDate *dates;
    LsmDateOut *date_out;
    LsmOptOut *opt_out;
    LsmPartial *partials;        // one slot per tile
    double *spots;
    double *values;
    const uint64_t *seeds;
    const double *normals1, *normals2;
    int32_t launch;              // 0: path simulation; r >= 1: backward round r
    int32_t n_tiles;
    int32_t n_opts;
    int32_t rng_mode;
};

// Start states of the reference stream for a group of paths (P the
// characteristic polynomial of the xoshiro256 transition). Per path
// (tile_ppt = 0): thread g of job j jumps the seeded state by
// g * chunk * draws draws with the polynomials polys[k] =
// x^(chunk * draws * 2^k) mod P(x), then writes `chunk` consecutive paths'
// states, stepping `draws` draws between them. Per consumer thread
// (tile_ppt > 0): see tile_ppt below.
struct LsmSeedJob {
    uint64_t base[4];            // xoshiro256++ state after seed_from_u64(seed)
    int64_t paths;
    int64_t draws;               // draws per path (steps, or 2 steps for Heston)
    int64_t seed_offset;         // into the seeds array ([path][word] per option)
    int64_t thread_first;        // first global thread of this job
    int32_t chunk;
    int32_t n_polys;
    int32_t poly_offset;         // into polys, 4 words per polynomial
    // > 0: one state per thread of a kernel whose 256-thread tiles hold
    // tile_ppt paths per thread, each thread's paths consecutive in the
    // stream (thread_stream_start); chunk then counts consumer threads per
    // seed thread and the polynomials jump by draws * 2^k
    int32_t tile_ppt;
};

// Start path in the stream of thread g of such a kernel and its path
// count: the valid paths of the tiles in g-major order. Tile t holds paths
// [t 256 ppt, (t + 1) 256 ppt) in storage order p = base + j 256 + lane;
// with R of them valid (R = q 256 + r), lane l has q + (l < r) of them.
__host__ __device__ inline int64_t thread_stream_start(int64_t g, int64_t paths, int32_t ppt,
                                                       int64_t *count) {
    const int64_t per_tile = 256 * (int64_t)ppt;
    const int64_t base = (g / 256) * per_tile, l = g % 256;
    const int64_t rem = paths - base;
    const int64_t R = rem < 0 ? 0 : (rem < per_tile ? rem : per_tile);
    const int64_t q = R / 256, r = R % 256;
    if (count) *count = q + (l < r ? 1 : 0);
    return base + l * q + (l < r ? l : r);
}

hipError_t launch_lsm_seeds(const LsmSeedJob *jobs, int32_t n_jobs, int64_t total_threads,
                            const uint64_t *polys, uint64_t *seeds, hipStream_t stream);
hipError_t launch_lsm(const LsmArgs &a, hipStream_t stream);

}  // namespace ofg
