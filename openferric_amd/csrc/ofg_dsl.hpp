// Device layouts and launch interface of the structured-product engine
// (ofg_dsl.hip): correlated multi-asset paths and the compiled product's
// program evaluated per path (src/dsl/engine.rs, src/dsl/eval.rs). Shared
// with ofg_dsl_front.cpp and ofg_dsl_api.cpp.
//
// Program model. The reference compiles each schedule's statements to stack
// bytecode with conditional jumps (eval.rs:1542-1772). Here the stack depth
// of every instruction is known when compiling, so stack slots become
// registers with fixed indices, and branches become predication: every lane
// of a wave runs the same instruction stream (fetched with scalar loads)
// under a per-lane `active` flag that if / else / end-if narrow and restore
// from a mask stack of fixed depth. A lane that redeems or skips clears its
// `alive` flag, which every restore re-applies; its later statements and
// observations run masked, as the reference returns from the path.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ofg {

constexpr int DSL_MAX_ASSETS = 16;
constexpr int DSL_MAX_REGS = 32;   // 64 KB of LDS per 256-thread block at the limit
constexpr int DSL_MAX_MASKS = 16;
constexpr int DSL_BLOCK = 256;
constexpr int DSL_FUSE_MAXS = 9;    // markets simulated together per path
constexpr int DSL_FUSE_MAXA = 4;    // assets per market in the fused simulation

enum DslOp : uint8_t {
    DOP_CONST = 0,    // r[dst] = consts[aux]
    DOP_NOTIONAL,     // r[dst] = notional
    DOP_DATE,         // r[dst] = observation date
    DOP_ISFINAL,      // r[dst] = is_final ? 1 : 0
    DOP_SET,          // active: r[dst] = r[a]   (let, set)
    DOP_ADD, DOP_SUB, DOP_MUL, DOP_DIV, DOP_MIN, DOP_MAX,
    DOP_EQ, DOP_NE, DOP_LT, DOP_LE, DOP_GT, DOP_GE, DOP_AND, DOP_OR,
    DOP_NEG, DOP_ABS, DOP_EXP, DOP_LOG, DOP_NOT,
    DOP_WOF_PERF,     // r[dst] = min_i S_i(t)/S_i(0) over assets with S_i(0) > 0
    DOP_BOF_PERF,     // r[dst] = max_i ...
    DOP_WORST,        // r[dst] = min of r[a .. a+aux-1], last argument first
    DOP_BEST,         // r[dst] = max of ...
    DOP_PRICE,        // r[dst] = S_idx(t), idx = r[a] as usize (active lanes check the range)
    DOP_PAY,          // active: pv += r[a] * df
    DOP_REDEEM,       // active: pv += r[a] * df; the path ends
    DOP_SKIP,         // active: the path ends
    DOP_IF,           // saved[aux] = active; cond[aux] = r[a] != 0; active &= cond[aux]
    DOP_ELSE,         // active = saved[aux] && !cond[aux] && alive
    DOP_ENDIF,        // active = saved[aux] && alive
};

struct DslIns {
    uint8_t op, dst, a, b;
    uint32_t aux;
};

// One schedule: its program and its observations.
struct DslSched {
    int32_t code_off, code_len;
    int32_t obs_first, n_obs;
};

struct DslObs {
    int32_t snap;       // snapshot index (eval.rs:288-302)
    int32_t is_final;
    double date;
};

enum DslStepper : int32_t { DSL_GBM = 0, DSL_SCHWARTZ = 1, DSL_VASICEK = 2 };

// One scenario: the product under one market (the base price or a bump).
struct alignas(16) DslScen {
    int32_t n_assets;
    int32_t tile_first;
    int32_t n_tiles;
    int32_t ppt;            // paths per thread
    int64_t paths;
    int64_t normals_offset; // external normals [step][asset][path]
    int64_t seed_offset;    // reference-stream start states [thread][word]: a thread's
                            // paths are consecutive in the stream
    uint64_t seed;          // fast draws: stream_seed(seed, path)
    int32_t df_off;         // discount factors of the observations, exp(-r t)
    int32_t own_corr;       // fused simulation: 1 when the Cholesky factor differs from market 0's
    int64_t snap_off;       // split form: this scenario's [snap][asset][path] snapshots
    // fused simulation: bit i set when asset i's step is market 0's GBM step
    // (same parameters and correlated normal), so its factor is shared
    uint32_t share_mask;
    uint32_t pad2;
    int32_t kind[DSL_MAX_ASSETS];
    double s0[DSL_MAX_ASSETS];
    // GBM: p0 drift, p1 diffusion; Schwartz: p0 long-run log mean,
    // p1 exp(-kappa dt), p2 vol sqrt(dt); Vasicek: p0 long-run mean,
    // p1 exp(-a dt), p2 vol sqrt(dt)   (engine.rs:51-91)
    double p0[DSL_MAX_ASSETS], p1[DSL_MAX_ASSETS], p2[DSL_MAX_ASSETS];
    double chol[DSL_MAX_ASSETS * (DSL_MAX_ASSETS + 1) / 2];
};

struct DslPartial {
    double sum, sum_sq;
    double err_idx;   // first out-of-range price() index of the tile, or -1
    double pad;
};

struct DslOut {
    double sum, sum_sq, err_idx, pad;
};

struct DslArgs {
    const DslScen *scen;
    const int32_t *tile_scen;
    const DslIns *code;
    const double *consts;
    const DslSched *sched;
    const DslObs *obs;
    const double *df;
    const int32_t *step_snap;   // [steps + 1]: snapshot index or -1
    const int32_t *snap_list;   // (step, snapshot index) pairs in step order,
                                // then the sentinel (steps + 1, -1)
    double *snaps;              // split: per scenario [snap][asset][path];
                                // fused: [tile][snap][asset][DSL_BLOCK]
    DslPartial *partials;
    DslOut *out;
    const double *normals;
    const uint64_t *seeds;
    double notional;
    int32_t steps;
    int32_t n_snaps;
    int32_t n_sched;
    int32_t n_locals;
    int32_t n_state;
    int32_t n_masks;
    int32_t n_tiles;
    int32_t n_scen;
    int32_t rng_mode;
    int32_t n_assets;           // of every scenario
    int32_t split;              // 1: simulation and evaluation kernels, 0: fused
    int32_t fuse_markets;       // split form: 1 = one simulation pass for all markets
    int32_t fuse_tiles;         // its grid: market 0's tiles
    int32_t n_regs;             // program registers (LDS per thread)
    int32_t gbm_only;           // every asset of every scenario steps as GBM
    const double *state_init;   // [n_state]
};

// Native evaluation kernel of one product (compiled at run time by
// ofg_dsl_jit_compile, ofg_dsl_jit.cpp): the layouts below are pasted into
// its generated source as text, so host and kernel agree on them by
// construction (one declaration per field, no commas).
#define OFG_DSL_JIT_LAYOUTS                                                           \
    struct DslJitScen {                                                               \
        long long paths;                                                              \
        long long snap_off;                                                           \
        int tile_first;                                                               \
        int ppt;                                                                      \
        int df_off;                                                                   \
        int pad;                                                                      \
        double s0[16];                                                                \
    };                                                                                \
    struct DslJitObs {                                                                \
        int snap;                                                                     \
        int is_final;                                                                 \
        double date;                                                                  \
    };                                                                                \
    struct DslJitPartial {                                                            \
        double sum;                                                                   \
        double sum_sq;                                                                \
        double err_idx;                                                               \
        double pad;                                                                   \
    };                                                                                \
    struct DslJitArgs {                                                               \
        const int *tile_scen;                                                         \
        const DslJitScen *scen;                                                       \
        const double *snaps;                                                          \
        const DslJitObs *obs;                                                         \
        const double *df;                                                             \
        DslJitPartial *partials;                                                      \
    };
OFG_DSL_JIT_LAYOUTS
#define OFG_DSL_STR_(...) #__VA_ARGS__
#define OFG_DSL_STR(...) OFG_DSL_STR_(__VA_ARGS__)
static_assert(sizeof(DslJitObs) == sizeof(DslObs) && offsetof(DslJitObs, date) == offsetof(DslObs, date),
              "observation layout");
static_assert(sizeof(DslJitPartial) == sizeof(DslPartial), "partial-sum layout");
static_assert(sizeof(((DslJitScen *)nullptr)->s0) == sizeof(((DslScen *)nullptr)->s0), "spots");

// Launches the split or fused form and the close kernel. jit_eval, when not
// null, is the product's native evaluation kernel, run with jit_args in place
// of the program interpreter (split form only).
hipError_t launch_dsl(const DslArgs &a, hipStream_t stream, hipFunction_t jit_eval = nullptr,
                      const DslJitArgs *jit_args = nullptr);

}  // namespace ofg
