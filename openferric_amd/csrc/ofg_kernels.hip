// Monte Carlo pricing kernels for MI355X (gfx950), FP64 on the VALU.
//
// Layout of the work:
//   * a persistent grid of wavefronts pops work items from a device queue
//     (one atomic per item, cost-descending order), so thousands of options
//     with uneven path and step counts share one launch;
//   * a work item is a sample range of ONE option, so every lane of a wave
//     runs the same step count and the same payoff (no step divergence, all
//     contract parameters are wave-uniform and live in SGPRs);
//   * a lane simulates one sample at a time with the whole path state in
//     registers: spot, RNG state, running Asian sums or barrier flag; nothing
//     per step touches memory except (wave-uniform) observation weights;
//   * per item the wave reduces its five FP64 sums with a fixed butterfly
//     and writes them to the item's own slot, and a segmented reduce sums the
//     slots of each option in item order, so prices do not depend on the
//     queue schedule.
//
// Reference semantics: MonteCarloEngine::run's fold body
// (src/mc/simulation.rs:212-262), GbmPathGenerator::generate_into
// (simulation.rs:58-71), HestonPathGenerator + Heston::step_euler
// (simulation.rs:98-116, src/models/stochastic.rs:56-71) and the payoffs of
// src/engines/monte_carlo/mc_engine.rs:59-121, 441-584.
#include "ofg_internal.hpp"
#include "ofg_math.hpp"
#include "ofg_paths.hpp"

#include <type_traits>

namespace ofg {

namespace {

constexpr int WAVE = 64;
// Reference draws per inverse-CDF block (acklam_block_wave). 16 draws per
// block halve the tail rounds per draw but take the GBM loop to 190 VGPRs
// (2 waves per SIMD), which measured slower.
constexpr int REF_AK = 8;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

// ------------------------------------------------------------------------
// Sample simulation: one (possibly antithetic) sample -> (x, y)
// ------------------------------------------------------------------------

// Observation weights: read-only for the kernel's lifetime and indexed by the
// wave-uniform step, so they are read through the constant address space,
// which makes them scalar loads into SGPRs instead of per-lane vector loads.
using cdouble = __attribute__((address_space(4))) const double;

template <int MODEL, int PK, int RNG, bool ANTI>
__device__ __forceinline__ void simulate_sample(const OptDev &o, int64_t sample,
                                                const cdouble *w,
                                                const LaunchArgs &a, const double *tab,
                                                double *slots, double &x, double &y) {
    const double barrier = o.barrier;
    const int steps = o.steps;
    Leg ga, gb;
    leg_init<PK>(ga, o, (PK == PK_ASIAN || PK == PK_ASIAN_LOG) ? w[0] : 0.0);
    if constexpr (ANTI) leg_init<PK>(gb, o, (PK == PK_ASIAN || PK == PK_ASIAN_LOG) ? w[0] : 0.0);

    // Reference draws use the reference's xoshiro256++ (or its Pcg64 kind);
    // the fast mode uses xoroshiro128+ with Box-Muller (four normals per three
    // 64-bit draws).
    constexpr bool REF = RNG == OFG_RNG_REFERENCE || RNG == OFG_RNG_PCG64;
    using Rng = typename std::conditional<
        RNG == OFG_RNG_FAST, Xoroshiro128p,
        typename std::conditional<RNG == OFG_RNG_PCG64, Pcg64, Xoshiro256pp>::type>::type;
    Rng rng;
    ExternalNormals ext;
    if constexpr (REF || RNG == OFG_RNG_FAST) {
        rng.seed(stream_seed(o.seed, (uint64_t)sample));
    } else {
        ext.p1 = a.normals1 + o.normals_offset + sample;
        ext.p2 = a.normals2 ? a.normals2 + o.normals_offset + sample : nullptr;
        ext.stride = o.samples;
    }
    constexpr bool ASIAN = PK == PK_ASIAN || PK == PK_ASIAN_LOG;

    if constexpr (MODEL == OFG_MODEL_GBM) {
        // one GBM step on both legs, then the observation of path index j+1
        auto advance_w = [&](double z, double wj) {
            gbm_step<PK, RNG == OFG_RNG_FAST, REF>(ga, o, z, tab);
            if constexpr (ANTI) gbm_step<PK, RNG == OFG_RNG_FAST, REF>(gb, o, -z, tab);
            leg_observe<PK>(ga, wj, barrier);
            if constexpr (ANTI) leg_observe<PK>(gb, wj, barrier);
        };
        auto advance = [&](double z, int j) { advance_w(z, ASIAN ? w[j + 1] : 0.0); };
        if constexpr (RNG == OFG_RNG_FAST) {
            int j = 0;
            for (; j + 3 < steps; j += 4) {
                double z0, z1, z2, z3;
                normals4(rng, z0, z1, z2, z3, tab);
                advance(z0, j);
                advance(z1, j + 1);
                advance(z2, j + 2);
                advance(z3, j + 3);
            }
            if (j < steps) {
                double z0, z1, z2, z3;
                normals4(rng, z0, z1, z2, z3, tab);
                advance(z0, j);
                if (j + 1 < steps) advance(z1, j + 1);
                if (j + 2 < steps) advance(z2, j + 2);
            }
        } else if constexpr (REF) {
            // the reference's draws AK at a time (the inverse CDF with
            // its tail region evaluated once per wave), applied in order; a
            // block's observation weights are read together (the weights
            // array is padded past its last option)
            constexpr int AK = REF_AK;
            for (int j = 0; j < steps; j += AK) {
                double z[AK], wv[AK];
                acklam_block_wave<AK>(
                    [&](int i) { return j + i < steps ? uniform_open01(rng.next_f64()) : 0.5; },
                    z, slots);
#pragma unroll
                for (int i = 0; i < AK; ++i) wv[i] = ASIAN ? w[j + i + 1] : 0.0;
#pragma unroll
                for (int i = 0; i < AK; ++i)
                    if (j + i < steps) advance_w(z[i], wv[i]);
            }
        } else {
            for (int j = 0; j < steps; ++j) advance(ext.z1(j), j);
        }
    } else {
        double va = o.v0, vb = o.v0;
        auto advance_w = [&](double z1, double z2, double wj) {
            heston_step<PK, RNG != OFG_RNG_FAST, REF>(ga, va, o, z1, z2, tab);
            if constexpr (ANTI) heston_step<PK, RNG != OFG_RNG_FAST, REF>(gb, vb, o, -z1, -z2, tab);
            leg_observe<PK>(ga, wj, barrier);
            if constexpr (ANTI) leg_observe<PK>(gb, wj, barrier);
        };
        auto advance = [&](double z1, double z2, int j) {
            advance_w(z1, z2, ASIAN ? w[j + 1] : 0.0);
        };
        if constexpr (RNG == OFG_RNG_FAST) {
            int j = 0;
            for (; j + 1 < steps; j += 2) {
                double z0, z1, z2, z3;
                normals4(rng, z0, z1, z2, z3, tab);
                advance(z0, z1, j);
                advance(z2, z3, j + 1);
            }
            if (j < steps) {
                double z0, z1, z2, z3;
                normals4(rng, z0, z1, z2, z3, tab);
                advance(z0, z1, j);
            }
        } else if constexpr (REF) {
            // z1, z2 of AK / 2 steps per block, in the reference's order
            constexpr int AK = REF_AK;
            for (int j = 0; j < steps; j += AK / 2) {
                double z[AK], wv[AK / 2];
                acklam_block_wave<AK>(
                    [&](int i) {
                        return j + i / 2 < steps ? uniform_open01(rng.next_f64()) : 0.5;
                    },
                    z, slots);
#pragma unroll
                for (int i = 0; i < AK / 2; ++i) wv[i] = ASIAN ? w[j + i + 1] : 0.0;
#pragma unroll
                for (int i = 0; i < AK / 2; ++i)
                    if (j + i < steps) advance_w(z[2 * i], z[2 * i + 1], wv[i]);
            }
        } else {
            for (int j = 0; j < steps; ++j) advance(ext.z1(j), ext.z2(j), j);
        }
    }

    leg_payoff<PK>(ga, o, x, y);
    if constexpr (ANTI) {
        double xa, ya;
        leg_payoff<PK>(gb, o, xa, ya);
        x = 0.5 * (x + xa);
        y = 0.5 * (y + ya);
    }
}

struct Acc {
    double sx, sx2, sy, sxy, sy2;
};

template <int MODEL, int PK, int RNG, bool ANTI>
__device__ void run_item(const OptDev &o, const WorkItem &wi, int lane, const LaunchArgs &a,
                         const double *tab, double *slots, Acc &acc) {
    const cdouble *w = (const cdouble *)(a.obs_w + o.obs_offset);
    for (int64_t base = wi.begin; base < wi.end; base += WAVE) {
        const int64_t sample = base + lane;
        if (sample < wi.end) {
            double x, y;
            simulate_sample<MODEL, PK, RNG, ANTI>(o, sample, w, a, tab, slots, x, y);
            acc.sx += x;
            acc.sx2 += x * x;
            acc.sy += y;
            acc.sxy += x * y;
            acc.sy2 += y * y;
            if (a.xs) a.xs[o.out_offset + sample] = x;
            if (a.ys) a.ys[o.out_offset + sample] = y;
        }
    }
}

template <int MODEL, int RNG, bool ANTI>
__device__ __forceinline__ void dispatch_kind(const OptDev &o, const WorkItem &wi, int lane,
                                              const LaunchArgs &a, const double *tab,
                                              double *slots, Acc &acc) {
    switch (o.kind) {
    case PK_VANILLA: run_item<MODEL, PK_VANILLA, RNG, ANTI>(o, wi, lane, a, tab, slots, acc); break;
    case PK_ASIAN: run_item<MODEL, PK_ASIAN, RNG, ANTI>(o, wi, lane, a, tab, slots, acc); break;
    case PK_ASIAN_LOG:
        run_item<MODEL, PK_ASIAN_LOG, RNG, ANTI>(o, wi, lane, a, tab, slots, acc);
        break;
    case PK_BARRIER: run_item<MODEL, PK_BARRIER, RNG, ANTI>(o, wi, lane, a, tab, slots, acc); break;
    default:
        run_item<MODEL, PK_BARRIER_DOWN, RNG, ANTI>(o, wi, lane, a, tab, slots, acc);
        break;
    }
}

// Minimum waves per SIMD asked of the register allocator: the plain fast GBM
// and Heston loops fit 128 VGPRs (4 waves) without spilling (Heston takes 129
// left free); the reference-draw Heston loop is held to 168 (3 waves), which
// measured faster than its unconstrained 177 (2 waves).
template <int RNG, int MODEL, bool ANTI>
constexpr int price_waves_per_eu() {
    constexpr bool REF = RNG == OFG_RNG_REFERENCE || RNG == OFG_RNG_PCG64;
    return (RNG == OFG_RNG_FAST && !ANTI) ? 4 : (REF && MODEL == OFG_MODEL_HESTON) ? 3 : 1;
}
template <int RNG, int MODEL, bool ANTI, int BLOCK>
constexpr int price_waves_target() {
    // one-wave blocks are held to fewer waves by the LDS table anyway
    return BLOCK >= 128 ? price_waves_per_eu<RNG, MODEL, ANTI>() : 1;
}

template <int RNG, int MODEL, bool ANTI, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(
    price_waves_target<RNG, MODEL, ANTI, BLOCK>()))) void price_kernel(LaunchArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    // the fast mode's tables (log, exp2, sin/cos), or the reference draws'
    // 2^(j/256) table of the path exponentials, staged once per block in LDS
    constexpr bool REF = RNG == OFG_RNG_REFERENCE || RNG == OFG_RNG_PCG64;
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : REF ? 256 : 1];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    } else if constexpr (REF) {
        for (int i = threadIdx.x; i < 256; i += BLOCK) tab[i] = OFG_FAST_TABLE[OFG_TAB_EXP2 + i];
        __syncthreads();
    }
    // reference draws: each wave's slots for the tail values of one
    // inverse-CDF block (acklam_block_wave)
    constexpr int AK = REF_AK;
    __shared__ double tail_slots[REF ? BLOCK * AK : 1];
    double *slots = REF ? tail_slots + (threadIdx.x / WAVE) * WAVE * AK : tail_slots;
    const uint32_t span = (uint32_t)(a.order_end - a.order_begin);
    for (;;) {
        // Every lane executes the atomic (lane 0 adds 1, the others 0) so the
        // loop holds no divergent branch on the loop-invariant `lane == 0`;
        // with one, the compiler may unswitch the loop and leave lanes 1..63
        // in a copy that never sees the popped index.
        const uint32_t q = __builtin_amdgcn_readfirstlane(
            atomicAdd(a.queue_head, lane == 0 ? 1u : 0u));
        if (q >= span) break;
        const int32_t it = __builtin_amdgcn_readfirstlane(a.order[a.order_begin + (int32_t)q]);
        const WorkItem wi = a.items[it];
        const OptDev o = a.opts[wi.option];
        Acc acc{0.0, 0.0, 0.0, 0.0, 0.0};
        dispatch_kind<MODEL, RNG, ANTI>(o, wi, lane, a, tab, slots, acc);
        const double sx = wave_sum(acc.sx);
        const double sx2 = wave_sum(acc.sx2);
        const double sy = wave_sum(acc.sy);
        const double sxy = wave_sum(acc.sxy);
        const double sy2 = wave_sum(acc.sy2);
        if (lane == 0) {
            Partial p;
            p.sx = sx;
            p.sx2 = sx2;
            p.sy = sy;
            p.sxy = sxy;
            p.sy2 = sy2;
            p.n = (double)(wi.end - wi.begin);
            a.partials[it] = p;
        }
    }
}

__global__ void reset_queue_kernel(uint32_t *head, int32_t n) {
    if ((int32_t)threadIdx.x < n) head[threadIdx.x] = 0u;
}

// One wavefront per option: lanes stride over the option's item slots, then a
// fixed butterfly; lane 0 closes the statistics.
__global__ __launch_bounds__(256) void reduce_kernel(const OptDev *__restrict__ opts,
                                                     const int64_t *__restrict__ item_first,
                                                     const int32_t *__restrict__ item_count,
                                                     const Partial *__restrict__ partials,
                                                     OptOut *__restrict__ out, int32_t n_opts) {
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    if (gw >= n_opts) return;
    const int64_t first = item_first[gw];
    const int32_t count = item_count[gw];
    double sx = 0, sx2 = 0, sy = 0, sxy = 0, sy2 = 0, n = 0;
    for (int32_t i = lane; i < count; i += WAVE) {
        const Partial p = partials[first + i];
        sx += p.sx;
        sx2 += p.sx2;
        sy += p.sy;
        sxy += p.sxy;
        sy2 += p.sy2;
        n += p.n;
    }
    sx = wave_sum(sx);
    sx2 = wave_sum(sx2);
    sy = wave_sum(sy);
    sxy = wave_sum(sxy);
    sy2 = wave_sum(sy2);
    n = wave_sum(n);
    if (lane == 0) {
        const OptDev &o = opts[gw];
        OptOut r;
        close_stats(sx, sx2, sy, sxy, sy2, n, o.has_cv, o.cv_expected, o.df, &r.price, &r.stderr_);
        r.sx = sx;
        r.sx2 = sx2;
        r.sy = sy;
        r.sxy = sxy;
        r.sy2 = sy2;
        r.n = n;
        out[gw] = r;
    }
}

template <int RNG, int MODEL, bool ANTI>
hipError_t launch_variant(const LaunchArgs &a, hipStream_t stream) {
    switch (a.block) {
    case 64: price_kernel<RNG, MODEL, ANTI, 64><<<a.grid, 64, 0, stream>>>(a); break;
    case 128: price_kernel<RNG, MODEL, ANTI, 128><<<a.grid, 128, 0, stream>>>(a); break;
    case 512: price_kernel<RNG, MODEL, ANTI, 512><<<a.grid, 512, 0, stream>>>(a); break;
    default: price_kernel<RNG, MODEL, ANTI, 256><<<a.grid, 256, 0, stream>>>(a); break;
    }
    return hipGetLastError();
}

template <int RNG, int MODEL, bool ANTI>
const void *variant_ptr(int block) {
    switch (block) {
    case 64: return (const void *)price_kernel<RNG, MODEL, ANTI, 64>;
    case 128: return (const void *)price_kernel<RNG, MODEL, ANTI, 128>;
    case 512: return (const void *)price_kernel<RNG, MODEL, ANTI, 512>;
    default: return (const void *)price_kernel<RNG, MODEL, ANTI, 256>;
    }
}

template <int RNG>
hipError_t launch_rng(const LaunchArgs &a, hipStream_t stream) {
    if (a.model == OFG_MODEL_GBM) {
        return a.antithetic ? launch_variant<RNG, OFG_MODEL_GBM, true>(a, stream)
                            : launch_variant<RNG, OFG_MODEL_GBM, false>(a, stream);
    }
    return a.antithetic ? launch_variant<RNG, OFG_MODEL_HESTON, true>(a, stream)
                        : launch_variant<RNG, OFG_MODEL_HESTON, false>(a, stream);
}

template <int RNG>
const void *kernel_ptr(int block, int model, int anti) {
    if (model == OFG_MODEL_GBM)
        return anti ? variant_ptr<RNG, OFG_MODEL_GBM, true>(block)
                    : variant_ptr<RNG, OFG_MODEL_GBM, false>(block);
    return anti ? variant_ptr<RNG, OFG_MODEL_HESTON, true>(block)
                : variant_ptr<RNG, OFG_MODEL_HESTON, false>(block);
}

}  // namespace

hipError_t launch_reset_queue(uint32_t *head, int32_t n, hipStream_t stream) {
    reset_queue_kernel<<<1, 64, 0, stream>>>(head, n);
    return hipGetLastError();
}

hipError_t launch_pricing(const LaunchArgs &a, hipStream_t stream) {
    switch (a.rng_mode) {
    case OFG_RNG_EXTERNAL: return launch_rng<OFG_RNG_EXTERNAL>(a, stream);
    case OFG_RNG_FAST: return launch_rng<OFG_RNG_FAST>(a, stream);
    case OFG_RNG_PCG64: return launch_rng<OFG_RNG_PCG64>(a, stream);
    default: return launch_rng<OFG_RNG_REFERENCE>(a, stream);
    }
}

hipError_t launch_reduce(const OptDev *opts, const int64_t *item_first, const int32_t *item_count,
                         const Partial *partials, OptOut *out, int32_t n_opts,
                         hipStream_t stream) {
    if (n_opts <= 0) return hipSuccess;
    const int waves_per_block = 4;
    const int grid = (n_opts + waves_per_block - 1) / waves_per_block;
    reduce_kernel<<<grid, 256, 0, stream>>>(opts, item_first, item_count, partials, out, n_opts);
    return hipGetLastError();
}

int persistent_grid(int block, int rng_mode, int model, int antithetic, int device) {
    const void *k = rng_mode == OFG_RNG_EXTERNAL ? kernel_ptr<OFG_RNG_EXTERNAL>(block, model, antithetic)
                    : rng_mode == OFG_RNG_FAST   ? kernel_ptr<OFG_RNG_FAST>(block, model, antithetic)
                    : rng_mode == OFG_RNG_PCG64  ? kernel_ptr<OFG_RNG_PCG64>(block, model, antithetic)
                                                 : kernel_ptr<OFG_RNG_REFERENCE>(block, model, antithetic);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return per_cu * cus;
}

}  // namespace ofg
