// Monte Carlo Greeks estimators of European GBM vanillas for MI355X (gfx950):
// the MonteCarloGreeksEngine of src/engines/monte_carlo/mc_greeks.rs:131-314.
//
// Per sample one standard normal z (exact terminal GBM, one step over T) gives
// five contributions (single_path_contribution, mc_greeks.rs:245-284):
//   pathwise delta      1{S_T > K} S_T/S0 (calls; puts mirrored),
//   pathwise gamma      CRN difference of the pathwise deltas at spot +- h,
//   likelihood ratio    payoff x score weights for delta, gamma and vega.
// The kernel reuses the pricing path's layout: a persistent grid pops work
// items (sample ranges of one option) from the device queue, a lane handles
// one sample at a time in registers, a wave reduces its five sums per item
// into the item's slot, and one wave per option sums its slots in item order
// (independent of the queue schedule). The host scales by discount / n.
#include "ofg_adjoint.hpp"
#include "ofg_internal.hpp"
#include "ofg_math.hpp"
#include "ofg_mcgreeks.hpp"
#include "ofg_paths.hpp"

#include <type_traits>

namespace ofg {

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

// pathwise_delta, mc_greeks.rs:293-310
__device__ __forceinline__ double pw_delta(bool put, double st, double strike, double growth) {
    return put ? (st < strike ? -growth : 0.0) : (st > strike ? growth : 0.0);
}

// Per-option constants of run_estimators (mc_greeks.rs:168-181), read once per
// item; with one step over T the pricing descriptor's drift and diffusion are
// (r - q - vol^2/2) T and vol sqrt(T).
struct McgConst {
    double spot, strike, vol, sqrt_t, spot_up, spot_dn, spot_span, inv_svt;
    bool put;
};

__device__ __forceinline__ McgConst mcg_const(const OptDev &o) {
    McgConst c;
    c.spot = o.s0;
    c.strike = o.strike;
    c.vol = o.sigma;
    c.sqrt_t = o.sqrt_dt;
    c.spot_up = o.spot_up;
    c.spot_dn = o.spot_dn;
    c.spot_span = o.spot_up - o.spot_dn;
    c.inv_svt = 1.0 / (o.s0 * o.sigma * o.sqrt_dt);
    c.put = o.flags & F_PUT;
    return c;
}

// single_path_contribution (mc_greeks.rs:245-284) in the reference's
// operation order (the file is compiled with -ffp-contract=off).
template <bool FAST>
__device__ __forceinline__ void contribution(const OptDev &o, const McgConst &c, double z,
                                             const double *tab, double e[5]) {
    double x_unused;
    const double growth = gbm_growth<FAST>(o, z, tab, x_unused);
    const double st = c.spot * growth;
    const double payoff = vanilla_payoff(c.put, st, c.strike);
    const double delta_pw = pw_delta(c.put, st, c.strike, growth);
    const double delta_up = pw_delta(c.put, c.spot_up * growth, c.strike, growth);
    const double delta_dn = pw_delta(c.put, c.spot_dn * growth, c.strike, growth);
    const double z2m1 = z * z - 1.0;
    const double w_delta = z * c.inv_svt;
    const double w_gamma = z2m1 * (c.inv_svt * c.inv_svt) - z * c.inv_svt / c.spot;
    const double w_vega = (z2m1 - c.vol * c.sqrt_t * z) / c.vol;
    e[0] = delta_pw;
    e[1] = (delta_up - delta_dn) / c.spot_span;
    e[2] = payoff * w_delta;
    e[3] = payoff * w_gamma;
    e[4] = payoff * w_vega;
}

template <int RNG, bool ANTI>
__device__ void mcg_item(const OptDev &o, const WorkItem &wi, int lane, const AdjArgs &a,
                         const double *tab, double acc[5]) {
    const McgConst c = mcg_const(o);
    for (int64_t base = wi.begin; base < wi.end; base += WAVE) {
        const int64_t sample = base + lane;
        if (sample >= wi.end) continue;
        double z;
        if constexpr (RNG == OFG_RNG_REFERENCE) {
            Xoshiro256pp rng;
            rng.seed(stream_seed(o.seed, (uint64_t)sample));
            z = reference_normal(rng);
        } else if constexpr (RNG == OFG_RNG_PCG64) {
            Pcg64 rng;
            rng.seed(stream_seed(o.seed, (uint64_t)sample));
            z = acklam_inv_cdf(uniform_open01(rng.next_f64()));
        } else if constexpr (RNG == OFG_RNG_FAST) {
            Xoroshiro128p rng;
            rng.seed(stream_seed(o.seed, (uint64_t)sample));
            const uint64_t wa = rng.next_u64();
            const uint64_t wb = rng.next_u64();
            double z1_unused;
            box_muller_pair(wa, (uint32_t)(wb >> 32), z, z1_unused, tab);
        } else {
            z = a.normals1[o.normals_offset + sample];
        }
        double e[5];
        contribution<RNG == OFG_RNG_FAST>(o, c, z, tab, e);
        if constexpr (ANTI) {
            double b[5];
            contribution<RNG == OFG_RNG_FAST>(o, c, -z, tab, b);
#pragma unroll
            for (int k = 0; k < 5; ++k) e[k] = 0.5 * (e[k] + b[k]);
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] += e[k];
    }
}

template <int RNG, bool ANTI, int BLOCK>
__global__ __launch_bounds__(BLOCK) void mc_greeks_kernel(AdjArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const uint32_t span = (uint32_t)(a.order_end - a.order_begin);
    for (;;) {
        // every lane takes part in the atomic (see price_kernel)
        const uint32_t q = __builtin_amdgcn_readfirstlane(
            atomicAdd(a.queue_head, lane == 0 ? 1u : 0u));
        if (q >= span) break;
        const int32_t it = __builtin_amdgcn_readfirstlane(a.order[a.order_begin + (int32_t)q]);
        const WorkItem wi = a.items[it];
        const OptDev o = a.opts[wi.option];
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        mcg_item<RNG, ANTI>(o, wi, lane, a, tab, acc);
        AdjPartial p;
        p.spv = 0.0;
        p.spv2 = 0.0;
#pragma unroll
        for (int k = 0; k < 5; ++k) p.g[k] = wave_sum(acc[k]);
#pragma unroll
        for (int k = 5; k < ADJ_NG; ++k) p.g[k] = 0.0;
        p.n = (double)(wi.end - wi.begin);
        if (lane == 0) a.partials[it] = p;
    }
}

// One wavefront per option: raw sums of the item slots in item order.
__global__ __launch_bounds__(256) void mc_greeks_reduce_kernel(
    const int64_t *__restrict__ item_first, const int32_t *__restrict__ item_count,
    const AdjPartial *__restrict__ partials, AdjOut *__restrict__ out, int32_t n_opts) {
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    if (gw >= n_opts) return;
    const int64_t first = item_first[gw];
    const int32_t count = item_count[gw];
    double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, n = 0.0;
    for (int32_t i = lane; i < count; i += WAVE) {
        const AdjPartial &p = partials[first + i];
#pragma unroll
        for (int k = 0; k < 5; ++k) g[k] += p.g[k];
        n += p.n;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) g[k] = wave_sum(g[k]);
    n = wave_sum(n);
    if (lane == 0) {
        AdjOut r{};
#pragma unroll
        for (int k = 0; k < 5; ++k) r.g[k] = g[k];
        r.n = n;
        out[gw] = r;
    }
}

template <int RNG, bool ANTI>
hipError_t launch_block(const AdjArgs &a, hipStream_t s) {
    switch (a.block) {
    case 64: mc_greeks_kernel<RNG, ANTI, 64><<<a.grid, 64, 0, s>>>(a); break;
    case 128: mc_greeks_kernel<RNG, ANTI, 128><<<a.grid, 128, 0, s>>>(a); break;
    default: mc_greeks_kernel<RNG, ANTI, 256><<<a.grid, 256, 0, s>>>(a); break;
    }
    return hipGetLastError();
}

template <int RNG, bool ANTI>
const void *block_ptr(int block) {
    switch (block) {
    case 64: return (const void *)mc_greeks_kernel<RNG, ANTI, 64>;
    case 128: return (const void *)mc_greeks_kernel<RNG, ANTI, 128>;
    default: return (const void *)mc_greeks_kernel<RNG, ANTI, 256>;
    }
}

template <int RNG>
hipError_t launch_anti(const AdjArgs &a, hipStream_t s) {
    return a.antithetic ? launch_block<RNG, true>(a, s) : launch_block<RNG, false>(a, s);
}

template <int RNG>
const void *anti_ptr(int block, int anti) {
    return anti ? block_ptr<RNG, true>(block) : block_ptr<RNG, false>(block);
}

}  // namespace

hipError_t launch_mc_greeks(const AdjArgs &a, hipStream_t s) {
    switch (a.rng_mode) {
    case OFG_RNG_EXTERNAL: return launch_anti<OFG_RNG_EXTERNAL>(a, s);
    case OFG_RNG_FAST: return launch_anti<OFG_RNG_FAST>(a, s);
    case OFG_RNG_PCG64: return launch_anti<OFG_RNG_PCG64>(a, s);
    default: return launch_anti<OFG_RNG_REFERENCE>(a, s);
    }
}

hipError_t launch_mc_greeks_reduce(const int64_t *item_first, const int32_t *item_count,
                                   const AdjPartial *partials, AdjOut *out, int32_t n_opts,
                                   hipStream_t s) {
    if (n_opts <= 0) return hipSuccess;
    const int grid = (n_opts + 3) / 4;
    mc_greeks_reduce_kernel<<<grid, 256, 0, s>>>(item_first, item_count, partials, out, n_opts);
    return hipGetLastError();
}

int mc_greeks_grid(int block, int rng_mode, int antithetic, int device) {
    const void *k = rng_mode == OFG_RNG_EXTERNAL ? anti_ptr<OFG_RNG_EXTERNAL>(block, antithetic)
                    : rng_mode == OFG_RNG_FAST   ? anti_ptr<OFG_RNG_FAST>(block, antithetic)
                    : rng_mode == OFG_RNG_PCG64  ? anti_ptr<OFG_RNG_PCG64>(block, antithetic)
                                                 : anti_ptr<OFG_RNG_REFERENCE>(block, antithetic);
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
