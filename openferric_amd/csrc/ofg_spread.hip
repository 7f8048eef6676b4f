// Spread options max(S1_T - S2_T - K, 0) under two correlated GBMs for
// MI355X (gfx950): SpreadMonteCarloEngine::price of
// src/engines/monte_carlo/spread_mc.rs:83-200.
//
// Every sample i draws z1 then z2 from its own stream seeded with
// stream_seed(seed, i) (resolve_stream_seed, reproducible), so samples are
// independent work: a thread takes SPREAD_PER samples of one option, 64
// apart (a wave covers 64 * SPREAD_PER consecutive samples and writes the
// optional per-sample payoffs coalesced). Reference and Pcg64 draws are
// inverted two samples (four uniforms) per acklam_block, whose rare tail
// branch runs once per block; fast draws are one Box-Muller pair per sample.
// Per-wave (sum, sum_sq) slots are summed per option in wave order, so the
// result does not depend on the schedule.
#include "../../include/openferric_amd.h"
#include "ofg_math.hpp"
#include "ofg_spread.hpp"

#include <type_traits>

namespace ofg {

namespace {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

// terminal_spread_payoff (spread_mc.rs:178-197)
__device__ __forceinline__ double spread_payoff(const SpreadJob &j, double z1, double z2) {
    const double w2 = fma(j.rho, z1, j.corr_tail * z2);
    const double s1t = j.s1 * exp_fp64(fma(j.vol_term1, z1, j.drift1));
    const double s2t = j.s2 * exp_fp64(fma(j.vol_term2, w2, j.drift2));
    return fmax(s1t - s2t - j.k, 0.0);
}

template <bool ANTI>
__device__ __forceinline__ double sample_value(const SpreadJob &j, double z1, double z2) {
    const double x = spread_payoff(j, z1, z2);
    if constexpr (ANTI) return 0.5 * (x + spread_payoff(j, -z1, -z2));
    return x;
}

template <int RNG, bool ANTI>
__global__ __launch_bounds__(BLOCK) void spread_kernel(const SpreadJob *__restrict__ jobs,
                                                       int32_t n_jobs, int64_t total_waves,
                                                       SpreadSums *__restrict__ wave_sums,
                                                       double *__restrict__ payoffs) {
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t wave = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) / WAVE;
    if (wave >= total_waves) return;   // whole waves only: the shuffles below stay uniform
    int lo = 0, hi = n_jobs - 1;        // last job with wave_first <= wave
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].wave_first <= wave) lo = mid;
        else hi = mid - 1;
    }
    const SpreadJob j = jobs[lo];
    const int64_t first = (wave - j.wave_first) * (WAVE * SPREAD_PER) + lane;
    double *pay = payoffs ? payoffs + j.pay_offset : nullptr;
    double sum = 0.0, sum_sq = 0.0;
    auto take = [&](int64_t i, double x) {
        sum += x;
        sum_sq += x * x;
        if (pay) pay[i] = x;
    };
    if constexpr (RNG == OFG_RNG_FAST) {
#pragma unroll
        for (int q = 0; q < SPREAD_PER; ++q) {
            const int64_t i = first + (int64_t)q * WAVE;
            if (i < j.samples) {
                Xoroshiro128p g;
                g.seed(stream_seed(j.seed, (uint64_t)i));
                const uint64_t wa = g.next_u64();
                const uint64_t wb = g.next_u64();
                double z1, z2;
                box_muller_pair(wa, (uint32_t)(wb >> 32), z1, z2, tab);
                take(i, sample_value<ANTI>(j, z1, z2));
            }
        }
    } else {
        using Gen = typename std::conditional<RNG == OFG_RNG_PCG64, Pcg64, Xoshiro256pp>::type;
#pragma unroll
        for (int q = 0; q < SPREAD_PER; q += 2) {
            const int64_t ia = first + (int64_t)q * WAVE, ib = ia + WAVE;
            const bool la = ia < j.samples, lb = ib < j.samples;
            double p[4] = {0.5, 0.5, 0.5, 0.5}, z[4];
            if (la) {
                Gen g;
                g.seed(stream_seed(j.seed, (uint64_t)ia));
                p[0] = uniform_open01(g.next_f64());
                p[1] = uniform_open01(g.next_f64());
            }
            if (lb) {
                Gen g;
                g.seed(stream_seed(j.seed, (uint64_t)ib));
                p[2] = uniform_open01(g.next_f64());
                p[3] = uniform_open01(g.next_f64());
            }
            acklam_block<4>(p, z);
            if (la) take(ia, sample_value<ANTI>(j, z[0], z[1]));
            if (lb) take(ib, sample_value<ANTI>(j, z[2], z[3]));
        }
    }
    sum = wave_sum(sum);
    sum_sq = wave_sum(sum_sq);
    if (lane == 0) wave_sums[wave] = SpreadSums{sum, sum_sq};
}

// One wave per job: its wave slots in wave order (fixed butterfly).
__global__ __launch_bounds__(BLOCK) void spread_reduce_kernel(
    const SpreadJob *__restrict__ jobs, int32_t n_jobs, const SpreadSums *__restrict__ wave_sums,
    SpreadSums *__restrict__ job_sums) {
    const int gw = (int)((blockIdx.x * BLOCK + threadIdx.x) / WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    if (gw >= n_jobs) return;
    const SpreadJob &j = jobs[gw];
    double s = 0.0, s2 = 0.0;
    for (int32_t i = lane; i < j.wave_count; i += WAVE) {
        const SpreadSums w = wave_sums[j.wave_first + i];
        s += w.sum;
        s2 += w.sum_sq;
    }
    s = wave_sum(s);
    s2 = wave_sum(s2);
    if (lane == 0) job_sums[gw] = SpreadSums{s, s2};
}

template <int RNG>
void launch_rng(unsigned grid, bool anti, const SpreadJob *jobs, int32_t n_jobs, int64_t waves,
                SpreadSums *wave_sums, double *payoffs, hipStream_t s) {
    if (anti) spread_kernel<RNG, true><<<grid, BLOCK, 0, s>>>(jobs, n_jobs, waves, wave_sums, payoffs);
    else spread_kernel<RNG, false><<<grid, BLOCK, 0, s>>>(jobs, n_jobs, waves, wave_sums, payoffs);
}

}  // namespace

hipError_t launch_spread(const SpreadJob *jobs, int32_t n_jobs, int64_t total_waves,
                         int rng_mode, bool antithetic, SpreadSums *wave_sums,
                         SpreadSums *job_sums, double *payoffs, hipStream_t stream) {
    if (n_jobs <= 0 || total_waves <= 0) return hipSuccess;
    const unsigned grid = (unsigned)((total_waves * WAVE + BLOCK - 1) / BLOCK);
    switch (rng_mode) {
    case OFG_RNG_FAST:
        launch_rng<OFG_RNG_FAST>(grid, antithetic, jobs, n_jobs, total_waves, wave_sums, payoffs,
                                 stream);
        break;
    case OFG_RNG_PCG64:
        launch_rng<OFG_RNG_PCG64>(grid, antithetic, jobs, n_jobs, total_waves, wave_sums, payoffs,
                                  stream);
        break;
    default:
        launch_rng<OFG_RNG_REFERENCE>(grid, antithetic, jobs, n_jobs, total_waves, wave_sums,
                                      payoffs, stream);
        break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    spread_reduce_kernel<<<(unsigned)((n_jobs + 3) / 4), BLOCK, 0, stream>>>(jobs, n_jobs,
                                                                             wave_sums, job_sums);
    return hipGetLastError();
}

}  // namespace ofg
