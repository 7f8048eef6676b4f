// Exact-terminal European pricers of the reference's default build
// (src/engines/monte_carlo/mc_parallel.rs:98-148, 264-434;
// mc_engine.rs:186-343) on the GPU, per path on the reference's draws: each
// pricer draws ONE standard normal per path from a single xoshiro256++
// stream (one per chunk for mc_european_parallel), so thread g of a stream
// jumps g * 64 draws ahead (GF(2) polynomial jump, host-computed powers of x)
// and prices 64 consecutive paths: S_T = S0 exp(fma(total_diffusion, z,
// total_drift)), payoff max(+-(S_T - K), 0). Sums are reduced per wave in a
// fixed order, then per stream.
#include "ofg_exact.hpp"
#include "ofg_math.hpp"

namespace ofg {

namespace {

constexpr int WAVE = 64;
constexpr int CHUNK = 64;   // paths per thread

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

__global__ __launch_bounds__(256) void exact_kernel(const ExactJob *__restrict__ jobs,
                                                    int32_t n_jobs, int64_t total,
                                                    const uint64_t *__restrict__ polys,
                                                    ExactSums *__restrict__ wave_sums,
                                                    double *__restrict__ payoffs) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (WAVE - 1);
    double sum = 0.0, sum_sq = 0.0;
    if (gid < total) {
        int lo = 0, hi = n_jobs - 1;   // last job with thread_first <= gid
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (jobs[mid].seed.thread_first <= gid) lo = mid;
            else hi = mid - 1;
        }
        const ExactJob &j = jobs[lo];
        const int64_t g = gid - j.seed.thread_first;
        const int64_t first = g * CHUNK;
        if (first < j.seed.paths) {
            Xoshiro256pp st;
            st.s0 = j.seed.base[0];
            st.s1 = j.seed.base[1];
            st.s2 = j.seed.base[2];
            st.s3 = j.seed.base[3];
            for (int k = 0; k < j.seed.n_polys; ++k)
                if ((g >> k) & 1) xoshiro_apply_poly(polys + 4 * (j.seed.poly_offset + k), st);
            const int cnt = (int)min<int64_t>(CHUNK, j.seed.paths - first);
            const double spot = j.spot, strike = j.strike, drift = j.drift, diff = j.diffusion;
            const bool put = j.put != 0;
            double *pay = payoffs ? payoffs + j.pay_offset + first : nullptr;
            reference_steps<8, 1>(st, cnt, [&](int q, double z, double) {
                const double s = spot * exp_fp64(fma(diff, z, drift));
                const double px = put ? fmax(strike - s, 0.0) : fmax(s - strike, 0.0);
                sum += px;
                sum_sq += px * px;
                if (pay) pay[q] = px;
            });
        }
    }
    sum = wave_sum(sum);
    sum_sq = wave_sum(sum_sq);
    if (lane == 0 && gid < total) wave_sums[gid / WAVE] = ExactSums{sum, sum_sq};
}

// One wave per job: its waves' sums in wave order (fixed butterfly).
__global__ __launch_bounds__(256) void exact_reduce_kernel(const ExactJob *__restrict__ jobs,
                                                           int32_t n_jobs,
                                                           const ExactSums *__restrict__ wave_sums,
                                                           ExactSums *__restrict__ job_sums) {
    const int gw = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    if (gw >= n_jobs) return;
    const ExactJob &j = jobs[gw];
    double s = 0.0, s2 = 0.0;
    for (int32_t i = lane; i < j.wave_count; i += WAVE) {
        const ExactSums w = wave_sums[j.wave_first + i];
        s += w.sum;
        s2 += w.sum_sq;
    }
    s = wave_sum(s);
    s2 = wave_sum(s2);
    if (lane == 0) job_sums[gw] = ExactSums{s, s2};
}

}  // namespace

hipError_t launch_exact(const ExactJob *jobs, int32_t n_jobs, int64_t total_threads,
                        const uint64_t *polys, ExactSums *wave_sums, ExactSums *job_sums,
                        double *payoffs, hipStream_t stream) {
    if (n_jobs <= 0 || total_threads <= 0) return hipSuccess;
    const int64_t grid = (total_threads + 255) / 256;
    exact_kernel<<<(unsigned)grid, 256, 0, stream>>>(jobs, n_jobs, total_threads, polys,
                                                      wave_sums, payoffs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    exact_reduce_kernel<<<(unsigned)((n_jobs + 3) / 4), 256, 0, stream>>>(jobs, n_jobs,
                                                                        wave_sums, job_sums);
    return hipGetLastError();
}

}  // namespace ofg
