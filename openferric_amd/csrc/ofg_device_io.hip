// Device-resident inputs and outputs (ofg_price_device): the staging kernel
// forms each option's descriptor fields from the caller's device arrays as
// the host plan builder does from ofg_option (ofg_plan.cpp build_plan:
// GbmPathGenerator::generate_into constants, simulation.rs:59-62; the
// Heston step constants; mu = r - q, mc_engine.rs:916; the discount
// exp(-rT), mc_engine.rs:945) and validates them with the reference's rules;
// the scatter kernels write the reduced results to the caller's arrays.
#include "ofg_device_io.hpp"
#include "ofg_math.hpp"

namespace ofg {

namespace {

constexpr int IO_BLOCK = 256;

__device__ __forceinline__ double pick(const double *p, int32_t i, double keep) {
    return p ? p[i] : keep;
}

__device__ __forceinline__ void report(int32_t *err, int32_t code, int32_t index) {
    if (atomicCAS(err, 0, code) == 0) err[1] = index;
}

__global__ __launch_bounds__(IO_BLOCK) void stage_kernel(StageArgs a, ofg_device_io io) {
    const int32_t k = blockIdx.x * IO_BLOCK + threadIdx.x;
    if (k >= a.n_plan) return;
    const int32_t i = a.active[k];
    OptDev d = a.opts[k];
    const double spot = pick(io.spot, i, d.s0);
    const double strike = pick(io.strike, i, d.strike);
    const double expiry = pick(io.expiry, i, a.host[2 * k + 1]);
    const double rate = pick(io.rate, i, d.rate);
    const double q = a.host[2 * k];
    const double vol = pick(io.vol, i, d.sigma);
    const double kappa = pick(io.kappa, i, d.kappa), theta = pick(io.theta, i, d.theta);
    const double xi = pick(io.xi, i, d.xi), rho = pick(io.rho, i, d.rho), v0 = pick(io.v0, i, d.v0);
    const bool heston = d.model == OFG_MODEL_HESTON;
    const double vals[10] = {spot, strike, expiry, rate, vol, kappa, theta, xi, rho, v0};
    for (int j = 0; j < 10; ++j)
        if (!isfinite(vals[j])) {
            report(a.err, DIE_NOT_FINITE, i);
            return;
        }
    if (!(strike > 0.0)) report(a.err, DIE_STRIKE, i);
    if (!(expiry > 0.0)) report(a.err, DIE_EXPIRY, i);
    if (!heston && !(vol > 0.0)) report(a.err, DIE_VOL, i);
    // Heston::validate (src/models/stochastic.rs:47-54)
    if (heston && !(kappa > 0.0 && theta >= 0.0 && xi >= 0.0 && v0 >= 0.0 && rho > -1.0 && rho < 1.0))
        report(a.err, DIE_HESTON, i);
    d.s0 = spot;
    d.strike = strike;
    d.mu = rate - q;
    d.sigma = vol;
    d.dt = expiry / (double)d.steps;
    d.sqrt_dt = sqrt(d.dt);
    d.drift = (d.mu - 0.5 * d.sigma * d.sigma) * d.dt;
    d.diffusion = d.sigma * d.sqrt_dt;
    d.kappa = kappa;
    d.theta = theta;
    d.xi = xi;
    d.rho = rho;
    d.rho_c = sqrt(1.0 - rho * rho);
    d.v0 = v0;
    d.log_s0 = log(fmax(spot, 1e-12));
    d.df = exp(-rate * expiry);
    d.rate = rate;
    d.kdt = kappa * d.dt;
    d.kthdt = kappa * theta * d.dt;
    d.xsdt = xi * d.sqrt_dt;
    d.mudt = d.mu * d.dt;
    d.hdt = 0.5 * d.dt;
    d.drift2 = d.drift * OFG_INV_LN2_256;
    d.diffusion2 = d.diffusion * OFG_INV_LN2_256;
    d.sqrt_dt2 = d.sqrt_dt * OFG_INV_LN2_256;
    d.mudt2 = d.mudt * OFG_INV_LN2_256;
    d.hdt2 = d.hdt * OFG_INV_LN2_256;
    if (a.rng_mode == OFG_RNG_FAST && !heston &&
        !(fabs(d.drift2) + 8.6 * d.diffusion2 < OFG_EXP2_256_YMAX))
        report(a.err, DIE_FAST_RANGE, i);
    a.opts[k] = d;
}

__global__ __launch_bounds__(IO_BLOCK) void scatter_price_kernel(const OptOut *__restrict__ outs,
                                                                 const int32_t *__restrict__ active,
                                                                 int32_t n_plan, ofg_device_io io,
                                                                 const int32_t *err) {
    const int32_t k = blockIdx.x * IO_BLOCK + threadIdx.x;
    if (k >= n_plan || err[0] != DIE_NONE) return;   // invalid inputs: nothing written
    const int32_t i = active[k];
    io.price[i] = outs[k].price;
    if (io.stderr_) io.stderr_[i] = outs[k].stderr_;
}

__global__ __launch_bounds__(IO_BLOCK) void scatter_adjoint_kernel(const AdjOut *__restrict__ outs,
                                                                   const int32_t *__restrict__ active,
                                                                   int32_t n_plan, ofg_device_io io,
                                                                   const int32_t *err) {
    const int32_t k = blockIdx.x * IO_BLOCK + threadIdx.x;
    if (k >= n_plan || err[0] != DIE_NONE) return;   // invalid inputs: nothing written
    const int32_t i = active[k];
    io.price[i] = outs[k].price;
    if (io.stderr_) io.stderr_[i] = outs[k].stderr_;
    if (io.grads)
        for (int g = 0; g < ADJ_NG; ++g) io.grads[(int64_t)i * ADJ_NG + g] = outs[k].g[g];
}

int blocks_for(int32_t n) { return (n + IO_BLOCK - 1) / IO_BLOCK; }

}  // namespace

hipError_t launch_stage(const StageArgs &a, const ofg_device_io &io, hipStream_t stream) {
    if (a.n_plan <= 0) return hipSuccess;
    hipLaunchKernelGGL(stage_kernel, dim3(blocks_for(a.n_plan)), dim3(IO_BLOCK), 0, stream, a, io);
    return hipGetLastError();
}

hipError_t launch_scatter_price(const OptOut *outs, const int32_t *active, int32_t n_plan,
                                const ofg_device_io &io, const int32_t *err, hipStream_t stream) {
    if (n_plan <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_price_kernel, dim3(blocks_for(n_plan)), dim3(IO_BLOCK), 0, stream,
                       outs, active, n_plan, io, err);
    return hipGetLastError();
}

hipError_t launch_scatter_adjoint(const AdjOut *outs, const int32_t *active, int32_t n_plan,
                                  const ofg_device_io &io, const int32_t *err, hipStream_t stream) {
    if (n_plan <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_adjoint_kernel, dim3(blocks_for(n_plan)), dim3(IO_BLOCK), 0, stream,
                       outs, active, n_plan, io, err);
    return hipGetLastError();
}

}  // namespace ofg
