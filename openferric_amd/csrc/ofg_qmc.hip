// Quasi-Monte Carlo European pricer under GBM for MI355X (gfx950), FP64 on
// the VALU: mc_european_qmc_with_seed (src/engines/monte_carlo/mc_qmc.rs:
// 40-144) with SobolSequence points (src/math/sobol.rs), each path's point
// formed directly from its index (Gray-code identity, ofg_qmc.hpp) instead of
// the reference's sequential next_into. The uniform map fma(x, 2^-64,
// 2^-65), uniform_open01, the Acklam inverse CDF and the step
// spot *= exp(fma(dt_vol, z, dt_drift)) follow the reference's operations
// (exp within about an ulp of the reference's).
//
// Work: a tile is one block of 256 threads over 256 * ppt consecutive point
// indices of one option; tile partial sums (sum, sum of squares) go to the
// tile's own slot and a close kernel per option sums them in tile order.
#include "ofg_qmc.hpp"
#include "ofg_math.hpp"

namespace ofg {

namespace {

constexpr int WAVE = 64;
constexpr int NWAVES = QMC_BLOCK / WAVE;
constexpr int QMC_AK = 8;   // inverse-CDF values per wave-tail block

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

__global__ __launch_bounds__(QMC_BLOCK) void qmc_kernel(QmcArgs a) {
    __shared__ uint64_t xhi[NWAVES][QMC_SEG];
    __shared__ double red[NWAVES][2];
    // each wave's slots for the tail values of an inverse-CDF block, and the
    // 2^(j/256) table of the path exponentials
    __shared__ double tail_slots[QMC_BLOCK * QMC_AK];
    __shared__ double exp2_tab[256];
    for (int i = threadIdx.x; i < 256; i += QMC_BLOCK) exp2_tab[i] = OFG_FAST_TABLE[OFG_TAB_EXP2 + i];
    __syncthreads();
    const int tile = blockIdx.x;
    const int opt = a.tile_opt[tile];
    const QmcDev o = a.opts[opt];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    double *slots = tail_slots + wave * WAVE * QMC_AK;
    const bool put = o.put != 0;
    const uint64_t *dirs = a.dirs + o.dim_offset * 64;
    const uint64_t *lows = a.lows + o.dim_offset * 64;
    const uint64_t *scr = a.scr + o.dim_offset;
    const double inv = 1.0 / 18446744073709551616.0;   // 2^-64
    const double half = 0.5 * inv;
    double sum = 0.0, sum_sq = 0.0;
    const int64_t base = (int64_t)(tile - o.tile_first) * QMC_BLOCK * o.ppt;
    for (int j = 0; j < o.ppt; ++j) {
        const uint64_t k = (uint64_t)(base + (int64_t)j * QMC_BLOCK + threadIdx.x);
        const bool valid = k >= 1 && k <= (uint64_t)o.paths;
        const uint64_t g = k ^ (k >> 1);
        // bits 6.. of g are the same on every lane of the wave
        const uint32_t hlo = __builtin_amdgcn_readfirstlane((uint32_t)(g >> 6));
        const uint32_t hhi = __builtin_amdgcn_readfirstlane((uint32_t)(g >> 38));
        const uint64_t hi = ((uint64_t)hhi << 32) | hlo;
        const int low = (int)(g & 63u);
        double spot = o.s0;
        for (int d0 = 0; d0 < o.steps; d0 += QMC_SEG) {
            const int seg = min(QMC_SEG, o.steps - d0);
            // the wave's high-bit XOR for each dimension of the segment
            for (int dd = lane; dd < seg; dd += WAVE) {
                const uint64_t *v = dirs + (int64_t)(d0 + dd) * 64 + 6;
                uint64_t x = 0;
                for (uint64_t h = hi; h; h &= h - 1) x ^= v[__builtin_ctzll(h)];
                xhi[wave][dd] = x;
            }
            __syncthreads();
            // QMC_AK dimensions at a time: their uniforms, the inverse CDF
            // with the tail values inverted per wave, then the steps in
            // order; exp by the 2^(j/256) table (within about an ulp, as
            // exp_fp64), its argument clamped to the table's domain
            for (int db = 0; db < seg; db += QMC_AK) {
                double z[QMC_AK];
                acklam_block_wave<QMC_AK>(
                    [&](int i) {
                        const int dd = min(db + i, seg - 1);
                        const int d = d0 + dd;
                        const uint64_t x = xhi[wave][dd] ^ lows[(int64_t)d * 64 + low] ^ scr[d];
                        return uniform_open01(fma((double)x, inv, half));
                    },
                    z, slots);
#pragma unroll
                for (int i = 0; i < QMC_AK; ++i) {
                    if (db + i < seg) {
                        const double y = fma(o.dt_vol, z[i], o.dt_drift) * OFG_INV_LN2_256;
                        spot *= exp2_256t(fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX),
                                          exp2_tab);
                    }
                }
            }
            __syncthreads();
        }
        if (valid) {
            const double px = put ? fmax(o.strike - spot, 0.0) : fmax(spot - o.strike, 0.0);
            sum += px;
            sum_sq += px * px;
        }
    }
    sum = wave_sum(sum);
    sum_sq = wave_sum(sum_sq);
    if (lane == 0) {
        red[wave][0] = sum;
        red[wave][1] = sum_sq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0][0], q = red[0][1];
        for (int w = 1; w < NWAVES; ++w) {
            s += red[w][0];
            q += red[w][1];
        }
        a.partials[tile].sum = s;
        a.partials[tile].sum_sq = q;
    }
}

// One block per option: tile slots in tile order, then the close of
// mc_qmc.rs:127-143 (variance clamped at 0, discount applied to both).
__global__ __launch_bounds__(QMC_BLOCK) void qmc_close_kernel(QmcArgs a) {
    __shared__ double red[NWAVES][2];
    const int opt = blockIdx.x;
    const QmcDev o = a.opts[opt];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    double s = 0.0, q = 0.0;
    for (int t = threadIdx.x; t < o.n_tiles; t += QMC_BLOCK) {
        s += a.partials[o.tile_first + t].sum;
        q += a.partials[o.tile_first + t].sum_sq;
    }
    s = wave_sum(s);
    q = wave_sum(q);
    if (lane == 0) {
        red[wave][0] = s;
        red[wave][1] = q;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double sum = red[0][0], sum_sq = red[0][1];
    for (int w = 1; w < NWAVES; ++w) {
        sum += red[w][0];
        sum_sq += red[w][1];
    }
    const double n = (double)o.paths;
    const double mean = sum / n;
    const double variance = o.paths > 1 ? fmax((sum_sq - sum * sum / n) / (n - 1.0), 0.0) : 0.0;
    a.out[opt].price = o.discount * mean;
    a.out[opt].stderr_ = o.discount * sqrt(variance / n);
}

}  // namespace

hipError_t launch_qmc(const QmcArgs &a, hipStream_t stream) {
    if (a.n_tiles <= 0) return hipSuccess;
    qmc_kernel<<<a.n_tiles, QMC_BLOCK, 0, stream>>>(a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    qmc_close_kernel<<<a.n_opts, QMC_BLOCK, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace ofg
