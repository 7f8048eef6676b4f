// Multi-asset path kernels for MI355X (gfx950), FP64 on the VALU: worst-of
// autocallables and Phoenix notes (simulate_autocallable_paths,
// src/pricing/autocallable.rs:441-579) and basket options under a Gaussian
// copula (simulate_basket_once_matrix, src/pricing/basket.rs:599-683).
//
// Per path, the underlyings' spots live in registers (up to MA_MAX); each
// step draws one normal per underlying, correlates them with the scenario's
// lower Cholesky factor (correlate_normals, :659-667: a running sum from 0
// in column order), moves each spot by exp(drift + vol_dt * corr) floored at
// 1e-12, takes the worst performance s_i / s0_i, records a knock-in
// (worst <= ki_barrier) and, on an observation step, pays Phoenix coupons
// (with memory) and calls the note when worst >= autocall_barrier. A called
// lane leaves the step loop; the wave continues with the rest.
//
// Draws: external normals [step][asset][path] of the contract (the same for
// all of its scenarios), or fast per-path xoroshiro128+ Box-Muller streams
// seeded by stream_seed(seed, path) (one pair per two underlyings per step).
// Reductions: per-tile (sum, sum of squares) of pv - shift in the tile's own
// slot, then one block per scenario sums them in tile order.
#include "ofg_multi.hpp"
#include "ofg_internal.hpp"
#include "ofg_math.hpp"

namespace ofg {

namespace {

constexpr int WAVE = 64;
constexpr int NWAVES = MA_BLOCK / WAVE;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}

// Fast mode: one step of an underlying's log performance Y = 256 log2(S / s0)
// (the spot's step S = max(S g, 1e-12) with g = 2^(y/256), taken in logs).
__device__ __forceinline__ double ac_log_step(double y, double drift2, double vol_dt2, double c,
                                              double floor) {
    return fmax(y + fma(vol_dt2, c, drift2), floor);
}

// Fast mode: the worst performance 2^(y/256) of the worst log performance.
__device__ __forceinline__ double ac_perf(double y, const double *tab) {
    return exp2_256(fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX), tab);
}

// One path of simulate_autocallable_paths. The parity modes step the spots
// as the reference does (S *= exp(drift + vol_dt c), performances S / s0);
// the fast mode steps the log performances (ac_log_step) and compares them
// with the barriers' logs, so the exponential is taken once per path, for the
// knock-in redemption, instead of once per underlying and step.
template <int RNG>
__device__ __forceinline__ double autocall_path(const AcDev &o, const AcObs *obs, int64_t p,
                                                const double *normals, const double *tab) {
    constexpr bool FAST = RNG == OFG_RNG_FAST;
    const int n = o.n;
    // spots, or log performances in the fast mode
    double s[MA_MAX];
#pragma unroll
    for (int i = 0; i < MA_MAX; ++i) s[i] = FAST ? 0.0 : o.s0[i];
    const double ki_b = FAST ? o.ki_y : o.ki_barrier;
    const double coupon_b = FAST ? o.coupon_y : o.coupon_barrier;
    const double call_b = FAST ? o.autocall_y : o.autocall_barrier;
    Xoroshiro128p rng;
    if constexpr (FAST) rng.seed(stream_seed(o.seed, (uint64_t)p));
    const double *zp = normals + o.normals_offset + p;
    int k = 0;
    int next_step = o.n_obs > 0 ? obs[0].step : 0x7fffffff;
    bool called = false, ki = false;
    double worst_final = FAST ? 0.0 : 1.0, pv = 0.0, pending = 0.0, prev_t = 0.0;
    for (int step = 1; step <= o.steps; ++step) {
        double z[MA_MAX];
        if constexpr (FAST) {
#pragma unroll
            for (int i = 0; i < MA_MAX; i += 2) {
                if (i < n) {
                    const uint64_t wu = rng.next_u64();
                    const uint64_t wt = rng.next_u64();
                    box_muller_pair(wu, (uint32_t)(wt >> 32), z[i], z[i + 1], tab);
                }
            }
        } else {
            const int64_t row = (int64_t)(step - 1) * n;
#pragma unroll
            for (int i = 0; i < MA_MAX; ++i)
                if (i < n) z[i] = zp[(row + i) * o.paths];
        }
        double worst = INFINITY;
#pragma unroll
        for (int i = 0; i < MA_MAX; ++i) {
            if (i < n) {
                const double *li = o.chol + i * (i + 1) / 2;
                double c = 0.0;
#pragma unroll
                for (int j = 0; j <= i; ++j) c += li[j] * z[j];
                if constexpr (FAST) {
                    s[i] = ac_log_step(s[i], o.drift2[i], o.vol_dt2[i], c, o.yfloor[i]);
                    worst = fmin(worst, s[i]);
                } else {
                    const double g = exp_fp64(o.drift[i] + o.vol_dt[i] * c);
                    s[i] = fmax(s[i] * g, 1.0e-12);
                    worst = fmin(worst, s[i] / o.s0[i]);
                }
            }
        }
        worst_final = worst;
        if (worst <= ki_b) ki = true;
        if (step == next_step) {
            const AcObs ob = obs[k];
            if (o.phoenix) {
                const double accrual = o.notional * o.coupon_rate * fmax(ob.time - prev_t, 0.0);
                if (worst >= coupon_b) {
                    const double due = o.memory ? pending + accrual : accrual;
                    pv += ob.df * due;
                    pending = 0.0;
                } else if (o.memory) {
                    pending += accrual;
                }
                prev_t = ob.time;
            }
            if (worst >= call_b) {
                if (o.phoenix) pv += ob.df * o.notional;
                else pv = ob.call_pv;
                called = true;
                break;
            }
            ++k;
            next_step = k < o.n_obs ? obs[k].step : 0x7fffffff;
        }
    }
    if (!called) {
        double redemption;
        if constexpr (FAST) worst_final = ac_perf(worst_final, tab);
        if (ki) redemption = o.notional * (worst_final / o.ki_strike);
        else if (!o.phoenix) redemption = o.redeem_plain;
        else redemption = o.notional;
        pv += o.df_maturity * redemption;
    }
    return pv;
}

template <int RNG>
__global__ __launch_bounds__(MA_BLOCK) void autocall_kernel(AcArgs a) {
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    __shared__ double red[NWAVES][2];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += MA_BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const int tile = blockIdx.x;
    const int sc = a.tile_scen[tile];
    const AcDev &o = a.scen[sc];
    const AcObs *obs = a.obs + o.obs_offset;
    const int64_t base = (int64_t)(tile - o.tile_first) * MA_BLOCK * o.ppt;
    double sum = 0.0, sq = 0.0;
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * MA_BLOCK + threadIdx.x;
        if (p >= o.paths) break;
        const double x = autocall_path<RNG>(o, obs, p, a.normals, tab) - o.shift;
        sum += x;
        sq += x * x;
    }
    sum = wave_sum(sum);
    sq = wave_sum(sq);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = sum;
        red[wave][1] = sq;
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

// One block per scenario: tile slots in tile order; mean and standard error
// of the discounted payoffs (autocallable.rs:566-578).
__global__ __launch_bounds__(MA_BLOCK) void autocall_close_kernel(AcArgs a) {
    __shared__ double red[NWAVES][2];
    const int sc = blockIdx.x;
    const AcDev &o = a.scen[sc];
    double s = 0.0, q = 0.0;
    for (int t = threadIdx.x; t < o.n_tiles; t += MA_BLOCK) {
        s += a.partials[o.tile_first + t].sum;
        q += a.partials[o.tile_first + t].sum_sq;
    }
    s = wave_sum(s);
    q = wave_sum(q);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = s;
        red[wave][1] = q;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double sum = red[0][0], sq = red[0][1];
    for (int w = 1; w < NWAVES; ++w) {
        sum += red[w][0];
        sq += red[w][1];
    }
    const double n = (double)o.paths;
    const double var = o.paths > 1 ? fmax((sq - sum * sum / n) / (n - 1.0), 0.0) : 0.0;
    a.out[sc].price = o.shift + sum / n;
    a.out[sc].stderr_ = sqrt(var / n);
}

// ------------------------------------------------------------------------
// Basket options: one exact terminal step per asset
// ------------------------------------------------------------------------

// I_x(a, b), statrs 0.18's beta_reg (the crate behind StudentsT::cdf in
// math/correlation.rs:564-568): the modified Lentz continued fraction with
// the symmetry transform when x >= (a + 1) / (a + b + 2), at most 140 terms,
// tolerance 2^-53, and a zero prefactor when x == 0 or x lies within 4 ulps
// of 1. lbeta carries the prefactor's log-gamma terms for the caller's (a, b).
__device__ __noinline__ double beta_reg_dev(double a, double b, double x, double lbeta) {
    const double bt = (x == 0.0 || x >= 1.0 - 4.0 * 0x1p-53)
                          ? 0.0
                          : exp_fp64(lbeta + a * log(x) + b * log(1.0 - x));
    const bool symm = x >= (a + 1.0) / (a + b + 2.0);
    const double eps = 0x1p-53, fpmin = 0x1p-1022 / eps;
    if (symm) {
        const double t = a;
        x = 1.0 - x;
        a = b;
        b = t;
    }
    const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
    double c = 1.0, d = 1.0 - qab * x / qap;
    if (fabs(d) < fpmin) d = fpmin;
    d = 1.0 / d;
    double h = d;
    for (int mi = 1; mi < 141; ++mi) {
        const double m = (double)mi, m2 = m * 2.0;
        double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
        d = 1.0 + aa * d;
        if (fabs(d) < fpmin) d = fpmin;
        c = 1.0 + aa / c;
        if (fabs(c) < fpmin) c = fpmin;
        d = 1.0 / d;
        h = h * d * c;
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
        d = 1.0 + aa * d;
        if (fabs(d) < fpmin) d = fpmin;
        c = 1.0 + aa / c;
        if (fabs(c) < fpmin) c = fpmin;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (fabs(del - 1.0) <= eps) break;
    }
    return symm ? 1.0 - bt * h / a : bt * h / a;
}

// StudentsT::cdf at location 0, scale 1: h = nu / (nu + x^2),
// ib = I_h(nu/2, 1/2) / 2, x <= 0 ? ib : 1 - ib.
__device__ __forceinline__ double student_t_cdf_dev(double x, double nu, double a, double lbeta) {
    const double h = nu / (nu + x * x);
    const double ib = 0.5 * beta_reg_dev(a, 0.5, h, lbeta);
    return x <= 0.0 ? ib : 1.0 - ib;
}

__device__ __forceinline__ double clamp_copula_uniform(double u) {   // clamp(1e-12, 1 - 1e-12)
    const double lo = 1.0e-12, hi = 1.0 - 1.0e-12;
    return u < lo ? lo : (u > hi ? hi : u);
}

// One path's payoff (terminal_payoff, basket.rs:759-784; outperformance_payoff,
// :786-800; times the quanto FX rate, :429). The correlated
// normal of asset i accumulates L[i][j] z_j column by column as each z_j is
// drawn, the same terms in the same order as correlate_normals' running sum.
// The seed kernel's state for thread g (reference draws only).
template <int RNG>
__device__ __forceinline__ Xoshiro256pp thread_seed_state(const uint64_t *seeds, int64_t offset,
                                                          int64_t g) {
    Xoshiro256pp st;
    st.s0 = st.s1 = st.s2 = st.s3 = 0;
    if constexpr (RNG == OFG_RNG_REFERENCE) {
        const uint64_t *sd = seeds + offset + 4 * g;
        st.s0 = sd[0];
        st.s1 = sd[1];
        st.s2 = sd[2];
        st.s3 = sd[3];
    }
    return st;
}

// ref: the thread's reference stream at this path's first draw (a thread's
// paths are consecutive in the basket's stream, ofg::thread_stream_start).
template <int RNG, bool TCOP>
__device__ __forceinline__ double basket_path(const BkDev &o, int64_t p, const double *normals,
                                              Xoshiro256pp &ref, const double *tab) {
    constexpr bool FAST = RNG == OFG_RNG_FAST;
    const int n = o.n;
    double c[BK_MAX];
#pragma unroll
    for (int i = 0; i < BK_MAX; ++i) c[i] = 0.0;
    Xoroshiro128p fast;
    if constexpr (FAST) fast.seed(stream_seed(o.seed, (uint64_t)p));
    double zpair = 0.0;
    // draw k of the path (external normals: [draw][path])
    auto draw = [&](int k) {
        double z;
        if constexpr (RNG == OFG_RNG_REFERENCE) {
            z = reference_normal(ref);
        } else if constexpr (RNG == OFG_RNG_EXTERNAL) {
            z = normals[o.normals_offset + (int64_t)k * o.paths + p];
        } else {
            if ((k & 1) == 0) {
                const uint64_t wu = fast.next_u64();
                const uint64_t wt = fast.next_u64();
                box_muller_pair(wu, (uint32_t)(wt >> 32), z, zpair, tab);
            } else {
                z = zpair;
            }
        }
        return z;
    };
    if (o.n_factors == 0) {
        // the assets' draws in blocks of BK_AK (reference draws: the inverse
        // CDF's tail region deferred over the block), each column added to
        // the correlated normals as it is drawn
        constexpr int BK_AK = 8;
#pragma unroll
        for (int h = 0; h < BK_MAX; h += BK_AK) {
            if (h < n) {
                double zr[BK_AK];
                if constexpr (RNG == OFG_RNG_REFERENCE) {
                    double pr[BK_AK];
#pragma unroll
                    for (int k = 0; k < BK_AK; ++k)
                        pr[k] = h + k < n ? uniform_open01(ref.next_f64()) : 0.5;
                    acklam_block<BK_AK>(pr, zr);
                }
#pragma unroll
                for (int k = 0; k < BK_AK; ++k) {
                    const int j = h + k;
                    if (j < n) {
                        double z;
                        if constexpr (RNG == OFG_RNG_REFERENCE) z = zr[k];
                        else z = draw(j);
#pragma unroll
                        for (int i = j; i < BK_MAX; ++i)
                            if (i < n) c[i] += o.chol[i * (i + 1) / 2 + j] * z;
                    }
                }
            }
        }
    } else {
        // FactorCorrelationModel::sample_correlated_normals (correlation.rs:
        // 173-212): the factors first, then one idiosyncratic draw per asset
        double f[BK_FMAX];
#pragma unroll
        for (int k = 0; k < BK_FMAX; ++k)
            if (k < o.n_factors) f[k] = draw(k);
#pragma unroll
        for (int i = 0; i < BK_MAX; ++i) {
            if (i < n) {
                const double eps = draw(o.n_factors + i);
                if (o.one_factor) {
                    c[i] = fma(o.load[i * BK_FMAX], f[0], o.idio[i] * eps);
                } else {
                    double sys = -0.0;
#pragma unroll
                    for (int k = 0; k < BK_FMAX; ++k)
                        if (k < o.n_factors) sys += o.load[i * BK_FMAX + k] * f[k];
                    c[i] = sys + o.idio[i] * eps;
                }
            }
        }
    }
    if (TCOP && o.t_dof > 0) {
        // Student-t copula: sample_chi_square_integer_dof, then
        // map_copula_normals_to_uniforms and copula_uniforms_to_normals
        // (math/correlation.rs:545-582, 624-633)
        const int k0 = n + o.n_factors;
        double chi2 = 0.0;
        for (int k = 0; k < o.t_dof; ++k) {
            const double z = draw(k0 + k);
            chi2 += z * z;
        }
        const double scale = sqrt(fmax(chi2 / o.t_nu, 1.0e-16));
#pragma unroll
        for (int i = 0; i < BK_MAX; ++i) {
            if (i < n) {
                const double u =
                    clamp_copula_uniform(student_t_cdf_dev(c[i] / scale, o.t_nu, o.t_a, o.t_lbeta));
                c[i] = acklam_inv_cdf(clamp_copula_uniform(u));
            }
        }
    }
    const bool outperf = o.payoff == BK_OUTPERFORMANCE;
    double u = (o.type == BK_AVERAGE || outperf) ? -0.0
               : (o.type == BK_BEST_OF ? -INFINITY : INFINITY);
    double leader = 0.0;
#pragma unroll
    for (int i = 0; i < BK_MAX; ++i) {
        if (i < n) {
            double g;
            if constexpr (FAST) {
                double y = o.drift2[i] + o.vol_t2[i] * c[i];
                y = fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX);
                g = exp2_256(y, tab);
            } else {
                g = exp_fp64(o.drift[i] + o.vol_t[i] * c[i]);
            }
            const double st = o.s0[i] * g;
            if (i == o.leader) leader = st;
            // the fast mode's performance multiplies by 1 / s0 (within an ulp
            // of the reference's division, which the parity modes keep)
            if (o.type == BK_AVERAGE || outperf)
                u += o.w[i] * st;
            else if (o.type == BK_BEST_OF)
                u = fmax(u, FAST ? st * o.inv_s0[i] : st / o.s0[i]);
            else
                u = fmin(u, FAST ? st * o.inv_s0[i] : st / o.s0[i]);
        }
    }
    if (outperf) u = leader / fmax(u, 1.0e-12);   // outperformance_payoff (basket.rs:786-800)
    const double x = o.is_call ? fmax(u - o.strike, 0.0) : fmax(o.strike - u, 0.0);
    return o.scale * x;
}

// TCOP: the batch holds Student-t copula baskets (a separate instantiation
// keeps the Gaussian kernel's register budget).
template <int RNG, bool TCOP>
__global__ __launch_bounds__(MA_BLOCK) void basket_kernel(BkArgs a) {
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    __shared__ double red[NWAVES][2];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += MA_BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const int tile = blockIdx.x;
    const int sc = a.tile_scen[tile];
    const BkDev &o = a.scen[sc];
    const int64_t base = (int64_t)(tile - o.tile_first) * MA_BLOCK * o.ppt;
    double sum = 0.0, sq = 0.0;
    Xoshiro256pp ref = thread_seed_state<RNG>(
        a.seeds, o.seed_offset, (int64_t)(tile - o.tile_first) * MA_BLOCK + threadIdx.x);
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * MA_BLOCK + threadIdx.x;
        if (p >= o.paths) break;
        const double x = basket_path<RNG, TCOP>(o, p, a.normals, ref, tab);
        sum += x;
        sq += x * x;
    }
    sum = wave_sum(sum);
    sq = wave_sum(sq);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = sum;
        red[wave][1] = sq;
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

// One block per scenario: discounted mean and standard error of the payoffs
// (basket.rs:676-682).
__global__ __launch_bounds__(MA_BLOCK) void basket_close_kernel(BkArgs a) {
    __shared__ double red[NWAVES][2];
    const int sc = blockIdx.x;
    const BkDev &o = a.scen[sc];
    double s = 0.0, q = 0.0;
    for (int t = threadIdx.x; t < o.n_tiles; t += MA_BLOCK) {
        s += a.partials[o.tile_first + t].sum;
        q += a.partials[o.tile_first + t].sum_sq;
    }
    s = wave_sum(s);
    q = wave_sum(q);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = s;
        red[wave][1] = q;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double sum = red[0][0], sq = red[0][1];
    for (int w = 1; w < NWAVES; ++w) {
        sum += red[w][0];
        sq += red[w][1];
    }
    const double n = (double)o.paths;
    const double var = o.paths > 1 ? fmax((sq - sum * sum / n) / (n - 1.0), 0.0) : 0.0;
    a.out[sc].price = o.discount * (sum / n);
    a.out[sc].stderr_ = o.discount * sqrt(var / n);
}

// ------------------------------------------------------------------------
// autocallable sensitivities fused per path
// ------------------------------------------------------------------------

template <int RNG, int MAXN>
__device__ __forceinline__ void acg_move(const AcGreeksDev &o, int set, const double *c,
                                         double *sp) {
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        if (i < o.n) {
            if constexpr (RNG == OFG_RNG_FAST) {
                sp[i] = ac_log_step(sp[i], o.drift2[set][i], o.vol_dt2[set][i], c[i], o.yfloor[i]);
            } else {
                const double g = exp_fp64(o.drift[set][i] + o.vol_dt[set][i] * c[i]);
                sp[i] = fmax(sp[i] * g, 1.0e-12);
            }
        }
    }
}

template <int MAXN>
__device__ __forceinline__ void acg_correlate(const AcGreeksDev &o, int set, const double *z,
                                              double *c) {
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j <= i; ++j)
            if (i < o.n) acc += o.chol[set][i * (i + 1) / 2 + j] * z[j];
        c[i] = acc;
    }
}

// Worst performance; in the fast mode the worst log performance, as the
// one-scenario kernel's fast mode forms it.
template <bool FAST, int MAXN>
__device__ __forceinline__ double acg_worst(const AcGreeksDev &o, const double *sp) {
    double w = INFINITY;
#pragma unroll
    for (int i = 0; i < MAXN; ++i)
        if (i < o.n) w = fmin(w, FAST ? sp[i] : sp[i] / o.s0[i]);
    return w;
}

// Waves per SIMD the register allocator is held to: up to four underlyings
// the kernel fits 3 (168 VGPRs); left free it takes a few more and drops to
// 2, where the per-step scalar reloads of the scenario parameters are no
// longer hidden.
template <int MAXN>
constexpr int acg_waves_target() {
    return MAXN <= 4 ? 3 : 1;
}

template <int RNG, int MAXN>
__global__ __launch_bounds__(MA_BLOCK) __attribute__((amdgpu_waves_per_eu(
    acg_waves_target<MAXN>()))) void autocall_greeks_kernel(AcgArgs a) {
    constexpr int MAXS = 5 + 2 * MAXN;
    constexpr bool FAST = RNG == OFG_RNG_FAST;
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    __shared__ double red[NWAVES][ACG_MAXS][2];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += MA_BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const int tile = blockIdx.x;
    const AcGreeksDev &o = a.ct[a.tile_ct[tile]];
    const AcObs *obs = a.obs + o.obs_offset;
    const int n = o.n, ns = o.n_scen;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    const int64_t base = (int64_t)(tile - o.tile_first) * MA_BLOCK * o.ppt;
    // scenarios of this contract: 0..4 always present in the layout, 3 and 4
    // only with correlation bumps, 5 + 2q / 6 + 2q for q < n
    uint32_t present = 0x7u | (o.has_cega ? 0x18u : 0u);
    for (int q = 0; q < n; ++q) present |= 3u << (5 + 2 * q);
    const double ki_b = FAST ? o.ki_y : o.ki_barrier;
    const double coupon_b = FAST ? o.coupon_y : o.coupon_barrier;
    const double call_b = FAST ? o.autocall_y : o.autocall_barrier;
    double acc = 0.0, acc_sq = 0.0;
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * MA_BLOCK + threadIdx.x;
        const bool valid = p < o.paths;
        if (!__any(valid)) break;
        double pv[MAXS], pend[MAXS];
#pragma unroll
        for (int s = 0; s < MAXS; ++s) {
            pv[s] = 0.0;
            pend[s] = 0.0;
        }
        uint32_t called = valid ? ~present : 0xffffffffu, ki = 0u;
        // spots, or log performances in the fast mode (ac_log_step)
        double sb[MAXN], sv[2][MAXN], sc[2][MAXN], su[MAXN], sd[MAXN];
#pragma unroll
        for (int i = 0; i < MAXN; ++i) {
            sb[i] = sv[0][i] = sv[1][i] = sc[0][i] = sc[1][i] = FAST ? 0.0 : o.s0[i];
            su[i] = FAST ? 0.0 : o.s0_up[i];
            sd[i] = FAST ? 0.0 : o.s0_dn[i];
        }
        Xoroshiro128p rng;
        if constexpr (RNG == OFG_RNG_FAST) rng.seed(stream_seed(o.seed, (uint64_t)p));
        const double *zp = a.normals + o.normals_offset + p;
        int k = 0;
        int next_step = o.n_obs > 0 ? obs[0].step : 0x7fffffff;
        double prev_t = 0.0;
        // every scenario's worst performance (worst log performance in the
        // fast mode), handed to f(s, w) in scenario order
        auto each = [&](auto &&f) {
            double rb[MAXN];
#pragma unroll
            for (int i = 0; i < MAXN; ++i)
                rb[i] = i < n ? (FAST ? sb[i] : sb[i] / o.s0[i]) : INFINITY;
            double w0 = INFINITY;
#pragma unroll
            for (int i = 0; i < MAXN; ++i) w0 = fmin(w0, rb[i]);
            f(0, w0);
            f(1, acg_worst<FAST, MAXN>(o, sv[0]));
            f(2, acg_worst<FAST, MAXN>(o, sv[1]));
            if (o.has_cega) {
                f(3, acg_worst<FAST, MAXN>(o, sc[0]));
                f(4, acg_worst<FAST, MAXN>(o, sc[1]));
            }
#pragma unroll
            for (int q = 0; q < MAXN; ++q) {
                if (q < n) {
                    double wu = INFINITY, wd = INFINITY;
#pragma unroll
                    for (int i = 0; i < MAXN; ++i) {
                        wu = fmin(wu, i == q ? (FAST ? su[i] : su[i] / o.s0_up[i]) : rb[i]);
                        wd = fmin(wd, i == q ? (FAST ? sd[i] : sd[i] / o.s0_dn[i]) : rb[i]);
                    }
                    f(5 + 2 * q, wu);
                    f(6 + 2 * q, wd);
                }
            }
        };
        // a lane leaves once every scenario of its path is called; the lanes
        // still stepping share the step, hence the observation index and the
        // observation branch
        for (int step = 1; step <= o.steps && called != 0xffffffffu; ++step) {
            double z[MAXN];
            if constexpr (RNG == OFG_RNG_FAST) {
#pragma unroll
                for (int i = 0; i < MAXN; i += 2) {
                    if (i < n) {
                        const uint64_t wu = rng.next_u64();
                        const uint64_t wt = rng.next_u64();
                        double z1;
                        box_muller_pair(wu, (uint32_t)(wt >> 32), z[i], z1, tab);
                        if (i + 1 < MAXN) z[i + 1] = z1;
                    }
                }
            } else {
                // only the contract's paths read normals
                const int64_t row = (int64_t)(step - 1) * n;
#pragma unroll
                for (int i = 0; i < MAXN; ++i)
                    z[i] = i < n && valid ? zp[(row + i) * o.paths] : 0.0;
            }
            double c[MAXN];
            acg_correlate<MAXN>(o, 0, z, c);
            // base factors, shared by the base spots and the spot-bumped ones
#pragma unroll
            for (int i = 0; i < MAXN; ++i) {
                if (i < n) {
                    if constexpr (FAST) {
                        const double d2 = o.drift2[0][i], v2 = o.vol_dt2[0][i];
                        sb[i] = ac_log_step(sb[i], d2, v2, c[i], o.yfloor[i]);
                        su[i] = ac_log_step(su[i], d2, v2, c[i], o.yfloor_up[i]);
                        sd[i] = ac_log_step(sd[i], d2, v2, c[i], o.yfloor_dn[i]);
                    } else {
                        const double f = exp_fp64(o.drift[0][i] + o.vol_dt[0][i] * c[i]);
                        sb[i] = fmax(sb[i] * f, 1.0e-12);
                        su[i] = fmax(su[i] * f, 1.0e-12);
                        sd[i] = fmax(sd[i] * f, 1.0e-12);
                    }
                }
            }
            acg_move<RNG, MAXN>(o, 1, c, sv[0]);
            acg_move<RNG, MAXN>(o, 2, c, sv[1]);
            if (o.has_cega) {
                acg_correlate<MAXN>(o, 1, z, c);
                acg_move<RNG, MAXN>(o, 0, c, sc[0]);
                acg_correlate<MAXN>(o, 2, z, c);
                acg_move<RNG, MAXN>(o, 0, c, sc[1]);
            }
            // knock-in of every scenario (a called scenario's bit is never
            // read) and, on an observation step (the same step on every
            // lane), the coupons and calls of the scenarios not called yet
            const bool obs_now = step == next_step;
            AcObs ob{};
            double accrual = 0.0;
            if (obs_now) {
                ob = obs[k];
                accrual = o.notional * o.coupon_rate * fmax(ob.time - prev_t, 0.0);
            }
            each([&](int sc_, double w) {
                const uint32_t bit = 1u << sc_;
                if (w <= ki_b) ki |= bit;
                if (!obs_now || (called & bit)) return;
                if (o.phoenix) {
                    if (w >= coupon_b) {
                        const double due = o.memory ? pend[sc_] + accrual : accrual;
                        pv[sc_] += ob.df * due;
                        pend[sc_] = 0.0;
                    } else if (o.memory) {
                        pend[sc_] += accrual;
                    }
                }
                if (w >= call_b) {
                    if (o.phoenix) pv[sc_] += ob.df * o.notional;
                    else pv[sc_] = ob.call_pv;
                    called |= bit;
                }
            });
            if (obs_now) {
                if (o.phoenix) prev_t = ob.time;
                ++k;
                next_step = k < o.n_obs ? obs[k].step : 0x7fffffff;
            }
        }
        // each scenario's present value (redemption of those never called, on
        // the last step's worst) into lane s's sums
        auto finish = [&](int s, double wf) {
            double x = 0.0;
            if (valid) {
                double v = pv[s];
                if (!((called >> s) & 1u)) {
                    double redemption;
                    if constexpr (FAST) wf = ac_perf(wf, tab);
                    if ((ki >> s) & 1u) redemption = o.notional * (wf / o.ki_strike);
                    else if (!o.phoenix) redemption = o.redeem_plain;
                    else redemption = o.notional;
                    v += o.df_maturity * redemption;
                }
                x = v - o.shift;
            }
            const double sx = wave_sum(x), sq = wave_sum(x * x);
            if (lane == s) {
                acc += sx;
                acc_sq += sq;
            }
        };
        each(finish);
    }
    if (lane < ACG_MAXS) {
        red[wave][lane][0] = acc;
        red[wave][lane][1] = acc_sq;
    }
    __syncthreads();
    if (threadIdx.x < ns) {
        double sm = red[0][threadIdx.x][0], q = red[0][threadIdx.x][1];
        for (int w2 = 1; w2 < NWAVES; ++w2) {
            sm += red[w2][threadIdx.x][0];
            q += red[w2][threadIdx.x][1];
        }
        a.partials[(int64_t)tile * ACG_MAXS + threadIdx.x].sum = sm;
        a.partials[(int64_t)tile * ACG_MAXS + threadIdx.x].sum_sq = q;
    }
}

// One block per (contract, scenario): the tile slots in tile order.
__global__ __launch_bounds__(MA_BLOCK) void autocall_greeks_close_kernel(AcgArgs a) {
    __shared__ double red[NWAVES][2];
    const AcGreeksDev &o = a.ct[blockIdx.x];
    const int sc = blockIdx.y;
    if (sc >= o.n_scen) return;
    double s = 0.0, q = 0.0;
    for (int t = threadIdx.x; t < o.n_tiles; t += MA_BLOCK) {
        const AcPartial pt = a.partials[(int64_t)(o.tile_first + t) * ACG_MAXS + sc];
        s += pt.sum;
        q += pt.sum_sq;
    }
    s = wave_sum(s);
    q = wave_sum(q);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = s;
        red[wave][1] = q;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double sum = red[0][0], sq = red[0][1];
    for (int w = 1; w < NWAVES; ++w) {
        sum += red[w][0];
        sq += red[w][1];
    }
    const double nn = (double)o.paths;
    const double var = o.paths > 1 ? fmax((sq - sum * sum / nn) / (nn - 1.0), 0.0) : 0.0;
    a.out[o.out_first + sc].price = o.shift + sum / nn;
    a.out[o.out_first + sc].stderr_ = sqrt(var / nn);
}

template <int MAXN>
void launch_acg(const AcgArgs &a, hipStream_t stream) {
    if (a.rng_mode == OFG_RNG_FAST)
        autocall_greeks_kernel<OFG_RNG_FAST, MAXN><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a);
    else
        autocall_greeks_kernel<OFG_RNG_EXTERNAL, MAXN><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a);
}

// ------------------------------------------------------------------------
// basket_sensitivities fused per path
// ------------------------------------------------------------------------

// terminal_payoff with asset k's terminal spot and initial spot replaced
// (k = -1: none; invk = 1 / s0k), the reference's fold order. The fast mode's
// performances multiply by 1 / s0 as the one-scenario kernel's fast mode
// does.
template <bool FAST, int MAXN>
__device__ __forceinline__ double bkg_payoff(const BkGreeksDev &o, const double *st, int k,
                                             double stk, double s0k, double invk) {
    const int n = o.n;
    double u = o.type == BK_AVERAGE ? -0.0 : (o.type == BK_BEST_OF ? -INFINITY : INFINITY);
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        if (i < n) {
            const double sti = i == k ? stk : st[i];
            if (o.type == BK_AVERAGE) {
                u += o.w[i] * sti;
            } else {
                const double perf =
                    FAST ? sti * (i == k ? invk : o.inv_s0[i]) : sti / (i == k ? s0k : o.s0[i]);
                u = o.type == BK_BEST_OF ? fmax(u, perf) : fmin(u, perf);
            }
        }
    }
    return o.is_call ? fmax(u - o.strike, 0.0) : fmax(o.strike - u, 0.0);
}

// exp(drift + vol_t c) per asset for parameter set `set`.
template <int RNG, int MAXN>
__device__ __forceinline__ void bkg_factors(const BkGreeksDev &o, int set, const double *c,
                                            double *g, const double *tab) {
#pragma unroll
    for (int i = 0; i < MAXN; ++i) {
        if (i < o.n) {
            if constexpr (RNG == OFG_RNG_FAST) {
                double y = o.drift2[set][i] + o.vol_t2[set][i] * c[i];
                y = fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX);
                g[i] = exp2_256(y, tab);
            } else {
                g[i] = exp_fp64(o.drift[set][i] + o.vol_t[set][i] * c[i]);
            }
        }
    }
}

template <int MAXN>
__device__ __forceinline__ void bkg_correlate(const BkGreeksDev &o, int set, const double *z,
                                              double *c) {
#pragma unroll
    for (int i = 0; i < MAXN; ++i) c[i] = 0.0;
#pragma unroll
    for (int j = 0; j < MAXN; ++j)
#pragma unroll
        for (int i = j; i < MAXN; ++i)
            if (i < o.n && j < o.n) c[i] += o.chol[set][i * (i + 1) / 2 + j] * z[j];
}

template <int RNG, int MAXN>
__global__ __launch_bounds__(MA_BLOCK) void basket_greeks_kernel(BkgArgs a) {
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    __shared__ double red[NWAVES][BKG_MAXS][2];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += MA_BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const int tile = blockIdx.x;
    const BkGreeksDev &o = a.bk[a.tile_bk[tile]];
    const int n = o.n;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    const int64_t base = (int64_t)(tile - o.tile_first) * MA_BLOCK * o.ppt;
    // lane s accumulates scenario s's wave sums
    double acc = 0.0, acc_sq = 0.0;
    auto add = [&](int s, double x) {
        const double sx = wave_sum(x), sq = wave_sum(x * x);
        if (lane == s) {
            acc += sx;
            acc_sq += sq;
        }
    };
    // reference draws: the thread's paths are consecutive in the basket's
    // stream (ofg::thread_stream_start), from one state per thread
    Xoshiro256pp ref = thread_seed_state<RNG>(
        a.seeds, o.seed_offset, (int64_t)(tile - o.tile_first) * MA_BLOCK + threadIdx.x);
    for (int j = 0; j < o.ppt; ++j) {
        const int64_t p = base + (int64_t)j * MA_BLOCK + threadIdx.x;
        const bool valid = p < o.paths;
        if (!__any(valid)) break;
        double z[MAXN], c[MAXN], g[MAXN], st[MAXN];
        if (valid) {
            Xoroshiro128p fast;
            if constexpr (RNG == OFG_RNG_FAST) fast.seed(stream_seed(o.seed, (uint64_t)p));
            double zpair = 0.0;
            if constexpr (RNG == OFG_RNG_REFERENCE) {
                // the path's n draws, tail region deferred (acklam_block)
                double pr[MAXN];
#pragma unroll
                for (int i = 0; i < MAXN; ++i) pr[i] = i < n ? uniform_open01(ref.next_f64()) : 0.5;
                acklam_block<MAXN>(pr, z);
            }
#pragma unroll
            for (int i = 0; i < MAXN; ++i) {
                if (i < n) {
                    if constexpr (RNG == OFG_RNG_EXTERNAL) {
                        z[i] = a.normals[o.normals_offset + (int64_t)i * o.paths + p];
                    } else if constexpr (RNG == OFG_RNG_FAST) {
                        if ((i & 1) == 0) {
                            const uint64_t wu = fast.next_u64();
                            const uint64_t wt = fast.next_u64();
                            box_muller_pair(wu, (uint32_t)(wt >> 32), z[i], zpair, tab);
                        } else {
                            z[i] = zpair;
                        }
                    }
                } else {
                    z[i] = 0.0;
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < MAXN; ++i) z[i] = 0.0;
        }
        const double keep = valid ? 1.0 : 0.0;   // invalid lanes add zeros
        bkg_correlate<MAXN>(o, 0, z, c);
        bkg_factors<RNG, MAXN>(o, 0, c, g, tab);
#pragma unroll
        for (int i = 0; i < MAXN; ++i) st[i] = i < n ? o.s0[i] * g[i] : 0.0;
        add(0, keep * bkg_payoff<RNG == OFG_RNG_FAST, MAXN>(o, st, -1, 0.0, 0.0, 0.0));
        for (int k = 0; k < n; ++k) {
            double gk = g[0];
#pragma unroll
            for (int i = 1; i < MAXN; ++i)
                if (i == k) gk = g[i];
            add(1 + 2 * k, keep * bkg_payoff<RNG == OFG_RNG_FAST, MAXN>(o, st, k, o.s0_up[k] * gk, o.s0_up[k],
                                                                o.inv_up[k]));
            add(2 + 2 * k, keep * bkg_payoff<RNG == OFG_RNG_FAST, MAXN>(o, st, k, o.s0_dn[k] * gk, o.s0_dn[k],
                                                                o.inv_dn[k]));
        }
        for (int v = 1; v <= 2; ++v) {   // vol up, vol down: the base correlated normals
            bkg_factors<RNG, MAXN>(o, v, c, g, tab);
#pragma unroll
            for (int i = 0; i < MAXN; ++i) st[i] = i < n ? o.s0[i] * g[i] : 0.0;
            add(2 * n + v, keep * bkg_payoff<RNG == OFG_RNG_FAST, MAXN>(o, st, -1, 0.0, 0.0, 0.0));
        }
        if (o.has_cega) {
            for (int v = 1; v <= 2; ++v) {   // correlation up, down: the same draws
                bkg_correlate<MAXN>(o, v, z, c);
                bkg_factors<RNG, MAXN>(o, 0, c, g, tab);
#pragma unroll
                for (int i = 0; i < MAXN; ++i) st[i] = i < n ? o.s0[i] * g[i] : 0.0;
                add(2 * n + 2 + v, keep * bkg_payoff<RNG == OFG_RNG_FAST, MAXN>(o, st, -1, 0.0, 0.0, 0.0));
            }
        }
    }
    if (lane < BKG_MAXS) {
        red[wave][lane][0] = acc;
        red[wave][lane][1] = acc_sq;
    }
    __syncthreads();
    if (threadIdx.x < o.n_scen) {
        double sm = red[0][threadIdx.x][0], q = red[0][threadIdx.x][1];
        for (int w = 1; w < NWAVES; ++w) {
            sm += red[w][threadIdx.x][0];
            q += red[w][threadIdx.x][1];
        }
        a.partials[(int64_t)tile * BKG_MAXS + threadIdx.x].sum = sm;
        a.partials[(int64_t)tile * BKG_MAXS + threadIdx.x].sum_sq = q;
    }
}

// One block per (basket, scenario): the tile slots in tile order.
__global__ __launch_bounds__(MA_BLOCK) void basket_greeks_close_kernel(BkgArgs a) {
    __shared__ double red[NWAVES][2];
    const BkGreeksDev &o = a.bk[blockIdx.x];
    const int sc = blockIdx.y;
    if (sc >= o.n_scen) return;
    double s = 0.0, q = 0.0;
    for (int t = threadIdx.x; t < o.n_tiles; t += MA_BLOCK) {
        const AcPartial pt = a.partials[(int64_t)(o.tile_first + t) * BKG_MAXS + sc];
        s += pt.sum;
        q += pt.sum_sq;
    }
    s = wave_sum(s);
    q = wave_sum(q);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    if (lane == 0) {
        red[wave][0] = s;
        red[wave][1] = q;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double sum = red[0][0], sq = red[0][1];
    for (int w = 1; w < NWAVES; ++w) {
        sum += red[w][0];
        sq += red[w][1];
    }
    const double nn = (double)o.paths;
    const double var = o.paths > 1 ? fmax((sq - sum * sum / nn) / (nn - 1.0), 0.0) : 0.0;
    a.out[o.out_first + sc].price = o.discount * (sum / nn);
    a.out[o.out_first + sc].stderr_ = o.discount * sqrt(var / nn);
}

template <int MAXN>
void launch_bkg(const BkgArgs &a, hipStream_t stream) {
    switch (a.rng_mode) {
    case OFG_RNG_FAST:
        basket_greeks_kernel<OFG_RNG_FAST, MAXN><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a);
        break;
    case OFG_RNG_EXTERNAL:
        basket_greeks_kernel<OFG_RNG_EXTERNAL, MAXN><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a);
        break;
    default: basket_greeks_kernel<OFG_RNG_REFERENCE, MAXN><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a);
    }
}

}  // namespace

hipError_t launch_autocall_greeks(const AcgArgs &a, hipStream_t stream) {
    if (a.n_tiles <= 0) return hipSuccess;
    if (a.max_n <= 2) launch_acg<2>(a, stream);
    else if (a.max_n <= 4) launch_acg<4>(a, stream);
    else launch_acg<MA_MAX>(a, stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    autocall_greeks_close_kernel<<<dim3(a.n_ct, ACG_MAXS), MA_BLOCK, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_basket_greeks(const BkgArgs &a, hipStream_t stream) {
    if (a.n_tiles <= 0) return hipSuccess;
    if (a.max_n <= 4) launch_bkg<4>(a, stream);
    else if (a.max_n <= 8) launch_bkg<8>(a, stream);
    else launch_bkg<BK_MAX>(a, stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    basket_greeks_close_kernel<<<dim3(a.n_bk, BKG_MAXS), MA_BLOCK, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_basket(const BkArgs &a, hipStream_t stream) {
    if (a.n_tiles <= 0) return hipSuccess;
    auto go = [&](auto tcop) {
        constexpr bool T = decltype(tcop)::value;
        switch (a.rng_mode) {
        case OFG_RNG_FAST: basket_kernel<OFG_RNG_FAST, T><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a); break;
        case OFG_RNG_EXTERNAL:
            basket_kernel<OFG_RNG_EXTERNAL, T><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a);
            break;
        default: basket_kernel<OFG_RNG_REFERENCE, T><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a); break;
        }
    };
    if (a.any_t) go(std::true_type{});
    else go(std::false_type{});
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    basket_close_kernel<<<a.n_scen, MA_BLOCK, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_autocall(const AcArgs &a, hipStream_t stream) {
    if (a.n_tiles <= 0) return hipSuccess;
    if (a.rng_mode == OFG_RNG_FAST)
        autocall_kernel<OFG_RNG_FAST><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a);
    else
        autocall_kernel<OFG_RNG_EXTERNAL><<<a.n_tiles, MA_BLOCK, 0, stream>>>(a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    autocall_close_kernel<<<a.n_scen, MA_BLOCK, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace ofg
