// Forward pathwise Greeks for MI355X (gfx950), FP64: GBM by running sums of
// the path (below), Heston by a forward tangent of the Euler step (further
// down). GBM: the same
// pathwise estimator as the adjoint kernel (positive-part derivative
// 1{x > 0}, zero derivative of barrier indicators, observation indices fixed
// relative to T, d/dT through dt = T/n and the discount), evaluated in ONE
// forward sweep with no checkpoints or reverse pass.
//
// Under GBM, ln S_j = ln S0 + j drift + diffusion Z_j with Z_j = z_1 + .. + z_j,
// so d ln S_j / d theta = a + b j + c Z_j with per-parameter constants:
//   S0: a = 1/S0;  sigma: b = -sigma dt, c = sqrt(dt);  r: b = dt;
//   T:  b = drift / T, c = diffusion / (2T).
// With the payoff seeds (d x/d S_n = ct, d x/d S_j += ca w_j + cg w_j / S_j)
//   d x / d theta = ct S_n L_n + ca sum_j w_j S_j L_j + cg sum_j w_j L_j,
// and the weighted sums follow from running sums by summation by parts
// (R_k = sum_{i<=k} w_i S_i, the running arithmetic sum):
//   sum_j j w_j S_j   = n R_n - sum_{k<n} R_k,
//   sum_j w_j S_j Z_j = R_n Z_n - sum_i z_i R_{i-1},
// so a vanilla or barrier path carries one extra add per step (Z) and an
// arithmetic Asian path three (Z, sum R, sum z R). The draws and path
// operations are the pricing kernel's, so prices equal the pricing pass.
#include "ofg_adjoint.hpp"
#include "ofg_internal.hpp"
#include "ofg_math.hpp"
#include "ofg_pathwise.hpp"
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

// Running state of one leg beyond the path state: sum of the running
// weighted sums (Q) and of z times the running weighted sum before the step
// (P); both zero for payoffs without an arithmetic average.
struct Acc2 {
    double q, p;
};

// Gradient of one leg's pv from its final state. Z is this leg's normal sum,
// G2 = sum_j w_j Z_j (geometric averages only). Returns pv.
template <int PK>
__device__ __forceinline__ double leg_gradient(const Leg &g, const OptDev &o, double Z,
                                               const Acc2 &s, double G2, double (&gr)[ADJ_NG]) {
    double ct, ca, cg;
    const double x = leg_payoff_seeds<PK>(g, o, ct, ca, cg);
    const double pv = o.df * x;
    const double n = (double)o.steps;
    const double T = o.dt * n;
    const double tS = ct * g.s;
    double ta = tS, tb = tS * n, tc = tS * Z;
    if constexpr (PK == PK_ASIAN || PK == PK_ASIAN_LOG) {
        const double A0 = g.sum;
        const double A1 = fma(n, A0, -s.q);
        const double A2 = fma(A0, Z, -s.p);
        ta = fma(ca, A0, ta);
        tb = fma(ca, A1, tb);
        tc = fma(ca, A2, tc);
    }
    if constexpr (PK == PK_ASIAN_LOG) {
        ta = fma(cg, o.nobs, ta);
        tb = fma(cg, o.obs_jsum, tb);
        tc = fma(cg, G2, tc);
    }
    gr[0] += o.df * ta / o.s0;
    gr[1] += o.df * fma(-o.sigma * o.dt, tb, o.sqrt_dt * tc);
    gr[2] += fma(o.df * o.dt, tb, -T * pv);
    gr[3] += fma(o.df, fma(o.drift, tb, 0.5 * o.diffusion * tc) / T, -o.rate * pv);
    return pv;
}

template <int PK, int RNG, bool ANTI>
__device__ __forceinline__ double pathwise_sample(const OptDev &o, int64_t sample,
                                                  const double *__restrict__ w, const AdjArgs &a,
                                                  const double *tab, double (&gr)[ADJ_NG]) {
    constexpr bool ASIAN = PK == PK_ASIAN || PK == PK_ASIAN_LOG;
    constexpr bool FAST = RNG == OFG_RNG_FAST;
    const double barrier = o.barrier;
    const int steps = o.steps;
    Leg ga, gb;
    leg_init<PK>(ga, o, ASIAN ? w[0] : 0.0);
    if constexpr (ANTI) leg_init<PK>(gb, o, ASIAN ? w[0] : 0.0);
    Acc2 sa{0.0, 0.0}, sb{0.0, 0.0};
    double Z = 0.0, G2 = 0.0;

    using Rng = typename std::conditional<FAST, Xoroshiro128p, Xoshiro256pp>::type;
    Rng rng;
    ExternalNormals ext;
    if constexpr (RNG == OFG_RNG_EXTERNAL) {
        ext.p1 = a.normals1 + o.normals_offset + sample;
        ext.p2 = nullptr;
        ext.stride = o.samples;
    } else {
        rng.seed(stream_seed(o.seed, (uint64_t)sample));
    }

    auto advance = [&](double z, int j) {
        if constexpr (ASIAN) {
            sa.q += ga.sum;
            sa.p = fma(z, ga.sum, sa.p);
            if constexpr (ANTI) {
                sb.q += gb.sum;
                sb.p = fma(-z, gb.sum, sb.p);
            }
        }
        Z += z;
        gbm_step<PK, FAST>(ga, o, z, tab);
        if constexpr (ANTI) gbm_step<PK, FAST>(gb, o, -z, tab);
        const double wj = ASIAN ? w[j + 1] : 0.0;
        leg_observe<PK>(ga, wj, barrier);
        if constexpr (ANTI) leg_observe<PK>(gb, wj, barrier);
        if constexpr (PK == PK_ASIAN_LOG) G2 = fma(wj, Z, G2);
    };
    if constexpr (FAST) {
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
    } else if constexpr (RNG == OFG_RNG_REFERENCE) {
        reference_steps<8, 1>(rng, steps, [&](int j, double z, double) { advance(z, j); });
    } else {
        for (int j = 0; j < steps; ++j) advance(ext.z1(j), j);
    }

    double pv = leg_gradient<PK>(ga, o, Z, sa, G2, gr);
    if constexpr (ANTI) {
        double gb_grad[ADJ_NG] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        pv += leg_gradient<PK>(gb, o, -Z, sb, -G2, gb_grad);
        pv *= 0.5;
#pragma unroll
        for (int k = 0; k < 4; ++k) gr[k] = 0.5 * (gr[k] + gb_grad[k]);
    }
    return pv;
}

// ---------------------------------------------------------------------------
// Heston: forward tangent of the Euler step (no closed form for the variance
// path), with the adjoint kernel's conventions: the clamp indicators
// 1{sn >= 1e-12} and 1{vn > 0}, d v_pos/dv = 1{v > 0}, d sqrt(v_pos)/d v_pos =
// 1/(2 sqrt) taken as 0 at v_pos = 0.
//
// Only v0, kappa, theta and xi move the variance, so only they carry a
// variance tangent dv and a log-spot tangent dl through each step:
//   dl' = dl + B dv,              B = d e / d v,
//   dv' = 1{vn > 0} (A dv + c),   A = d vn / d v, c = d vn / d parameter.
// The recursions run in rescaled units that keep constant factors out of the
// step (each factor is applied once, after the last step): B' = B / sqrt(dt)
// = zs / (2 sqrt(v+)) - sqrt(dt) / 2, and four variance tangents
//   v0 (c = 0), Y (c' = v+), U (c' = 1), X (c' = sqrt(v+) z1)
// with dv_v0 = v0's, dv_kappa = dt (theta U - Y), dv_theta = kappa dt U and
// dv_xi = sqrt(dt) X; m_k = sum B' tangent_k gives d ln S (v0: sqrt(dt) m_v0,
// kappa: dt sqrt(dt) (theta m_U - m_Y), theta: kappa dt sqrt(dt) m_U,
// xi: dt m_X).
// The other slots need no per-step tangent:
//   S0   d ln S_j / d S0 = 1/S0 until the spot floor is first hit, then 0;
//   mu   d ln S_j / d mu = dt x (steps since the last spot floor);
//   rho  d ln S_j / d rho = sqrt(dt) (P1 - (rho / rho_c) P2) with P1, P2 the
//        sums of sqrt(v+) z1 and sqrt(v+) z2 over the steps since the last floor;
//   T    the Euler path depends on T only through kappa dt, theta dt, xi dt,
//        mu dt and v0 dt (in the per-step variance w = v dt), so path by path
//        T dP/dT = kappa P_kappa + theta P_theta + xi P_xi + mu P_mu + v0 P_v0
//        and the T slot follows from the others after the last step.
// A spot-floor hit (rare) resets the log-spot tangents in a divergent branch.
// Asian legs also carry every slot's observation sum: sum_j w_j S_j d ln S_j
// (arithmetic average) or sum_j w_j d ln S_j (geometric).
// ---------------------------------------------------------------------------
struct HTan {
    double dv[4], dl[4];          // v0, Y, U, X (rescaled as above): tangent and m
    double p1, p2;                // rho running sums
    double nf;                    // steps since the last spot floor (mu; Asian only)
    int jf;                       // path index of the last spot floor (0: none)
    bool s0_live;                 // the spot floor has not been hit (S0)
    double obs_s0, obs_s0_frozen; // S0 observation sum, running and at the first floor
    double a[4], ap1, ap2, amu;   // observation sums of the other slots (Asian)
};

template <int PK>
__device__ __forceinline__ bool htan_geometric(const OptDev &o) {
    if constexpr (PK == PK_ASIAN_LOG) return (o.flags & F_GEOMETRIC) != 0;
    return false;
}

template <int PK>
__device__ __forceinline__ void htan_init(HTan &t, const OptDev &o, double w0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) t.dv[k] = t.dl[k] = t.a[k] = 0.0;
    t.dv[0] = 1.0;
    t.p1 = t.p2 = t.nf = t.ap1 = t.ap2 = t.amu = 0.0;
    t.jf = 0;
    t.s0_live = true;
    t.obs_s0_frozen = 0.0;
    // path index 0 is S0 itself: only the S0 slot has a tangent there
    t.obs_s0 = htan_geometric<PK>(o) ? w0 : w0 * o.s0;
}

// zero_start: this is the first step of a path with v0 = 0 (the only step
// at which v = 0 while a variance tangent is nonzero: after a clamp to 0
// every variance tangent is 0 as well, so elsewhere d v_pos / d v = 1{v > 0}
// may be taken as 1).
template <int PK, bool REF>
__device__ __forceinline__ void heston_tangent_step(Leg &g, double &v, HTan &t, const OptDev &o,
                                                    double z1, double z2, const double *tab,
                                                    int j, bool zero_start) {
    constexpr bool ASIAN = PK == PK_ASIAN || PK == PK_ASIAN_LOG;
    double s_next, v_next, growth, e;
    HestonParts hp;
    heston_euler<REF>(g.s, v, o, z1, z2, s_next, v_next, growth, e, tab, hp);
    if constexpr (PK == PK_ASIAN_LOG) {
        const bool floor_hit = !(g.s * growth >= 1e-12);
        g.ls = floor_hit ? -27.631021115928547 : g.ls + e;
    }
    g.s = s_next;
    v = v_next;

    const double v_pos = hp.vp;
    const bool live = v_pos > 0.0;  // v > 0
    double hinv;  // d sqrt(v_pos) / d v_pos
    if constexpr (REF) {
        hinv = live ? 0.5 / hp.sv : 0.0;
    } else {
        // 1 / (2 sv) by one Newton step from the rsqrt seed that sqrt_pos
        // starts from (relative error 2^-23 -> 2^-46). At v_pos = 0 the seed
        // is +inf: it is capped so that hinv stays finite, which is all that
        // step needs -- every variance tangent is 0 there (the clamp that
        // put v at 0 zeroed them; a path starting at v0 = 0 takes the
        // zero_start branch), so hinv only ever multiplies zeros.
        const double y = fmin(rsq_seed(v_pos), 1e150);
        hinv = (0.5 * y) * fma(-hp.sv, y, 2.0);
    }
    double A, Bs;  // d vn / d v and (d e / d v) / sqrt(dt)
    if (zero_start) {
        A = 1.0;
        Bs = 0.0;
    } else {
        A = fma(o.xsdt * z1, hinv, 1.0 - o.kdt);
        Bs = fma(hp.zs, hinv, -0.5 * o.sqrt_dt);
    }
    // the variance clamp: 1{vn > 0} (0 for NaN)
    const double iv = hp.vn > 0.0 ? 1.0 : 0.0;
    const double Ai = A * iv;
    const double svz1 = hp.sv * z1;
    // log-spot tangents through e, at the variance tangents of v_old
#pragma unroll
    for (int k = 0; k < 4; ++k) t.dl[k] = fma(Bs, t.dv[k], t.dl[k]);
    t.p1 += svz1;
    t.p2 = fma(hp.sv, z2, t.p2);
    if constexpr (ASIAN) t.nf += 1.0;
    // variance tangents: v0 (no direct term), Y (v+), U (1), X (sqrt(v+) z1)
    t.dv[0] *= Ai;
    t.dv[1] = fma(t.dv[1], Ai, v_pos * iv);
    t.dv[2] = fma(t.dv[2], Ai, iv);
    t.dv[3] = fma(t.dv[3], Ai, svz1 * iv);
    // spot floor (or NaN): d ln S = 0 for every slot from here. Rare, so the
    // wave branches on whether any lane hit it and pays nothing otherwise.
    const bool floored = !(hp.sn >= 1e-12);
    if (__builtin_expect(__ballot(floored) != 0, 0)) {
        if (floored) {
#pragma unroll
            for (int k = 0; k < 4; ++k) t.dl[k] = 0.0;
            t.p1 = t.p2 = t.nf = 0.0;
            t.jf = j + 1;
            if (t.s0_live) t.obs_s0_frozen = t.obs_s0;
            t.s0_live = false;
        }
    }
}

template <int PK>
__device__ __forceinline__ void htan_observe(HTan &t, const Leg &g, const OptDev &o, double wj) {
    if constexpr (PK == PK_ASIAN || PK == PK_ASIAN_LOG) {
        const double q = htan_geometric<PK>(o) ? wj : wj * g.s;
#pragma unroll
        for (int k = 0; k < 4; ++k) t.a[k] = fma(q, t.dl[k], t.a[k]);
        t.ap1 = fma(q, t.p1, t.ap1);
        t.ap2 = fma(q, t.p2, t.ap2);
        t.amu = fma(q, t.nf, t.amu);
        t.obs_s0 += q;
    }
}

template <int PK>
__device__ __forceinline__ double heston_leg_gradient(const Leg &g, const OptDev &o,
                                                      const HTan &t, double (&gr)[ADJ_NG]) {
    constexpr bool ASIAN = PK == PK_ASIAN || PK == PK_ASIAN_LOG;
    double ct, ca, cg;
    const double x = leg_payoff_seeds<PK>(g, o, ct, ca, cg);
    const double pv = o.df * x;
    const double T = o.dt * (double)o.steps;
    const double tS = ct * g.s;
    const double cobs = htan_geometric<PK>(o) ? cg : ca;
    // d pv / d parameter from the terminal log-spot tangent and the
    // observation sum
    auto dpv = [&](double dln, double obs) {
        double d = tS * dln;
        if constexpr (ASIAN) d = fma(cobs, obs, d);
        return o.df * d;
    };
    const double rr = o.rho / o.rho_c;
    const double inv_s0 = 1.0 / o.s0;
    const double sdt = o.sqrt_dt, dt = o.dt;
    const double dt15 = dt * sdt;
    const double g_s0 = dpv(t.s0_live ? inv_s0 : 0.0,
                            (t.s0_live ? t.obs_s0 : t.obs_s0_frozen) * inv_s0);
    const double g_v0 = dpv(sdt * t.dl[0], sdt * t.a[0]);
    const double g_kappa = dpv(dt15 * fma(o.theta, t.dl[2], -t.dl[1]),
                               dt15 * fma(o.theta, t.a[2], -t.a[1]));
    const double g_theta = dpv(o.kappa * dt15 * t.dl[2], o.kappa * dt15 * t.a[2]);
    const double g_xi = dpv(dt * t.dl[3], dt * t.a[3]);
    const double nf = ASIAN ? t.nf : (double)(o.steps - t.jf);
    const double g_mu = dpv(nf * dt, t.amu * dt);
    const double g_rho = dpv(sdt * fma(-rr, t.p2, t.p1), sdt * fma(-rr, t.ap2, t.ap1));
    gr[0] += g_s0;
    gr[1] += g_v0;
    gr[2] += g_mu - T * pv;   // mu = r - q and the discount exp(-r T)
    const double tdp = o.kappa * g_kappa + o.theta * g_theta + o.xi * g_xi + o.mu * g_mu +
                       o.v0 * g_v0;
    gr[3] += tdp / T - o.rate * pv;
    gr[4] += g_kappa;
    gr[5] += g_theta;
    gr[6] += g_xi;
    gr[7] += g_rho;
    return pv;
}

template <int PK, int RNG, bool ANTI>
__device__ __forceinline__ double heston_tangent_sample(const OptDev &o, int64_t sample,
                                                        const double *__restrict__ w,
                                                        const AdjArgs &a, const double *tab,
                                                        double (&gr)[ADJ_NG]) {
    constexpr bool ASIAN = PK == PK_ASIAN || PK == PK_ASIAN_LOG;
    constexpr bool REF = RNG != OFG_RNG_FAST;
    const double barrier = o.barrier;
    const int steps = o.steps;
    Leg ga, gb;
    HTan ta, tb;
    double va = o.v0, vb = o.v0;
    leg_init<PK>(ga, o, ASIAN ? w[0] : 0.0);
    htan_init<PK>(ta, o, ASIAN ? w[0] : 0.0);
    if constexpr (ANTI) {
        leg_init<PK>(gb, o, ASIAN ? w[0] : 0.0);
        htan_init<PK>(tb, o, ASIAN ? w[0] : 0.0);
    }
    using Rng = typename std::conditional<RNG == OFG_RNG_FAST, Xoroshiro128p, Xoshiro256pp>::type;
    Rng rng;
    ExternalNormals ext;
    if constexpr (RNG == OFG_RNG_EXTERNAL) {
        ext.p1 = a.normals1 + o.normals_offset + sample;
        ext.p2 = a.normals2 + o.normals_offset + sample;
        ext.stride = o.samples;
    } else {
        rng.seed(stream_seed(o.seed, (uint64_t)sample));
    }
    const bool v0_zero = o.v0 == 0.0;
    auto advance = [&](double z1, double z2, int j) {
        const bool zs = v0_zero && j == 0;
        heston_tangent_step<PK, REF>(ga, va, ta, o, z1, z2, tab, j, zs);
        if constexpr (ANTI) heston_tangent_step<PK, REF>(gb, vb, tb, o, -z1, -z2, tab, j, zs);
        const double wj = ASIAN ? w[j + 1] : 0.0;
        leg_observe<PK>(ga, wj, barrier);
        htan_observe<PK>(ta, ga, o, wj);
        if constexpr (ANTI) {
            leg_observe<PK>(gb, wj, barrier);
            htan_observe<PK>(tb, gb, o, wj);
        }
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
    } else if constexpr (RNG == OFG_RNG_REFERENCE) {
        reference_steps<8, 2>(rng, steps, [&](int j, double z1, double z2) { advance(z1, z2, j); });
    } else {
        for (int j = 0; j < steps; ++j) advance(ext.z1(j), ext.z2(j), j);
    }
    double pv = heston_leg_gradient<PK>(ga, o, ta, gr);
    if constexpr (ANTI) {
        double g2[ADJ_NG] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        pv += heston_leg_gradient<PK>(gb, o, tb, g2);
        pv *= 0.5;
#pragma unroll
        for (int k = 0; k < ADJ_NG; ++k) gr[k] = 0.5 * (gr[k] + g2[k]);
    }
    return pv;
}

template <int MODEL, int PK, int RNG, bool ANTI>
__device__ void pathwise_item(const OptDev &o, const WorkItem &wi, int lane, const AdjArgs &a,
                              const double *tab, double &spv, double &spv2, double (&acc)[ADJ_NG]) {
    const double *__restrict__ w = a.obs_w + o.obs_offset;
    for (int64_t base = wi.begin; base < wi.end; base += WAVE) {
        const int64_t sample = base + lane;
        if (sample < wi.end) {
            double g[ADJ_NG] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            double pv;
            if constexpr (MODEL == OFG_MODEL_GBM) pv = pathwise_sample<PK, RNG, ANTI>(o, sample, w, a, tab, g);
            else pv = heston_tangent_sample<PK, RNG, ANTI>(o, sample, w, a, tab, g);
            spv += pv;
            spv2 += pv * pv;
#pragma unroll
            for (int k = 0; k < ADJ_NG; ++k) acc[k] += g[k];
        }
    }
}

// Waves per SIMD asked of the register allocator. The single-leg fast-draw
// Heston tangent keeps its vanilla and barrier step loops free of spills at
// three waves (168 VGPRs; only the Asian paths spill) and runs 4 % faster
// than at two; at four the step loop itself spills. The reference-draw
// variants stay at two, two antithetic legs at one. The fast-draw GBM
// tangent takes 169-171 VGPRs left free, just over the 168 of three waves,
// and fits them.
template <int RNG, int MODEL, bool ANTI>
constexpr int pathwise_waves_target() {
    if (MODEL == OFG_MODEL_HESTON) return ANTI ? 1 : (RNG == OFG_RNG_FAST ? 3 : 2);
    return RNG == OFG_RNG_FAST ? 3 : 1;
}

template <int RNG, int MODEL, bool ANTI, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(
    pathwise_waves_target<RNG, MODEL, ANTI>()))) void pathwise_kernel(AdjArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    __shared__ double tab[RNG == OFG_RNG_FAST ? OFG_FAST_TABLE_LEN : 1];
    if constexpr (RNG == OFG_RNG_FAST) {
        for (int i = threadIdx.x; i < OFG_FAST_TABLE_LEN; i += BLOCK) tab[i] = OFG_FAST_TABLE[i];
        __syncthreads();
    }
    const uint32_t span = (uint32_t)(a.order_end - a.order_begin);
    for (;;) {
        // every lane takes part in the atomic (see the pricing kernel)
        const uint32_t q = __builtin_amdgcn_readfirstlane(
            atomicAdd(a.queue_head, lane == 0 ? 1u : 0u));
        if (q >= span) break;
        const int32_t it = __builtin_amdgcn_readfirstlane(a.order[a.order_begin + (int32_t)q]);
        const WorkItem wi = a.items[it];
        const OptDev o = a.opts[wi.option];
        // per-lane sums in the adjoint kernel's order, so both return the
        // same price bit for bit
        double spv = 0.0, spv2 = 0.0;
        double acc[ADJ_NG] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        switch (o.kind) {
        case PK_VANILLA: pathwise_item<MODEL, PK_VANILLA, RNG, ANTI>(o, wi, lane, a, tab, spv, spv2, acc); break;
        case PK_ASIAN: pathwise_item<MODEL, PK_ASIAN, RNG, ANTI>(o, wi, lane, a, tab, spv, spv2, acc); break;
        case PK_ASIAN_LOG: pathwise_item<MODEL, PK_ASIAN_LOG, RNG, ANTI>(o, wi, lane, a, tab, spv, spv2, acc); break;
        case PK_BARRIER: pathwise_item<MODEL, PK_BARRIER, RNG, ANTI>(o, wi, lane, a, tab, spv, spv2, acc); break;
        default: pathwise_item<MODEL, PK_BARRIER_DOWN, RNG, ANTI>(o, wi, lane, a, tab, spv, spv2, acc); break;
        }
        AdjPartial p;
        p.spv = wave_sum(spv);
        p.spv2 = wave_sum(spv2);
#pragma unroll
        for (int k = 0; k < ADJ_NG; ++k)
            p.g[k] = (MODEL == OFG_MODEL_HESTON || k < 4) ? wave_sum(acc[k]) : 0.0;
        p.n = (double)(wi.end - wi.begin);
        if (lane == 0) a.partials[it] = p;
    }
}

template <int RNG, int MODEL, bool ANTI>
hipError_t launch_block(const AdjArgs &a, hipStream_t s) {
    switch (a.block) {
    case 64: pathwise_kernel<RNG, MODEL, ANTI, 64><<<a.grid, 64, 0, s>>>(a); break;
    case 128: pathwise_kernel<RNG, MODEL, ANTI, 128><<<a.grid, 128, 0, s>>>(a); break;
    default: pathwise_kernel<RNG, MODEL, ANTI, 256><<<a.grid, 256, 0, s>>>(a); break;
    }
    return hipGetLastError();
}

template <int RNG, int MODEL, bool ANTI>
const void *block_ptr(int block) {
    switch (block) {
    case 64: return (const void *)pathwise_kernel<RNG, MODEL, ANTI, 64>;
    case 128: return (const void *)pathwise_kernel<RNG, MODEL, ANTI, 128>;
    default: return (const void *)pathwise_kernel<RNG, MODEL, ANTI, 256>;
    }
}

template <int RNG>
hipError_t launch_model(const AdjArgs &a, hipStream_t s) {
    if (a.model == OFG_MODEL_GBM)
        return a.antithetic ? launch_block<RNG, OFG_MODEL_GBM, true>(a, s)
                            : launch_block<RNG, OFG_MODEL_GBM, false>(a, s);
    return a.antithetic ? launch_block<RNG, OFG_MODEL_HESTON, true>(a, s)
                        : launch_block<RNG, OFG_MODEL_HESTON, false>(a, s);
}

template <int RNG>
const void *model_ptr(int block, int model, int anti) {
    if (model == OFG_MODEL_GBM)
        return anti ? block_ptr<RNG, OFG_MODEL_GBM, true>(block)
                    : block_ptr<RNG, OFG_MODEL_GBM, false>(block);
    return anti ? block_ptr<RNG, OFG_MODEL_HESTON, true>(block)
                : block_ptr<RNG, OFG_MODEL_HESTON, false>(block);
}

}  // namespace

hipError_t launch_pathwise(const AdjArgs &a, hipStream_t s) {
    switch (a.rng_mode) {
    case OFG_RNG_EXTERNAL: return launch_model<OFG_RNG_EXTERNAL>(a, s);
    case OFG_RNG_FAST: return launch_model<OFG_RNG_FAST>(a, s);
    default: return launch_model<OFG_RNG_REFERENCE>(a, s);
    }
}

int pathwise_grid(int block, int rng_mode, int model, int antithetic, int device) {
    const void *k = rng_mode == OFG_RNG_EXTERNAL ? model_ptr<OFG_RNG_EXTERNAL>(block, model, antithetic)
                    : rng_mode == OFG_RNG_FAST   ? model_ptr<OFG_RNG_FAST>(block, model, antithetic)
                                                 : model_ptr<OFG_RNG_REFERENCE>(block, model, antithetic);
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
