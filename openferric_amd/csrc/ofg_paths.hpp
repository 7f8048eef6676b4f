// Path-state building blocks shared by the pricing and adjoint kernels: one
// GBM step (src/mc/simulation.rs:58-71), one Heston Euler step with the
// generator clamps (src/models/stochastic.rs:56-71, simulation.rs:98-116),
// observation bookkeeping and the payoffs of
// src/engines/monte_carlo/mc_engine.rs:59-121, 441-584.
#pragma once

#include "ofg_internal.hpp"
#include "ofg_math.hpp"

namespace ofg {

__device__ __forceinline__ double vanilla_payoff(bool put, double spot, double strike) {
    return put ? fmax(strike - spot, 0.0) : fmax(spot - strike, 0.0);
}

// Running state of one path leg, held in registers.
struct Leg {
    double s;      // spot
    double sum;    // weighted sum of observed spots (Asian)
    double ls;     // log spot tracked additively (geometric Asian)
    double lsum;   // weighted sum of observed log spots
    bool hit;      // barrier touched (inclusive, path[0] included)
};

template <int PK>
__device__ __forceinline__ void leg_init(Leg &g, const OptDev &o, double w0) {
    g.s = o.s0;
    if constexpr (PK == PK_ASIAN || PK == PK_ASIAN_LOG) g.sum = w0 * o.s0;
    if constexpr (PK == PK_ASIAN_LOG) {
        g.ls = o.log_s0;
        g.lsum = w0 * o.log_s0;
    }
    if constexpr (PK == PK_BARRIER) g.hit = o.s0 >= o.barrier;
    if constexpr (PK == PK_BARRIER_DOWN) g.hit = o.s0 <= o.barrier;
}

// Observation bookkeeping after the spot of path index j+1 is known.
template <int PK>
__device__ __forceinline__ void leg_observe(Leg &g, double w, double barrier) {
    if constexpr (PK == PK_ASIAN || PK == PK_ASIAN_LOG) g.sum = fma(w, g.s, g.sum);
    if constexpr (PK == PK_ASIAN_LOG) g.lsum = fma(w, g.ls, g.lsum);
    if constexpr (PK == PK_BARRIER) g.hit = g.hit || (g.s >= barrier);
    if constexpr (PK == PK_BARRIER_DOWN) g.hit = g.hit || (g.s <= barrier);
}

// Growth factor exp(x) of one GBM step (simulation.rs:67-70), x = diffusion z
// + drift, and x itself. FAST evaluates it as 2^(y/256) with y = x 256/ln2
// formed from the pre-scaled constants (one rounding, as for x). Otherwise x
// is the reference's fma and exp(x) is exp_fp64(x), or with TABLE
// 2^(y/256) of y = x 256/ln2 (tab: the 256 entries 2^(j/256)); both are
// within about an ulp of exp(x).
template <bool FAST, bool TABLE = false>
__device__ __forceinline__ double gbm_growth(const OptDev &o, double z, const double *tab,
                                             double &x) {
    if constexpr (FAST) {
        const double y = fma(o.diffusion2, z, o.drift2);
        x = y * OFG_LN2_256;
        return exp2_256(y, tab);
    } else if constexpr (TABLE) {
        x = fma(o.diffusion, z, o.drift);
        // past |y| = 2^30 the growth is 0 or inf either way
        return exp2_256t(fmin(fmax(x * OFG_INV_LN2_256, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX),
                         tab);
    } else {
        x = fma(o.diffusion, z, o.drift);
        return exp_fp64(x);
    }
}

// One GBM step: s *= exp(fma(diffusion, z, drift)).
template <int PK, bool FAST, bool TABLE = false>
__device__ __forceinline__ void gbm_step(Leg &g, const OptDev &o, double z, const double *tab) {
    double x;
    g.s *= gbm_growth<FAST, TABLE>(o, z, tab, x);
    if constexpr (PK == PK_ASIAN_LOG) g.ls += x;
}

// One Heston Euler step (stochastic.rs:56-71) plus the generator clamps
// (simulation.rs:112-113), returning the spot growth factor exp(e).
//   REF = true : the reference's operation order (parity modes; the file is
//                compiled with -ffp-contract=off);
//   REF = false: the same scheme re-associated into fused multiply-adds with
//                per-option constants (fast mode): kappa dt, kappa theta dt,
//                xi sqrt(dt), and the log-spot increment's mu dt, dt / 2,
//                sqrt(dt) pre-scaled by 256/ln2 for exp2_256.
// Intermediates of one Heston step that its tangent needs: the positive
// part of the variance and its sqrt, the correlated normal, and the
// unclamped spot and variance.
struct HestonParts {
    double vp, sv, zs, sn, vn;
};

template <bool REF, bool TABLE = false>
__device__ __forceinline__ void heston_euler(double s, double v, const OptDev &o, double z1,
                                             double z2, double &s_next, double &v_next,
                                             double &growth, double &e, const double *tab,
                                             HestonParts &hp) {
    const double v_pos = fmax(v, 0.0);
    double vn, sv, zs;
    if constexpr (REF) {
        sv = sqrt(v_pos);
        zs = o.rho * z1 + o.rho_c * z2;
        vn = v + o.kappa * (o.theta - v_pos) * o.dt + o.xi * sv * o.sqrt_dt * z1;
        e = (o.mu - 0.5 * v_pos) * o.dt + sv * o.sqrt_dt * zs;
        // e is bounded by the clamp below only through the growth: keep the
        // table exp's argument inside its domain (past |y| = 2^30 the growth
        // is 0 or inf either way)
        if constexpr (TABLE)
            growth = exp2_256t(fmin(fmax(e * OFG_INV_LN2_256, -OFG_EXP2_256_YMAX),
                                    OFG_EXP2_256_YMAX), tab);
        else
            growth = exp_fp64(e);
    } else {
        sv = v_pos > 0.0 ? sqrt_pos(v_pos) : 0.0;
        zs = fma(o.rho, z1, o.rho_c * z2);
        vn = fma(o.xsdt * sv, z1, fma(-o.kdt, v_pos, v + o.kthdt));
        // v has no a-priori bound: clamp y into exp2_256's domain (past
        // |y| = 2^30 the growth is 0 or inf either way)
        double y = fma(sv * o.sqrt_dt2, zs, fma(-o.hdt2, v_pos, o.mudt2));
        y = fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX);
        e = y * OFG_LN2_256;
        growth = exp2_256(y, tab);
    }
    const double sn = s * growth;
    s_next = fmax(sn, 1e-12);   // f64::max(sn, 1e-12): maxNum, NaN -> 1e-12
    v_next = fmax(vn, 0.0);
    hp.vp = v_pos;
    hp.sv = sv;
    hp.zs = zs;
    hp.sn = sn;
    hp.vn = vn;
}

template <bool REF, bool TABLE = false>
__device__ __forceinline__ void heston_euler(double s, double v, const OptDev &o, double z1,
                                             double z2, double &s_next, double &v_next,
                                             double &growth, double &e, const double *tab) {
    HestonParts hp;
    heston_euler<REF, TABLE>(s, v, o, z1, z2, s_next, v_next, growth, e, tab, hp);
}

template <int PK, bool REF, bool TABLE = false>
__device__ __forceinline__ void heston_step(Leg &g, double &v, const OptDev &o, double z1,
                                            double z2, const double *tab) {
    double s_next, v_next, growth, e;
    heston_euler<REF, TABLE>(g.s, v, o, z1, z2, s_next, v_next, growth, e, tab);
    if constexpr (PK == PK_ASIAN_LOG) {
        const bool floor_hit = !(g.s * growth >= 1e-12);
        g.ls = floor_hit ? -27.631021115928547 : g.ls + e;   // ln(1e-12) at the floor
    }
    g.s = s_next;
    v = v_next;
}

template <int PK>
__device__ __forceinline__ void leg_payoff(const Leg &g, const OptDev &o, double &x, double &y) {
    const bool put = o.flags & F_PUT;
    if constexpr (PK == PK_VANILLA) {
        x = vanilla_payoff(put, g.s, o.strike);
    } else if constexpr (is_barrier(PK)) {
        const bool active = (o.flags & F_KNOCK_IN) ? g.hit : !g.hit;
        x = active ? vanilla_payoff(put, g.s, o.strike) : o.rebate;
    } else {
        double avg;
        if constexpr (PK == PK_ASIAN_LOG) {
            avg = (o.flags & F_GEOMETRIC) ? exp_fp64(g.lsum / o.nobs) : g.sum / o.nobs;
        } else {
            avg = g.sum / o.nobs;
        }
        x = (o.flags & F_FLOATING) ? vanilla_payoff(put, g.s, avg) : vanilla_payoff(put, avg, o.strike);
    }
    y = 0.0;
    if (o.flags & F_CV) {
        if constexpr (PK == PK_ASIAN_LOG) {
            if (o.flags & F_CV_GEO) {
                y = vanilla_payoff(put, exp_fp64(g.lsum / o.nobs), o.strike);
                return;
            }
        }
        y = vanilla_payoff(put, g.s, o.strike);
    }
}

// Payoff value and derivative seeds of one leg:
//   dx/dS_n = ct, and per observation j: dx/dS_j += ca w_j + cg w_j / S_j.
template <int PK>
__device__ __forceinline__ double leg_payoff_seeds(const Leg &g, const OptDev &o, double &ct,
                                                   double &ca, double &cg) {
    const bool put = o.flags & F_PUT;
    ct = ca = cg = 0.0;
    if constexpr (PK == PK_VANILLA || is_barrier(PK)) {
        const double x = vanilla_payoff(put, g.s, o.strike);
        const double d = put ? (o.strike - g.s > 0.0 ? -1.0 : 0.0) : (g.s - o.strike > 0.0 ? 1.0 : 0.0);
        if constexpr (is_barrier(PK)) {
            const bool active = (o.flags & F_KNOCK_IN) ? g.hit : !g.hit;
            ct = active ? d : 0.0;
            return active ? x : o.rebate;
        } else {
            ct = d;
            return x;
        }
    } else {
        bool geo = false;
        if constexpr (PK == PK_ASIAN_LOG) geo = (o.flags & F_GEOMETRIC) != 0;
        const double avg = geo ? exp_fp64(g.lsum / o.nobs) : g.sum / o.nobs;
        double davg;
        double x;
        if (o.flags & F_FLOATING) {
            x = vanilla_payoff(put, g.s, avg);
            const bool on = put ? (avg - g.s > 0.0) : (g.s - avg > 0.0);
            ct = on ? (put ? -1.0 : 1.0) : 0.0;
            davg = on ? (put ? 1.0 : -1.0) : 0.0;
        } else {
            x = vanilla_payoff(put, avg, o.strike);
            const bool on = put ? (o.strike - avg > 0.0) : (avg - o.strike > 0.0);
            davg = on ? (put ? -1.0 : 1.0) : 0.0;
        }
        if (geo) cg = davg * avg / o.nobs;
        else ca = davg / o.nobs;
        return x;
    }
}

// ------------------------------------------------------------------------
// Normal sources
// ------------------------------------------------------------------------

struct ExternalNormals {
    const double *p1, *p2;
    int64_t stride;
    __device__ __forceinline__ double z1(int j) const { return p1[(int64_t)j * stride]; }
    __device__ __forceinline__ double z2(int j) const { return p2[(int64_t)j * stride]; }
};



}  // namespace ofg
