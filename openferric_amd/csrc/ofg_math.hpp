// Device math for the Monte Carlo path kernels (gfx950, FP64 on the VALU).
//
// Everything here is __host__ __device__ so the same code can be exercised on
// the CPU by the unit tests (tests/test_device_math_cpu.py builds a host-only
// shim of this header).
//
// Contents:
//   * SplitMix64 / xoshiro256++ streams seeded exactly as the reference
//     (src/math/fast_rng.rs:35-73, 181-214) so a path's draws are the
//     reference's draws.
//   * Acklam inverse normal CDF (src/math/fast_norm.rs:53-139), restated with
//     the two tails merged into one branch (bitwise the same values: the upper
//     tail is the negated lower-tail formula applied to 1 - p).
//   * exp: Cody-Waite reduction by ln2 into |r| <= ln2/2 and a degree-11
//     polynomial (Chebyshev-node interpolant of e^r, max relative error
//     1.7e-17 in exact arithmetic), scaled with ldexp.
//   * fast RNG mode: xoshiro256+ streams and a branch-free Box-Muller
//     transform with its own log, sqrt and sin/cos on [-pi/4, pi/4].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OFG_HD __host__ __device__ __forceinline__

namespace ofg {

// ---------------------------------------------------------------- RNG -----

OFG_HD uint64_t splitmix64_next(uint64_t &state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

OFG_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// Per-path xoshiro256++ stream (fast_rng.rs:30-73). State lives in 8 VGPRs.
struct Xoshiro256pp {
    uint64_t s0, s1, s2, s3;

    // seed_from_u64 (fast_rng.rs:35-48): four SplitMix64 outputs; an all-zero
    // state is replaced by s0 = 1.
    OFG_HD void seed(uint64_t seed) {
        uint64_t sm = seed;
        s0 = splitmix64_next(sm);
        s1 = splitmix64_next(sm);
        s2 = splitmix64_next(sm);
        s3 = splitmix64_next(sm);
        if ((s0 | s1 | s2 | s3) == 0) s0 = 1;
    }

    OFG_HD uint64_t next_u64() {
        const uint64_t result = rotl64(s0 + s3, 23) + s0;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl64(s3, 45);
        return result;
    }

    // next_f64 (fast_rng.rs:69-73): (x >> 11) * 2^-53, exact in double.
    OFG_HD double next_f64() {
        const uint64_t x = next_u64() >> 11;
        return (double)x * (1.0 / 9007199254740992.0);
    }

    // Steps back over the last draw and returns it again (the adjoint's
    // backward sweep regenerates a path's draws in reverse order).
    OFG_HD uint64_t prev_u64();
};

// Inverse of the xoshiro256 state transition (shared by ++ and +): the
// transition is an invertible GF(2)-linear map, so a stream can be walked
// backwards. With old state (a,b,c,d) and new (A,B,C,D):
//   d ^ b = rotr(D, 45); a = A ^ (d ^ b); B ^ C = b ^ (b << 17), solved for b
//   by b = x ^ x<<17 ^ x<<34 ^ x<<51; c = C ^ a ^ (b << 17); d = (d ^ b) ^ b.
OFG_HD void xoshiro256_unstep(uint64_t &s0, uint64_t &s1, uint64_t &s2, uint64_t &s3) {
    const uint64_t d1 = (s3 >> 45) | (s3 << 19);
    const uint64_t a = s0 ^ d1;
    const uint64_t x = s1 ^ s2;
    const uint64_t b = x ^ (x << 17) ^ (x << 34) ^ (x << 51);
    const uint64_t c = s2 ^ a ^ (b << 17);
    s0 = a;
    s1 = b;
    s2 = c;
    s3 = d1 ^ b;
}

// stream_seed (fast_rng.rs:202-205), wrapping.
OFG_HD uint64_t stream_seed(uint64_t base, uint64_t index) { return base + index * 7919ull; }

// uniform_open01 (fast_rng.rs:218-221): clamp to [eps, 1 - eps].
OFG_HD double uniform_open01(double u) {
    const double lo = 2.220446049250313e-16;
    const double hi = 1.0 - 2.220446049250313e-16;
    u = u < lo ? lo : u;
    return u > hi ? hi : u;
}

// ------------------------------------------------------- elementary math --

// exp(x) for finite x in the double range. |x| well inside +-700 for all
// path increments; outside that ldexp saturates to 0 / inf as exp does.
OFG_HD double exp_fp64(double x) {
    const double LOG2E = 1.4426950408889634;
    const double LN2_HI = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
    const double LN2_LO = 1.90821492927058770002e-10;  // 0x3dea39ef35793c76
    const double k = rint(x * LOG2E);
    double r = fma(-k, LN2_HI, x);
    r = fma(-k, LN2_LO, r);
    double p = 2.5110037605963777e-08;
    p = fma(p, r, 2.763263963904103e-07);
    p = fma(p, r, 2.755724091857897e-06);
    p = fma(p, r, 2.4801485482328494e-05);
    p = fma(p, r, 0.00019841269890047113);
    p = fma(p, r, 0.0013888888952314775);
    p = fma(p, r, 0.008333333333319601);
    p = fma(p, r, 0.0416666666664881);
    p = fma(p, r, 0.1666666666666668);
    p = fma(p, r, 0.5000000000000019);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

// --------------------------------------------------- inverse normal CDF --

// Acklam coefficients as published (identical to fast_norm.rs:66-94).
#define OFG_AK_A0 -3.969683028665376e1
#define OFG_AK_A1 2.209460984245205e2
#define OFG_AK_A2 -2.759285104469687e2
#define OFG_AK_A3 1.383577518672690e2
#define OFG_AK_A4 -3.066479806614716e1
#define OFG_AK_A5 2.506628277459239
#define OFG_AK_B0 -5.447609879822406e1
#define OFG_AK_B1 1.615858368580409e2
#define OFG_AK_B2 -1.556989798598866e2
#define OFG_AK_B3 6.680131188771972e1
#define OFG_AK_B4 -1.328068155288572e1
#define OFG_AK_C0 -7.784894002430293e-3
#define OFG_AK_C1 -3.223964580411365e-1
#define OFG_AK_C2 -2.400758277161838
#define OFG_AK_C3 -2.549732539343734
#define OFG_AK_C4 4.374664141464968
#define OFG_AK_C5 2.938163982698783
#define OFG_AK_D0 7.784695709041462e-3
#define OFG_AK_D1 3.224671290700398e-1
#define OFG_AK_D2 2.445134137142996
#define OFG_AK_D3 3.754408661907416

// Acklam on p already clamped to [eps, 1 - eps] by uniform_open01, so the
// NaN / 0 / 1 guards of the reference never fire on this path.
OFG_HD double acklam_inv_cdf(double p) {
    const double P_LOW = 0.02425;
    const double P_HIGH = 1.0 - P_LOW;
    double z;
    if (p < P_LOW || p > P_HIGH) {
        const bool lower = p < P_LOW;
        const double pt = lower ? p : 1.0 - p;
        const double q = sqrt(-2.0 * log(pt));
        double num = fma(OFG_AK_C0, q, OFG_AK_C1);
        num = fma(num, q, OFG_AK_C2);
        num = fma(num, q, OFG_AK_C3);
        num = fma(num, q, OFG_AK_C4);
        num = fma(num, q, OFG_AK_C5);
        double den = fma(OFG_AK_D0, q, OFG_AK_D1);
        den = fma(den, q, OFG_AK_D2);
        den = fma(den, q, OFG_AK_D3);
        den = fma(den, q, 1.0);
        const double r = num / den;
        z = lower ? r : -r;
    } else {
        const double q = p - 0.5;
        const double r = q * q;
        double num = fma(OFG_AK_A0, r, OFG_AK_A1);
        num = fma(num, r, OFG_AK_A2);
        num = fma(num, r, OFG_AK_A3);
        num = fma(num, r, OFG_AK_A4);
        num = fma(num, r, OFG_AK_A5);
        double den = fma(OFG_AK_B0, r, OFG_AK_B1);
        den = fma(den, r, OFG_AK_B2);
        den = fma(den, r, OFG_AK_B3);
        den = fma(den, r, OFG_AK_B4);
        den = fma(den, r, 1.0);
        z = num * q / den;
    }
    return z;
}

// sample_standard_normal (fast_rng.rs:223-226) on a xoshiro stream.
OFG_HD double reference_normal(Xoshiro256pp &rng) {
    return acklam_inv_cdf(uniform_open01(rng.next_f64()));
}

// ------------------------------------------------ fast-mode generator ---

// xoshiro256+ (Blackman & Vigna 2018): the same state and transition as
// xoshiro256++ with the cheaper "+" scrambler (s0 + s3), the authors'
// recommendation for floating-point output from the upper bits. Seeded by the
// same SplitMix64 expansion of stream_seed(seed, sample).
struct Xoshiro256p {
    uint64_t s0, s1, s2, s3;

    OFG_HD void seed(uint64_t seed) {
        uint64_t sm = seed;
        s0 = splitmix64_next(sm);
        s1 = splitmix64_next(sm);
        s2 = splitmix64_next(sm);
        s3 = splitmix64_next(sm);
        if ((s0 | s1 | s2 | s3) == 0) s0 = 1;
    }

    // Steps back over the last draw and returns it again.
    OFG_HD uint64_t prev_u64();

    OFG_HD uint64_t next_u64() {
        const uint64_t result = s0 + s3;
        const uint64_t t = s1 << 17;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = rotl64(s3, 45);
        return result;
    }
};

OFG_HD uint64_t Xoshiro256pp::prev_u64() {
    xoshiro256_unstep(s0, s1, s2, s3);
    return rotl64(s0 + s3, 23) + s0;
}

OFG_HD uint64_t Xoshiro256p::prev_u64() {
    xoshiro256_unstep(s0, s1, s2, s3);
    return s0 + s3;
}

OFG_HD double bits_as_double(uint64_t b) {
    double d;
    __builtin_memcpy(&d, &b, sizeof d);
    return d;
}

OFG_HD double rcp_seed(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcp(d);
#else
    return 1.0 / d;
#endif
}

OFG_HD double rsq_seed(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsq(d);
#else
    return 1.0 / sqrt(d);
#endif
}

// Natural log of a positive normal double, branch-free.
// x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
// evaluated as 2s + 2s z Q(z), z = s^2 <= 0.0295, Q a degree-6 Chebyshev-node
// interpolant of (atanh(sqrt z)/sqrt z - 1)/z (max relative error of the
// reconstructed log 4.6e-18 in exact arithmetic). The division is a
// reciprocal seed with one Newton step and a final residual correction.
OFG_HD double log_fp64(double x) {
    int e;
    double m = frexp(x, &e);
    const bool lo = m < 0.70710678118654752;
    m = lo ? m + m : m;
    const double ed = (double)(lo ? e - 1 : e);
    const double f = m - 1.0;
    const double d = f + 2.0;
    double r = rcp_seed(d);
    r = fma(r, fma(-d, r, 1.0), r);
    double s = f * r;
    s = fma(fma(-d, s, f), r, s);
    const double z = s * s;
    double q = 0.07308224842521703;
    q = fma(q, z, 0.07665860800278021);
    q = fma(q, z, 0.09091444562630861);
    q = fma(q, z, 0.1111110556739754);
    q = fma(q, z, 0.14285714312987743);
    q = fma(q, z, 0.19999999999949752);
    q = fma(q, z, 0.3333333333333335);
    const double s2 = s + s;
    const double lm = fma(s2 * z, q, s2);
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    return fma(ed, LN2_HI, fma(ed, LN2_LO, lm));
}

// sqrt of a positive finite double from the reciprocal square-root seed and
// one Goldschmidt refinement.
OFG_HD double sqrt_pos(double t) {
    const double y = rsq_seed(t);
    double g = t * y;
    double h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, t);
    return fma(d, h, g);
}

// sin and cos of |t| <= pi/4: degree-5 Chebyshev-node interpolants in
// w = t^2 of (sin t / t - 1)/w and (cos t - 1 + w/2)/w^2 (max relative error
// 2.0e-17 and 1.3e-18 in exact arithmetic).
OFG_HD void sincos_octant(double t, double &sn, double &cs) {
    const double w = t * t;
    double ps = 1.5918129294866608e-10;
    ps = fma(ps, w, -2.5051131845003624e-08);
    ps = fma(ps, w, 2.755731610255244e-06);
    ps = fma(ps, w, -0.00019841269836758574);
    ps = fma(ps, w, 0.008333333333330948);
    ps = fma(ps, w, -0.16666666666666666);
    sn = fma(t * w, ps, t);
    double pc = -1.1382632425521717e-11;
    pc = fma(pc, w, 2.08761462684032e-09);
    pc = fma(pc, w, -2.7557317271729793e-07);
    pc = fma(pc, w, 2.480158729876569e-05);
    pc = fma(pc, w, -0.0013888888888887398);
    pc = fma(pc, w, 0.041666666666666664);
    cs = fma(w * w, pc, fma(-0.5, w, 1.0));
}

// Box-Muller pair from 52 bits of `wu` (radius), 3 bits of `wu` (symmetry)
// and 32 bits `wt` (angle), all branch-free:
//   u  = (k + 1/2) 2^-52 in (0, 1), k = wu >> 12, so log u is finite and < 0;
//   R  = sqrt(-2 log u);
//   an angle uniform on [-pi/4, pi/4) from wt, mapped to the whole circle by
//   one of the 8 symmetries of the square (swap, two sign flips) chosen by
//   bits 8..10 of wu: the image is uniform on the circle.
OFG_HD void box_muller_pair(uint64_t wu, uint32_t wt, double &z0, double &z1) {
    const double one_k = bits_as_double(0x3FF0000000000000ull | (wu >> 12));  // 1 + k 2^-52
    const double u = one_k - 0.99999999999999988898;                        // exact
    const double rad = sqrt_pos(-2.0 * log_fp64(u));
    const double t = fma((double)wt, 3.657295198167899e-10, -0.7853981633974483);  // pi / 2^33, pi / 4
    double sn, cs;
    sincos_octant(t, sn, cs);
    const bool swap = (wu >> 8) & 1u;
    double a = swap ? sn : cs;
    double b = swap ? cs : sn;
    a = ((wu >> 9) & 1u) ? -a : a;
    b = ((wu >> 10) & 1u) ? -b : b;
    z0 = rad * a;
    z1 = rad * b;
}

// Four standard normals from three 64-bit draws (two Box-Muller pairs; the
// third draw supplies both angles).
template <class Rng>
OFG_HD void normals4(Rng &rng, double &z0, double &z1, double &z2, double &z3) {
    const uint64_t wa = rng.next_u64();
    const uint64_t wb = rng.next_u64();
    const uint64_t wc = rng.next_u64();
    box_muller_pair(wa, (uint32_t)(wc >> 32), z0, z1);
    box_muller_pair(wb, (uint32_t)wc, z2, z3);
}

}  // namespace ofg
