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

// Log table for log_fp64: for m in [0.5, 1), 128 intervals of width 1/256
// with midpoints c_i = 0.5 + (i + 1/2)/256; entry i holds T1 = 1/c_i rounded to
// double and T2 = -log(T1) (computed to 50 digits, then rounded), so that
// log m = T2 + log1p(m T1 - 1) with |m T1 - 1| <= 0.00391 (values generated
// with mpmath for this project). The last interval uses c = 1 (T1 = 1,
// T2 = 0): r = m - 1 is then exact and log x keeps full relative accuracy as
// x -> 1 from below.
#define OFG_LOG_TABLE_N 128
static constexpr double OFG_LOG_TABLE[2 * OFG_LOG_TABLE_N] = {
    1.9922178988326849, -0.689248540144288,
    1.9768339768339769, -0.6814965633399701,
    1.9616858237547892, -0.6738042177168143,
    1.946768060836502, -0.6661705928617432,
    1.9320754716981132, -0.6585947990532856,
    1.9176029962546817, -0.6510759666392583,
    1.903345724907063, -0.6436132454376686,
    1.8892988929889298, -0.6362058041598069,
    1.8754578754578755, -0.628852829854548,
    1.8618181818181818, -0.6215535273729365,
    1.848375451263538, -0.6143071188521694,
    1.8351254480286738, -0.6071128432181422,
    1.8220640569395017, -0.599969955705762,
    1.8091872791519434, -0.5928777273962701,
    1.7964912280701755, -0.5858354447708572,
    1.78397212543554, -0.5788424092798867,
    1.7716262975778547, -0.5718979369270757,
    1.7594501718213058, -0.5650013578680152,
    1.7474402730375427, -0.5581520160224405,
    1.735593220338983, -0.5513492686996879,
    1.7239057239057238, -0.5445924862368081,
    1.7123745819397993, -0.5378810516488214,
    1.7009966777408638, -0.5312143602906321,
    1.6897689768976898, -0.5245918195301387,
    1.678688524590164, -0.5180128484320962,
    1.6677524429967427, -0.5114768774523106,
    1.656957928802589, -0.5049833481417624,
    1.6463022508038585, -0.49853171286027365,
    1.6357827476038338, -0.4921214344993545,
    1.6253968253968254, -0.4857519862138747,
    1.615141955835962, -0.47942285116222716,
    1.6050156739811912, -0.47313352225466315,
    1.5950155763239875, -0.4668835019094919,
    1.585139318885449, -0.4606723018168513,
    1.5753846153846154, -0.4544994427097703,
    1.5657492354740061, -0.4483644541422544,
    1.5562310030395137, -0.4422668742741359,
    1.5468277945619335, -0.43620624966244487,
    1.5375375375375375, -0.43018213505906394,
    1.528358208955224, -0.42419409321444135,
    1.5192878338278932, -0.4182416946871461,
    1.5103244837758112, -0.41232451765905753,
    1.501466275659824, -0.40644214775599097,
    1.4927113702623906, -0.4005941778735679,
    1.4840579710144928, -0.39478020800814806,
    1.4755043227665705, -0.3889998450926485,
    1.4670487106017192, -0.38325270283708057,
    1.4586894586894588, -0.37753840157364205,
    1.4504249291784703, -0.37185656810621104,
    1.4422535211267606, -0.36620683556409206,
    1.4341736694677871, -0.36058884325986873,
    1.426183844011142, -0.3550022365512289,
    1.4182825484764543, -0.3494466667066269,
    1.4104683195592287, -0.343921790774657,
    1.4027397260273973, -0.3384272714570163,
    1.3950953678474114, -0.3329627769849375,
    1.3875338753387534, -0.3275279809989806,
    1.3800539083557952, -0.3221225624320727,
    1.3726541554959786, -0.31674620539569226,
    1.3653333333333333, -0.31139859906909695,
    1.3580901856763925, -0.306079437591497,
    1.3509234828496042, -0.3007884199570814,
    1.3438320209973753, -0.2955252499128068,
    1.3368146214099217, -0.2902896358588618,
    1.3298701298701299, -0.28508129075172356,
    1.322997416020672, -0.279899932009726,
    1.3161953727506426, -0.2747452814210614,
    1.3094629156010231, -0.26961706505414207,
    1.3027989821882953, -0.2645150131702466,
    1.2962025316455696, -0.2594388601383859,
    1.2896725440806045, -0.25438834435231733,
    1.2832080200501252, -0.24936320814964427,
    1.2768079800498753, -0.24436319773293858,
    1.2704714640198511, -0.23938806309282482,
    1.2641975308641975, -0.23443755793296864,
    1.257985257985258, -0.22951143959691278,
    1.2518337408312958, -0.22460946899670603,
    1.245742092457421, -0.2197314105432732,
    1.2397094430992737, -0.21487703207847508,
    1.2337349397590363, -0.21004610480880959,
    1.2278177458033572, -0.20523840324070627,
    1.2219570405727924, -0.2004537051173701,
    1.2161520190023754, -0.19569179135712642,
    1.210401891252955, -0.1909524459932298,
    1.204705882352941, -0.18623545611509087,
    1.199063231850117, -0.18154061181088324,
    1.1934731934731935, -0.1768677061114908,
    1.1879350348027842, -0.17221653493575995,
    1.1824480369515011, -0.16758689703701793,
    1.1770114942528735, -0.16297859395082367,
    1.17162471395881, -0.15839142994391764,
    1.1662870159453302, -0.15382521196433638,
    1.1609977324263039, -0.14927974959266183,
    1.1557562076749435, -0.14475485499437207,
    1.150561797752809, -0.14025034287326765,
    1.145413870246085, -0.13576603042593893,
    1.1403118040089086, -0.13130173729725345,
    1.1352549889135255, -0.12685728553682943,
    1.130242825607064, -0.12243249955647377,
    1.1252747252747253, -0.11802720608855737,
    1.1203501094091903, -0.11364123414530306,
    1.1154684095860568, -0.10927441497896273,
    1.1106290672451193, -0.10492658204285929,
    1.1058315334773219, -0.10059757095327378,
    1.1010752688172043, -0.09628721945215148,
    1.0963597430406853, -0.09199536737061052,
    1.091684434968017, -0.0877218565932284,
    1.0870488322717622, -0.08346653102309001,
    1.0824524312896406, -0.07922923654757486,
    1.0778947368421052, -0.07500982100486656,
    1.0733752620545074, -0.07080813415116662,
    1.068893528183716, -0.06662402762859244,
    1.0644490644490645, -0.06245735493374666,
    1.060041407867495, -0.05830797138693517,
    1.0556701030927835, -0.054175734102024614,
    1.051334702258727, -0.05006050195691803,
    1.047034764826176, -0.04596213556463585,
    1.0427698574338085, -0.04188049724498711,
    1.0385395537525355, -0.037815450996817664,
    1.0343434343434343, -0.033766862470817484,
    1.0301810865191148, -0.029734598942879144,
    1.0260521042084167, -0.025718529287989036,
    1.0219560878243512, -0.021718523954642903,
    1.0178926441351888, -0.017734454939768475,
    1.0138613861386139, -0.01376619576414797,
    1.009861932938856, -0.00981362144832467,
    1.005893909626719, -0.005876608488984971,
    1.0, 0.0,
};

// Natural log of x in (0, 1] (the Box-Muller argument), branch-free:
// x = m 2^e (frexp, m in [0.5, 1)), log x = e ln2 + T2[i] + log1p(r),
// r = fma(m, T1[i], -1), log1p by its Taylor series through r^7 (truncation
// 1.7e-18 relative). `tab` points at OFG_LOG_TABLE, or at a copy in LDS.
OFG_HD double log_fp64(double x, const double *tab) {
    int e;
    const double m = frexp(x, &e);
    uint64_t mb;
    __builtin_memcpy(&mb, &m, sizeof mb);
    const int i = (int)((mb >> 45) & 127u);  // top 7 fraction bits of m in [0.5, 1)
    const double t1 = tab[2 * i];
    const double t2 = tab[2 * i + 1];
    const double r = fma(m, t1, -1.0);
    double p = 1.0 / 7.0;
    p = fma(p, r, -1.0 / 6.0);
    p = fma(p, r, 1.0 / 5.0);
    p = fma(p, r, -0.25);
    p = fma(p, r, 1.0 / 3.0);
    p = fma(p, r, -0.5);
    const double l1p = fma(r * r, p, r);
    const double ed = (double)e;
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    return fma(ed, LN2_HI, t2 + fma(ed, LN2_LO, l1p));
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
OFG_HD void box_muller_pair(uint64_t wu, uint32_t wt, double &z0, double &z1, const double *tab) {
    const double one_k = bits_as_double(0x3FF0000000000000ull | (wu >> 12));  // 1 + k 2^-52
    const double u = one_k - 0.99999999999999988898;                        // exact
    const double rad = sqrt_pos(-2.0 * log_fp64(u, tab));
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
OFG_HD void normals4(Rng &rng, double &z0, double &z1, double &z2, double &z3, const double *tab) {
    const uint64_t wa = rng.next_u64();
    const uint64_t wb = rng.next_u64();
    const uint64_t wc = rng.next_u64();
    box_muller_pair(wa, (uint32_t)(wc >> 32), z0, z1, tab);
    box_muller_pair(wb, (uint32_t)wc, z2, z3, tab);
}

}  // namespace ofg
