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
//   * fast RNG mode: xoroshiro128+ streams and a branch-free Box-Muller
//     transform with table-driven -2 log, sin/cos of a full turn and a
//     Newton sqrt; a table-driven 2^(y/256) for the path steps.
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

// 64-bit rotate left by a constant 0 < k < 64, k != 32. On the device it is
// two v_alignbit_b32 (funnel shifts of the 32-bit halves): for k < 32 the new
// high word is ({hi, lo} >> (32 - k))[31:0] and the new low word
// ({lo, hi} >> (32 - k))[31:0]; for k > 32 the halves swap roles.
OFG_HD uint64_t rotl64(uint64_t x, int k) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    uint32_t nh, nl;
    if (k < 32) {
        nh = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)(32 - k));
        nl = __builtin_amdgcn_alignbit(lo, hi, (uint32_t)(32 - k));
    } else {
        nh = __builtin_amdgcn_alignbit(lo, hi, (uint32_t)(64 - k));
        nl = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)(64 - k));
    }
    return ((uint64_t)nh << 32) | nl;
#else
    return (x << k) | (x >> (64 - k));
#endif
}

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

// Inverse of the xoshiro256 state transition: the
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

// uniform_open01 (fast_rng.rs:218-221): clamp to [eps, 1 - eps]. u is a
// next_f64 draw (finite, in [0, 1)), so the reference's two compare-selects
// are exactly a max and a min.
OFG_HD double uniform_open01(double u) {
    const double lo = 2.220446049250313e-16;
    const double hi = 1.0 - 2.220446049250313e-16;
    return fmin(fmax(u, lo), hi);
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

// Acklam's two regions (p already clamped to [eps, 1 - eps] by
// uniform_open01, so the NaN / 0 / 1 guards of the reference never fire on
// this path): the tail rational in q = sqrt(-2 ln p') and the central one in
// q = p - 1/2.
constexpr double OFG_AK_P_LOW = 0.02425;
constexpr double OFG_AK_P_HIGH = 1.0 - OFG_AK_P_LOW;

OFG_HD bool acklam_is_tail(double p) { return p < OFG_AK_P_LOW || p > OFG_AK_P_HIGH; }

// n / d, correctly rounded, for the operands of the Acklam rationals: d in
// [2.6e-3, 1] (central) or [36, 447] (tail), n finite with |n| in
// [2^-56, 2^12] or n = +0. On the device this is the Newton-Markstein
// sequence of the compiler's IEEE f64 division (reciprocal, two Newton
// refinements, quotient and one residual correction) without its
// v_div_scale range scaling, v_div_fmas rescale and v_div_fixup special
// cases, which leave operands in these ranges unchanged: the same
// instructions on the same values, so the same bits as n / d.
OFG_HD double acklam_div(double n, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    const double q = n * r;
    return fma(fma(-d, q, n), r, q);
#else
    return n / d;
#endif
}

OFG_HD double acklam_tail(double p) {
    const bool lower = p < OFG_AK_P_LOW;
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
    const double r = acklam_div(num, den);
    return lower ? r : -r;
}

OFG_HD double acklam_central(double p) {
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
    return acklam_div(num * q, den);
}

OFG_HD double acklam_inv_cdf(double p) {
    return acklam_is_tail(p) ? acklam_tail(p) : acklam_central(p);
}

// K inverse-CDF values at once, bit for bit acklam_inv_cdf of each p[k],
// with the tail region deferred. About 4.9 % of uniforms fall in a tail, so
// a 64-lane wave almost always holds a tail lane and a draw-by-draw branch
// costs both rationals per draw. Here every lane evaluates the central
// rational for its K values, then the tail rounds run one tail value per
// lane per round: the wave runs as many rounds as its busiest lane has tail
// values (about 2.3 for K = 8, against about 7.7 tail evaluations per 8 draws
// draw by draw). The selects keep p and z in registers.
template <int K>
__device__ __forceinline__ void acklam_block(const double (&p)[K], double (&z)[K]) {
    uint32_t tails = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        tails |= (uint32_t)acklam_is_tail(p[k]) << k;
        z[k] = acklam_central(p[k]);
    }
    while (tails) {
        const int j = __builtin_ctz(tails);
        tails &= tails - 1;
        double pj = p[0];
#pragma unroll
        for (int k = 1; k < K; ++k) pj = j == k ? p[k] : pj;
        const double t = acklam_tail(pj);
#pragma unroll
        for (int k = 0; k < K; ++k) z[k] = j == k ? t : z[k];
    }
}

#if defined(__HIP__)  // wave-level code: HIP compilations only (not the host tests' g++ builds)
// Orders this wave's LDS accesses: its writes before the barrier are seen by
// its reads after it (LDS executes one wave's instructions in order; this
// keeps the compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acklam_block with the tail region evaluated once per wave instead of once
// per lane: every lane evaluates the central rational for its K values; the
// p of the whole wave's tail values go to consecutive slots of the wave's
// LDS buffer (value k of every lane, lane order, then k + 1, ...), the
// calling lanes invert them one each per round, and each lane reads its
// values back from their slots. A block of K = 8 holds about 25 tail values
// per wave, so one round of tail evaluations nearly always serves the block,
// against about 2.3 rounds when each lane inverts its own tail values in
// turn. Bit for bit acklam_inv_cdf of each p[k]. draw(k) supplies p[k],
// k = 0 .. K-1 in order; slots is this wave's 64 K doubles of LDS. Any
// subset of the wave's lanes may make the call, all of them together.
template <int K, class Draw>
__device__ __forceinline__ void acklam_block_wave(Draw &&draw, double (&z)[K], double *slots) {
    uint32_t slot[K];     // this lane's slot of value k, ~0u for a central value
    uint32_t total = 0;   // wave-uniform count of tail values
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double p = draw(k);
        const bool t = acklam_is_tail(p);
        z[k] = acklam_central(p);
        const uint64_t m = __ballot(t);
        // tail lanes below this one at k, after the earlier positions' values
        const uint32_t s = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, total));
        slot[k] = t ? s : ~0u;
        if (t) slots[s] = p;
        total += (uint32_t)__popcll(m);
    }
    if (total == 0) return;
    wave_lds_sync();
    // the lanes executing this call (the tail of a sample range leaves a
    // wave partly idle) share the rounds by their rank among themselves
    const uint64_t act = __ballot(true);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
    const uint32_t n_act = (uint32_t)__popcll(act);
    for (uint32_t r = 0; r < total; r += n_act) {
        const uint32_t i = r + rank;
        if (i < total) slots[i] = acklam_tail(slots[i]);
    }
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (slot[k] != ~0u) z[k] = slots[slot[k]];
    wave_lds_sync();   // the buffer is reused by the next block
}

// reference_steps (below) with each block's tail values inverted per wave
// (acklam_block_wave): the same draws in the same order, the same normals.
// n_steps must be the same on every lane of the wave; slots holds the
// wave's 64 K doubles.
template <int K, int PER, class RNG, class F>
__device__ __forceinline__ void reference_steps_wave(RNG &rng, int n_steps, F &&f,
                                                     double *slots) {
    static_assert(K % PER == 0, "whole steps per block");
    constexpr int SPB = K / PER;
    for (int j = 0; j < n_steps; j += SPB) {
        double z[K];
        acklam_block_wave<K>(
            [&](int i) { return j + i / PER < n_steps ? uniform_open01(rng.next_f64()) : 0.5; },
            z, slots);
#pragma unroll
        for (int i = 0; i < SPB; ++i) {
            if (j + i < n_steps) {
                if constexpr (PER == 2) f(j + i, z[2 * i], z[2 * i + 1]);
                else f(j + i, z[i], 0.0);
            }
        }
    }
}
#endif  // __HIP__

// n_steps path steps of PER reference normals each (Heston: z1 then z2),
// drawn K at a time from the stream and inverted by acklam_block; f(j, z1,
// z2) runs for j = 0 .. n_steps - 1 in order (z2 = 0 when PER is 1). The
// stream advances by exactly PER * n_steps draws.
template <int K, int PER, class RNG, class F>
__device__ __forceinline__ void reference_steps(RNG &rng, int n_steps, F &&f) {
    static_assert(K % PER == 0, "whole steps per block");
    constexpr int SPB = K / PER;
    for (int j = 0; j < n_steps; j += SPB) {
        double p[K], z[K];
#pragma unroll
        for (int i = 0; i < K; ++i)
            p[i] = j + i / PER < n_steps ? uniform_open01(rng.next_f64()) : 0.5;
        acklam_block<K>(p, z);
#pragma unroll
        for (int i = 0; i < SPB; ++i) {
            if (j + i < n_steps) {
                if constexpr (PER == 2) f(j + i, z[2 * i], z[2 * i + 1]);
                else f(j + i, z[i], 0.0);
            }
        }
    }
}

// state <- r(T) state for a GF(2) polynomial r (bit i = coefficient of x^i,
// four words): the XOR over the set bits i of T^i state, T the xoshiro256
// transition. With r = x^k mod P(x), P the transition's characteristic
// polynomial, this jumps the stream k draws ahead.
OFG_HD void xoshiro_apply_poly(const uint64_t *r, Xoshiro256pp &st) {
    Xoshiro256pp acc;
    acc.s0 = acc.s1 = acc.s2 = acc.s3 = 0;
    Xoshiro256pp cur = st;
    for (int w = 0; w < 4; ++w) {
        const uint64_t word = r[w];
        for (int b = 0; b < 64; ++b) {
            if ((word >> b) & 1u) {
                acc.s0 ^= cur.s0;
                acc.s1 ^= cur.s1;
                acc.s2 ^= cur.s2;
                acc.s3 ^= cur.s3;
            }
            (void)cur.next_u64();
        }
    }
    st = acc;
}

// sample_standard_normal (fast_rng.rs:223-226) on a xoshiro stream.
OFG_HD double reference_normal(Xoshiro256pp &rng) {
    return acklam_inv_cdf(uniform_open01(rng.next_f64()));
}

// ------------------------------------------------ fast-mode generator ---

// xoroshiro128+ (Blackman & Vigna, "Scrambled linear pseudorandom number
// generators", 2018; parameters a = 24, b = 16, c = 37): 128 bits of state
// and the "+" scrambler, the authors' generator for floating-point output from
// the upper bits. One draw is 12 VALU operations (xoshiro256: 14) and the
// state is 4 VGPRs. Seeded by two SplitMix64 outputs of stream_seed(seed,
// sample); an all-zero state is replaced by s0 = 1.
struct Xoroshiro128p {
    uint64_t s0, s1;

    OFG_HD void seed(uint64_t seed) {
        uint64_t sm = seed;
        s0 = splitmix64_next(sm);
        s1 = splitmix64_next(sm);
        if ((s0 | s1) == 0) s0 = 1;
    }

    OFG_HD uint64_t next_u64() {
        const uint64_t result = s0 + s1;
        const uint64_t t = s1 ^ s0;
        s0 = rotl64(s0, 24) ^ t ^ (t << 16);
        s1 = rotl64(t, 37);
        return result;
    }

    // Steps back over the last draw and returns it again. With t = old s0 ^
    // old s1: t = rotr(s1, 37) and old s0 = rotr(s0 ^ t ^ (t << 16), 24).
    OFG_HD uint64_t prev_u64() {
        const uint64_t t = rotl64(s1, 27);
        const uint64_t a = s0 ^ t ^ (t << 16);
        s0 = rotl64(a, 40);
        s1 = t ^ s0;
        return s0 + s1;
    }
};

OFG_HD uint64_t Xoshiro256pp::prev_u64() {
    xoshiro256_unstep(s0, s1, s2, s3);
    return rotl64(s0 + s3, 23) + s0;
}

// ------------------------------------------------ reference Pcg64 stream ---

OFG_HD uint64_t mul_hi_u64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Pcg64 of fast_rng.rs:93-137 (PCG XSL-RR 128/64): the 128-bit state held as
// two words, advanced by state * M + inc mod 2^128 (one 64x64 full product
// and two low products), each draw the XSL-RR permutation of the state
// BEFORE the step: (hi ^ lo) rotated right by hi >> 58. Seeded from three
// SplitMix64 outputs (state high, state low, stream; inc = stream << 1 | 1
// as a 128-bit value), then one draw discarded.
struct Pcg64 {
    static constexpr uint64_t M_HI = 0x2360ED051FC65DA4ull;
    static constexpr uint64_t M_LO = 0x4385DF649FCCF645ull;
    uint64_t lo, hi, inc_lo, inc_hi;

    OFG_HD void seed(uint64_t seed) {
        uint64_t sm = seed;
        hi = splitmix64_next(sm);
        lo = splitmix64_next(sm);
        const uint64_t stream = splitmix64_next(sm);
        inc_lo = (stream << 1) | 1ull;
        inc_hi = stream >> 63;
        (void)next_u64();
    }

    OFG_HD uint64_t next_u64() {
        const uint64_t olo = lo, ohi = hi;
        const uint64_t plo = olo * M_LO;
        const uint64_t phi = mul_hi_u64(olo, M_LO) + olo * M_HI + ohi * M_LO;
        lo = plo + inc_lo;
        hi = phi + inc_hi + (lo < plo ? 1ull : 0ull);
        const uint64_t x = ohi ^ olo;
        const uint32_t r = (uint32_t)(ohi >> 58);
        return (x >> r) | (x << ((64u - r) & 63u));
    }

    OFG_HD double next_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
};

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

// Constant tables of the fast mode (log, exp2, sin/cos), generated by
// gen_fast_tables.py. The kernels stage a copy in LDS; `tab` below points at
// OFG_FAST_TABLE itself or at that copy.
#include "ofg_fast_tables.inc"
#define OFG_TAB_LOG 0
#define OFG_TAB_EXP2 512
#define OFG_TAB_SINCOS 768

// -2 log x for x in (0, 1] (the Box-Muller radius argument), branch-free:
// x = m 2^e (frexp, m in [0.5, 1)), interval i of m from its top 8 fraction
// bits, r = m T1[i] - 1 (|r| <= 0.00196), and
//   -2 log x = -2 e ln2 + 2 log T1[i] + r q(r),
// q(r) = -2 log1p(r) / r by its Taylor series through r^5 (truncation
// 8.1e-18 relative). ln2 is taken as the double nearest to it: the error,
// |e| 4.6e-17, stays below 3.4e-17 of the result because |result| >= 1.38 |e|.
OFG_HD double neg2log_fp64(double x, const double *tab) {
    int e;
    const double m = frexp(x, &e);
    uint64_t mb;
    __builtin_memcpy(&mb, &m, sizeof mb);
    const int i = (int)((mb >> 44) & 255u);
    const double t1 = tab[OFG_TAB_LOG + 2 * i];
    const double t2 = tab[OFG_TAB_LOG + 2 * i + 1];
    const double r = fma(m, t1, -1.0);
    double q = 1.0 / 3.0;
    q = fma(q, r, -0.4);
    q = fma(q, r, 0.5);
    q = fma(q, r, -2.0 / 3.0);
    q = fma(q, r, 1.0);
    q = fma(q, r, -2.0);
    return fma((double)e, -1.3862943611198906, t2 + r * q);
}

// sqrt of a positive finite double: the reciprocal square-root seed y
// (relative error ~2^-23 on gfx950), then two Newton corrections
// g <- g + (t - g^2) y/2 (errors ~1.5 eps^2, then ~1.5 eps^3).
OFG_HD double sqrt_pos(double t) {
    const double y = rsq_seed(t);
    const double h = 0.5 * y;
    double g = t * y;
    g = fma(fma(-g, g, t), h, g);
    return fma(fma(-g, g, t), h, g);
}

// (sin, cos) of the angle 2 pi (wt + 1/2) / 2^32: the table holds the pair at
// a_i = 2 pi (i + 1/2) / 512 for the top 9 bits i of wt; the remaining 23
// bits give d = angle - a_i, |d| <= pi/512, and the pair is rotated by d:
//   sin d = d (1 - w/6 + w^2/120)  (relative truncation 1.1e-17),
//   cos d = 1 + w (-1/2 + c w)     (c fitted, absolute error 7.9e-18),
// w = d^2.
OFG_HD void sincos_turn(uint32_t wt, double &sn, double &cs, const double *tab) {
    const uint32_t i = wt >> 23;
    const double sa = tab[OFG_TAB_SINCOS + 2 * i];
    const double ca = tab[OFG_TAB_SINCOS + 2 * i + 1];
    // 2 pi / 2^32 and 2 pi (1/2 - 2^22) / 2^32
    const double d = fma((double)(wt & 0x7FFFFFu), 1.4629180792671596e-09, -0.0061359224200835255);
    const double w = d * d;
    const double sd = fma(d * w, fma(w, 1.0 / 120.0, -1.0 / 6.0), d);
    const double cm = w * fma(w, 0.041666619912862315, -0.5);
    sn = fma(ca, sd, fma(sa, cm, sa));
    cs = fma(-sa, sd, fma(ca, cm, ca));
}

// Box-Muller pair from the top 52 bits of `wu` (radius) and 32 bits `wt`
// (angle), branch-free:
//   u = (k + 1/2) 2^-52 in (0, 1), k = wu >> 12, so log u is finite and < 0;
//   (z0, z1) = sqrt(-2 log u) (cos, sin)(2 pi (wt + 1/2) / 2^32).
OFG_HD void box_muller_pair(uint64_t wu, uint32_t wt, double &z0, double &z1, const double *tab) {
    const double one_k = bits_as_double(0x3FF0000000000000ull | (wu >> 12));  // 1 + k 2^-52
    const double u = one_k - 0.99999999999999988898;                        // exact
    const double rad = sqrt_pos(neg2log_fp64(u, tab));
    double sn, cs;
    sincos_turn(wt, sn, cs, tab);
    z0 = rad * cs;
    z1 = rad * sn;
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

// 2^(y/256) for |y| < 2^31, the fast mode's exp: callers scale the exponent
// by 256/ln2 once per option (x = y ln2/256, so |x| < 5.8e6; beyond |x| ~ 745
// ldexp already returns 0 or inf as exp does). Adding 1.5 2^52 rounds y to
// the integer k held in the low 32 bits; k splits into the table index
// j = k mod 256 and the binary exponent floor(k/256), and f = y - k is exact
// (|f| <= 1/2); 2^(f/256) - 1 is its Taylor series through f^4 (truncation
// 3.8e-17 relative).
// t256: the 256 entries 2^(j/256).
OFG_HD double exp2_256t(double y, const double *t256) {
    const double kd = y + 6755399441055744.0;
    const double k = kd - 6755399441055744.0;
    const double f = y - k;
    uint64_t kb;
    __builtin_memcpy(&kb, &kd, sizeof kb);
    const int ki = (int)(uint32_t)kb;
    const double t = t256[ki & 255];
    // (ln2/256)^n / n!, n = 4, 3, 2, 1
    double q = fma(f, 2.239395190875157e-12, 3.3083026805413713e-09);
    q = fma(q, f, 3.6655655969101062e-06);
    q = fma(q, f, 0.0027076061740622863);
    return ldexp(fma(t, f * q, t), ki >> 8);
}

OFG_HD double exp2_256(double y, const double *tab) { return exp2_256t(y, tab + OFG_TAB_EXP2); }

// ln2 / 256 and its inverse: the fast mode's exponent scale; the bound on
// |y| that exp2_256 accepts.
#define OFG_LN2_256 0.0027076061740622863
#define OFG_INV_LN2_256 369.3299304675746
#define OFG_EXP2_256_YMAX 1073741824.0

}  // namespace ofg
