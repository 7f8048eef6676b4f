// Host-only build of the device math header so the unit tests can check the
// exact functions the kernels use (exp_fp64, Acklam, xoshiro streams) on the
// CPU against the oracle and high-precision references.
#include "../../openferric_amd/csrc/ofg_math.hpp"

extern "C" {
double shim_exp(double x) { return ofg::exp_fp64(x); }
double shim_acklam(double p) { return ofg::acklam_inv_cdf(p); }
void shim_reference_normals(uint64_t seed, uint64_t sample, int64_t n, double *out) {
    ofg::Xoshiro256pp r;
    r.seed(ofg::stream_seed(seed, sample));
    for (int64_t j = 0; j < n; ++j) out[j] = ofg::reference_normal(r);
}
}

extern "C" {
double shim_neg2log(double x) { return ofg::neg2log_fp64(x, ofg::OFG_FAST_TABLE); }
double shim_sqrt(double x) { return ofg::sqrt_pos(x); }
double shim_exp2_256(double y) { return ofg::exp2_256(y, ofg::OFG_FAST_TABLE); }
void shim_sincos_turn(uint32_t wt, double *s, double *c) {
    ofg::sincos_turn(wt, *s, *c, ofg::OFG_FAST_TABLE);
}
void shim_fast_normals(uint64_t seed, uint64_t sample, int64_t n4, double *out) {
    ofg::Xoroshiro128p r;
    r.seed(ofg::stream_seed(seed, sample));
    for (int64_t k = 0; k < n4; ++k)
        ofg::normals4(r, out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3],
                      ofg::OFG_FAST_TABLE);
}
void shim_box_muller(uint64_t wu, uint32_t wt, double *z0, double *z1) {
    ofg::box_muller_pair(wu, wt, *z0, *z1, ofg::OFG_FAST_TABLE);
}
// Terminal spot of one fast-mode GBM sample as the pricing kernel steps it:
// normals in groups of four (the last group's unused normals discarded),
// s *= 2^(y/256), y = diffusion2 z + drift2 with the host's constants.
double shim_fast_gbm_terminal(uint64_t seed, uint64_t sample, double s0, double mu, double sigma,
                              double expiry, int steps) {
    const double dt = expiry / (double)steps;
    const double drift2 = ((mu - 0.5 * sigma * sigma) * dt) * OFG_INV_LN2_256;
    const double diffusion2 = (sigma * sqrt(dt)) * OFG_INV_LN2_256;
    ofg::Xoroshiro128p r;
    r.seed(ofg::stream_seed(seed, sample));
    double s = s0;
    for (int j = 0; j < steps; j += 4) {
        double z[4];
        ofg::normals4(r, z[0], z[1], z[2], z[3], ofg::OFG_FAST_TABLE);
        for (int k = 0; k < 4 && j + k < steps; ++k)
            s *= ofg::exp2_256(fma(diffusion2, z[k], drift2), ofg::OFG_FAST_TABLE);
    }
    return s;
}
// Terminal spot of one fast-mode Heston sample: normals in groups of four
// (two steps each, (z1, z2) = (n0, n1) then (n2, n3)), the re-associated
// Euler step with the host's per-option constants and the generator clamps.
double shim_fast_heston_terminal(uint64_t seed, uint64_t sample, double s0, double mu,
                                 double kappa, double theta, double xi, double rho, double v0,
                                 double expiry, int steps) {
    const double dt = expiry / (double)steps;
    const double sqrt_dt = sqrt(dt);
    const double rho_c = sqrt(1.0 - rho * rho);
    const double kdt = kappa * dt, kthdt = kappa * theta * dt, xsdt = xi * sqrt_dt;
    const double sqrt_dt2 = sqrt_dt * OFG_INV_LN2_256;
    const double mudt2 = (mu * dt) * OFG_INV_LN2_256;
    const double hdt2 = (0.5 * dt) * OFG_INV_LN2_256;
    ofg::Xoroshiro128p r;
    r.seed(ofg::stream_seed(seed, sample));
    double s = s0, v = v0;
    for (int j = 0; j < steps; j += 2) {
        double z[4];
        ofg::normals4(r, z[0], z[1], z[2], z[3], ofg::OFG_FAST_TABLE);
        for (int k = 0; k < 2 && j + k < steps; ++k) {
            const double z1 = z[2 * k], z2 = z[2 * k + 1];
            const double v_pos = fmax(v, 0.0);
            const double sv = v_pos > 0.0 ? ofg::sqrt_pos(v_pos) : 0.0;
            const double zs = fma(rho, z1, rho_c * z2);
            const double vn = fma(xsdt * sv, z1, fma(-kdt, v_pos, v + kthdt));
            double y = fma(sv * sqrt_dt2, zs, fma(-hdt2, v_pos, mudt2));
            y = fmin(fmax(y, -OFG_EXP2_256_YMAX), OFG_EXP2_256_YMAX);
            const double sn = s * ofg::exp2_256(y, ofg::OFG_FAST_TABLE);
            s = !(sn >= 1e-12) ? 1e-12 : sn;
            v = fmax(vn, 0.0);
        }
    }
    return s;
}
}

extern "C" {
// n outputs of the device Pcg64 seeded with seed, and their next_f64 values
// from a second stream of the same seed.
void shim_pcg64(uint64_t seed, int64_t n, uint64_t *out, double *u) {
    ofg::Pcg64 a, b;
    a.seed(seed);
    b.seed(seed);
    for (int64_t i = 0; i < n; ++i) {
        out[i] = a.next_u64();
        u[i] = b.next_f64();
    }
}
// n outputs of xoroshiro128+ from an explicit state (no seeding).
void shim_xoroshiro_from_state(uint64_t s0, uint64_t s1, int64_t n, uint64_t *out) {
    ofg::Xoroshiro128p r;
    r.s0 = s0;
    r.s1 = s1;
    for (int64_t i = 0; i < n; ++i) out[i] = r.next_u64();
}
// Draws n values forward, then walks back n steps; writes both sequences and
// returns 1 if the state after walking back equals the initial state.
int shim_xoshiro_roundtrip(uint64_t seed, int plus, int64_t n, uint64_t *fwd, uint64_t *bwd) {
    if (plus) {
        ofg::Xoroshiro128p r;
        r.seed(seed);
        const ofg::Xoroshiro128p r0 = r;
        for (int64_t i = 0; i < n; ++i) fwd[i] = r.next_u64();
        for (int64_t i = 0; i < n; ++i) bwd[i] = r.prev_u64();
        return r.s0 == r0.s0 && r.s1 == r0.s1;
    }
    ofg::Xoshiro256pp r;
    r.seed(seed);
    const ofg::Xoshiro256pp r0 = r;
    for (int64_t i = 0; i < n; ++i) fwd[i] = r.next_u64();
    for (int64_t i = 0; i < n; ++i) bwd[i] = r.prev_u64();
    return r.s0 == r0.s0 && r.s1 == r0.s1 && r.s2 == r0.s2 && r.s3 == r0.s3;
}
}
