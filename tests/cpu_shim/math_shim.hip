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
double shim_log(double x) { return ofg::log_fp64(x, ofg::OFG_LOG_TABLE); }
double shim_sqrt(double x) { return ofg::sqrt_pos(x); }
void shim_sincos(double t, double *s, double *c) { ofg::sincos_octant(t, *s, *c); }
void shim_fast_normals(uint64_t seed, uint64_t sample, int64_t n4, double *out) {
    ofg::Xoshiro256p r;
    r.seed(ofg::stream_seed(seed, sample));
    for (int64_t k = 0; k < n4; ++k)
        ofg::normals4(r, out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3], ofg::OFG_LOG_TABLE);
}
void shim_box_muller(uint64_t wu, uint32_t wt, double *z0, double *z1) {
    ofg::box_muller_pair(wu, wt, *z0, *z1, ofg::OFG_LOG_TABLE);
}
}

extern "C" {
// Draws n values forward, then walks back n steps; writes both sequences and
// returns 1 if the state after walking back equals the initial state.
int shim_xoshiro_roundtrip(uint64_t seed, int plus, int64_t n, uint64_t *fwd, uint64_t *bwd) {
    if (plus) {
        ofg::Xoshiro256p r;
        r.seed(seed);
        const ofg::Xoshiro256p r0 = r;
        for (int64_t i = 0; i < n; ++i) fwd[i] = r.next_u64();
        for (int64_t i = 0; i < n; ++i) bwd[i] = r.prev_u64();
        return r.s0 == r0.s0 && r.s1 == r0.s1 && r.s2 == r0.s2 && r.s3 == r0.s3;
    }
    ofg::Xoshiro256pp r;
    r.seed(seed);
    const ofg::Xoshiro256pp r0 = r;
    for (int64_t i = 0; i < n; ++i) fwd[i] = r.next_u64();
    for (int64_t i = 0; i < n; ++i) bwd[i] = r.prev_u64();
    return r.s0 == r0.s0 && r.s1 == r0.s1 && r.s2 == r0.s2 && r.s3 == r0.s3;
}
}
