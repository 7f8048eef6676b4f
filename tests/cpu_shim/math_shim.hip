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
