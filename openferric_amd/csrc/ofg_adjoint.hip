// Pathwise adjoint Greeks kernels.
#include "ofg_adjoint.hpp"
#include "ofg_internal.hpp"
#include "ofg_math.hpp"

namespace ofg {

int32_t adjoint_price(ofg_ctx *, const ofg_option *, size_t, const ofg_engine_cfg *,
                      ofg_result *) {
    host_fail(OFG_INVALID_INPUT, "adjoint Greeks are not available in this build");
}

}  // namespace ofg
