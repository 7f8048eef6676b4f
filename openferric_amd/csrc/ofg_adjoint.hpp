// Pathwise adjoint Greeks (forward pass with sqrt(N)-step checkpoints, then a
// backward sweep that recomputes each segment from its checkpoint).
#pragma once

#include "../../include/openferric_amd.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

namespace ofg {

// Prices n options and fills their Greeks by the adjoint kernels.
int32_t adjoint_price(ofg_ctx *ctx, const ofg_option *opts, size_t n, const ofg_engine_cfg *cfg,
                      ofg_result *out);

// Host helpers shared with ofg_api.cpp.
[[noreturn]] void host_fail(int32_t code, const std::string &msg);
void host_hip_check(hipError_t e, const char *what);
void host_validate(const ofg_option &o);
double host_payoff_at_spot(const ofg_option &o);
hipStream_t ctx_stream(ofg_ctx *ctx);
int32_t ctx_device(ofg_ctx *ctx);
std::mutex &ctx_mutex(ofg_ctx *ctx);
void *ctx_adjoint_scratch(ofg_ctx *ctx, size_t bytes);

}  // namespace ofg
