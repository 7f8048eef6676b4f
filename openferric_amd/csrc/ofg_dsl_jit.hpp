// A product's device programs as one image (the per-schedule programs
// concatenated, constants rebased), and their run-time compilation to a
// native gfx950 evaluation kernel (hipRTC): the counterpart of the
// reference's cranelift JIT of the product bytecode (src/dsl/jit.rs, behind
// the "jit" feature). The generated kernel runs the same operations in the
// same order as the predicated-program interpreter (dsl_eval_kernel in
// ofg_dsl.hip) with the registers, masks and constants resolved when
// compiling, so its prices equal the interpreter's bit for bit.
#pragma once

#include "ofg_dsl.hpp"
#include "ofg_dsl_front.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace ofg_dsl_jit {

struct ProgramImage {
    std::vector<ofg::DslSched> sched;   // code range and observation range per schedule
    std::vector<ofg::DslIns> code;      // every schedule's program; DOP_CONST indexes consts
    std::vector<double> consts;
    std::vector<double> state_init;
    int32_t n_regs = 0, n_masks = 0, n_locals = 0;
};

// build_execution_plan's programs (eval.rs:268-330) of a compiled product;
// fails with OFG_NUMERICAL_ERROR / OFG_INVALID_INPUT as ofg_dsl_price does.
ProgramImage assemble(const ofg_dsl::Product &prod);

// Source text of the native evaluation kernel `ofg_dsl_eval_jit` for markets
// of n_assets assets.
std::string kernel_source(const ProgramImage &im, double notional, int32_t n_assets);

struct Kernel {
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
};

// Compiles the source for gfx950 (host work only, no device needed); fails
// with OFG_DEVICE_ERROR and hipRTC's log when compiling fails.
std::vector<char> compile(const std::string &source);

// Loads a compiled code object on the current device.
Kernel load(const std::vector<char> &code);

}  // namespace ofg_dsl_jit
