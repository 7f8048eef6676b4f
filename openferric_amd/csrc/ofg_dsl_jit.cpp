// Run-time compilation of a product's programs to a native evaluation
// kernel (see ofg_dsl_jit.hpp). Each program instruction becomes the C++
// statement the interpreter executes for it (ofg_dsl.hip evaluate_path),
// with register indices, mask slots, constants, the notional and the asset
// count written in: registers become named locals the compiler keeps in
// VGPRs, and no instruction is fetched or decoded at run time.
#include "ofg_dsl_jit.hpp"

#include "../../include/openferric_amd.h"
#include "ofg_host.hpp"

#include <dlfcn.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace ofg_dsl_jit {

using ofg_host::fail;

ProgramImage assemble(const ofg_dsl::Product &prod) {
    ProgramImage im;
    int n_locals = 0;
    std::vector<ofg_dsl::Program> progs;
    try {
        progs = ofg_dsl::build_programs(prod, n_locals);
    } catch (const ofg_dsl::DslError &e) {
        fail(OFG_NUMERICAL_ERROR, e.what());
    }
    im.n_locals = n_locals;
    int32_t n_obs = 0;
    for (size_t k = 0; k < prod.schedules.size(); ++k) {
        ofg::DslSched d{};
        d.obs_first = n_obs;
        d.n_obs = (int32_t)prod.schedules[k].dates.size();
        n_obs += d.n_obs;
        ofg_dsl::Program &pg = progs[k];
        const int32_t c0 = (int32_t)im.consts.size();
        for (ofg::DslIns in : pg.code) {
            if (in.op == ofg::DOP_CONST) in.aux += (uint32_t)c0;
            im.code.push_back(in);
        }
        d.code_off = (int32_t)(im.code.size() - pg.code.size());
        d.code_len = (int32_t)pg.code.size();
        im.consts.insert(im.consts.end(), pg.consts.begin(), pg.consts.end());
        im.n_regs = std::max(im.n_regs, pg.n_regs);
        im.n_masks = std::max(im.n_masks, pg.n_masks);
        im.sched.push_back(d);
    }
    im.n_regs = std::max<int32_t>(im.n_regs, n_locals + (int32_t)prod.state_vars.size());
    if (im.n_regs > ofg::DSL_MAX_REGS || im.n_masks > ofg::DSL_MAX_MASKS)
        fail(OFG_INVALID_INPUT, "product program exceeds the device limits (32 registers, 16 nested ifs)");
    for (const auto &sv : prod.state_vars) im.state_init.push_back(sv.initial);
    return im;
}

namespace {

// A double as an exact C++ hexadecimal literal, formatted from its bits
// (independent of the host's locale): 0x1.<52-bit fraction>p<exponent> for
// normal numbers, 0x0.<fraction>p-1022 for subnormals.
std::string lit(double v) {
    if (std::isnan(v)) return "__builtin_nan(\"\")";
    if (std::isinf(v)) return v > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
    uint64_t bits;
    std::memcpy(&bits, &v, sizeof bits);
    const bool neg = (bits >> 63) != 0;
    const int e = (int)((bits >> 52) & 0x7FF);
    const uint64_t frac = bits & 0xFFFFFFFFFFFFFull;
    char buf[48];
    if (e == 0)
        std::snprintf(buf, sizeof buf, "0x0.%013llxp-1022", (unsigned long long)frac);
    else
        std::snprintf(buf, sizeof buf, "0x1.%013llxp%+d", (unsigned long long)frac, e - 1023);
    return neg ? "(-" + std::string(buf) + ")" : std::string(buf);
}

std::string R(int i) { return "r" + std::to_string(i); }

// One instruction as the interpreter's statement for it.
std::string statement(const ofg::DslIns &in, const ProgramImage &im, double notional,
                      int32_t n_assets) {
    const std::string d = R(in.dst), a = R(in.a), b = R(in.b);
    const std::string k = std::to_string(in.aux);
    switch (in.op) {
    case ofg::DOP_CONST: return d + " = " + lit(im.consts[in.aux]) + ";";
    case ofg::DOP_NOTIONAL: return d + " = " + lit(notional) + ";";
    case ofg::DOP_DATE: return d + " = ob.date;";
    case ofg::DOP_ISFINAL: return d + " = ob.is_final ? 1.0 : 0.0;";
    case ofg::DOP_SET: return in.b ? d + " = " + a + ";" : "if (active) " + d + " = " + a + ";";
    case ofg::DOP_ADD: return d + " = " + a + " + " + b + ";";
    case ofg::DOP_SUB: return d + " = " + a + " - " + b + ";";
    case ofg::DOP_MUL: return d + " = " + a + " * " + b + ";";
    case ofg::DOP_DIV:
        return "{ const double den = " + b + "; " + d + " = den == 0.0 ? __builtin_nan(\"\") : " +
               a + " / den; }";
    case ofg::DOP_MIN: return d + " = fmin(" + a + ", " + b + ");";
    case ofg::DOP_MAX: return d + " = fmax(" + a + ", " + b + ");";
    case ofg::DOP_EQ:
        return d + " = fabs(" + a + " - " + b + ") < 2.220446049250313e-16 ? 1.0 : 0.0;";
    case ofg::DOP_NE:
        return d + " = fabs(" + a + " - " + b + ") >= 2.220446049250313e-16 ? 1.0 : 0.0;";
    case ofg::DOP_LT: return d + " = " + a + " < " + b + " ? 1.0 : 0.0;";
    case ofg::DOP_LE: return d + " = " + a + " <= " + b + " ? 1.0 : 0.0;";
    case ofg::DOP_GT: return d + " = " + a + " > " + b + " ? 1.0 : 0.0;";
    case ofg::DOP_GE: return d + " = " + a + " >= " + b + " ? 1.0 : 0.0;";
    case ofg::DOP_AND: return d + " = (" + a + " != 0.0 && " + b + " != 0.0) ? 1.0 : 0.0;";
    case ofg::DOP_OR: return d + " = (" + a + " != 0.0 || " + b + " != 0.0) ? 1.0 : 0.0;";
    case ofg::DOP_NEG: return d + " = -" + a + ";";
    case ofg::DOP_ABS: return d + " = fabs(" + a + ");";
    case ofg::DOP_EXP: return d + " = exp(" + a + ");";
    case ofg::DOP_LOG: return d + " = log(" + a + ");";
    case ofg::DOP_NOT: return d + " = " + a + " != 0.0 ? 0.0 : 1.0;";
    case ofg::DOP_WOF_PERF:
    case ofg::DOP_BOF_PERF: {
        const bool worst = in.op == ofg::DOP_WOF_PERF;
        return std::string("{ double v = ") + (worst ? "__builtin_inf()" : "(-__builtin_inf())") +
               "; for (int i = 0; i < " + std::to_string(n_assets) +
               "; ++i) { if (o.s0[i] > 0.0) { const double perf = row[i * stride] / o.s0[i]; "
               "if (perf " + (worst ? "<" : ">") + " v) v = perf; } } " + d + " = v; }";
    }
    case ofg::DOP_WORST:
    case ofg::DOP_BEST: {
        const bool worst = in.op == ofg::DOP_WORST;
        std::string s = std::string("{ double v = ") +
                        (worst ? "__builtin_inf()" : "(-__builtin_inf())") + ";";
        for (int j = (int)in.aux - 1; j >= 0; --j)   // popped last first
            s += " if (" + R(in.a + j) + (worst ? " < " : " > ") + "v) v = " + R(in.a + j) + ";";
        return s + " " + d + " = v; }";
    }
    case ofg::DOP_PRICE:
        return "{ int idx; double shown; const bool ok = ofg_index_in_range(" + a + ", " +
               std::to_string(n_assets) +
               ", idx, shown); if (!ok && active) err = fmax(err, shown); " + d +
               " = ok ? row[idx * stride] : 0.0; }";
    case ofg::DOP_PAY: return "if (active) pv += " + a + " * df;";
    case ofg::DOP_REDEEM:
        return "if (active) { pv += " + a + " * df; alive = false; active = false; }";
    case ofg::DOP_SKIP: return "if (active) { alive = false; active = false; }";
    case ofg::DOP_IF:
        return "sv" + k + " = active; cd" + k + " = " + a + " != 0.0; active = active && cd" + k + ";";
    case ofg::DOP_ELSE: return "active = sv" + k + " && !cd" + k + " && alive;";
    case ofg::DOP_ENDIF: return "active = sv" + k + " && alive;";
    default: return "";
    }
}

// hipRTC, opened on first use: the library itself loads without it (only
// ofg_dsl_jit_compile needs the compiler).
struct Rtc {
    hiprtcResult (*create)(hiprtcProgram *, const char *, const char *, int, const char **,
                           const char **) = nullptr;
    hiprtcResult (*compile)(hiprtcProgram, int, const char **) = nullptr;
    hiprtcResult (*log_size)(hiprtcProgram, size_t *) = nullptr;
    hiprtcResult (*log)(hiprtcProgram, char *) = nullptr;
    hiprtcResult (*code_size)(hiprtcProgram, size_t *) = nullptr;
    hiprtcResult (*code)(hiprtcProgram, char *) = nullptr;
    hiprtcResult (*destroy)(hiprtcProgram *) = nullptr;
    const char *(*error)(hiprtcResult) = nullptr;
    std::string problem;   // why it is unavailable, if it is
};

template <class F>
bool bind(void *lib, const char *name, F &f) {
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

const Rtc &rtc() {
    static const Rtc r = [] {
        Rtc x;
        void *lib = dlopen("libhiprtc.so", RTLD_NOW | RTLD_LOCAL);
        if (!lib) lib = dlopen("/opt/rocm/lib/libhiprtc.so", RTLD_NOW | RTLD_LOCAL);
        if (!lib) {
            const char *e = dlerror();
            x.problem = std::string("hipRTC is not available: ") + (e ? e : "libhiprtc.so");
            return x;
        }
        if (!(bind(lib, "hiprtcCreateProgram", x.create) &&
              bind(lib, "hiprtcCompileProgram", x.compile) &&
              bind(lib, "hiprtcGetProgramLogSize", x.log_size) &&
              bind(lib, "hiprtcGetProgramLog", x.log) &&
              bind(lib, "hiprtcGetCodeSize", x.code_size) && bind(lib, "hiprtcGetCode", x.code) &&
              bind(lib, "hiprtcDestroyProgram", x.destroy) &&
              bind(lib, "hiprtcGetErrorString", x.error)))
            x.problem = "hipRTC is missing an entry point";
        return x;
    }();
    return r;
}

std::string rtc_log(const Rtc &r, hiprtcProgram p) {
    size_t n = 0;
    if (r.log_size(p, &n) != HIPRTC_SUCCESS || n == 0) return "";
    std::string s(n, '\0');
    if (r.log(p, &s[0]) != HIPRTC_SUCCESS) return "";
    return s;
}

}  // namespace

std::string kernel_source(const ProgramImage &im, double notional, int32_t n_assets) {
    std::string s;
    s.reserve(8192);
    s += OFG_DSL_STR(OFG_DSL_JIT_LAYOUTS);
    s += R"(
__device__ __forceinline__ bool ofg_index_in_range(double v, int n, int &idx, double &shown) {
    if (!(v >= 1.0)) { idx = 0; shown = 0.0; return n > 0; }
    if (v >= (double)n) {
        idx = 0;
        shown = v >= 1.8446744073709552e19 ? 1.8446744073709552e19 : trunc(v);
        return false;
    }
    idx = (int)v;
    shown = (double)idx;
    return true;
}
__device__ __forceinline__ double ofg_wave_sum(double v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ __forceinline__ double ofg_wave_max(double v) {
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
}
extern "C" __global__ void __launch_bounds__(256) ofg_dsl_eval_jit(DslJitArgs a) {
    __shared__ double red[4][3];
    const int tile = blockIdx.x;
    const DslJitScen &o = a.scen[a.tile_scen[tile]];
    const long long base = (long long)(tile - o.tile_first) * 256 * o.ppt;
    const long long stride = o.paths;
    double sum = 0.0, sq = 0.0, err = -1.0;
    for (int j = 0; j < o.ppt; ++j) {
        const long long p = base + (long long)j * 256 + threadIdx.x;
        if (p >= o.paths) break;
        const double *snap = a.snaps + o.snap_off + p;
)";
    for (int i = 0; i < im.n_regs; ++i) s += "        double " + R(i) + " = 0.0;\n";
    for (int k = 0; k < std::max(im.n_masks, 1); ++k)
        s += "        bool sv" + std::to_string(k) + " = false, cd" + std::to_string(k) + " = false;\n";
    for (size_t k = 0; k < im.state_init.size(); ++k)
        s += "        " + R(im.n_locals + (int)k) + " = " + lit(im.state_init[k]) + ";\n";
    s += "        bool alive = true;\n        double pv = 0.0;\n";
    for (const ofg::DslSched &sc : im.sched) {
        s += "        for (int oi = 0; oi < " + std::to_string(sc.n_obs) + "; ++oi) {\n";
        s += "            if (!__any(alive)) break;\n";
        s += "            const DslJitObs ob = a.obs[" + std::to_string(sc.obs_first) + " + oi];\n";
        s += "            const double df = a.df[o.df_off + " + std::to_string(sc.obs_first) +
             " + oi];\n";
        s += "            const double *row = snap + (long long)ob.snap * " +
             std::to_string(n_assets) + " * stride;\n";
        for (int i = 0; i < im.n_locals; ++i) s += "            " + R(i) + " = 0.0;\n";
        s += "            bool active = alive;\n";
        for (int32_t pc = sc.code_off; pc < sc.code_off + sc.code_len; ++pc)
            s += "            " + statement(im.code[(size_t)pc], im, notional, n_assets) + "\n";
        s += "            (void)active;\n        }\n";
    }
    s += R"(        sum += pv;
        sq += pv * pv;
    }
    sum = ofg_wave_sum(sum);
    sq = ofg_wave_sum(sq);
    err = ofg_wave_max(err);
    const int lane = threadIdx.x & 63, wave = threadIdx.x / 64;
    if (lane == 0) {
        red[wave][0] = sum;
        red[wave][1] = sq;
        red[wave][2] = err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0][0], q = red[0][1], e = red[0][2];
        for (int w = 1; w < 4; ++w) {
            s += red[w][0];
            q += red[w][1];
            e = fmax(e, red[w][2]);
        }
        a.partials[tile].sum = s;
        a.partials[tile].sum_sq = q;
        a.partials[tile].err_idx = e;
    }
}
)";
    return s;
}

std::vector<char> compile(const std::string &source) {
    const Rtc &r = rtc();
    if (!r.problem.empty()) fail(OFG_DEVICE_ERROR, r.problem);
    hiprtcProgram prog = nullptr;
    if (r.create(&prog, source.c_str(), "ofg_dsl_eval_jit.hip", 0, nullptr, nullptr) !=
        HIPRTC_SUCCESS)
        fail(OFG_DEVICE_ERROR, "hiprtcCreateProgram failed");
    // the library's own build flags: contraction off (the interpreter's
    // separate multiplies and adds), no packed FP32 instructions
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = r.compile(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    if (rc != HIPRTC_SUCCESS) {
        const std::string log = rtc_log(r, prog);
        r.destroy(&prog);
        fail(OFG_DEVICE_ERROR, std::string("product kernel compilation failed: ") + r.error(rc) +
                                   "\n" + log.substr(0, 2000));
    }
    size_t n = 0;
    std::vector<char> code;
    if (r.code_size(prog, &n) == HIPRTC_SUCCESS && n > 0) {
        code.resize(n);
        if (r.code(prog, code.data()) != HIPRTC_SUCCESS) code.clear();
    }
    r.destroy(&prog);
    if (code.empty()) fail(OFG_DEVICE_ERROR, "product kernel compilation gave no code");
    return code;
}

Kernel load(const std::vector<char> &code) {
    Kernel k;
    ofg_host::hip_check(hipModuleLoadData(&k.module, code.data()), "hipModuleLoadData");
    const hipError_t e = hipModuleGetFunction(&k.fn, k.module, "ofg_dsl_eval_jit");
    if (e != hipSuccess) {
        (void)hipModuleUnload(k.module);
        ofg_host::hip_check(e, "hipModuleGetFunction");
    }
    return k;
}

}  // namespace ofg_dsl_jit
