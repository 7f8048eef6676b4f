// Host-side pieces shared by the C ABI translation units: error handling
// (status codes + thread-local message), device buffers and the per-GPU
// context object.
#pragma once

#include "../../include/openferric_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

namespace ofg_host {

inline thread_local std::string g_last_error;

// OFG_TRACE=1 prints each host-side stage to stderr (flushed), so a run that
// stalls shows where.
inline bool trace_on() {
    static const bool on = [] {
        const char *e = std::getenv("OFG_TRACE");
        return e && e[0] == '1';
    }();
    return on;
}
inline void trace(const char *what) {
    if (trace_on()) {
        std::fprintf(stderr, "[ofg] %s\n", what);
        std::fflush(stderr);
    }
}

struct OfgError : std::runtime_error {
    int32_t code;
    OfgError(int32_t c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int32_t code, const std::string &msg) { throw OfgError(code, msg); }

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        fail(OFG_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Armed by ofg_inject_alloc_failure: the k-th device allocation from then on
// (1 = the next) fails as an out-of-memory hipMalloc would; 0 = disarmed.
inline std::atomic<int32_t> g_alloc_fault{0};

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void ensure(size_t bytes) {
        if (bytes <= cap) return;
        if (p) hip_check(hipFree(p), "hipFree");
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes, 256);
        int32_t k = g_alloc_fault.load();
        while (k > 0 && !g_alloc_fault.compare_exchange_weak(k, k - 1)) {
        }
        const hipError_t e = k == 1 ? hipErrorOutOfMemory : hipMalloc(&p, want);
        if (e != hipSuccess) {
            // the caller unwinds and frees the host arrays of copies it has
            // already enqueued: let those finish first
            p = nullptr;
            (void)hipDeviceSynchronize();
            hip_check(e, k == 1 ? "hipMalloc (injected failure)" : "hipMalloc");
        }
        cap = want;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};


struct Tuning {
    int32_t block = 256;
    int32_t item_steps = 8192;
    // item_steps may shrink for a call too small to give the queue
    // ofg_plan::MIN_ITEMS items (not when the caller fixed the size)
    bool auto_items = true;
};

template <class F>
inline int32_t guarded(F &&f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const OfgError &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return OFG_NUMERICAL_ERROR;
    } catch (...) {
        g_last_error = "unknown error";
        return OFG_NUMERICAL_ERROR;
    }
}

struct LsmBuffers;   // ofg_lsm_api.cpp
struct QmcBuffers;   // ofg_qmc_api.cpp
struct MultiBuffers; // ofg_multi_api.cpp
struct BasketBuffers; // ofg_basket_api.cpp
struct DslBuffers;    // ofg_dsl_api.cpp
struct ExactBuffers;  // ofg_exact_api.cpp
struct SpreadBuffers; // ofg_spread_api.cpp

// Instrument validation with the reference's messages (instruments/*.rs
// validate()); throws OfgError.
void validate_instrument(const ofg_option &o);

// Launch-shape overrides of an engine configuration. The kernels exist for
// 64, 128, 256 and 512 threads per block (the Greeks kernels for 64, 128 and
// 256: a 512 request runs them at 256), so any other block size is refused
// before a grid or scratch size is derived from it.
inline void check_tuning(const ofg_engine_cfg &cfg) {
    const int32_t b = cfg.tune_block;
    if (!(b == 0 || b == 64 || b == 128 || b == 256 || b == 512))
        fail(OFG_INVALID_INPUT, "tune_block must be 0, 64, 128, 256 or 512");
    if (cfg.tune_item_steps < 0) fail(OFG_INVALID_INPUT, "tune_item_steps must be >= 0");
}

}  // namespace ofg_host

struct ofg_ctx {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    std::mutex mu;
    ofg_host::DevBuf opts, obs_w, items, order, item_first, item_count, partials, outs, head, xs, ys, n1,
        n2, adj, fill, stage, active, err;
    std::map<std::string, ofg_host::Tuning> tuned;
    ofg_host::LsmBuffers *lsm = nullptr;   // Longstaff-Schwartz work buffers
    ofg_host::QmcBuffers *qmc = nullptr;   // quasi-Monte Carlo tables and partials
    ofg_host::MultiBuffers *multi = nullptr;   // multi-asset scenarios and partials
    ofg_host::BasketBuffers *basket = nullptr; // basket scenarios, seeds and partials
    ofg_host::DslBuffers *dsl = nullptr;       // structured-product programs and snapshots
    ofg_host::ExactBuffers *exact = nullptr;   // exact-terminal pricers' streams and sums
    ofg_host::SpreadBuffers *spread = nullptr; // spread options' jobs and sums
    void *noalloc = nullptr;   // this GPU's fixed region of ofg_price_noalloc (not owned)
    int cus = 256;
};

namespace ofg_host {
// Releases the context's Longstaff-Schwartz buffers (ofg_lsm_api.cpp).
void lsm_release(ofg_ctx *ctx);
// Releases the context's quasi-Monte Carlo buffers (ofg_qmc_api.cpp).
void qmc_release(ofg_ctx *ctx);
// Releases the context's multi-asset buffers (ofg_multi_api.cpp).
void multi_release(ofg_ctx *ctx);
// Releases the context's basket buffers (ofg_basket_api.cpp).
void basket_release(ofg_ctx *ctx);
// Releases the context's structured-product buffers (ofg_dsl_api.cpp).
void dsl_release(ofg_ctx *ctx);
// Releases the context's exact-terminal pricer buffers (ofg_exact_api.cpp).
void exact_release(ofg_ctx *ctx);
// Releases the context's spread-option buffers (ofg_spread_api.cpp).
void spread_release(ofg_ctx *ctx);
}  // namespace ofg_host
