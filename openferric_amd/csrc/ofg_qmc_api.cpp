// Host side of ofg_qmc_price: mc_european_qmc_with_seed
// (src/engines/monte_carlo/mc_qmc.rs:40-144) for n options per call.
//
// The reference returns results instead of errors for its degenerate inputs
// (NaN price and no stderr for zero paths or steps and for a vol that is not
// finite and positive; the payoff at the spot with stderr 0 for expiry <= 0);
// those are reproduced here on the host. SobolSequence::new asserts
// 1 <= dimensions <= 21201 (math/sobol.rs:14, :26-29); that limit is an
// InvalidInput here. The direction numbers and scramblers are built on the
// host exactly as build_direction_numbers / new (sobol.rs:26-43, 121-152).
#include "../../include/openferric_amd.h"
#include "ofg_host.hpp"
#include "ofg_qmc.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace ofg_host;

namespace ofg_host {

struct QmcBuffers {
    DevBuf opts, tile_opt, dirs, lows, scr, partials, out;
};

void qmc_release(ofg_ctx *ctx) {
    if (!ctx || !ctx->qmc) return;
    QmcBuffers *b = ctx->qmc;
    for (DevBuf *d : {&b->opts, &b->tile_opt, &b->dirs, &b->lows, &b->scr, &b->partials, &b->out})
        d->release();
    delete b;
    ctx->qmc = nullptr;
}

}  // namespace ofg_host

namespace {

constexpr int64_t SOBOL_MAX_DIMENSIONS = 21201;

// splitmix64 of sobol.rs:147-152 (a stateless mix of x)
uint64_t sob_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// build_direction_numbers (sobol.rs:121-144): dimension 0 is van der Corput,
// the others take odd m_j < 2^(j+1) from the hash, left-aligned.
void directions(uint64_t dim, uint64_t seed, uint64_t *v) {
    for (int j = 0; j < 64; ++j) {
        if (dim == 0) {
            v[j] = 1ull << (63 - j);
            continue;
        }
        const uint64_t hash = sob_mix(seed ^ ((dim + 1) << 40) ^ (uint64_t)j);
        const uint64_t mask = j == 63 ? ~0ull : ((1ull << (j + 1)) - 1);
        v[j] = ((hash | 1ull) & mask) << (63 - j);
    }
}

int32_t qmc_impl(ofg_ctx *ctx, const ofg_option *opts, size_t n, ofg_result *out) {
    if (!ctx || (n && (!opts || !out))) fail(OFG_INVALID_INPUT, "null argument");
    std::vector<ofg::QmcDev> dev;
    std::vector<int32_t> dev_of(n, -1), tile_opt;
    std::vector<uint64_t> dirs, lows, scr;
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        if (o.model != OFG_MODEL_GBM || o.payoff != OFG_PAYOFF_VANILLA)
            fail(OFG_INVALID_INPUT, "quasi-Monte Carlo prices European vanillas under GBM only");
        if (o.option_type != OFG_CALL && o.option_type != OFG_PUT)
            fail(OFG_INVALID_INPUT, "unknown option type");
    }
    for (size_t i = 0; i < n; ++i) {
        const ofg_option &o = opts[i];
        ofg_result &r = out[i];
        std::memset(&r, 0, sizeof r);
        r.num_paths = r.num_steps = r.vol = NAN;
        // mc_qmc.rs:47-81, in the reference's order
        if (o.num_paths <= 0 || o.num_steps <= 0) {
            r.price = NAN;
            continue;
        }
        if (o.expiry <= 0.0) {
            r.price = o.option_type == OFG_CALL ? std::fmax(o.spot - o.strike, 0.0)
                                                : std::fmax(o.strike - o.spot, 0.0);
            r.has_stderr = 1;
            continue;
        }
        if (o.vol <= 0.0 || !std::isfinite(o.vol)) {
            r.price = NAN;
            continue;
        }
        if (o.num_steps > SOBOL_MAX_DIMENSIONS)
            fail(OFG_INVALID_INPUT, "Sobol dimensions must be in [1, 21201]");
        ofg::QmcDev d{};
        d.put = o.option_type == OFG_PUT;
        d.steps = (int32_t)o.num_steps;
        d.paths = o.num_paths;
        d.dim_offset = (int64_t)scr.size();
        const double dt = o.expiry / (double)o.num_steps;
        d.dt_drift = (o.rate - o.div_yield - 0.5 * o.vol * o.vol) * dt;
        d.dt_vol = o.vol * std::sqrt(dt);
        d.discount = std::exp(-o.rate * o.expiry);
        d.s0 = o.spot;
        d.strike = o.strike;
        // tiles over point indices 0..paths (index 0 is not a point)
        // the tile size depends on the option alone (its summation order, and
        // so every bit of its price, is the same in any batch)
        const int64_t idx = o.num_paths + 1, blk = ofg::QMC_BLOCK;
        const int64_t ppt =
            std::max<int64_t>(1, (idx + blk * ofg::QMC_MAX_TILES - 1) / (blk * ofg::QMC_MAX_TILES));
        d.ppt = (int32_t)ppt;
        d.n_tiles = (int32_t)((idx + blk * ppt - 1) / (blk * ppt));
        d.tile_first = (int32_t)tile_opt.size();
        dev_of[i] = (int32_t)dev.size();
        for (int32_t t = 0; t < d.n_tiles; ++t) tile_opt.push_back((int32_t)dev.size());
        for (int64_t dim = 0; dim < o.num_steps; ++dim) {
            uint64_t v[64];
            directions((uint64_t)dim, o.seed, v);
            dirs.insert(dirs.end(), v, v + 64);
            for (int pat = 0; pat < 64; ++pat) {
                uint64_t x = 0;
                for (int b = 0; b < 6; ++b)
                    if ((pat >> b) & 1) x ^= v[b];
                lows.push_back(x);
            }
            scr.push_back(sob_mix(o.seed ^ ((uint64_t)(dim + 1) << 32)));
        }
        dev.push_back(d);
    }

    std::vector<ofg::QmcOut> res(dev.size());
    float ms = 0.f;
    if (!dev.empty()) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        if (!ctx->qmc) ctx->qmc = new QmcBuffers();
        QmcBuffers &b = *ctx->qmc;
        hipStream_t s = ctx->stream;
        auto up = [&](DevBuf &buf, const void *src, size_t bytes, const char *what) {
            buf.ensure(std::max<size_t>(bytes, 1));
            hip_check(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, s), what);
        };
        up(b.opts, dev.data(), dev.size() * sizeof(dev[0]), "H2D qmc options");
        up(b.tile_opt, tile_opt.data(), tile_opt.size() * sizeof(int32_t), "H2D qmc tiles");
        up(b.dirs, dirs.data(), dirs.size() * sizeof(uint64_t), "H2D sobol directions");
        up(b.lows, lows.data(), lows.size() * sizeof(uint64_t), "H2D sobol low tables");
        up(b.scr, scr.data(), scr.size() * sizeof(uint64_t), "H2D sobol scramblers");
        b.partials.ensure(tile_opt.size() * sizeof(ofg::QmcPartial));
        b.out.ensure(dev.size() * sizeof(ofg::QmcOut));
        ofg::QmcArgs a{};
        a.opts = b.opts.as<ofg::QmcDev>();
        a.tile_opt = b.tile_opt.as<int32_t>();
        a.dirs = b.dirs.as<uint64_t>();
        a.lows = b.lows.as<uint64_t>();
        a.scr = b.scr.as<uint64_t>();
        a.partials = b.partials.as<ofg::QmcPartial>();
        a.out = b.out.as<ofg::QmcOut>();
        a.n_tiles = (int32_t)tile_opt.size();
        a.n_opts = (int32_t)dev.size();
        hip_check(hipEventRecord(ctx->ev0, s), "event");
        hip_check(ofg::launch_qmc(a, s), "qmc kernel launch");
        hip_check(hipEventRecord(ctx->ev1, s), "event");
        hip_check(hipMemcpyAsync(res.data(), a.out, res.size() * sizeof(res[0]),
                                 hipMemcpyDeviceToHost, s),
                  "D2H qmc results");
        hip_check(hipStreamSynchronize(s), "qmc kernel execution");
        hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "event time");
    }
    for (size_t i = 0; i < n; ++i) {
        out[i].kernel_ms = ms;
        if (dev_of[i] < 0) continue;
        const ofg::QmcOut &q = res[(size_t)dev_of[i]];
        out[i].price = q.price;
        out[i].stderr_ = q.stderr_;
        out[i].has_stderr = 1;
        // mc_qmc.rs:132-135
        out[i].num_paths = (double)opts[i].num_paths;
        out[i].num_steps = (double)opts[i].num_steps;
        out[i].vol = opts[i].vol;
    }
    return OFG_OK;
}

}  // namespace

extern "C" int32_t ofg_qmc_price(ofg_ctx *ctx, const ofg_option *opts, size_t n,
                                 ofg_result *out) {
    return guarded([&]() { return qmc_impl(ctx, opts, n, out); });
}
