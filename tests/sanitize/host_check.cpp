// Host-code driver for the address and undefined-behaviour sanitizer build
// (tests/test_sanitizers.py): the library's host translation units that take
// caller data -- the structured-product front end (lexer, parser, compiler,
// device-program builder), the correlation repair and Cholesky factor, and
// the pricing launch-plan builder -- compiled from the library's own sources
// with -fsanitize=address,undefined and driven here without a GPU.
//
//   host_check dsl FILE...   compile each product source; one line per file:
//                            "ok <schedules> <locals> <instructions>" or
//                            "err <message>"
//   host_check corr          validate / repair / factor seeded matrices
//   host_check plan          build launch plans of seeded option batches
//
// Any sanitizer report aborts the process with a non-zero status.
#include "../../openferric_amd/csrc/ofg_corr.hpp"
#include "../../openferric_amd/csrc/ofg_dsl_front.hpp"
#include "../../openferric_amd/csrc/ofg_plan.hpp"

#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Rng {  // SplitMix64: a seeded, portable case generator
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u() { return (double)(next() >> 11) * 0x1.0p-53; }
    double range(double a, double b) { return a + (b - a) * u(); }
    int64_t pick(int64_t n) { return (int64_t)(next() % (uint64_t)n); }
};

int run_dsl(int argc, char **argv) {
    for (int i = 2; i < argc; ++i) {
        std::ifstream f(argv[i], std::ios::binary);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string src = ss.str();
        try {
            auto prod = ofg_dsl::parse_and_compile(src);
            int n_locals = 0;
            auto progs = ofg_dsl::build_programs(*prod, n_locals);
            size_t ins = 0;
            for (const auto &p : progs) ins += p.code.size();
            std::printf("ok %zu %d %zu\n", prod->schedules.size(), n_locals, ins);
        } catch (const ofg_dsl::DslError &e) {
            std::printf("err %s\n", e.what());
        } catch (const ofg_host::OfgError &e) {
            std::printf("err %s\n", e.what());
        }
    }
    return 0;
}

int run_corr() {
    Rng r{12345};
    int repaired = 0, refused = 0, factored = 0;
    for (int c = 0; c < 400; ++c) {
        const int n = 1 + (int)r.pick(12);
        ofg_corr::Mat m((size_t)n * n, 0.0);
        const int kind = (int)r.pick(4);
        for (int i = 0; i < n; ++i) {
            m[(size_t)i * n + i] = 1.0;
            for (int j = 0; j < i; ++j) {
                double v = r.range(-1.0, 1.0);
                if (kind == 0) v *= 0.3;                       // usually PSD
                m[(size_t)i * n + j] = m[(size_t)j * n + i] = v;
            }
        }
        if (kind == 2 && n > 1) m[1] = std::nan("");          // refused: not finite
        if (kind == 3 && n > 1) m[1] += 1e-6;                 // refused: asymmetric
        try {
            bool rep = false;
            const ofg_corr::Mat out = ofg_corr::validate_or_repair(m, n, rep);
            ofg_corr::Mat l;
            if (ofg_corr::cholesky_psd(out, n, l)) ++factored;
            repaired += rep ? 1 : 0;
        } catch (const ofg_host::OfgError &) {
            ++refused;
        }
    }
    std::printf("corr repaired %d refused %d factored %d\n", repaired, refused, factored);
    return 0;
}

int run_plan() {
    Rng r{777};
    int built = 0, refused = 0;
    for (int c = 0; c < 300; ++c) {
        const size_t n = 1 + (size_t)r.pick(24);
        std::vector<ofg_option> opts(n);
        std::vector<std::vector<double>> obs(n);
        for (size_t i = 0; i < n; ++i) {
            ofg_option &o = opts[i];
            o = ofg_option{};
            o.model = r.pick(3) == 0 ? OFG_MODEL_HESTON : OFG_MODEL_GBM;
            o.payoff = (int32_t)r.pick(3);
            o.option_type = (int32_t)r.pick(2);
            o.averaging = (int32_t)r.pick(2);
            o.strike_type = (int32_t)r.pick(2);
            o.barrier_dir = (int32_t)r.pick(2);
            o.barrier_style = (int32_t)r.pick(2);
            o.spot = r.range(50, 150);
            o.strike = r.range(50, 150);
            o.expiry = r.pick(10) == 0 ? 0.0 : r.range(0.01, 5.0);
            o.rate = r.range(0, 0.08);
            o.div_yield = r.range(0, 0.03);
            o.vol = r.range(0.05, 0.8);
            o.kappa = r.range(0.2, 5);
            o.theta = r.range(0.01, 0.1);
            o.xi = r.range(0.1, 1);
            o.rho = r.range(-0.9, 0.5);
            o.v0 = r.range(0.01, 0.1);
            o.barrier_level = r.range(60, 180);
            o.num_paths = 1 + r.pick(300000);
            o.num_steps = 1 + r.pick(800);
            o.seed = r.next();
            if (o.payoff == OFG_PAYOFF_ASIAN) {
                const int64_t k = r.pick(40);  // 0: refused (empty observations)
                for (int64_t j = 1; j <= k; ++j) obs[i].push_back(o.expiry * (double)j / (double)k);
                o.obs_times = obs[i].empty() ? nullptr : obs[i].data();
                o.n_obs = (int32_t)obs[i].size();
            }
            switch (r.pick(160)) {  // occasional invalid fields
            case 0: o.strike = -1.0; break;
            case 1: o.num_steps = 0; break;
            case 2: o.num_paths = -5; break;
            case 3: o.payoff = 7; break;
            case 4: o.rho = 1.0; break;
            case 5: o.expiry = -0.5; break;
            default: break;
            }
        }
        ofg_engine_cfg cfg{};
        cfg.antithetic = (int32_t)r.pick(2);
        cfg.control_variate = (int32_t)r.pick(2);
        cfg.rng_mode = (int32_t)r.pick(4);
        ofg_host::Tuning t;
        t.item_steps = 1024 << r.pick(5);
        try {
            const ofg_plan::Plan p = ofg_plan::build_plan(opts.data(), n, cfg, t);
            // every plan option's samples are covered by its items, in order
            std::set<int32_t> seen;
            for (size_t k = 0; k < p.opts.size(); ++k) {
                int64_t next = 0;
                for (int32_t j = 0; j < p.item_count[k]; ++j) {
                    const ofg::WorkItem &w = p.items[(size_t)(p.item_first[k] + j)];
                    if (w.option != (int32_t)k || w.begin != next || w.end <= w.begin) return 3;
                    next = w.end;
                }
                if (next != p.opts[k].samples) return 4;
            }
            for (int32_t x : p.order) seen.insert(x);
            if (seen.size() != p.items.size() || p.order.size() != p.items.size()) return 5;
            ++built;
        } catch (const ofg_host::OfgError &) {
            ++refused;
        }
    }
    std::printf("plan built %d refused %d\n", built, refused);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "dsl") return run_dsl(argc, argv);
    if (mode == "corr") return run_corr();
    if (mode == "plan") return run_plan();
    return 2;
}
