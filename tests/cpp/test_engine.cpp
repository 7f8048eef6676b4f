// The reference's Monte Carlo engine tests, ported onto the C++ mirror of the
// pricing contract (include/openferric_amd/engine.hpp) and run on the GPU.
// Each check cites the reference test it restates; tolerances are the
// reference's. Prints one line per check and exits non-zero on failure.
#include <openferric_amd/engine.hpp>

#include <cmath>
#include <cstdio>
#include <string>

using namespace openferric_amd;

static int failures = 0;

static void expect(bool ok, const std::string &name, const std::string &detail = "") {
    std::printf("%s %s %s\n", ok ? "PASS" : "FAIL", name.c_str(), detail.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

static const double BS_CALL = 10.450583572185565;  // S=K=100, r=0.05, sigma=0.2, T=1

static Market flat(double s, double r, double q, double v) {
    return Market::builder().spot(s).rate(r).dividend_yield(q).flat_vol(v).build();
}

int main() {
    const Market m = flat(100.0, 0.05, 0.0, 0.2);
    const auto call = VanillaOption::european_call(100.0, 1.0);

    {  // mc_engine.rs:998-1021
        auto r = GpuMonteCarloPricingEngine(100000, 252, 42).price(call, m);
        const double rel = std::fabs((r.price - BS_CALL) / BS_CALL);
        expect(rel <= 0.01, "mc_european_call_matches_black_scholes_within_one_percent",
               std::to_string(r.price));
        expect(r.diagnostics.get("num_paths").value_or(0) == 100000.0 &&
                   r.diagnostics.get("num_steps").value_or(0) == 252.0 &&
                   r.diagnostics.get("vol").value_or(0) == 0.2,
               "diagnostics_num_paths_num_steps_vol");
    }
    {  // mc_engine.rs:1024-1046
        auto plain = GpuMonteCarloPricingEngine(100000, 252, 42).price(call, m);
        auto anti = GpuMonteCarloPricingEngine(100000, 252, 42)
                        .with_variance_reduction(VarianceReduction::Antithetic)
                        .price(call, m);
        expect(*anti.stderr_ < *plain.stderr_, "mc_antithetic_has_lower_stderr_than_plain_mc");
    }
    {  // mc_engine.rs:1049-1073
        auto r = GpuMonteCarloPricingEngine(100000, 252, 42)
                     .with_variance_reduction(VarianceReduction::ControlVariate)
                     .price(call, m);
        expect(std::fabs((r.price - BS_CALL) / BS_CALL) <= 0.005,
               "mc_european_control_variate_is_within_half_percent_of_bs", std::to_string(r.price));
    }
    {  // mc_engine.rs:1097-1132 (control variate, 100k x 252, 12 fixings)
        AsianOption a{OptionType::Call, 100.0, 1.0, {Averaging::Arithmetic, StrikeType::Fixed, {}}};
        for (int k = 0; k < 12; ++k) a.asian.observation_times.push_back(k / 11.0);
        auto r = GpuMonteCarloPricingEngine(100000, 252, 42)
                     .with_variance_reduction(VarianceReduction::ControlVariate)
                     .price(a, m);
        expect(std::fabs((r.price - 5.73) / 5.73) <= 0.02,
               "arithmetic_asian_mc_kemna_turnbull_reference_within_two_percent",
               std::to_string(r.price));
    }
    {  // mc_engine.rs:1135-1167
        auto e = GpuMonteCarloPricingEngine(20000, 64, 123);
        auto a = e.price(call, m), b = e.price(call, m);
        expect(std::fabs(a.price - b.price) < 1e-12 && std::fabs(*a.stderr_ - *b.stderr_) < 1e-12,
               "mc_seeded_xoshiro_is_reproducible");
    }
    {  // validation (instruments/vanilla.rs:69-96, mc_engine.rs:876-912)
        bool threw = false;
        try {
            GpuMonteCarloPricingEngine(1000, 10, 1).price(VanillaOption::european_call(-1.0, 1.0), m);
        } catch (const PricingError &e) {
            threw = e.kind() == PricingError::Kind::InvalidInput &&
                    std::string(e.what()) == "invalid input: vanilla strike must be > 0";
        }
        expect(threw, "invalid_strike_is_invalid_input");
        threw = false;
        try {
            GpuMonteCarloPricingEngine(0, 10, 1).price(call, m);
        } catch (const PricingError &e) {
            threw = std::string(e.what()) == "invalid input: num_paths must be > 0";
        }
        expect(threw, "zero_paths_is_invalid_input");
        auto z = GpuMonteCarloPricingEngine(1000, 10, 1).price(VanillaOption::european_put(110.0, 0.0), m);
        expect(z.price == 10.0 && z.stderr_ && *z.stderr_ == 0.0 && z.diagnostics.is_empty(),
               "maturity_zero_returns_intrinsic");
    }
    {  // barrier in + out = vanilla, path by path (mc_engine.rs:488-514)
        BarrierOption in{OptionType::Call, 100.0, 1.0, {BarrierDirection::Up, BarrierStyle::In, 120.0, 0.0}};
        BarrierOption out = in;
        out.barrier.style = BarrierStyle::Out;
        auto e = GpuMonteCarloPricingEngine(50000, 128, 9);
        const double s = e.price(in, m).price + e.price(out, m).price;
        const double v = e.price(call, m).price;
        expect(std::fabs(s - v) <= 1e-12 * v, "barrier_in_plus_out_equals_vanilla");
    }
    {  // mc_aad.rs:353-409 via price_with_greeks_aad
        const Market mq = flat(100.0, 0.03, 0.01, 0.2);
        auto r = GpuMonteCarloPricingEngine(120000, 64, 42)
                     .with_variance_reduction(VarianceReduction::Antithetic)
                     .price_with_greeks_aad(call, mq);
        // BS Greeks, q = 0.01, r = 0.03, sigma = 0.2, T = 1, S = K = 100
        const double d1 = (std::log(1.0) + (0.03 - 0.01 + 0.02) * 1.0) / 0.2, d2 = d1 - 0.2;
        auto N = [](double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); };
        const double pdf = std::exp(-0.5 * d1 * d1) / std::sqrt(2 * M_PI);
        const double delta = std::exp(-0.01) * N(d1);
        const double vega = 100.0 * std::exp(-0.01) * pdf;
        const double rho = 100.0 * std::exp(-0.03) * N(d2);
        const double theta = -100.0 * std::exp(-0.01) * pdf * 0.2 / 2.0 + 0.01 * 100.0 * std::exp(-0.01) * N(d1) -
                             0.03 * 100.0 * std::exp(-0.03) * N(d2);
        expect(r.greeks.has_value(), "pricing_engine_trait_has_aad_entry_point_for_mc");
        auto f = GpuMonteCarloPricingEngine(120000, 64, 42)
                     .with_variance_reduction(VarianceReduction::Antithetic)
                     .with_greek_method(GreekMethod::Forward)
                     .price_with_greeks_aad(call, mq);
        expect(f.greeks && std::fabs(f.greeks->delta - r.greeks->delta) < 1e-10 &&
                   std::fabs(f.greeks->vega - r.greeks->vega) < 1e-9 &&
                   std::fabs(f.greeks->rho - r.greeks->rho) < 1e-9 &&
                   std::fabs(f.greeks->theta - r.greeks->theta) < 1e-9,
               "forward_pathwise_greeks_equal_adjoint");
        expect(std::fabs(r.greeks->delta - delta) < 6e-3 && std::fabs(r.greeks->vega - vega) < 8e-2 &&
                   std::fabs(r.greeks->rho - rho) < 7e-2 && std::fabs(r.greeks->theta - theta) < 9e-2,
               "mc_pathwise_aad_recovers_bs_first_order_greeks");
    }
    {  // mc_aad.rs:438-480: Heston AAD delta vs CRN central bump (h = 1e-3)
        const Heston h{0.03, 1.8, 0.04, 0.6, -0.5, 0.04};
        const auto k70 = VanillaOption::european_call(70.0, 1.0);
        auto e = HestonMonteCarloPricingEngine(h, 60000, 64, 7);
        auto aad = e.price_with_greeks_aad(k70, flat(100.0, 0.03, 0.0, 0.2));
        const double up = e.price(k70, flat(100.001, 0.03, 0.0, 0.2)).price;
        const double dn = e.price(k70, flat(99.999, 0.03, 0.0, 0.2)).price;
        const double bump = (up - dn) / 2e-3;
        expect(std::fabs(aad.greeks->delta - bump) < 1e-4,
               "heston_aad_delta_matches_common_random_number_bump_to_1e4",
               std::to_string(aad.greeks->delta) + " vs " + std::to_string(bump));
        expect(aad.heston_gradient && aad.heston_gradient->size() == 6, "heston_gradient_present");
    }
    {  // mc_greeks.rs:330-413: pathwise delta (2 %), LR vega (5 %), QuantLib values
        auto e = MonteCarloGreeksEngine(100000, 42);
        auto pw = e.estimate_pathwise(call, m);
        auto lr = e.estimate_likelihood_ratio(call, m);
        const double d1 = (std::log(1.0) + (0.05 + 0.5 * 0.04) * 1.0) / 0.2;
        auto N = [](double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); };
        const double vega = 100.0 * std::exp(-0.5 * d1 * d1) / std::sqrt(2 * M_PI);
        expect(std::fabs((pw.delta - N(d1)) / N(d1)) <= 0.02,
               "pathwise_delta_converges_to_closed_form_call_delta", std::to_string(pw.delta));
        expect(std::fabs((lr.vega - vega) / vega) <= 0.05,
               "likelihood_ratio_vega_converges_to_closed_form_call_vega", std::to_string(lr.vega));
        expect(std::fabs(pw.delta - 0.6368) / 0.6368 < 0.02 &&
                   std::fabs(pw.gamma - 0.01876) / 0.01876 < 0.20 &&
                   std::fabs(lr.vega - 37.524) / 37.524 < 0.05,
               "quantlib_reference_values_are_reasonable");
    }
    {  // mc_parallel.rs: exact-terminal European against Black-Scholes
        auto r = mc_european_parallel(call, m, 2000000, 252);
        expect(std::fabs(r.price - BS_CALL) <= 3.0 * *r.stderr_ &&
                   r.diagnostics.get("num_steps").value_or(0) == 1.0,
               "mc_european_parallel_within_3se_of_bs", std::to_string(r.price));
        expect(std::isnan(mc_european_sequential(call, m, 0).price),
               "mc_european_sequential_zero_paths_is_nan");
    }
    {  // pricing/asian.rs:207-297 and pricing/barrier.rs:313-355
        auto [mc, se] = geometric_asian_price_mc(OptionType::Call, AsianStrike::Fixed(100.0), 100.0,
                                                 0.05, 0.2, 1.0, 252, 80000, 44);
        const double mm = std::log(100.0) + (0.05 - 0.02) * 1.0 / 2.0, v = 0.04 / 3.0;
        const double dd1 = (mm - std::log(100.0) + v) / std::sqrt(v);
        auto N = [](double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); };
        const double closed =
            std::exp(-0.05) * (std::exp(mm + 0.5 * v) * N(dd1) - 100.0 * N(dd1 - std::sqrt(v)));
        expect(std::fabs(mc - closed) <= 2.0 * se + 2e-2, "geometric_asian_mc_converges_to_closed_form",
               std::to_string(mc) + " vs " + std::to_string(closed));
        auto a = arithmetic_asian_price_mc(OptionType::Call, AsianStrike::Fixed(100.0), 100.0, 0.02,
                                           0.2, 1.0, 128, 30000, 11);
        auto g = geometric_asian_price_mc(OptionType::Call, AsianStrike::Fixed(100.0), 100.0, 0.02,
                                          0.2, 1.0, 128, 30000, 11);
        expect(a.first >= g.first - 0.05, "arithmetic_call_is_above_geometric_call_for_fixed_strike");
        auto out = barrier_price_mc(OptionType::Call, BarrierStyle::Out, BarrierDirection::Up, 100.0,
                                    100.0, 110.0, 0.01, 0.2, 1.0, 252, 30000, 10);
        auto inn = barrier_price_mc(OptionType::Call, BarrierStyle::In, BarrierDirection::Up, 100.0,
                                    100.0, 110.0, 0.01, 0.2, 1.0, 252, 30000, 10);
        const double bd1 = (0.01 + 0.02) / 0.2;
        const double vanilla = 100.0 * N(bd1) - 100.0 * std::exp(-0.01) * N(bd1 - 0.2);
        expect(std::fabs(out.first + inn.first - vanilla) <= 2.5 * (out.second + inn.second),
               "barrier_mc_parity_matches_vanilla");
    }
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
