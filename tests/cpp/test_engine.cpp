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
    {  // tests/lsm_reference.rs: Longstaff-Schwartz Table 1 (100k x 50, seed 42, 0.10)
        const double table[5][5] = {{36.0, 4.474, 4.831, 7.081, 8.480},
                                    {38.0, 3.242, 3.746, 6.141, 7.647},
                                    {40.0, 2.319, 2.875, 5.318, 6.912},
                                    {42.0, 1.615, 2.216, 4.586, 6.241},
                                    {44.0, 1.100, 1.678, 3.942, 5.642}};
        const LongstaffSchwartzEngine lsm(100000, 50, 42);
        double worst = 0.0;
        for (const auto &row : table) {
            // columns: (sigma, T) = (0.2, 1), (0.2, 2), (0.4, 1), (0.4, 2)
            for (int c = 1; c <= 4; ++c) {
                const double v = c <= 2 ? 0.2 : 0.4;
                const double T = (c % 2) ? 1.0 : 2.0;
                const Market mk = flat(row[0], 0.06, 0.0, v);
                const double px = lsm.price(VanillaOption::american_put(40.0, T), mk).price;
                worst = std::fmax(worst, std::fabs(px - row[c]));
            }
        }
        expect(worst <= 0.10, "lsm_table1_american_puts_within_0_10", std::to_string(worst));
        const Market mk = flat(40.0, 0.06, 0.0, 0.2);
        const auto am = lsm.price(VanillaOption::american_put(40.0, 1.0), mk);
        const auto eu = lsm.price(VanillaOption::european_put(40.0, 1.0), mk);
        expect(am.price > eu.price, "lsm_american_put_above_european");
        expect(am.diagnostics.get("num_paths").value_or(0) == 100000.0 &&
                   am.diagnostics.get("vol").value_or(0) == 0.2,
               "lsm_diagnostics");
    }
    {  // tests/bermudan_quantlib.rs: step-down strikes and Heston dynamics
        const Market mk = flat(100.0, 0.03, 0.0, 0.25);
        BermudanOption step{OptionType::Put, 1.0, {0.25, 0.5, 0.75, 1.0}, {100.0, 97.5, 95.0, 92.5}};
        auto cst = BermudanOption::with_constant_strike(OptionType::Put, 100.0, 1.0,
                                                        {0.25, 0.5, 0.75, 1.0});
        const LongstaffSchwartzEngine lsm(250000, 100, 42);
        expect(lsm.price(step, mk).price < lsm.price(cst, mk).price,
               "bermudan_step_down_strike_prices_below_constant_strike");
        const Market mh = flat(100.0, 0.02, 0.01, 0.20);
        auto inst = BermudanOption::with_constant_strike(OptionType::Put, 100.0, 1.0,
                                                         {0.2, 0.4, 0.6, 0.8, 1.0});
        const auto out = LongstaffSchwartzEngine(220000, 100, 321)
                             .with_heston_dynamics(Heston{0.0, 2.0, 0.04, 0.6, -0.5, 0.04})
                             .price_bermudan_with_boundary(inst, mh);
        bool any = false;
        for (const auto &pt : out.exercise_boundary) any = any || pt.boundary_spot.has_value();
        expect(std::isfinite(out.result.price) && out.result.price > 0.0 && any,
               "bermudan_lsm_heston_dynamics_boundary");
        expect(out.result.diagnostics.get("exercise_dates").value_or(0) == 5.0 &&
                   !out.result.diagnostics.get("vol"),
               "bermudan_lsm_diagnostics");
    }
    {  // mc_qmc.rs:184-197: QMC within 0.1 % of Black-Scholes at 10k points
        const auto q = mc_european_qmc_with_seed(call, m, 10000, 1, 11);
        expect(std::fabs(q.price - BS_CALL) / BS_CALL <= 0.001,
               "qmc_price_matches_black_scholes_with_tight_error_at_10k_paths",
               std::to_string(q.price));
        const auto am = mc_european_qmc(VanillaOption::american_put(100.0, 1.0), m, 100, 4);
        expect(std::isnan(am.price) && !am.stderr_, "qmc_non_european_is_nan");
    }
    {  // basket.rs:1067-1099 and the multi-asset mirrors
        const auto one = price_basket_mc(BasketOption{{1.0}, 100.0, 1.0, true, BasketType::Average},
                                         {100.0}, {0.2}, {1.0}, 0.03, {0.0}, 120000);
        auto N = [](double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); };
        const double d1 = (0.03 + 0.5 * 0.04) / 0.2, d2 = d1 - 0.2;
        const double bs = 100.0 * N(d1) - 100.0 * std::exp(-0.03) * N(d2);
        expect(std::fabs(one.price - bs) <= 3.0 * *one.stderr_ + 0.15,
               "basket_with_one_asset_reduces_to_vanilla_option", std::to_string(one.price));
        const auto bad = price_basket_mc(BasketOption{{0.5}, 100.0, 1.0, true, BasketType::Average},
                                         {100.0, 90.0}, {0.2, 0.2}, {1, 0.3, 0.3, 1}, 0.03,
                                         {0.0, 0.0}, 1000);
        expect(std::isnan(bad.price) && !bad.stderr_, "basket_invalid_weights_is_nan");
        const auto t1 = price_basket_mc_with_copula(
            BasketOption{{1.0}, 100.0, 1.0, true, BasketType::Average}, {100.0}, {0.2}, {1.0},
            0.03, {0.0}, 120000, BasketCopula::StudentT(4));
        expect(std::fabs(t1.price - bs) <= 4.0 * *t1.stderr_,
               "student_t_copula_one_asset_marginal_is_black_scholes", std::to_string(t1.price));
        const auto tbad = price_basket_mc_with_copula(
            BasketOption{{1.0}, 100.0, 1.0, true, BasketType::Average}, {100.0}, {0.2}, {1.0},
            0.03, {0.0}, 100, BasketCopula::StudentT(1));
        expect(std::isnan(tbad.price), "student_t_copula_requires_two_degrees_of_freedom");
        ofg_corr_stress shift{};
        shift.kind = OFG_STRESS_ADDITIVE_SHIFT;
        shift.value = 0.35;
        const auto stressed = stressed_correlation_matrix(
            {1.0, 0.7, 0.6, 0.7, 1.0, 0.65, 0.6, 0.65, 1.0}, 3, {shift});
        expect(stressed[0] == 1.0 && stressed[1] > 0.9 && stressed[1] == stressed[3],
               "stressed_correlation_is_repaired_to_psd");
        const auto sens = basket_sensitivities(
            BasketOption{{0.5, 0.5}, 100.0, 1.0, true, BasketType::Average}, {100.0, 90.0},
            {0.2, 0.25}, {1, 0.3, 0.3, 1}, 0.03, {0.0, 0.0}, 50000);
        expect(sens.delta.size() == 2 && sens.delta[0] > 0.0 && sens.delta[1] > 0.0 &&
                   sens.vega > 0.0,
               "basket_sensitivities_signs");
        Autocallable note;
        note.underlyings = {0, 1};
        note.autocall_dates = {0.25, 0.5, 0.75, 1.0};
        note.coupon_rate = 0.08;
        const auto ac = price_autocallable(note, {100.0, 100.0}, {0.2, 0.25}, {1, 0.5, 0.5, 1},
                                           0.02, 0.0, 20000, 64);
        expect(std::isfinite(ac.price) && ac.greeks.has_value() &&
                   ac.diagnostics.get("observation_count").value_or(0) == 4.0,
               "price_autocallable_with_greeks");
    }
    {  // dsl/mod.rs end_to_end_simple_forward, engine.rs greeks_produce_sensible_delta
        const auto prod = dsl::parse_and_compile(
            "product \"Forward\"\n    notional: 100\n    maturity: 1.0\n    underlyings\n"
            "        SPX = asset(0)\n    schedule annual from 1.0 to 1.0\n"
            "        redeem notional * price(0) / 100\n");
        const auto mk = dsl::MultiAssetMarket::single(100.0, 0.20, 0.05, 0.0);
        const auto r = dsl::DslMonteCarloEngine(10000, 100, 42).price_multi_asset(prod, mk);
        expect(std::fabs(r.price - 100.0) < 1.0, "dsl_forward_product_prices_near_forward",
               std::to_string(r.price));
        const auto g = dsl::DslMonteCarloEngine(20000, 50, 42).greeks_multi_asset(prod, mk, 0);
        expect(std::fabs(g.delta - 1.0) < 0.05, "dsl_greeks_produce_sensible_delta",
               std::to_string(g.delta));
        bool threw = false;
        try {
            dsl::parse_and_compile("product \"x\"\n    notional: 1\n");
        } catch (const PricingError &e) {
            threw = e.kind() == PricingError::Kind::InvalidInput &&
                    std::string(e.what()).find("missing maturity") != std::string::npos;
        }
        expect(threw, "dsl_compile_error_is_invalid_input");
    }
        {  // spread_mc.rs:206-241: spread MC within 1 % of Kirk (300k paths, seed 13)
        const SpreadOption sp{100.0, 96.0, 3.0, 0.20, 0.15, 0.5, 0.0, 0.0, 0.05, 0.5};
        auto r = SpreadMonteCarloEngine(300000, 13).price(sp);
        const double kirk = 5.57702191;   // kirk_reference_case, spread.rs:193-209
        expect(std::fabs((r.price - kirk) / kirk) <= 0.01,
               "spread_mc_matches_kirk_within_one_percent_for_near_zero_strike_case",
               std::to_string(r.price));
        expect(r.diagnostics.get("num_paths").value_or(0) == 300000.0 &&
                   r.diagnostics.get("rho").value_or(0) == 0.5,
               "spread_diagnostics_num_paths_rho");
        auto p = SpreadMonteCarloEngine(300000, 13).with_rng_kind(FastRngKind::Pcg64).price(sp);
        expect(std::fabs((p.price - kirk) / kirk) <= 0.01 && p.price != r.price,
               "spread_mc_pcg64_kind_within_one_percent_of_kirk", std::to_string(p.price));
        bool threw = false;
        try {
            SpreadMonteCarloEngine(10, 1).price(SpreadOption{100.0, 96.0, -1.0, 0.2, 0.15, 0.5,
                                                             0.0, 0.0, 0.05, 0.5});
        } catch (const PricingError &e) {
            threw = std::string(e.what()) == "invalid input: spread strike k must be >= 0";
        }
        expect(threw, "spread_negative_strike_is_invalid_input");
    }
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
