// Probability-analysis entry point: engine build, burn-in, fixed-N or
// early-stop passes, tallies. Restates the control flow of
// ProbabilityAnalyzer<mc::DirectEval>::ComputeTallies
// (src/mc/core/probability_analysis.cc:98-209) and
// convergence_controller (src/mc/scheduler/convergence_controller.h:294-397,
// 438-491) with the Wald half-width rule (convergence_policy.h:28-61).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "mcscram_amd.h"

#include "status.h"

namespace mcs {

// Two-sided normal quantile z with P(|Z| <= z) = confidence, by Acklam's
// rational approximation of the inverse normal CDF (P. J. Acklam, "An
// algorithm for computing the inverse normal cumulative distribution
// function", 2003; relative error < 1.15e-9). The reference uses the same
// method (src/mc/stats/ci_utils.h:30-92).
double normal_quantile_two_sided(double confidence) {
    const double eps = 1.0e-12;
    const double c = std::clamp(confidence, eps, 1.0 - eps);
    const double q = 1.0 - (1.0 - c) / 2.0;
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                                -2.759285104469687e+02, 1.383577518672690e+02,
                                -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                                -1.556989798598866e+02, 6.680131188771972e+01,
                                -1.328068155288572e+01};
    static const double cc[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                                 -2.400758277161838e+00, -2.549732539343734e+00,
                                 4.374664141464968e+00, 2.938163982698783e+00};
    static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                                2.445134137142996e+00, 3.754408661907416e+00};
    const double lo = 0.02425, hi = 1.0 - lo;
    if (q < lo) {
        const double u = std::sqrt(-2.0 * std::log(q));
        return (((((cc[0] * u + cc[1]) * u + cc[2]) * u + cc[3]) * u + cc[4]) * u + cc[5]) /
               ((((d[0] * u + d[1]) * u + d[2]) * u + d[3]) * u + 1.0);
    }
    if (q > hi) {
        const double u = std::sqrt(-2.0 * std::log(1.0 - q));
        return -(((((cc[0] * u + cc[1]) * u + cc[2]) * u + cc[3]) * u + cc[4]) * u + cc[5]) /
               ((((d[0] * u + d[1]) * u + d[2]) * u + d[3]) * u + 1.0);
    }
    const double u = q - 0.5, t = u * u;
    return (((((a[0] * t + a[1]) * t + a[2]) * t + a[3]) * t + a[4]) * t + a[5]) * u /
           (((((b[0] * t + b[1]) * t + b[2]) * t + b[3]) * t + b[4]) * t + 1.0);
}

}  // namespace mcs

extern "C" {

double mcs_normal_quantile_two_sided(double confidence) {
    return mcs::normal_quantile_two_sided(confidence);
}

mcs_status mcs_probability_analysis(const mcs_pdag *pdag, const mcs_settings *settings,
                                    const mcs_sample_shape *shape, const int32_t *observe,
                                    int32_t n_observe, mcs_tally *out, int32_t n_out,
                                    double *p_total, uint32_t *passes_run) {
    if (!pdag || !settings || !shape || !out || !p_total || n_observe < 0 ||
        (n_observe > 0 && !observe)) {
        mcs::set_error("null argument");
        return MCS_ERR_INVALID_ARGUMENT;
    }
    std::vector<int32_t> tallies;
    tallies.push_back(pdag->root);
    for (int32_t i = 0; i < n_observe; ++i) tallies.push_back(observe[i]);
    mcs_engine *eng = nullptr;
    mcs_status st = mcs_engine_create_ex(pdag, shape, tallies.data(), int32_t(tallies.size()),
                                         settings->key0, settings->key1, -1, &eng);
    if (st != MCS_OK) return st;
    mcs_engine_info info{};
    mcs_engine_get_info(eng, &info);
    if (n_out < info.num_tallies) {
        mcs_engine_destroy(eng);
        mcs::set_error("output needs one record for the root and each distinct observed node");
        return MCS_ERR_INVALID_ARGUMENT;
    }
    uint64_t rounded = 0, total_passes = 0;
    if (settings->num_trials) mcs_plan_passes(settings->num_trials, shape, &rounded, &total_passes);
    const bool early_stop = settings->num_trials == 0;
    const uint64_t per_pass = info.num_bitpacks * 64ull;
    const uint32_t cap = settings->max_passes ? settings->max_passes : (1u << 20);
    const double z = mcs::normal_quantile_two_sided(settings->ci_confidence);
    const double delta = settings->ci_rel_margin;
    std::vector<mcs_tally> rec(info.num_tallies);
    uint64_t iterations = 0;
    bool eps_ok = false, eps_log_ok = false;

    // Burn-in: passes until the trial count reaches burn_in_trials; the
    // reference does not check the iteration limit here
    // (convergence_controller.h:351-365, 489-491).
    while (iterations * per_pass < settings->burn_in_trials && iterations < cap) {
        st = mcs_engine_pass_wait_collect(eng, 1, 0, rec.data(), int32_t(rec.size()));
        if (st != MCS_OK) break;
        ++iterations;
    }
    // Steps: stop at the fixed pass count, or once the worst tally meets the
    // Wald targets in linear and log10 space (convergence_policy.h:28-61).
    auto converged = [&]() {
        double max_eps = 0.0, max_eps_log = 0.0, target = 0.0;
        for (const mcs_tally &t : rec) {
            const double p = std::max(t.mean, 1.0e-12);
            const double eps = z * t.std_err;
            if (eps > max_eps) {
                max_eps = eps;
                target = delta * p;
            }
            max_eps_log = std::max(max_eps_log, eps / (p * std::log(10.0)));
        }
        eps_ok = eps_ok || (max_eps > 0 && max_eps <= target);
        eps_log_ok = eps_log_ok || (max_eps_log > 0 && max_eps_log <= delta);
        return eps_ok && eps_log_ok;
    };
    if (st == MCS_OK && early_stop && iterations > 0) converged();
    while (st == MCS_OK && iterations < cap) {
        if (early_stop && eps_ok && eps_log_ok) break;
        if (!early_stop && iterations >= total_passes) break;
        st = mcs_engine_pass_wait_collect(eng, 1, 0, rec.data(), int32_t(rec.size()));
        if (st != MCS_OK) break;
        ++iterations;
        if (early_stop) converged();
    }
    if (st == MCS_OK && iterations == 0)
        st = mcs_engine_collect(eng, rec.data(), int32_t(rec.size()));
    if (st == MCS_OK) {
        std::copy(rec.begin(), rec.end(), out);
        *p_total = rec[0].mean;
        if (passes_run) *passes_run = uint32_t(iterations);
    }
    mcs_engine_destroy(eng);
    return st;
}

}  // extern "C"
