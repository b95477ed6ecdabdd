// Probability-analysis entry point: engine build, burn-in, fixed-N or
// early-stop passes, tallies. Restates the control flow of
// ProbabilityAnalyzer<mc::DirectEval>::ComputeTallies
// (src/mc/core/probability_analysis.cc:98-209) and
// convergence_controller (src/mc/scheduler/convergence_controller.h:191-250,
// 294-397, 438-491) with both stopping rules of
// src/mc/scheduler/convergence_policy.h: Wald (:28-61) and Bayes with the
// Jeffreys prior (:66-143).
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

// ---------------------------------------------------------------------------
// Beta distribution quantile, for the Bayes rule's central credible interval.
//
// I_x(a, b) is evaluated by the continued fraction of the incomplete beta
// function (modified Lentz; Press et al., Numerical Recipes, 3rd ed., 6.4),
// using I_x(a, b) = 1 - I_{1-x}(b, a) above x = (a + 1) / (a + b + 2). The
// prefactor x^a (1-x)^b / B(a, b) is expanded about the mean a / (a + b)
// with Stirling's series so that it keeps its precision at the trial counts
// of a Monte Carlo run (a + b up to ~1e12), where the plain logarithms would
// cancel to a few digits. The quantile is found by Newton's method on
// I_x - q, safeguarded by bisection inside a bracket that always holds the
// root.
// ---------------------------------------------------------------------------
namespace {

constexpr double kLn2Pi = 1.8378770664093454836;

// s(z) = ln Gamma(z) - [(z - 1/2) ln z - z + ln(2 pi) / 2]
double stirling_remainder(double z) {
    if (z >= 10.0) {
        const double r = 1.0 / z, r2 = r * r;
        return r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 / 1680.0)));
    }
    return std::lgamma(z) - ((z - 0.5) * std::log(z) - z + 0.5 * kLn2Pi);
}

// ln[x^a (1-x)^b / B(a, b)]
double log_power_terms(double a, double b, double x) {
    const double n = a + b, mu = a / n, nu = b / n;
    const double t = a * std::log1p((x - mu) / mu) + b * std::log1p((mu - x) / nu);
    return t + 0.5 * std::log(a * b / n) - 0.5 * kLn2Pi -
           (stirling_remainder(a) + stirling_remainder(b) - stirling_remainder(n));
}

double beta_continued_fraction(double a, double b, double x) {
    const double tiny = 1e-300, eps = 1e-16;
    const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
    double c = 1.0, d = 1.0 - qab * x / qap;
    if (std::fabs(d) < tiny) d = tiny;
    d = 1.0 / d;
    double h = d;
    for (int m = 1; m < 50000000; ++m) {
        const double m2 = 2.0 * m;
        double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
        d = 1.0 + aa * d;
        if (std::fabs(d) < tiny) d = tiny;
        c = 1.0 + aa / c;
        if (std::fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        h *= d * c;
        aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
        d = 1.0 + aa * d;
        if (std::fabs(d) < tiny) d = tiny;
        c = 1.0 + aa / c;
        if (std::fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) < eps) break;
    }
    return h;
}

// Regularized incomplete beta I_x(a, b).
double incomplete_beta(double a, double b, double x) {
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    if (x < (a + 1.0) / (a + b + 2.0))
        return std::exp(log_power_terms(a, b, x)) * beta_continued_fraction(a, b, x) / a;
    return 1.0 - std::exp(log_power_terms(b, a, 1.0 - x)) * beta_continued_fraction(b, a, 1.0 - x) / b;
}

double beta_quantile(double a, double b, double q) {
    if (!(a > 0.0) || !(b > 0.0) || !(q >= 0.0 && q <= 1.0)) return std::nan("");
    if (q == 0.0) return 0.0;
    if (q == 1.0) return 1.0;
    const double n = a + b, mean = a / n;
    const double sd = std::sqrt(a * b / (n * n * (n + 1.0)));
    double lo = 0.0, hi = 1.0;
    // A bracket of +-12 standard deviations holds every quantile used here
    // when the distribution is near normal; widen to [0, 1] otherwise.
    const double blo = std::max(0.0, mean - 12.0 * sd), bhi = std::min(1.0, mean + 12.0 * sd);
    if (incomplete_beta(a, b, blo) <= q) lo = blo;
    if (incomplete_beta(a, b, bhi) >= q) hi = bhi;
    double x = std::clamp(mean, lo, hi);
    if (!(x > lo && x < hi)) x = 0.5 * (lo + hi);
    for (int it = 0; it < 400; ++it) {
        const double f = incomplete_beta(a, b, x) - q;
        if (f == 0.0) return x;
        if (f < 0.0) lo = x; else hi = x;
        const double pdf = std::exp(log_power_terms(a, b, x)) / (x * (1.0 - x));
        double nx = x - f / pdf;
        if (!(nx > lo && nx < hi) || !std::isfinite(nx)) nx = 0.5 * (lo + hi);
        if (std::fabs(nx - x) <= 4e-16 * std::max(x, 1e-300) || nx == lo || nx == hi) return nx;
        x = nx;
    }
    return x;
}

double normal_cdf(double z) { return 0.5 * std::erfc(-z / std::sqrt(2.0)); }

constexpr double kDeltaEpsilon = 1.0e-12;  // src/mc/stats/ci_utils.h:11

uint64_t ceil_to_u64(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 1.8e19) return UINT64_MAX;
    return uint64_t(std::ceil(v));
}

// One tally's update under the Wald rule (convergence_policy.h:28-61, with
// ci_utils.h:118-143 for the half-widths and trial counts) or the Bayes rule
// (convergence_policy.h:66-143).
mcs_convergence convergence_update(uint32_t policy, uint64_t ones, uint64_t total, double delta,
                                   double z) {
    mcs_convergence c{};
    const double ln10 = std::log(10.0);
    const double n = double(total);
    const double mean = total ? double(ones) / n : 0.0;
    if (policy == MCS_CI_WALD) {
        const double se = total ? std::sqrt(mean * (1.0 - mean) / n) : 0.0;
        const double p = std::max(mean, kDeltaEpsilon);
        c.epsilon = z * se;
        c.target_epsilon = delta * p;
        c.target_trials = ceil_to_u64(z * z * p * (1.0 - p) / (c.target_epsilon * c.target_epsilon));
        c.epsilon_log10 = z * se / (p * ln10);
        c.target_epsilon_log10 = delta;
        c.target_trials_log10 = ceil_to_u64(z * z * (1.0 - p) / (p * delta * delta * ln10 * ln10));
        return c;
    }
    const double alpha = 0.5 + double(ones), beta = 0.5 + (n - double(ones));
    const double p = alpha / (alpha + beta);
    const double cred = std::clamp(2.0 * normal_cdf(z) - 1.0, 0.0, 1.0);
    const double tail = (1.0 - cred) / 2.0;
    const double lower = beta_quantile(alpha, beta, tail);
    const double upper = beta_quantile(alpha, beta, 1.0 - tail);
    c.epsilon = (upper - lower) / 2.0;
    c.epsilon_log10 = c.epsilon / (p * ln10);
    c.target_epsilon = delta * p;
    c.target_epsilon_log10 = delta;
    const double rhs_lin = z * z * p * (1.0 - p) / (c.target_epsilon * c.target_epsilon);
    const double rhs_log = z * z * (1.0 - p) / (p * delta * delta * ln10 * ln10);
    constexpr double kPriorSize = 0.5 + 0.5 + 1.0;
    c.target_trials = std::max(ceil_to_u64(std::max(0.0, rhs_lin - kPriorSize)), total);
    c.target_trials_log10 = std::max(ceil_to_u64(std::max(0.0, rhs_log - kPriorSize)), total);
    return c;
}

}  // namespace

extern "C" {

double mcs_normal_quantile_two_sided(double confidence) {
    return mcs::normal_quantile_two_sided(confidence);
}

double mcs_beta_quantile(double a, double b, double q) { return beta_quantile(a, b, q); }

mcs_status mcs_convergence_update(uint32_t ci_policy, uint64_t ones, uint64_t total,
                                  double rel_margin, double z, mcs_convergence *out) {
    if (!out || ones > total || (ci_policy != MCS_CI_BAYES && ci_policy != MCS_CI_WALD)) {
        mcs::set_error("bad convergence arguments");
        return MCS_ERR_INVALID_ARGUMENT;
    }
    *out = convergence_update(ci_policy, ones, total, rel_margin, z);
    return MCS_OK;
}

mcs_status mcs_probability_analysis(const mcs_pdag *pdag, const mcs_settings *settings,
                                    const mcs_sample_shape *shape, const int32_t *observe,
                                    int32_t n_observe, mcs_tally *out, int32_t n_out,
                                    double *p_total, uint32_t *passes_run) {
    return mcs::probability_analysis(pdag, settings, shape, observe, n_observe, out, n_out,
                                     p_total, passes_run, nullptr);
}

}  // extern "C"

mcs_status mcs::probability_analysis(const mcs_pdag *pdag, const mcs_settings *settings,
                                     const mcs_sample_shape *shape, const int32_t *observe,
                                     int32_t n_observe, mcs_tally *out, int32_t n_out,
                                     double *p_total, uint32_t *passes_run,
                                     const ExternalMemory *ext) {
    if (!pdag || !settings || !shape || !out || !p_total || n_observe < 0 ||
        (n_observe > 0 && !observe)) {
        mcs::set_error("null argument");
        return MCS_ERR_INVALID_ARGUMENT;
    }
    if (settings->ci_policy != MCS_CI_BAYES && settings->ci_policy != MCS_CI_WALD) {
        mcs::set_error("ci_policy must be MCS_CI_BAYES or MCS_CI_WALD");
        return MCS_ERR_INVALID_ARGUMENT;
    }
    std::vector<int32_t> tallies;
    tallies.push_back(pdag->root);
    for (int32_t i = 0; i < n_observe; ++i) tallies.push_back(observe[i]);
    mcs_engine *eng = nullptr;
    mcs_status st = mcs::create_engine(pdag, shape, tallies.data(), int32_t(tallies.size()),
                                       settings->key0, settings->key1, -1, ext, &eng);
    if (st != MCS_OK) return st;
    mcs_engine_info info{};
    mcs_engine_get_info(eng, &info);
    if (n_out < info.num_tallies) {
        mcs_engine_destroy(eng);
        mcs::set_error("output needs one record for the root and each distinct observed node");
        return MCS_ERR_INVALID_ARGUMENT;
    }
    // total_passes is the shaper's TOTAL_ITERATIONS: 0 (no limit) when no
    // trial count is given (src/mc/queue/sample_shaper.h:139-146). A zero
    // trial count turns early stop on (src/settings.cc:105-112); otherwise
    // the caller's flag decides, so a trial count with early stop stops at
    // whichever comes first (targets/scram/utils/ConstructSettings.h:103-109).
    uint64_t rounded = 0, total_passes = 0;
    if (settings->num_trials) mcs_plan_passes(settings->num_trials, shape, &rounded, &total_passes);
    const bool early_stop = settings->num_trials == 0 || settings->early_stop != 0;
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
    // Steps (convergence_controller.h:294-328): before each pass, stop if
    // early stop is on and both flags are set, or if the fixed pass count
    // (burn-in passes included) is reached; after each pass, set the flags
    // once the worst tally meets the policy's targets in linear and log10
    // space. Burn-in never sets them (:351-365), so with early stop at least
    // one pass follows burn-in. The worst tally is the one with the widest
    // linear half-width; the linear target is the relative margin times that
    // tally's sample mean; both flags stay set once met
    // (convergence_controller.h:191-250, 313-323, 471-482).
    auto converged = [&]() {
        double max_eps = 0.0, max_eps_log = 0.0, target = 0.0;
        for (const mcs_tally &t : rec) {
            const mcs_convergence c =
                convergence_update(settings->ci_policy, t.num_one_bits, t.total_bits, delta, z);
            if (c.epsilon > max_eps) {
                max_eps = c.epsilon;
                target = delta * std::max(t.mean, kDeltaEpsilon);
            }
            max_eps_log = std::max(max_eps_log, c.epsilon_log10);
        }
        eps_ok = eps_ok || (max_eps > 0 && max_eps <= target);
        eps_log_ok = eps_log_ok || (max_eps_log > 0 && max_eps_log <= delta);
        return eps_ok && eps_log_ok;
    };
    while (st == MCS_OK && iterations < cap) {
        if (early_stop && eps_ok && eps_log_ok) break;
        if (total_passes && iterations >= total_passes) break;
        st = mcs_engine_pass_wait_collect(eng, 1, 0, rec.data(), int32_t(rec.size()));
        if (st != MCS_OK) break;
        ++iterations;
        converged();
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
