// AxionStringsParams.hpp -- readers for the `axion_strings.*`, `amr.*`,
// `geometry.*` and `evolution.*` keys the hot path needs, with the
// reference's defaults and validation messages
// (AxionStrings/AxionStringsParams.hpp:30-1208; read sites of the level
// driver AxionStringsLevel.cpp:137-208).
//
// Keys added by this implementation (documented extensions; absent = the
// reference's behaviour):
//   axion_strings.precision           fp64 | fp32c | fp32   (default fp64)
//   axion_strings.output_cadence_steps  diagnostics every N coarse steps
//                                       (replaces the log(m_r/H) cadence)
//   axion_strings.tagging.amplitude_threshold  tag where |psi|/R < value
//   axion_strings.spectrum_field      a_dot | pi   (default a_dot)
//   axion_strings.use_graph           replay a captured hipGraph per step
//   axion_strings.placement_trials    re-allocate the level's arrays up to N
//                                     times and keep a fast placement
#pragma once

#include "Background.hpp"
#include "ParmParse.hpp"

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace axs_host
{

enum class MaskingScheme
{
    None,
    A,
    B
};

struct MaskingParams
{
    MaskingScheme scheme{MaskingScheme::A};
    double threshold{0.8};
};

struct StringTaggerParams
{
    double gradient_threshold{DBL_MAX};
    double radial_gradient_threshold{DBL_MAX};
    bool force_full_refinement{false};
    double amplitude_threshold{0.0}; // extension, 0 = off
};

struct FourierICParams
{
    double k_max_over_mr{};
    double mean_square_variance{};
    std::uint64_t seed{};
};

// The adaptive xi-check cadence of the relaxing phase
// (AxionStringsParams.hpp:227-234).
struct XiCheckCadence
{
    long coarse_interval{50};
    long medium_interval{10};
    long fine_interval{2};
    double medium_threshold{1.5};
    double fine_threshold{1.1};
};

struct OutputCadence
{
    double first_log_mr_over_h{};
    double delta_log_mr_over_h{};
    long every_steps{0}; // extension: > 0 switches to a step cadence
};

namespace Params
{

// read_background, AxionStringsParams.hpp:30-103
inline Background read_background(const ParamTable &t)
{
    ParmParse pp(t, "axion_strings");
    double a_inv = 2.0;
    pp.get("a_inv", a_inv);
    if (a_inv <= 1.0)
        pp.error("a_inv", "must be > 1");
    CTauSchedule sched{};
    pp.get("c0", sched.c0);
    const bool has_log_switch = pp.contains("log_mr_over_h_switch");
    const bool has_c1 = pp.contains("c1");
    const bool has_tau_switch = pp.contains("tau_switch");
    if (has_log_switch && (has_c1 || has_tau_switch))
        pp.error("log_mr_over_h_switch",
                 "cannot be combined with c1/tau_switch -- pick one way to "
                 "specify the switch");
    if (has_log_switch)
    {
        if (std::abs(a_inv - sched.c0) < 1.0e-9)
            pp.error("log_mr_over_h_switch",
                     "cannot derive tau_switch when axion_strings.c0 = a_inv "
                     "already -- there is no fat phase to switch from");
        double l{};
        pp.get("log_mr_over_h_switch", l);
        CTauSchedule pre{};
        pre.c0 = sched.c0;
        const Background pre_bg{a_inv, pre};
        sched.has_switch = true;
        sched.c1 = a_inv;
        sched.tau_switch = pre_bg.tau_from_log_mr_over_h(l);
    }
    else
    {
        if (has_c1 != has_tau_switch)
            pp.error("c1", "c1 and tau_switch must both be set, or both left "
                           "unset (at most one switch)");
        sched.has_switch = has_c1;
        if (sched.has_switch)
        {
            pp.get("c1", sched.c1);
            pp.get("tau_switch", sched.tau_switch);
        }
    }
    if (sched.has_switch && sched.tau_switch <= 0.0)
        pp.error("tau_switch", "must be > 0");
    return {a_inv, sched};
}

// read_tau_i, AxionStringsParams.hpp:111-128
inline double read_tau_i(const ParamTable &t, const Background &b)
{
    ParmParse pp(t, "axion_strings");
    double l{};
    pp.get("log_mr_over_h_i", l);
    if (std::abs(b.a_inv - b.c_sched.c0) < 1.0e-9)
        pp.error("log_mr_over_h_i",
                 "cannot derive tau_i when axion_strings.c0 = a_inv (Moore "
                 "mode from the very start) -- reach Moore only via a switch "
                 "(axion_strings.c1/tau_switch), not as the starting c0");
    return b.tau_from_log_mr_over_h(l);
}

// read_pre_evolution_background, AxionStringsParams.hpp:135-145
inline PreEvolutionBackground read_pre_evolution_background(const ParamTable &t)
{
    ParmParse pp(t, "axion_strings.pre_evolution");
    double gamma_pre = 1.0;
    pp.get("gamma", gamma_pre);
    if (gamma_pre <= 0.0)
        pp.error("gamma", "must be > 0");
    return PreEvolutionBackground(gamma_pre);
}

// read_xi_target, AxionStringsParams.hpp:152-162
inline double read_xi_target(const ParamTable &t)
{
    ParmParse pp(t, "axion_strings.pre_evolution");
    double xi_target{};
    pp.get("xi_target", xi_target);
    if (xi_target <= 0.0)
        pp.error("xi_target", "must be > 0");
    return xi_target;
}

// read_handoff_transition_n_periods, AxionStringsParams.hpp:185-195
inline double read_handoff_transition_n_periods(const ParamTable &t)
{
    ParmParse pp(t, "axion_strings.pre_evolution");
    double n_periods = 3.0;
    pp.queryAdd("handoff_transition_n_periods", n_periods);
    if (n_periods < 0.0)
        pp.error("handoff_transition_n_periods", "must be >= 0");
    return n_periods;
}

// read_xi_check_cadence, AxionStringsParams.hpp:240-261
inline XiCheckCadence read_xi_check_cadence(const ParamTable &t)
{
    ParmParse pp(t, "axion_strings.pre_evolution");
    XiCheckCadence c{};
    pp.queryAdd("check_interval_coarse", c.coarse_interval);
    pp.queryAdd("check_interval_medium", c.medium_interval);
    pp.queryAdd("check_interval_fine", c.fine_interval);
    pp.queryAdd("check_threshold_medium", c.medium_threshold);
    pp.queryAdd("check_threshold_fine", c.fine_threshold);
    if (c.coarse_interval < 1 || c.medium_interval < 1 || c.fine_interval < 1)
        pp.error("check_interval_coarse", "all check intervals must be >= 1");
    if (!(c.fine_threshold > 1.0 && c.fine_threshold < c.medium_threshold))
        pp.error("check_threshold_medium",
                 "must have 1 < check_threshold_fine < check_threshold_medium");
    return c;
}

// read_fourier_ic_params, AxionStringsParams.hpp:207-225
inline FourierICParams read_fourier_ic_params(const ParamTable &t,
                                              const std::string &prefix)
{
    ParmParse pp(t, prefix);
    FourierICParams p{};
    pp.get("k_max_over_mr", p.k_max_over_mr);
    pp.get("mean_square_variance", p.mean_square_variance);
    if (p.k_max_over_mr <= 0.0)
        pp.error("k_max_over_mr", "must be > 0");
    if (p.mean_square_variance < 0.0)
        pp.error("mean_square_variance", "must be >= 0");
    long long seed = 0;
    pp.get("seed", seed);
    p.seed = static_cast<std::uint64_t>(seed);
    return p;
}

// read_output_cadence, AxionStringsParams.hpp:276-289 (+ step extension)
inline OutputCadence read_output_cadence(const ParamTable &t)
{
    ParmParse pp(t, "axion_strings");
    OutputCadence c{};
    if (pp.query("output_cadence_steps", c.every_steps))
    {
        if (c.every_steps < 1)
            pp.error("output_cadence_steps", "must be >= 1");
        return c;
    }
    pp.get("output_first_log_mr_over_h", c.first_log_mr_over_h);
    pp.get("output_delta_log_mr_over_h", c.delta_log_mr_over_h);
    if (c.delta_log_mr_over_h <= 0.0)
        pp.error("output_delta_log_mr_over_h", "must be > 0");
    return c;
}

// read_tagging_params, AxionStringsParams.hpp:334-350 (+ amplitude ext.)
inline StringTaggerParams read_tagging_params(const ParamTable &t)
{
    ParmParse pp(t, "axion_strings.tagging");
    StringTaggerParams p{};
    pp.queryAdd("gradient_threshold", p.gradient_threshold);
    if (p.gradient_threshold <= 0.0)
        pp.error("gradient_threshold", "must be > 0");
    pp.queryAdd("radial_gradient_threshold", p.radial_gradient_threshold);
    if (p.radial_gradient_threshold <= 0.0)
        pp.error("radial_gradient_threshold", "must be > 0");
    pp.queryAdd("force_full_refinement", p.force_full_refinement);
    pp.queryAdd("amplitude_threshold", p.amplitude_threshold);
    if (p.amplitude_threshold < 0.0 || p.amplitude_threshold >= 1.0)
        pp.error("amplitude_threshold", "must be in [0, 1)");
    return p;
}

// read_masking_params, AxionStringsParams.hpp:500-530
inline MaskingParams read_masking_params(const ParamTable &t,
                                         const std::string &prefix)
{
    ParmParse pp(t, prefix);
    MaskingParams p{};
    std::string scheme = "A";
    pp.queryAdd("scheme", scheme);
    if (scheme == "none" || scheme == "None")
        p.scheme = MaskingScheme::None;
    else if (scheme == "A")
        p.scheme = MaskingScheme::A;
    else if (scheme == "B")
        p.scheme = MaskingScheme::B;
    else
        pp.error("scheme", "must be \"A\", \"B\" or \"none\"");
    pp.queryAdd("threshold", p.threshold);
    if (p.scheme == MaskingScheme::B &&
        (p.threshold <= 0.0 || p.threshold >= 1.0))
        pp.error("threshold", "must be in (0, 1) for scheme B");
    return p;
}

// check_params, AxionStringsParams.hpp:1142-1208 (the parts that apply to
// a single-level run without box planning), plus the masking / tagging
// reads that AxionStringsLevel::variableSetUp performs at level set-up
// (AxionStringsLevel.cpp:186-190), so that every key error surfaces here.
inline void check_params(const ParamTable &t)
{
    read_masking_params(t, "axion_strings.masking");
    read_tagging_params(t);
    bool flat_space = false;
    ParmParse(t, "axion_strings").queryAdd("flat_space", flat_space);
    if (flat_space)
    {
        read_masking_params(t, "axion_strings.masking");
        read_tagging_params(t);
        return;
    }
    const Background b = read_background(t);
    read_tau_i(t, b);
    std::string ic_mode = "homogeneous";
    ParmParse(t, "axion_strings").queryAdd("ic_mode", ic_mode);
    if (ic_mode == "fourier_relaxed")
    {
        read_pre_evolution_background(t);
        read_xi_target(t);
        read_fourier_ic_params(t, "axion_strings.pre_evolution");
        read_xi_check_cadence(t);
        read_handoff_transition_n_periods(t);
    }
    else if (ic_mode == "fourier")
        read_fourier_ic_params(t, "axion_strings");
    else if (ic_mode != "homogeneous")
        ParmParse(t, "axion_strings")
            .error("ic_mode", "this build supports \"fourier_relaxed\", "
                              "\"fourier\" and \"homogeneous\"");
    read_output_cadence(t);
    bool compute_spectrum = false;
    ParmParse sp(t, "axion_strings");
    sp.queryAdd("compute_spectrum", compute_spectrum);
    if (compute_spectrum &&
        read_masking_params(t, "axion_strings.masking").scheme !=
            MaskingScheme::B)
        sp.error("compute_spectrum",
                 "requires axion_strings.masking.scheme = B -- the spectrum "
                 "saved to axion_spectrum.dat must be the genuinely screened "
                 "field");
}

} // namespace Params
} // namespace axs_host
