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
//   axion_strings.tuning_table        the autotuner's launch-shape table
//                                     (axs_tuning_load), relative to the
//                                     parameter file
#pragma once

#include "Background.hpp"
#include "BoxPlan.hpp"
#include "ParmParse.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <ostream>
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

// A double written so that reading it back gives the same value.
inline std::string exact_str(double v)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

// Derive-if-absent, cross-check-if-present for geometry.prob_extent
// (AxionStringsParams.hpp:808-830 and the two later copies).
inline void plan_prob_extent(ParamTable &t, double L_tilde, std::ostream *log)
{
    ParmParse geom(t, "geometry");
    if (geom.contains("prob_extent"))
    {
        std::vector<double> ext;
        geom.getarr("prob_extent", ext);
        for (double L : ext)
            if (std::abs(L - L_tilde) > 1.0e-6 * L_tilde)
                geom.error("prob_extent",
                           "does not match the box-planning-derived L_tilde "
                           "-- either remove geometry.prob_extent to let it "
                           "be derived, or fix axion_strings.N/N1/N2 to match");
        return;
    }
    const std::string v = exact_str(L_tilde);
    t.set("geometry.prob_extent", v + " " + v + " " + v);
    if (log)
        *log << "  -> geometry.prob_extent set to " << v << " " << v << " "
             << v << "\n";
}

// The same for amr.n_cell (AxionStringsParams.hpp:957-978, :1038-1058).
inline void plan_n_cell(ParamTable &t, int n, const char *what,
                        std::ostream *log)
{
    ParmParse amr(t, "amr");
    if (amr.contains("n_cell"))
    {
        std::vector<long> nc;
        amr.getarr("n_cell", nc);
        for (long v : nc)
            if (v != n)
                amr.error("n_cell", std::string("does not match ") + what +
                                        " -- either remove amr.n_cell to let "
                                        "it be derived, or fix "
                                        "axion_strings.N/amr.max_level to "
                                        "match");
        return;
    }
    const std::string v = std::to_string(n);
    t.set("amr.n_cell", v + " " + v + " " + v);
    if (log)
        *log << "  -> amr.n_cell set to " << v << " " << v << " " << v
             << "\n";
}

// The derived stop time, and evolution.stop_time = -1 (no limit) when it is
// absent, so that tau >= derived_tau_f is what ends the run
// (AxionStringsParams.hpp:1117-1129).
inline void plan_stop(ParamTable &t, double tau_f, std::ostream *log)
{
    t.set("axion_strings.derived_tau_f", exact_str(tau_f));
    if (!t.contains("evolution.stop_time"))
    {
        t.set("evolution.stop_time", "-1");
        if (log)
            *log << "  -> evolution.stop_time set to -1: the run stops once "
                    "tau >= axion_strings.derived_tau_f = "
                 << exact_str(tau_f) << "\n";
    }
}

// apply_pre_evolution_gamma_default, AxionStringsParams.hpp:638-688: a
// relaxing run relaxes at 1/dx of the level it starts on unless
// axion_strings.pre_evolution.gamma says otherwise (a differing value only
// warns).
inline void apply_pre_evolution_gamma_default(ParamTable &t,
                                              const std::string &ic_mode,
                                              double dx_base, std::ostream *log)
{
    if (ic_mode != "fourier_relaxed")
        return;
    const double expect = 1.0 / dx_base;
    ParmParse pre(t, "axion_strings.pre_evolution");
    if (pre.contains("gamma"))
    {
        double g = 0.0;
        pre.get("gamma", g);
        if (std::abs(g - expect) > 1.0e-6 * expect && log)
            *log << "  WARNING: axion_strings.pre_evolution.gamma = " << g
                 << " differs from 1/dx_base = " << expect
                 << " of the box plan; relaxing at the value given\n";
        return;
    }
    t.set("axion_strings.pre_evolution.gamma", exact_str(expect));
    if (log)
        *log << "  -> axion_strings.pre_evolution.gamma set to "
             << exact_str(expect) << " (= 1/dx_base)\n";
}

// apply_box_plan, AxionStringsParams.hpp:690-1130.  Nothing happens unless
// axion_strings.N is set; then N, N1, N2 fix amr.n_cell, geometry.prob_extent
// and the stop time axion_strings.derived_tau_f (and for refined runs the
// level schedule axion_strings.derived_level_add_log_mr_over_h), each
// derived when absent and checked when present.  Runs that reach Moore
// (c = a_inv) through a switch get the Moore plan (tau_f = its tau_end); a run
// with c0 = a_inv from the start only gets the dynamic-range check.  Safe to
// apply twice: the second pass finds its own values and accepts them.
inline void apply_box_plan(ParamTable &t, const Background &b, double tau_i,
                           std::ostream *log)
{
    ParmParse pp(t, "axion_strings");
    if (!pp.contains("N"))
        return;
    int N = 0;
    double N1 = 0.0, N2 = 0.0;
    pp.get("N", N);
    pp.get("N1", N1);
    pp.get("N2", N2);
    if (N <= 0 || N1 <= 0.0 || N2 <= 0.0)
        pp.error("N", "N, N1 and N2 must all be > 0");
    const double a_inv = b.a_inv;
    const bool moore = b.c_sched.has_switch ||
                       std::abs(b.c_sched.c0 - a_inv) < 1.0e-9;
    int max_level = 0;
    ParmParse(t, "amr").queryAdd("max_level", max_level);
    if (log)
        *log << "box plan: a_inv = " << a_inv << ", N = " << N
             << ", N1 = " << N1 << ", N2 = " << N2
             << ", amr.max_level = " << max_level << "\n";
    if (moore && max_level > 0)
        pp.error("N", "amr.max_level > 0 is not supported reaching Moore -- "
                      "set amr.max_level = 0 for a fat->Moore run");

    if (moore && !b.c_sched.has_switch)
    {
        double gamma = 0.0, target = 0.0;
        pp.get("gamma", gamma);
        pp.get("target_log_range", target);
        if (gamma <= 0.0)
            pp.error("gamma", "must be > 0");
        const double D = moore_max_log_dynamic_range(N, N1, N2, gamma);
        if (log)
            *log << "  Moore from tau_i: achievable log range " << D
                 << ", requested " << target
                 << " (geometry and stop time are set by hand)\n";
        if (D < target)
            pp.error("target_log_range",
                     "cannot be reached with the requested N, N1, N2, gamma");
        return;
    }

    if (moore)
    {
        const double tau_sw = b.c_sched.tau_switch;
        const double gamma = 1.0 / b.H_over_mr_direct(tau_sw);
        const MooreBoxPlan plan = compute_moore_box_plan(
            N, N1, N2, gamma, a_inv, b.b_inv, b.R(tau_sw),
            std::sqrt(b.lambda(tau_sw)), tau_sw);
        if (log)
            *log << "  Moore through a switch at tau = " << tau_sw
                 << " (m_r/H = " << gamma << "): L_tilde = " << plan.L_tilde
                 << ", dx = " << plan.dx << ", D = " << plan.D
                 << ", tau_end = " << plan.tau_end << "\n";
        double target = 0.0;
        if (pp.query("target_log_range", target) && plan.D < target)
            pp.error("target_log_range",
                     "cannot be reached with the requested N, N1, N2 -- "
                     "increase N or reduce N1/N2/the switch's log(m_r/H)");
        plan_prob_extent(t, plan.L_tilde, log);
        plan_stop(t, plan.tau_end, log);
        return;
    }

    const double c = b.c_sched.c0;
    std::string ic_mode = "homogeneous";
    pp.queryAdd("ic_mode", ic_mode);

    if (max_level > 0)
    {
        const AmrBoxPlan plan =
            compute_amr_box_plan(N, N1, N2, a_inv, c, max_level);
        if (log)
        {
            *log << "  refined: tau_f = " << plan.tau_f << " (tau_i = "
                 << tau_i << "), L_tilde = " << plan.L_tilde
                 << ", dx_finest = " << plan.dx_finest
                 << ", dx_base = " << plan.dx_base << "\n";
            for (int ell = 1; ell <= max_level; ++ell)
                *log << "    level " << ell << " from log(m_r/H) = "
                     << plan.log_add[size_t(ell - 1)] << "\n";
        }
        if (plan.N_base <= 0)
            pp.error("N", "N is not evenly divisible by 2^amr.max_level -- "
                          "choose an effective resolution N that is a clean "
                          "multiple of 2^max_level");
        if (plan.tau_f <= tau_i)
            pp.error("N", "the requested N, N1, N2, c configuration gives "
                          "tau_f <= tau_i; cannot evolve forward from tau_i");
        if (ic_mode == "fourier_relaxed" &&
            !t.contains("axion_strings.pre_evolution.gamma") &&
            -std::log(b.H_over_mr_direct(tau_i)) >= plan.log_add[0] && log)
            *log << "  WARNING: log_mr_over_h_i is at or past the level-1 "
                    "threshold; relaxing at 1/dx_base all the same\n";
        apply_pre_evolution_gamma_default(t, ic_mode, plan.dx_base, log);
        plan_n_cell(t, plan.N_base,
                    "the box-planning-derived base-level grid size "
                    "(N/2^max_level)",
                    log);
        plan_prob_extent(t, plan.L_tilde, log);
        std::string sched;
        for (double v : plan.log_add)
            sched += (sched.empty() ? "" : " ") + exact_str(v);
        t.set("axion_strings.derived_level_add_log_mr_over_h", sched);
        plan_stop(t, plan.tau_f, log);
        return;
    }

    plan_n_cell(t, N, "axion_strings.N", log);
    const GeneralBoxPlan plan = compute_general_box_plan(N, N1, N2, a_inv, c);
    if (log)
        *log << "  tau_f = " << plan.tau_f << " (tau_i = " << tau_i
             << "), L_tilde = " << plan.L_tilde << ", dx = " << plan.delta_x
             << ", dx/3 = " << plan.delta_tau_leapfrog_bound << "\n"
             << "  (pre-evolution box for this plan, reported only: "
             << pre_evolution_L_tilde(plan.L_tilde, a_inv, c, tau_i)
             << ")\n";
    if (plan.tau_f <= tau_i)
        pp.error("N", "the requested N, N1, N2, c configuration gives tau_f "
                      "<= tau_i; cannot evolve forward from tau_i");
    apply_pre_evolution_gamma_default(t, ic_mode, plan.delta_x, log);
    plan_prob_extent(t, plan.L_tilde, log);
    plan_stop(t, plan.tau_f, log);
}

// The amr.* keys of a refined run, read by AMReX's AmrMesh / Amr in the
// reference; absent keys take AMReX's defaults (n_error_buf 1,
// blocking_factor 8, max_grid_size 32 in 3-D, regrid_int 1).  The box
// levels here hold boxes of 8..64 cells per edge on corners that are
// multiples of 8, so blocking_factor is 8, 16, 32 or 64 and max_grid_size a
// multiple of it up to 64.
struct AmrKeys
{
    int max_level{0};
    int n_error_buf{1};
    int blocking_factor{8};
    int max_grid_size{32};
    int regrid_int{1};
};

inline AmrKeys read_amr_keys(const ParamTable &t)
{
    ParmParse pp(t, "amr");
    AmrKeys k{};
    pp.queryAdd("max_level", k.max_level);
    if (k.max_level < 0 || k.max_level > 8)
        pp.error("max_level", "must be in 0..8");
    pp.queryAdd("n_error_buf", k.n_error_buf);
    pp.queryAdd("blocking_factor", k.blocking_factor);
    pp.queryAdd("max_grid_size", k.max_grid_size);
    pp.queryAdd("regrid_int", k.regrid_int);
    if (k.max_level == 0)
        return k;
    const int bf = k.blocking_factor;
    if (bf != 8 && bf != 16 && bf != 32 && bf != 64)
        pp.error("blocking_factor", "must be 8, 16, 32 or 64 here");
    if (k.max_grid_size < bf || k.max_grid_size > 64 ||
        k.max_grid_size % bf != 0)
        pp.error("max_grid_size", "must be a multiple of amr.blocking_factor "
                                  "and at most 64 here");
    if (k.n_error_buf < 0 || k.n_error_buf > bf / 2)
        pp.error("n_error_buf", "must be in 0..amr.blocking_factor/2 here");
    return k;
}

// apply_regrid_buffer_policy, AxionStringsParams.hpp:403-494: asked for by
// axion_strings.tagging.regrid_interval_steps = N > 0 on a refined run.
// Between two regrids N steps apart a string moving at v = 1 crosses
// N * evolution.dt_multiplier cells, so the tag buffer must be at least
// ceil(buffer_safety_factor * N * dt_multiplier) cells (and 1): amr.n_error_buf
// is set to that when absent and must not be smaller when given, and
// amr.regrid_int is set to N when absent and must equal N when given.
inline void apply_regrid_buffer_policy(ParamTable &t, std::ostream *log)
{
    ParmParse amr(t, "amr");
    int max_level = 0;
    amr.queryAdd("max_level", max_level);
    if (max_level <= 0)
        return;
    ParmParse tag(t, "axion_strings.tagging");
    int every = 0;
    tag.queryAdd("regrid_interval_steps", every);
    if (every == 0)
        return;
    if (every < 0)
        tag.error("regrid_interval_steps", "must be > 0");
    double safety = 1.0;
    tag.queryAdd("buffer_safety_factor", safety);
    if (safety < 1.0)
        tag.error("buffer_safety_factor", "must be >= 1.0");
    double dtm = 0.0;
    ParmParse(t, "evolution").get("dt_multiplier", dtm);
    const int need =
        std::max(1, static_cast<int>(std::ceil(safety * every * dtm)));
    if (log)
        *log << "regrid buffer: " << every << " steps between regrids, "
             << "safety factor " << safety << ", dt_multiplier " << dtm
             << " -> at least " << need << " buffer cells\n";
    if (amr.contains("n_error_buf"))
    {
        int given = 0;
        amr.get("n_error_buf", given);
        if (given < need)
            amr.error("n_error_buf",
                      "is below the minimum buffer for axion_strings.tagging."
                      "regrid_interval_steps / buffer_safety_factor (" +
                          std::to_string(need) +
                          " cells): remove it or raise it");
    }
    else
    {
        t.set("amr.n_error_buf", std::to_string(need));
        if (log)
            *log << "  amr.n_error_buf = " << need << "\n";
    }
    if (amr.contains("regrid_int"))
    {
        int given = 0;
        amr.get("regrid_int", given);
        if (given != every)
            amr.error("regrid_int",
                      "differs from axion_strings.tagging."
                      "regrid_interval_steps: remove one or make them equal");
    }
    else
    {
        t.set("amr.regrid_int", std::to_string(every));
        if (log)
            *log << "  amr.regrid_int = " << every << "\n";
    }
}

// The rest of check_params, AxionStringsParams.hpp:1142-1208, plus the
// masking / tagging reads that AxionStringsLevel::variableSetUp performs at
// level set-up (AxionStringsLevel.cpp:186-190), so that every key error
// surfaces here.  Reads only; the box plan is applied before it.
inline void validate_params(const ParamTable &t)
{
    read_masking_params(t, "axion_strings.masking");
    read_tagging_params(t);
    read_amr_keys(t);
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

// check_params, AxionStringsParams.hpp:1142-1208: the box plan (which adds
// the derived keys to the table; log = where to print it, or nullptr), then
// every read and check.
inline void check_params(ParamTable &t, std::ostream *log = nullptr)
{
    bool flat_space = false;
    ParmParse(t, "axion_strings").queryAdd("flat_space", flat_space);
    if (!flat_space)
    {
        const Background b = read_background(t);
        apply_box_plan(t, b, read_tau_i(t, b), log);
    }
    apply_regrid_buffer_policy(t, log);
    validate_params(t);
}

} // namespace Params
} // namespace axs_host
