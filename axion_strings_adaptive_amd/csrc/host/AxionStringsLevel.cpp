// AxionStringsLevel.cpp -- see AxionStringsLevel.hpp.
#include "AxionStringsLevel.hpp"

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <filesystem>
#include <functional>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <vector>

namespace axs_host
{

static int precision_from(const std::string &s, const ParmParse &pp)
{
    if (s == "fp64" || s == "double" || s == "DOUBLE")
        return AXS_FP64;
    if (s == "fp32c")
        return AXS_FP32C;
    if (s == "fp32" || s == "float" || s == "FLOAT")
        return AXS_FP32;
    pp.error("precision", "must be fp64, fp32c or fp32");
}

AxionStringsLevel::AxionStringsLevel(const ParamTable &params,
                                     const std::string &out_dir)
    : m_params(params), m_out_dir(out_dir)
{
    // the output folder is made if it does not exist
    std::error_code ec;
    std::filesystem::create_directories(m_out_dir, ec);
    Params::check_params(m_params);
    ParmParse amr(m_params, "amr"), geom(m_params, "geometry"),
        evo(m_params, "evolution"), as(m_params, "axion_strings");
    std::vector<int> n_cell;
    amr.getarr("n_cell", n_cell);
    if (n_cell.size() != 3)
        amr.error("n_cell", "needs 3 values");
    m_amr = Params::read_amr_keys(m_params);
    std::vector<double> extent;
    geom.getarr("prob_extent", extent);
    if (extent.size() != 3)
        geom.error("prob_extent", "needs 3 values");
    for (int d = 0; d < 3; ++d)
        m_n[d] = n_cell[d];
    m_dx = extent[0] / n_cell[0];
    for (int d = 1; d < 3; ++d)
        if (std::abs(extent[d] / n_cell[d] - m_dx) > 1e-12 * m_dx)
            geom.error("prob_extent", "cells must be cubic (equal dx)");
    m_L = extent[0];
    double dt_multiplier = 0.0;
    evo.get("dt_multiplier", dt_multiplier);
    m_dt = dt_multiplier * m_dx; // AxionStringsParams.hpp:359-361
    evo.queryAdd("max_steps", m_max_steps);
    evo.queryAdd("stop_time", m_stop_time);

    std::string prec = "fp64";
    as.queryAdd("precision", prec);
    m_precision = precision_from(prec, as);
    as.queryAdd("use_graph", m_use_graph);
    as.queryAdd("flat_space", m_flat);
    // the box plan's stop time (AxionStringsLevel.cpp:168-175); flat-space
    // runs stop on evolution.stop_time / max_steps only
    m_has_tau_f = !m_flat && as.query("derived_tau_f", m_tau_f);
    m_masking = Params::read_masking_params(m_params, "axion_strings.masking");
    m_tagging = Params::read_tagging_params(m_params);
    if (m_flat)
    {
        ParmParse(m_params, "axion_strings.flat").queryAdd("m_r", m_flat_mr);
        long every = 0;
        ParmParse(m_params, "axion_strings.flat")
            .queryAdd("output_cadence_steps", every);
        m_cadence.every_steps = every > 0 ? every : 1;
        m_tau_i = 0.0;
        m_cbg = axs_background{2.0, 0, 0, 0, 0, 1, m_flat_mr};
    }
    else
    {
        m_bg = Params::read_background(m_params);
        m_tau_i = Params::read_tau_i(m_params, m_bg);
        m_cadence = Params::read_output_cadence(m_params);
        m_next_log = m_cadence.first_log_mr_over_h;
        m_cbg = axs_background{m_bg.a_inv,           m_bg.c_sched.c0,
                               m_bg.c_sched.c1,      m_bg.c_sched.tau_switch,
                               m_bg.c_sched.has_switch, 0, 1.0};
    }
    m_tau = m_tau_i;
    m_tau_offset = m_tau_i;
    std::string ic_mode = "homogeneous";
    as.queryAdd("ic_mode", ic_mode);
    if (!m_flat && ic_mode == "fourier_relaxed")
    {
        // AxionStringsLevel.cpp:192-203
        m_pre = Params::read_pre_evolution_background(m_params);
        m_xi_target = Params::read_xi_target(m_params);
        m_xi_cadence = Params::read_xi_check_cadence(m_params);
        m_xi_check_interval = m_xi_cadence.coarse_interval;
        m_handoff_n_periods = Params::read_handoff_transition_n_periods(m_params);
        m_phase = Phase::Relaxing;
    }
    amr.queryAdd("check_int", m_check_int);
    amr.queryAdd("check_file", m_check_file);
    amr.query("restart", m_restart);
    if (!m_restart.empty() && !std::ifstream(m_restart).good())
        amr.error("restart", "cannot open checkpoint " + m_restart);
    if (m_amr.max_level > 0)
    {
        if (m_check_int > 0 || !m_restart.empty())
            amr.error("max_level", "checkpoints of refined runs are not "
                                   "supported here (amr.check_int, "
                                   "amr.restart)");
        if (m_n[0] != m_n[1] || m_n[0] != m_n[2])
            amr.error("n_cell", "refined runs need a cubic base grid here");
        // the box plan's level schedule (AxionStringsLevel.cpp:178-184)
        if (!m_flat &&
            as.contains("derived_level_add_log_mr_over_h"))
            as.getarr("derived_level_add_log_mr_over_h", m_level_add);
    }
    as.queryAdd("compute_spectrum", m_compute_spectrum);
    std::string field = "a_dot";
    as.queryAdd("spectrum_field", field);
    if (field == "pi")
        m_spectrum_kind = -1;
    else if (field != "a_dot")
        as.error("spectrum_field", "must be a_dot or pi");

    // extension: the autotuner's launch-shape table, read before the level
    // is created so that the level takes its tuned tiles
    std::string tuning;
    as.queryAdd("tuning_table", tuning);
    if (!tuning.empty())
        axs_check(axs_tuning_load(tuning.c_str()), "axs_tuning_load");
    axs_check(axs_ctx_create(0, &m_ctx), "axs_ctx_create");
    axs_check(axs_level_create(m_ctx, m_n[0], m_n[1], m_n[2], m_dx,
                               m_precision, &m_level),
              "axs_level_create");
    // extension: placement trials of the level's arrays before initData
    int trials = 0;
    as.queryAdd("placement_trials", trials);
    if (trials < 0 || trials > 16)
        as.error("placement_trials", "must be in 0..16");
    if (trials > 0)
    {
        std::vector<double> ms(trials, 0.0);
        int used = 0;
        axs_check(axs_level_place(m_level, trials, ms.data(), &used),
                  "axs_level_place");
        std::cout << "  [AxionStrings] placement trials (ms/step):";
        for (int i = 0; i < used; ++i)
            std::cout << " " << ms[i];
        std::cout << "\n";
    }
}

AxionStringsLevel::~AxionStringsLevel()
{
    axs_hierarchy_destroy(m_hier);
    if (m_scalars)
        std::fclose(m_scalars);
    if (m_spectrum)
        std::fclose(m_spectrum);
    if (m_pre_scalars)
        std::fclose(m_pre_scalars);
    if (m_loop_scalars)
        std::fclose(m_loop_scalars);
    axs_level_destroy(m_level);
    axs_ctx_destroy(m_ctx);
}

// AxionStringsLevel::initData, AxionStringsLevel.cpp:786-847
void AxionStringsLevel::initData()
{
    if (!m_restart.empty())
    {
        read_checkpoint(m_restart);
        return;
    }
    std::string ic_mode = "homogeneous";
    ParmParse(m_params, "axion_strings").queryAdd("ic_mode", ic_mode);
    const bool relaxed = m_phase == Phase::Relaxing;
    if (ic_mode == "fourier" || relaxed)
    {
        const FourierICParams f = Params::read_fourier_ic_params(
            m_params,
            relaxed ? "axion_strings.pre_evolution" : "axion_strings");
        const double m_r =
            relaxed ? std::sqrt(m_pre.lambda(0.0))
                    : (m_flat ? m_flat_mr : std::sqrt(m_bg.lambda(m_tau_i)));
        const double k_max_cells = f.k_max_over_mr * m_r * m_dx;
        FourierIC::generate(m_level, 0, k_max_cells, f.mean_square_variance,
                            f.seed, 0);
        FourierIC::generate(m_level, 1, k_max_cells, f.mean_square_variance,
                            f.seed, 1);
        axs_check(axs_level_zero_components(m_level, 2, 2), "setVal(Pi)");
        return;
    }
    const double R = m_flat ? 1.0 : m_bg.R(m_tau_i);
    const double Rp = m_flat ? 0.0 : m_bg.R_prime(m_tau_i);
    axs_check(axs_level_set_homogeneous(m_level, R, Rp), "initData");
}

// The log(m_r/H) (or step) output cadence of AxionStringsLevel.cpp:
// 1198-1248, evaluated after a step that ends at tau_after / step_after.
// `commit` advances the thresholds like the reference's while loops.
struct CadenceState
{
    double next_log, next_moore;
    bool moore_init;
    double H_switch;
};

static bool cadence_due(const Background &bg, const OutputCadence &c,
                        CadenceState &st, double tau, long step)
{
    if (c.every_steps > 0)
        return step % c.every_steps == 0;
    const double log_mr_over_h = -std::log(bg.H_over_mr_direct(tau));
    const bool in_moore =
        bg.c_sched.has_switch && tau >= bg.c_sched.tau_switch;
    if (in_moore)
    {
        if (!st.moore_init)
        {
            st.H_switch = bg.H(bg.c_sched.tau_switch);
            st.next_moore = 0.0;
            st.moore_init = true;
        }
        const double d = std::log(st.H_switch / bg.H(tau));
        if (d < st.next_moore)
            return false;
        while (st.next_moore <= d)
            st.next_moore += c.delta_log_mr_over_h;
        return true;
    }
    if (log_mr_over_h < st.next_log)
        return false;
    while (st.next_log <= log_mr_over_h)
        st.next_log += c.delta_log_mr_over_h;
    return true;
}

long AxionStringsLevel::steps_until_next_output()
{
    CadenceState st{m_next_log, m_next_moore_range, m_moore_init,
                    m_moore_H_switch};
    double tau = m_tau;
    for (long k = 1; k < 100000000; ++k)
    {
        tau += m_dt;
        if (cadence_due(m_bg, m_cadence, st, tau, m_step + k))
            return k;
    }
    return 100000000;
}

bool AxionStringsLevel::output_due()
{
    CadenceState st{m_next_log, m_next_moore_range, m_moore_init,
                    m_moore_H_switch};
    const bool due = cadence_due(m_bg, m_cadence, st, m_tau, m_step);
    m_next_log = st.next_log;
    m_next_moore_range = st.next_moore;
    m_moore_init = st.moore_init;
    m_moore_H_switch = st.H_switch;
    return due;
}

// Steps left of k under evolution.max_steps and evolution.stop_time (the
// latter on the level clock a_time, which counts through the relaxing
// phase like the reference's curTime()).
static long clamp_steps(long k, long step, long max_steps, double a_time,
                        double stop_time, double dt)
{
    if (max_steps >= 0)
        k = std::min(k, max_steps - step);
    if (stop_time >= 0.0)
    {
        long left = 0;
        double t = a_time;
        while (t < stop_time - 1e-12 * dt && left < k)
        {
            t += dt;
            ++left;
        }
        k = std::min(k, left);
    }
    return k;
}

// Main-phase stage scalars at conformal time tau: the Background's
// (AxionStringsLevel.cpp:878-885), smoothly blended from the frozen
// pre-evolution values during the post-handoff window (:889-907).
void AxionStringsLevel::stage_coeffs_at(double tau, axs_rhs_coeffs *c) const
{
    axs_check(axs_background_rhs_coeffs(&m_cbg, tau, c),
              "axs_background_rhs_coeffs");
    if (m_handoff_dtau > 0.0)
    {
        const double t = (tau - m_handoff_tau_start) / m_handoff_dtau;
        if (t < 1.0)
        {
            const double tc = std::clamp(t, 0.0, 1.0);
            const double s = tc * tc * (3.0 - 2.0 * tc);
            c->lambda = (1.0 - s) * m_handoff_lambda_start + s * c->lambda;
            c->curvature_coeff = (1.0 - s) * m_handoff_curv_start +
                                 s * c->curvature_coeff;
        }
    }
}

void AxionStringsLevel::evolve()
{
    if (m_phase == Phase::Relaxing)
        evolve_relaxing();
    if (m_phase == Phase::Evolving && m_amr.max_level > 0 && !m_hier)
        build_hierarchy();
    while (m_phase == Phase::Evolving)
    {
        long k = std::min(steps_until_next_output(), steps_to_checkpoint());
        k = clamp_steps(k, m_step, m_max_steps, m_a_time, m_stop_time, m_dt);
        if (m_has_tau_f)
        {
            // okToContinue (AxionStringsLevel.cpp:2007-2025): another step
            // while tau < tau_f
            long left = 0;
            for (double t = m_tau; t < m_tau_f && left < k; t += m_dt)
                ++left;
            k = std::min(k, left);
        }
        if (k <= 0)
            break;
        const bool blending =
            m_handoff_dtau > 0.0 &&
            m_tau < m_handoff_tau_start + m_handoff_dtau + m_dt;
        if (m_hier)
        {
            // coarse steps of the hierarchy (its stage scalars carry the
            // handoff blend themselves)
            for (long j = 0; j < k; ++j)
                axs_check(axs_hierarchy_step(m_hier), "axs_hierarchy_step");
        }
        else if (blending)
        {
            std::vector<axs_rhs_coeffs> table((size_t)k * 4);
            double tau = m_tau;
            for (long j = 0; j < k; ++j)
            {
                const double ts[4] = {tau, tau + 0.5 * m_dt, tau + 0.5 * m_dt,
                                      tau + m_dt};
                for (int st = 0; st < 4; ++st)
                    stage_coeffs_at(ts[st], &table[(size_t)j * 4 + st]);
                tau += m_dt;
            }
            axs_check(axs_rk4_advance_coeffs(m_level, table.data(), m_dt,
                                             (int)k, m_use_graph ? 1 : 0),
                      "axs_rk4_advance_coeffs");
        }
        else
            axs_check(axs_rk4_advance(m_level, &m_cbg, m_tau, m_dt, (int)k,
                                      m_use_graph ? 1 : 0),
                      "axs_rk4_advance");
        for (long j = 0; j < k; ++j)
        {
            m_tau += m_dt;
            m_a_time += m_dt;
        }
        m_step += k;
        if (output_due())
            timed_diagnostics([this] { diagnostics(); });
        if (m_check_int > 0 && m_step % m_check_int == 0)
            write_checkpoint();
    }
    axs_check(axs_ctx_synchronize(m_ctx), "synchronize");
}

void AxionStringsLevel::timed_diagnostics(const std::function<void()> &f)
{
    axs_check(axs_ctx_synchronize(m_ctx), "synchronize");
    const auto t0 = std::chrono::steady_clock::now();
    f();
    axs_check(axs_ctx_synchronize(m_ctx), "synchronize");
    m_diag_seconds += std::chrono::duration<double>(
                          std::chrono::steady_clock::now() - t0)
                          .count();
}

long AxionStringsLevel::steps_to_checkpoint() const
{
    if (m_check_int <= 0)
        return 100000000;
    return m_check_int - m_step % m_check_int;
}

// The relaxing phase: RK4 steps on the pre-evolution background with its
// own clock tau_pre = a_time (AxionStringsLevel.cpp:869-876), a xi check
// every m_xi_check_interval steps (:1044-1191).
void AxionStringsLevel::evolve_relaxing()
{
    while (m_phase == Phase::Relaxing)
    {
        const long want = m_xi_check_interval - m_steps_since_check;
        const long k = clamp_steps(std::min(want, steps_to_checkpoint()),
                                   m_step, m_max_steps, m_a_time, m_stop_time,
                                   m_dt);
        if (k <= 0)
            return;
        std::vector<axs_rhs_coeffs> table((size_t)k * 4);
        double t = m_a_time;
        for (long j = 0; j < k; ++j)
        {
            const double ts[4] = {t, t + 0.5 * m_dt, t + 0.5 * m_dt, t + m_dt};
            for (int st = 0; st < 4; ++st)
            {
                axs_rhs_coeffs &c = table[(size_t)j * 4 + st];
                const double R = m_pre.R(ts[st]);
                c.curvature_coeff = m_pre.curvature_term_coeff(ts[st]);
                c.lambda = m_pre.lambda(ts[st]);
                c.R_squared = R * R;
            }
            t += m_dt;
        }
        axs_check(axs_rk4_advance_coeffs(m_level, table.data(), m_dt, (int)k,
                                         m_use_graph ? 1 : 0),
                  "axs_rk4_advance_coeffs");
        for (long j = 0; j < k; ++j)
            m_a_time += m_dt;
        m_step += k;
        m_steps_since_check += k;
        const bool checkpoint_due = m_check_int > 0 && m_step % m_check_int == 0;
        if (m_steps_since_check >= m_xi_check_interval)
        {
            m_steps_since_check = 0;
            timed_diagnostics([this] { relax_check(); });
        }
        if (checkpoint_due)
            write_checkpoint();
        if (m_steps_since_check > 0 && !checkpoint_due)
            return; // stopped by max_steps / stop_time
    }
}

void AxionStringsLevel::relax_check()
{
    const PlaquetteCounts counts = count_plaquettes(m_level);
    const double tau_pre = m_a_time;
    // xi at the main run's start tau_i (AxionStringsLevel.cpp:1063-1071)
    const double xi = xi_from_plaquette_count((double)counts.n_p_plain, m_dx,
                                              m_L, m_bg.a_inv, m_tau_i);
    const double ratio = xi / m_xi_target;
    const double r_pre_now = m_pre.R(tau_pre);
    const double lambda_pre_now = m_pre.lambda(tau_pre);
    const TotalEnergyResult e = compute_composite_total_energy(
        {EnergyLevelInput{m_level, m_dx, nullptr}}, r_pre_now, lambda_pre_now,
        1.0, m_masking);
    const double rad_total = e.rho_radial_kinetic_unscreened +
                             e.rho_radial_gradient_unscreened +
                             e.rho_radial_mass_unscreened;
    const double axion_total =
        e.rho_axion_kinetic_unscreened + e.rho_axion_gradient_unscreened;
    std::cout << "  [AxionStrings pre-evolution RADIAL] tau_pre = " << tau_pre
              << "  R_pre = " << r_pre_now
              << "  m_r_pre = " << std::sqrt(lambda_pre_now)
              << "  rho_radial_total = " << rad_total
              << "  rho_axion_total = " << axion_total
              << "  radial/axion = " << rad_total / axion_total
              << "  rho_tot = " << e.rho_tot_unscreened << "\n";
    if (!m_pre_scalars)
        // column names of AxionStringsLevel.cpp:1122-1125 after the time
        m_pre_scalars = open_dat("pre_evolution_scalars.dat",
                                 "# time tau_pre R_pre m_r_pre N_p xi "
                                 "xi_target rho_radial_kin rho_radial_grad "
                                 "rho_radial_mass rho_axion_kin "
                                 "rho_axion_grad rho_tot\n",
                                 tau_pre);
    const double row[13] = {tau_pre,
                            tau_pre,
                            r_pre_now,
                            std::sqrt(lambda_pre_now),
                            (double)counts.n_p_plain,
                            xi,
                            m_xi_target,
                            e.rho_radial_kinetic_unscreened,
                            e.rho_radial_gradient_unscreened,
                            e.rho_radial_mass_unscreened,
                            e.rho_axion_kinetic_unscreened,
                            e.rho_axion_gradient_unscreened,
                            e.rho_tot_unscreened};
    for (int i = 0; i < 13; ++i)
        std::fprintf(m_pre_scalars, i ? " %.10e" : "%.10e", row[i]);
    std::fprintf(m_pre_scalars, "\n");
    std::fflush(m_pre_scalars);
    const double n_p_required =
        xi > 0.0 ? (double)counts.n_p_plain * (m_xi_target / xi) : INFINITY;
    std::cout << "  [AxionStrings pre-evolution] tau_pre = " << tau_pre
              << "  N_p = " << counts.n_p_plain << "  xi(at tau_i) = " << xi
              << "  target = " << m_xi_target << "  ratio = " << ratio << "\n"
              << "  [AxionStrings pre-evolution] RELAXATION PROGRESS: N_p / "
                 "N_p_required = "
              << counts.n_p_plain << " / " << n_p_required << " = " << ratio
              << "\n";
    if (ratio <= 1.0)
    {
        std::cout << "  [AxionStrings pre-evolution] target xi reached -- "
                     "transitioning in place to the main evolution.\n";
        apply_pre_evolution_to_main_rescale();
        m_phase = Phase::Evolving;
        return;
    }
    if (ratio <= m_xi_cadence.fine_threshold)
        m_xi_check_interval = m_xi_cadence.fine_interval;
    else if (ratio <= m_xi_cadence.medium_threshold)
        m_xi_check_interval = m_xi_cadence.medium_interval;
    else
        m_xi_check_interval = m_xi_cadence.coarse_interval;
}

// AxionStringsLevel::apply_pre_evolution_to_main_rescale,
// AxionStringsLevel.cpp:2027-2125
void AxionStringsLevel::apply_pre_evolution_to_main_rescale()
{
    const double tau_pre_end = m_a_time;
    const double R_pre = m_pre.R(tau_pre_end);
    const double R_pre_prime = m_pre.R_prime(tau_pre_end);
    const double R_main = m_bg.R(m_tau_i);
    const double R_main_prime = m_bg.R_prime(m_tau_i);
    const double kappa = R_main / R_pre;
    const double common_term = R_main_prime - kappa * kappa * R_pre_prime;
    std::cout << "  [AxionStrings] pre-evolution -> main handoff: "
              << "tau_pre_end = " << tau_pre_end << ", R_pre = " << R_pre
              << ", R_main(tau_i) = " << R_main << ", kappa = " << kappa
              << "\n";
    if (m_handoff_n_periods > 0.0)
    {
        const double lambda_start = m_pre.lambda(tau_pre_end);
        const double curvature_start = m_pre.curvature_term_coeff(tau_pre_end);
        const double lambda_main = m_bg.lambda(m_tau_i);
        const double m_r_for_period = lambda_main > 0.0
                                          ? std::sqrt(lambda_main)
                                          : std::sqrt(lambda_start);
        m_handoff_lambda_start = lambda_start;
        m_handoff_curv_start = curvature_start;
        m_handoff_tau_start = m_tau_i;
        m_handoff_dtau = m_handoff_n_periods * 2.0 * M_PI / m_r_for_period;
        std::cout << "  [AxionStrings] handoff transition: lambda "
                  << lambda_start << " -> " << lambda_main
                  << ", curvature_term_coeff " << curvature_start << " -> "
                  << m_bg.curvature_term_coeff(m_tau_i) << ", over d(tau) = "
                  << m_handoff_dtau << " (" << m_handoff_n_periods
                  << " main-schedule core periods)\n";
    }
    // the level clock keeps counting; tau = offset + a_time = tau_i now
    m_tau_offset = m_tau_i - tau_pre_end;
    m_tau = m_tau_i;
    axs_check(axs_level_handoff_rescale(m_level, kappa, common_term, R_pre),
              "axs_level_handoff_rescale");
}

// specific_post_timestep's snapshot, AxionStringsLevel.cpp:1256-1764
void AxionStringsLevel::diagnostics()
{
    const double tau = m_tau;
    const long n_cells = (long)m_n[0] * m_n[1] * m_n[2];
    double R = 1.0, RpR = 0.0, lambda = m_flat_mr * m_flat_mr,
           m_r_over_h = 0.0, H = 0.0;
    if (!m_flat)
    {
        R = m_bg.R(tau);
        RpR = 1.0 / (m_bg.b_inv * tau); // AxionStringsLevel.cpp:1283
        lambda = m_bg.lambda(tau);
        m_r_over_h = 1.0 / m_bg.H_over_mr_direct(tau);
        H = m_bg.H(tau);
    }
    const double m_r = std::sqrt(lambda);

    PlaquetteCounts counts;
    double ell = 0.0, ell_w = 0.0;
    TotalEnergyResult composite{};
    if (m_hier)
    {
        // composite counts and energies over the hierarchy; the level-0
        // collapse into m_level for the velocities and the spectrum
        composite_sums(R, m_flat ? m_r * m_r : lambda, RpR, &counts, &ell,
                       &ell_w, &composite);
        axs_check(axs_hierarchy_collapse(m_hier, m_level),
                  "axs_hierarchy_collapse");
    }
    else
    {
        counts = count_plaquettes(m_level);
        ell = (2.0 / 3.0) * (double)counts.n_p_plain * m_dx;
        ell_w = (2.0 / 3.0) * (double)counts.n_p_weighted * m_dx;
    }
    if (m_flat)
    {
        flat_diagnostics(counts, ell, m_r, m_hier ? &composite : nullptr);
        return;
    }
    const double xi = m_flat ? 0.0 : xi_from_comoving_length(ell, m_L, m_bg.a_inv, tau);
    const double xi_w =
        m_flat ? 0.0 : xi_from_comoving_length(ell_w, m_L, m_bg.a_inv, tau);

    const VelocityResult vel =
        compute_velocity_at_pierced_corners(m_level, R, RpR, m_r);
    const double vw_sum = vel.sum_gamma_sq_v_sq * m_dx;
    const double vw_cnt = (double)vel.count * m_dx;

    const TotalEnergyResult e =
        m_hier ? composite
               : compute_composite_total_energy(
                     {EnergyLevelInput{m_level, m_dx, nullptr}}, R, lambda,
                     RpR, m_masking);

    const double sum_core = e.n_total * e.rho_tot_unscreened -
                            e.n_unmasked * e.rho_tot_screened;
    const double tail = e.n_unmasked * (e.rho_axion_gradient_screened -
                                        e.rho_axion_kinetic_screened);
    const double t_core = ell > 0.0 ? (R * R) * sum_core / ell : 0.0;
    const double t_tail = ell > 0.0 ? (R * R) * (sum_core + tail) / ell : 0.0;
    VelocityResult vr{vw_sum, vel.count};
    (void)vw_cnt;
    write_network_scalars_row(counts, xi, xi_w, m_r_over_h, e, vr, t_core,
                              t_tail);
    // progress against the box plan's tau_f (AxionStringsLevel.cpp:1444-1468)
    if (m_has_tau_f)
        std::cout << "  ************ tau/tau_f = " << tau / m_tau_f
                  << " ************\n";
    std::cout << "  [AxionStrings] tau = " << tau << "  step = " << m_step;
    if (m_has_tau_f)
        std::cout << "  (" << 100.0 * (tau - m_tau_i) / (m_tau_f - m_tau_i)
                  << "% of conformal time elapsed)";
    std::cout << "  N_p = " << counts.n_p_plain
              << "  N_p_W = " << counts.n_p_weighted << "  xi = " << xi
              << "  m_r/H = " << m_r_over_h
              << "  rho_tot unscreened = " << e.rho_tot_unscreened
              << "  screened = " << e.rho_tot_screened << "\n";
    if (m_hier)
    {
        int nl = 0;
        int64_t nreg = 0;
        axs_check(axs_hierarchy_info(m_hier, &nl, nullptr, nullptr, &nreg,
                                     nullptr),
                  "axs_hierarchy_info");
        std::cout << "  [AxionStrings AMR] levels = " << nl
                  << "  regrids = " << nreg << "  cells:";
        for (int l = 0; l < nl; ++l)
        {
            axs_boxlevel *bl = nullptr;
            int nb = 0;
            axs_check(axs_hierarchy_level(m_hier, l, &bl, nullptr, nullptr,
                                          &nb),
                      "axs_hierarchy_level");
            std::cout << " " << axs_boxlevel_num_cells(bl) << " (" << nb
                      << " boxes)";
        }
        std::cout << "\n";
    }

    if (!m_compute_spectrum)
        return;
    const int n_max = std::max(m_n[0], std::max(m_n[1], m_n[2]));
    MaskingParams none{};
    none.scheme = MaskingScheme::None;
    const Spectrum s_scr = compute_spectrum(m_level, m_spectrum_kind, m_masking,
                                            R, n_max);
    const Spectrum s_uns =
        compute_spectrum(m_level, m_spectrum_kind, none, R, n_max);
    const double nsq = (double)n_cells * (double)n_cells;
    if (!m_spectrum)
        m_spectrum = open_dat("axion_spectrum.dat",
                     "# tau mode_index k_over_H shell_average_screened "
                     "shell_average_unscreened full_cube_energy_screened "
                     "full_cube_energy_unscreened "
                     "inscribed_sphere_energy_screened "
                     "inscribed_sphere_energy_unscreened "
                     "real_space_mean_sq_screened real_space_mean_sq_unscreened "
                     "parseval_full_screened parseval_full_unscreened "
                     "parseval_inscribed_screened "
                     "parseval_inscribed_unscreened\n", tau);
    const double k_over_h_per_mode = H > 0.0 ? 2.0 * M_PI / (m_L * H) : 0.0;
    for (size_t i = 0; i < s_scr.shell_average.size(); ++i)
        std::fprintf(m_spectrum,
                     "%.10e %zu %.10e %.10e %.10e %.10e %.10e %.10e %.10e "
                     "%.10e %.10e %.10e %.10e %.10e %.10e\n",
                     tau, i, k_over_h_per_mode * (double)i,
                     s_scr.shell_average[i],
                     i < s_uns.shell_average.size() ? s_uns.shell_average[i] : 0.0,
                     s_scr.full_cube_energy, s_uns.full_cube_energy,
                     s_scr.inscribed_sphere_energy, s_uns.inscribed_sphere_energy,
                     s_scr.real_space_sum_sq / n_cells,
                     s_uns.real_space_sum_sq / n_cells,
                     s_scr.full_cube_energy / nsq, s_uns.full_cube_energy / nsq,
                     s_scr.inscribed_sphere_energy / nsq,
                     s_uns.inscribed_sphere_energy / nsq);
    std::fflush(m_spectrum);
}

// The flat-space branch of specific_post_timestep (AxionStringsLevel.cpp:
// 932-1040): string length, energy with R = 1 (n_total * rho_tot is the
// conserved total energy), mean gamma^2 v^2 at pierced corners, one
// loop_scalars.dat row per output; no xi, tension or spectrum.
void AxionStringsLevel::flat_diagnostics(const PlaquetteCounts &counts,
                                         double ell, double m_r,
                                         const TotalEnergyResult *composite)
{
    const double t = m_tau;
    const VelocityResult vel =
        compute_velocity_at_pierced_corners(m_level, 1.0, 0.0, m_r);
    const double vw_sum = vel.sum_gamma_sq_v_sq * m_dx;
    const double vw_cnt = (double)vel.count * m_dx;
    const TotalEnergyResult e =
        composite ? *composite
                  : compute_composite_total_energy(
                        {EnergyLevelInput{m_level, m_dx, nullptr}}, 1.0,
                        m_r * m_r, 0.0, m_masking);
    const double total_energy = e.n_total * e.rho_tot_unscreened;
    const double mean_g2v2 = vw_cnt > 0.0 ? vw_sum / vw_cnt : 0.0;
    const double mean_gamma = std::sqrt(1.0 + mean_g2v2);
    std::cout << "  [AxionStrings flat-space] t = " << t
              << "  N_p = " << counts.n_p_plain << "  string length = " << ell
              << "  rho_tot(unscreened) = " << e.rho_tot_unscreened
              << "  total energy = " << total_energy
              << "  <gamma^2 v^2> = " << mean_g2v2
              << "  <gamma> = " << mean_gamma
              << "  N_corners = " << vel.count << "\n";
    if (!m_loop_scalars)
        // column names of AxionStringsLevel.cpp:1027-1030 after the time
        m_loop_scalars = open_dat("loop_scalars.dat",
                                  "# time N_p string_length "
                                  "rho_tot_unscreened total_energy "
                                  "mean_gamma_sq_v_sq mean_gamma "
                                  "n_velocity_corners\n",
                                  t);
    const double row[8] = {t,
                           (double)counts.n_p_plain,
                           ell,
                           e.rho_tot_unscreened,
                           total_energy,
                           mean_g2v2,
                           mean_gamma,
                           (double)vel.count};
    for (int i = 0; i < 8; ++i)
        std::fprintf(m_loop_scalars, i ? " %.10e" : "%.10e", row[i]);
    std::fprintf(m_loop_scalars, "\n");
    std::fflush(m_loop_scalars);
}

// The hierarchy from the uniform level at the start of the main phase
// (tag_cells never refines while relaxing, AxionStringsLevel.cpp:1924-1934).
void AxionStringsLevel::build_hierarchy()
{
    axs_amr_params p{};
    p.max_level = m_amr.max_level;
    p.n_error_buf = m_amr.n_error_buf;
    p.blocking_factor = m_amr.blocking_factor;
    p.max_grid_size = m_amr.max_grid_size;
    p.regrid_int = m_amr.regrid_int;
    p.force_full_refinement = m_tagging.force_full_refinement ? 1 : 0;
    p.gradient_threshold = m_tagging.gradient_threshold;
    p.radial_gradient_threshold = m_tagging.radial_gradient_threshold;
    p.amplitude_threshold = m_tagging.amplitude_threshold;
    p.n_schedule = (int)std::min<size_t>(m_level_add.size(), AXS_MAX_LEVELS);
    for (int i = 0; i < p.n_schedule; ++i)
        p.level_add_log_mr_over_h[i] = m_level_add[(size_t)i];
    p.use_graph = m_use_graph ? 1 : 0;
    const axs_coeff_blend blend{m_handoff_tau_start, m_handoff_dtau,
                                m_handoff_lambda_start, m_handoff_curv_start};
    axs_check(axs_hierarchy_create(m_ctx, m_level, &p, &m_cbg, &blend, m_tau,
                                   m_dt, &m_hier),
              "axs_hierarchy_create");
    int nl = 0;
    axs_check(axs_hierarchy_info(m_hier, &nl, nullptr, nullptr, nullptr,
                                 nullptr),
              "axs_hierarchy_info");
    std::cout << "  [AxionStrings AMR] hierarchy at tau = " << m_tau << ": "
              << nl << " level(s), max_level " << p.max_level
              << ", blocking_factor " << p.blocking_factor
              << ", max_grid_size " << p.max_grid_size << ", n_error_buf "
              << p.n_error_buf << ", regrid_int " << p.regrid_int << "\n";
}

// The composite xi counts and energies of AxionStringsLevel.cpp:1299-1361:
// each level over the cells no finer level covers, lengths weighted by the
// level's dx, energies by dx^3.
void AxionStringsLevel::composite_sums(double R, double lambda, double RpR,
                                       PlaquetteCounts *counts, double *ell,
                                       double *ell_w, TotalEnergyResult *e)
{
    int nl = 0;
    axs_check(axs_hierarchy_info(m_hier, &nl, nullptr, nullptr, nullptr,
                                 nullptr),
              "axs_hierarchy_info");
    std::vector<int64_t> c((size_t)2 * nl, 0);
    axs_check(axs_hierarchy_plaquette_counts(m_hier, c.data()),
              "axs_hierarchy_plaquette_counts");
    *counts = PlaquetteCounts{};
    *ell = *ell_w = 0.0;
    for (int l = 0; l < nl; ++l)
    {
        axs_boxlevel *bl = nullptr;
        double dxl = 0.0;
        axs_check(axs_hierarchy_level(m_hier, l, &bl, nullptr, &dxl, nullptr),
                  "axs_hierarchy_level");
        counts->n_p_plain += (long)c[(size_t)2 * l];
        counts->n_p_weighted += (long)c[(size_t)2 * l + 1];
        *ell += (2.0 / 3.0) * (double)c[(size_t)2 * l] * dxl;
        *ell_w += (2.0 / 3.0) * (double)c[(size_t)2 * l + 1] * dxl;
    }
    const axs_masking m = to_c(m_masking);
    double acc[14];
    axs_check(axs_hierarchy_energy_sums(m_hier, R, lambda, RpR, &m, acc),
              "axs_hierarchy_energy_sums");
    TotalEnergyResult o{};
    o.n_total = acc[13];
    o.n_unmasked = acc[12];
    o.rho_tot_unscreened = acc[0] / o.n_total;
    o.rho_axion_kinetic_unscreened = acc[2] / o.n_total;
    o.rho_axion_gradient_unscreened = acc[4] / o.n_total;
    o.rho_radial_kinetic_unscreened = acc[6] / o.n_total;
    o.rho_radial_gradient_unscreened = acc[8] / o.n_total;
    o.rho_radial_mass_unscreened = acc[10] / o.n_total;
    if (o.n_unmasked > 0.0)
    {
        o.rho_tot_screened = acc[1] / o.n_unmasked;
        o.rho_axion_kinetic_screened = acc[3] / o.n_unmasked;
        o.rho_axion_gradient_screened = acc[5] / o.n_unmasked;
        o.rho_radial_kinetic_screened = acc[7] / o.n_unmasked;
        o.rho_radial_gradient_screened = acc[9] / o.n_unmasked;
        o.rho_radial_mass_screened = acc[11] / o.n_unmasked;
    }
    *e = o;
}

void AxionStringsLevel::write_network_scalars_row(
    const PlaquetteCounts &c, double xi, double xi_w, double m_r_over_h,
    const TotalEnergyResult &e, const VelocityResult &v, double tension_core,
    double tension_tail)
{
    if (!m_scalars)
        // column names of AxionStringsLevel.cpp:1541-1551
        m_scalars = open_dat("network_scalars.dat",
                     "# tau N_p N_p_weighted xi xi_weighted m_r_over_H "
                     "rho_tot_unscreened rho_tot_screened "
                     "rho_axion_kin_unscreened rho_axion_kin_screened "
                     "rho_axion_grad_unscreened rho_axion_grad_screened "
                     "rho_radial_kin_unscreened rho_radial_kin_screened "
                     "rho_radial_grad_unscreened rho_radial_grad_screened "
                     "rho_radial_mass_unscreened rho_radial_mass_screened "
                     "n_total n_unmasked mean_gamma_sq_v_sq mean_gamma "
                     "n_velocity_corners tension_core_only "
                     "tension_core_plus_tail\n", m_tau);
    const double wcount = (double)v.count * m_dx;
    const double mean_g2v2 = wcount > 0.0 ? v.sum_gamma_sq_v_sq / wcount : 0.0;
    const double row[25] = {m_tau,
                            (double)c.n_p_plain,
                            (double)c.n_p_weighted,
                            xi,
                            xi_w,
                            m_r_over_h,
                            e.rho_tot_unscreened,
                            e.rho_tot_screened,
                            e.rho_axion_kinetic_unscreened,
                            e.rho_axion_kinetic_screened,
                            e.rho_axion_gradient_unscreened,
                            e.rho_axion_gradient_screened,
                            e.rho_radial_kinetic_unscreened,
                            e.rho_radial_kinetic_screened,
                            e.rho_radial_gradient_unscreened,
                            e.rho_radial_gradient_screened,
                            e.rho_radial_mass_unscreened,
                            e.rho_radial_mass_screened,
                            e.n_total,
                            e.n_unmasked,
                            mean_g2v2,
                            std::sqrt(1.0 + mean_g2v2),
                            (double)v.count,
                            tension_core,
                            tension_tail};
    for (int i = 0; i < 25; ++i)
        std::fprintf(m_scalars, i ? " %.10e" : "%.10e", row[i]);
    std::fprintf(m_scalars, "\n");
    std::fflush(m_scalars);
}

// A .dat file opened the way AxionStringsLevel.cpp:1510-1537 opens its
// SmallDataIO files: a fresh run moves an existing file aside to
// <name>.old.<seconds> and writes the header; a restarted run (amr.restart
// set) keeps the file but drops its rows at or after the time t of the row
// about to be written, i.e. the provisional rows of the segment being redone.
FILE *AxionStringsLevel::open_dat(const std::string &name, const char *header,
                                  double t)
{
    const std::string path = m_out_dir + "/" + name;
    std::ifstream probe(path);
    const bool exists = probe.good();
    if (exists && !m_restart.empty())
    {
        std::vector<std::string> keep;
        std::string line;
        const double eps = 1e-8 * std::max(1.0, std::fabs(t));
        while (std::getline(probe, line))
        {
            if (line.empty())
                continue;
            if (line[0] == '#')
            {
                keep.push_back(line);
                continue;
            }
            const double ti = std::strtod(line.c_str(), nullptr);
            if (ti < t - eps)
                keep.push_back(line);
        }
        probe.close();
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f)
            throw AxsError("cannot write " + name);
        for (const std::string &l : keep)
            std::fprintf(f, "%s\n", l.c_str());
        std::fflush(f);
        return f;
    }
    probe.close();
    if (exists)
    {
        const std::string aside =
            path + ".old." + std::to_string((long long)std::time(nullptr));
        std::rename(path.c_str(), aside.c_str());
    }
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f)
        throw AxsError("cannot write " + name);
    std::fputs(header, f);
    return f;
}

// Checkpoint: a text header of key = value lines (doubles with 17
// significant digits, so a restart continues bit for bit), then the raw
// stored words of the state (FP32C: hi, then lo).
void AxionStringsLevel::write_checkpoint()
{
    char num[32];
    std::snprintf(num, sizeof num, "%05ld", m_step);
    const std::string path = m_out_dir + "/" + m_check_file + num + ".axs";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
        throw AxsError("cannot write checkpoint " + path);
    auto put = [&](const char *k, double v) {
        std::fprintf(f, "%s = %.17g\n", k, v);
    };
    std::fprintf(f, "axion_strings checkpoint 1\n");
    put("nx", m_n[0]);
    put("ny", m_n[1]);
    put("nz", m_n[2]);
    put("dx", m_dx);
    put("precision", m_precision);
    put("step", (double)m_step);
    put("tau", m_tau);
    put("a_time", m_a_time);
    put("tau_offset", m_tau_offset);
    put("phase", m_phase == Phase::Relaxing ? 0 : 1);
    put("next_log", m_next_log);
    put("next_moore", m_next_moore_range);
    put("moore_init", m_moore_init ? 1 : 0);
    put("H_switch", m_moore_H_switch);
    put("xi_check_interval", (double)m_xi_check_interval);
    put("steps_since_check", (double)m_steps_since_check);
    put("handoff_lambda_start", m_handoff_lambda_start);
    put("handoff_curv_start", m_handoff_curv_start);
    put("handoff_tau_start", m_handoff_tau_start);
    put("handoff_dtau", m_handoff_dtau);
    std::fprintf(f, "end\n");
    const size_t n4 = (size_t)4 * m_n[0] * m_n[1] * m_n[2];
    const size_t word = m_precision == AXS_FP64 ? 8 : 4;
    std::vector<char> hi(n4 * word);
    std::vector<float> lo(m_precision == AXS_FP32C ? n4 : 0);
    axs_check(axs_level_download_raw(m_level, hi.data(),
                                     lo.empty() ? nullptr : lo.data()),
              "axs_level_download_raw");
    bool ok = std::fwrite(hi.data(), 1, hi.size(), f) == hi.size();
    if (!lo.empty())
        ok = ok && std::fwrite(lo.data(), sizeof(float), lo.size(), f) ==
                       lo.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok)
        throw AxsError("short write of checkpoint " + path);
    std::cout << "  [AxionStrings] checkpoint " << path << "\n";
}

void AxionStringsLevel::read_checkpoint(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f)
        throw ParamError("amr.restart: cannot open " + path);
    std::map<std::string, double> kv;
    char line[256];
    if (!std::fgets(line, sizeof line, f) ||
        std::strncmp(line, "axion_strings checkpoint 1", 26) != 0)
    {
        std::fclose(f);
        throw ParamError("amr.restart: " + path + " is not a checkpoint");
    }
    while (std::fgets(line, sizeof line, f))
    {
        if (std::strncmp(line, "end", 3) == 0)
            break;
        char key[128];
        double v = 0.0;
        if (std::sscanf(line, "%127s = %lf", key, &v) == 2)
            kv[key] = v;
    }
    auto get = [&](const char *k) -> double {
        auto it = kv.find(k);
        if (it == kv.end())
        {
            std::fclose(f);
            throw ParamError(std::string("amr.restart: checkpoint lacks ") + k);
        }
        return it->second;
    };
    if ((int)get("nx") != m_n[0] || (int)get("ny") != m_n[1] ||
        (int)get("nz") != m_n[2] || (int)get("precision") != m_precision)
    {
        std::fclose(f);
        throw ParamError("amr.restart: checkpoint grid or precision differs "
                         "from amr.n_cell / axion_strings.precision");
    }
    m_step = (long)get("step");
    m_tau = get("tau");
    m_a_time = get("a_time");
    m_tau_offset = get("tau_offset");
    m_phase = get("phase") == 0 ? Phase::Relaxing : Phase::Evolving;
    m_next_log = get("next_log");
    m_next_moore_range = get("next_moore");
    m_moore_init = get("moore_init") != 0;
    m_moore_H_switch = get("H_switch");
    m_xi_check_interval = (long)get("xi_check_interval");
    m_steps_since_check = (long)get("steps_since_check");
    m_handoff_lambda_start = get("handoff_lambda_start");
    m_handoff_curv_start = get("handoff_curv_start");
    m_handoff_tau_start = get("handoff_tau_start");
    m_handoff_dtau = get("handoff_dtau");
    const size_t n4 = (size_t)4 * m_n[0] * m_n[1] * m_n[2];
    const size_t word = m_precision == AXS_FP64 ? 8 : 4;
    std::vector<char> hi(n4 * word);
    std::vector<float> lo(m_precision == AXS_FP32C ? n4 : 0);
    bool ok = std::fread(hi.data(), 1, hi.size(), f) == hi.size();
    if (!lo.empty())
        ok = ok && std::fread(lo.data(), sizeof(float), lo.size(), f) ==
                       lo.size();
    std::fclose(f);
    if (!ok)
        throw ParamError("amr.restart: " + path + " is truncated");
    axs_check(axs_level_upload_raw(m_level, hi.data(),
                                   lo.empty() ? nullptr : lo.data()),
              "axs_level_upload_raw");
    std::cout << "  [AxionStrings] restart from " << path << " at step "
              << m_step << ", tau = " << m_tau << "\n";
}

} // namespace axs_host
