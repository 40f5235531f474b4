// AxionStringsLevel.hpp -- the single-level driver around the hot path:
// initial data, the RK4 evolution, and the cadence diagnostics with their
// output rows.
//
// Follows the reference's AxionStringsLevel (AxionStringsLevel.cpp):
//   variableSetUp       :136-208  (phase selection :192-207)
//   initData            :210-848  (ic_mode fourier / fourier_relaxed
//                                  :784-822, homogeneous :824-847)
//   specific_eval_rhs   :850-921  (stage scalars of the relaxing phase,
//                                  the main schedule and the post-handoff
//                                  blend; here one device table)
//   relaxing-phase xi check :1044-1191 (pre_evolution_scalars.dat)
//   apply_pre_evolution_to_main_rescale :2027-2125
//   restart-safe .dat output :1510-1535 (SmallDataIO: a fresh run moves
//                                  an existing file aside, a restart
//                                  drops the rows of the redone segment)
// plus checkpoints of the level state (amr.check_int / amr.check_file /
// amr.restart, the keys of the reference's AMReX driver).
//   specific_post_timestep :923-1914  (log(m_r/H) cadence :1198-1248,
//                                  counts/xi/energy/velocity :1299-1361,
//                                  network_scalars.dat :1536-1579,
//                                  spectrum + axion_spectrum.dat :1593-1764)
// on one uniform periodic level, the level-0 path of the reference, or, with
// amr.max_level > 0, on the AMR hierarchy (axs_hierarchy_*) built from that
// level once the main phase starts: subcycled steps, regrids every
// amr.regrid_int coarse steps, and the composite diagnostics of
// AxionStringsLevel.cpp:1256-1361 (counts and energies over every level,
// each masked by the coverage of the next finer; velocities and the
// spectrum on the level-0 collapse, collapse_to_level0 :110-134).
#pragma once

#include "AxionStringsOps.hpp"
#include "AxionStringsParams.hpp"

#include <cstdio>
#include <functional>
#include <string>
#include <vector>

namespace axs_host
{

class AxionStringsLevel
{
  public:
    AxionStringsLevel(const ParamTable &params, const std::string &out_dir);
    ~AxionStringsLevel();

    void initData();
    // Runs until evolution.max_steps / evolution.stop_time / tau_f.
    void evolve();
    // One diagnostics snapshot at the current time (also writes rows).
    void diagnostics();

    double tau() const { return m_tau; }
    long step() const { return m_step; }
    // wall seconds spent in the output diagnostics and the relaxing phase's
    // xi checks (each measured from a synchronised device)
    double diagnostics_seconds() const { return m_diag_seconds; }
    axs_level *level() { return m_level; }
    axs_hierarchy *hierarchy() { return m_hier; }

    enum class Phase
    {
        Relaxing,
        Evolving
    };
    Phase phase() const { return m_phase; }

    void write_checkpoint();

  private:
    void read_checkpoint(const std::string &path);
    FILE *open_dat(const std::string &name, const char *header, double t);
    long steps_to_checkpoint() const;
    void evolve_relaxing();
    void relax_check();
    void apply_pre_evolution_to_main_rescale();
    void stage_coeffs_at(double a_time, axs_rhs_coeffs *c) const;
    bool output_due();
    long steps_until_next_output();
    void flat_diagnostics(const PlaquetteCounts &counts, double ell,
                          double m_r, const TotalEnergyResult *composite);
    void build_hierarchy();
    void composite_sums(double R, double lambda, double RpR,
                        PlaquetteCounts *counts, double *ell, double *ell_w,
                        TotalEnergyResult *e);
    void timed_diagnostics(const std::function<void()> &f);
    void write_network_scalars_row(const PlaquetteCounts &c, double xi,
                                   double xi_w, double m_r_over_h,
                                   const TotalEnergyResult &e,
                                   const VelocityResult &v, double tension_core,
                                   double tension_tail);

    ParamTable m_params; // the caller's table with the box plan applied
    std::string m_out_dir;
    axs_ctx *m_ctx{nullptr};
    axs_level *m_level{nullptr};
    // refined runs: the hierarchy (from the main phase on) and its keys
    axs_hierarchy *m_hier{nullptr};
    Params::AmrKeys m_amr;
    std::vector<double> m_level_add; // the box plan's level schedule
    int m_n[3]{};
    double m_dx{}, m_L{}, m_dt{};
    int m_precision{AXS_FP64};
    bool m_flat{false};
    double m_flat_mr{1.0};
    Background m_bg;
    axs_background m_cbg{};
    double m_tau_i{}, m_tau{}, m_stop_time{-1.0};
    bool m_has_tau_f{false}; // axion_strings.derived_tau_f of the box plan
    double m_tau_f{0.0};
    long m_step{0}, m_max_steps{-1};
    bool m_use_graph{true};
    MaskingParams m_masking;
    StringTaggerParams m_tagging;
    OutputCadence m_cadence;
    double m_next_log{};
    bool m_moore_init{false};
    double m_moore_H_switch{}, m_next_moore_range{};
    bool m_compute_spectrum{false};
    int m_spectrum_kind{AXS_FIELD_MASKED_A_DOT};
    FILE *m_scalars{nullptr};
    FILE *m_spectrum{nullptr};
    FILE *m_pre_scalars{nullptr};
    FILE *m_loop_scalars{nullptr};
    double m_diag_seconds{0.0}; // wall time in diagnostics and xi checks
    // pre-evolution (ic_mode = fourier_relaxed)
    Phase m_phase{Phase::Evolving};
    PreEvolutionBackground m_pre;
    double m_xi_target{};
    XiCheckCadence m_xi_cadence;
    long m_xi_check_interval{50};
    long m_steps_since_check{0};
    double m_handoff_n_periods{0.0};
    double m_a_time{0.0};     // the level clock: tau_pre, then on through
    double m_tau_offset{0.0}; // tau = m_tau_offset + m_a_time (main phase)
    double m_handoff_lambda_start{}, m_handoff_curv_start{};
    double m_handoff_tau_start{}, m_handoff_dtau{0.0};
    // checkpoint / restart
    long m_check_int{0};
    std::string m_check_file{"chk"};
    std::string m_restart;
};

} // namespace axs_host
