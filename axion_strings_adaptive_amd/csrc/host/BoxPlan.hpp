// BoxPlan.hpp -- box planning: the grid size, box length and final time a
// run needs so that, at its last time, the box spans N1 Hubble lengths and a
// string core (1/m_r) spans N2 cells.
//
// Same function names, arguments and results as the reference's
// AxionStrings/BoxPlan.hpp (compute_general_box_plan :36-50,
// compute_amr_box_plan :96-125, pre_evolution_L_tilde :130-137,
// moore_max_log_dynamic_range :144-149, compute_moore_box_plan :183-194).
// Units: tau0 = 1, lambda0 = 1, R0 = 1/(a_inv - 1), so that with
// R = R0 tau^(1/b_inv) and m_r = sqrt(lambda), m_r/H = tau^((a_inv-c)/b_inv).
#pragma once

#include <cmath>
#include <vector>

namespace axs_host
{

struct GeneralBoxPlan
{
    double tau_f{};
    double L_tilde{};
    double delta_x{};
    double delta_tau_leapfrog_bound{}; // dx/3
};

// H L_phys = N1 and 1/(dx_phys m_r) = N2 at tau_f with N = L/dx cells:
// N = N1 N2 m_r/H = N1 N2 tau_f^((a_inv-c)/b_inv), and L_tilde = N1/(H R)
// = (N1/R0) tau_f^(...) with the same power.
[[nodiscard]] inline GeneralBoxPlan
compute_general_box_plan(int N, double N1, double N2, double a_inv, double c)
{
    GeneralBoxPlan p;
    const double grow =
        std::pow(double(N) / (N1 * N2), (a_inv - 1.0) / (a_inv - c));
    p.tau_f = grow;
    const double R0 = 1.0 / (a_inv - 1.0);
    p.L_tilde = (N1 / R0) * grow;
    p.delta_x = p.L_tilde / N;
    p.delta_tau_leapfrog_bound = p.delta_x / 3.0;
    return p;
}

struct AmrBoxPlan
{
    double L_tilde{};
    double dx_finest{};
    double dx_base{};
    double tau_f{};
    int N_base{-1};               // N / 2^max_level, or -1 when not exact
    std::vector<double> log_add;  // log(m_r/H) at which level ell (1-based,
                                  // entry ell-1) must be present
};

// N is the effective resolution of the finest level.  Level ell halves the
// spacing of level ell-1; N2 falls as m_r/H grows, and halving dx restores it
// after log(m_r/H) advances by ln2 (a_inv - c)/(1 - c), so the levels come
// online that many log units apart, the finest one spacing before tau_f.
[[nodiscard]] inline AmrBoxPlan
compute_amr_box_plan(int N_effective, double N1, double N2, double a_inv,
                     double c, int max_level)
{
    const GeneralBoxPlan g =
        compute_general_box_plan(N_effective, N1, N2, a_inv, c);
    const int refine = 1 << max_level;
    AmrBoxPlan p;
    p.L_tilde = g.L_tilde;
    p.dx_finest = g.delta_x;
    p.dx_base = g.delta_x * double(refine);
    p.tau_f = g.tau_f;
    p.N_base = (N_effective % refine) == 0 ? N_effective / refine : -1;
    const double log_f = (a_inv - c) / (a_inv - 1.0) * std::log(g.tau_f);
    const double spacing = std::log(2.0) * (a_inv - c) / (1.0 - c);
    p.log_add.resize(size_t(max_level > 0 ? max_level : 0));
    for (int ell = 1; ell <= max_level; ++ell)
        p.log_add[size_t(ell - 1)] =
            log_f - double(max_level - ell + 1) * spacing;
    return p;
}

// Box length a pre-evolution run would need (it is reported, never used:
// relaxation runs on the main run's grid).
[[nodiscard]] inline double pre_evolution_L_tilde(double L_tilde_main,
                                                  double a_inv, double c,
                                                  double tau_i)
{
    return L_tilde_main * (1.0 / (a_inv - 1.0)) *
           std::pow(tau_i, (1.0 - c) / (a_inv - 1.0));
}

// Moore phase (c = a_inv): N = N2 N1 gamma e^(D/2) <=> D = 2 ln(N/(N2 N1 gamma))
[[nodiscard]] inline double moore_max_log_dynamic_range(int N, double N1,
                                                        double N2, double gamma)
{
    return 2.0 * std::log(double(N) / (N2 * N1 * gamma));
}

struct MooreBoxPlan
{
    double dx{};
    double L_tilde{};
    double D{};       // achievable log(H_switch / H_end)
    double tau_end{};
};

// dx from N2 at the switch (R m_r is constant through the fat phase), the
// box from N cells, and tau_end from H ~ tau^(-a_inv/b_inv).
[[nodiscard]] inline MooreBoxPlan
compute_moore_box_plan(int N, double N1, double N2, double gamma, double a_inv,
                       double b_inv, double R_switch, double m_r_switch,
                       double tau_switch)
{
    MooreBoxPlan p;
    p.dx = 1.0 / (N2 * R_switch * m_r_switch);
    p.L_tilde = double(N) * p.dx;
    p.D = moore_max_log_dynamic_range(N, N1, N2, gamma);
    p.tau_end = tau_switch * std::exp(p.D * b_inv / a_inv);
    return p;
}

} // namespace axs_host
