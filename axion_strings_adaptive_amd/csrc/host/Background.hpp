// Background.hpp -- host-side FP64 FRW background of the axion-string
// simulator: the analytic R(tau), R'(tau), curvature coefficient and
// lambda(tau) the RHS kernels take as plain scalars.
//
// Same class name, members and formulas as the reference's
// AxionStrings/Background.hpp:21-134 (CTauSchedule :21-32, R :55-58,
// R_prime :60-64, curvature_term_coeff :67-70, lambda :75-86,
// H :104-107, H_over_mr_direct :112-116, H_over_mr_closed_form :92-96,
// tau_from_log_mr_over_h :129-133), FlatBackground.hpp:26-44 and
// PreEvolutionBackground.hpp:45-85, so that a caller of the reference finds
// the same interface.  Evaluated on the host
// in FP64 only (never in a kernel).
#pragma once

#include <cmath>

struct CTauSchedule
{
    double c0{0.0};
    double c1{0.0};
    double tau_switch{0.0};
    bool has_switch{false};

    [[nodiscard]] double c(double tau) const
    {
        return (has_switch && tau >= tau_switch) ? c1 : c0;
    }
};

class Background
{
  public:
    static constexpr double lambda0 = 1.0;
    static constexpr double tau0 = 1.0;

    double a_inv{2.0};
    double b_inv{1.0};
    double R0{1.0};
    CTauSchedule c_sched;

    Background() = default;
    Background(double a_a_inv, CTauSchedule a_c_sched)
        : a_inv(a_a_inv), b_inv(a_a_inv - 1.0), R0(1.0 / (a_a_inv - 1.0)),
          c_sched(a_c_sched)
    {
    }

    [[nodiscard]] double R(double tau) const
    {
        return R0 * std::pow(tau / tau0, 1.0 / b_inv);
    }
    [[nodiscard]] double R_prime(double tau) const
    {
        return (R0 / (b_inv * tau0)) * std::pow(tau / tau0, 1.0 / b_inv - 1.0);
    }
    [[nodiscard]] double curvature_term_coeff(double tau) const
    {
        return (1.0 - b_inv) / (b_inv * b_inv * tau * tau);
    }
    [[nodiscard]] double lambda(double tau) const
    {
        const double c_now = c_sched.c(tau);
        if (!c_sched.has_switch || tau < c_sched.tau_switch)
            return lambda0 * std::pow(R(tau) / R0, -2.0 * c_now);
        const double tau_s = c_sched.tau_switch;
        const double R_s = R(tau_s);
        const double lambda_s = lambda0 * std::pow(R_s / R0, -2.0 * c_sched.c0);
        return lambda_s * std::pow(R(tau) / R_s, -2.0 * c_now);
    }
    [[nodiscard]] double H_over_mr_closed_form(double tau) const
    {
        const double c_now = c_sched.c(tau);
        return std::pow(tau / tau0, (c_now - a_inv) / (a_inv - 1.0));
    }
    [[nodiscard]] double H(double tau) const
    {
        return R_prime(tau) / (R(tau) * R(tau));
    }
    [[nodiscard]] double H_over_mr_direct(double tau) const
    {
        const double m_r = std::sqrt(lambda(tau));
        return H(tau) / m_r;
    }
    [[nodiscard]] double tau_from_log_mr_over_h(double log_mr_over_h) const
    {
        const double c0 = c_sched.c0;
        return tau0 * std::exp(log_mr_over_h * (a_inv - 1.0) / (a_inv - c0));
    }
};

class FlatBackground
{
  public:
    double m_r{1.0};
    FlatBackground() = default;
    explicit FlatBackground(double a_m_r) : m_r(a_m_r) {}
    [[nodiscard]] double R(double) const { return 1.0; }
    [[nodiscard]] double R_prime(double) const { return 0.0; }
    [[nodiscard]] double curvature_term_coeff(double) const { return 0.0; }
    [[nodiscard]] double lambda(double) const { return m_r * m_r; }
};

// The relaxation background: in the pre-evolution conformal time tau_pre,
// R = exp(tau_pre) (R0 = alpha = 1), a constant curvature coefficient 1 and
// lambda = gamma_pre^2 / R^2, so that m_r/H = gamma_pre and R m_r are both
// constant (PreEvolutionBackground.hpp:45-85).
class PreEvolutionBackground
{
  public:
    double gamma_pre{1.0};
    PreEvolutionBackground() = default;
    explicit PreEvolutionBackground(double a_gamma_pre) : gamma_pre(a_gamma_pre)
    {
    }
    [[nodiscard]] double R(double tau_pre) const { return std::exp(tau_pre); }
    [[nodiscard]] double R_prime(double tau_pre) const
    {
        return std::exp(tau_pre);
    }
    [[nodiscard]] double curvature_term_coeff(double) const { return 1.0; }
    [[nodiscard]] double lambda(double tau_pre) const
    {
        const double r = R(tau_pre);
        return gamma_pre * gamma_pre / (r * r);
    }
    [[nodiscard]] double m_r_over_H_direct(double tau_pre) const
    {
        const double hubble = R_prime(tau_pre) / (R(tau_pre) * R(tau_pre));
        return std::sqrt(lambda(tau_pre)) / hubble;
    }
};
