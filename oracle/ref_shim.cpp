// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Exposes the reference's own AMReX-free headers, compiled from source where
// they lie (AxionStrings/ of the reference checkout, passed with -I by
// oracle/Makefile), behind a C ABI, so tests/ can check the CPU oracle
// (axs_oracle.c) and the product's host Background against the reference
// code itself.  Only headers that compile with nothing but <cmath> are used
// (Background.hpp, BoxPlan.hpp, PlaquetteWinding.hpp, Masking.hpp,
// Energy.hpp, XiFormula.hpp, Velocity.hpp, PreEvolutionBackground.hpp);
// everything else in the reference needs the absent
// GRTeclyn/AMReX submodules and is not built.  The output library goes to
// oracle/_ref/ and is never shipped in git.
#include "Background.hpp"
#include "BoxPlan.hpp"
#include "Energy.hpp"
#include "Masking.hpp"
#include "PlaquetteWinding.hpp"
#include "PreEvolutionBackground.hpp"
#include "Velocity.hpp"
#include "XiFormula.hpp"

namespace
{
Background make_bg(double a_inv, double c0, double c1, double tau_switch,
                   int has_switch)
{
    CTauSchedule s{};
    s.c0 = c0;
    s.c1 = c1;
    s.tau_switch = tau_switch;
    s.has_switch = has_switch != 0;
    return Background(a_inv, s);
}
} // namespace

extern "C"
{
    void ref_general_box_plan(int N, double N1, double N2, double a_inv,
                              double c, double out[4])
    {
        const GeneralBoxPlan p = compute_general_box_plan(N, N1, N2, a_inv, c);
        out[0] = p.tau_f;
        out[1] = p.L_tilde;
        out[2] = p.delta_x;
        out[3] = p.delta_tau_leapfrog_bound;
    }

    void ref_amr_box_plan(int N_effective, double N1, double N2, double a_inv,
                          double c, int max_level, double out[4], int *N_base,
                          double *log_add)
    {
        const AmrBoxPlan p =
            compute_amr_box_plan(N_effective, N1, N2, a_inv, c, max_level);
        out[0] = p.L_tilde;
        out[1] = p.dx_finest;
        out[2] = p.dx_base;
        out[3] = p.tau_f;
        *N_base = p.N_base;
        for (int i = 0; i < max_level; ++i)
            log_add[i] = p.log_add[(size_t)i];
    }

    double ref_pre_evolution_L_tilde(double L, double a_inv, double c,
                                     double tau_i)
    {
        return pre_evolution_L_tilde(L, a_inv, c, tau_i);
    }

    double ref_moore_max_log_dynamic_range(int N, double N1, double N2,
                                           double gamma)
    {
        return moore_max_log_dynamic_range(N, N1, N2, gamma);
    }

    void ref_moore_box_plan(int N, double N1, double N2, double gamma,
                            double a_inv, double b_inv, double R_switch,
                            double m_r_switch, double tau_switch,
                            double out[4])
    {
        const MooreBoxPlan p =
            compute_moore_box_plan(N, N1, N2, gamma, a_inv, b_inv, R_switch,
                                   m_r_switch, tau_switch);
        out[0] = p.dx;
        out[1] = p.L_tilde;
        out[2] = p.D;
        out[3] = p.tau_end;
    }

    // which: 0 R, 1 R', 2 curvature, 3 lambda, 4 m_r/H direct
    double ref_pre_background(double gamma_pre, int which, double tau_pre)
    {
        const PreEvolutionBackground b(gamma_pre);
        switch (which)
        {
        case 0:
            return b.R(tau_pre);
        case 1:
            return b.R_prime(tau_pre);
        case 2:
            return b.curvature_term_coeff(tau_pre);
        case 3:
            return b.lambda(tau_pre);
        default:
            return b.m_r_over_H_direct(tau_pre);
        }
    }

    // which: 0 R, 1 R', 2 curvature, 3 lambda, 4 H, 5 H/m_r direct,
    //        6 H/m_r closed form, 7 tau_from_log_mr_over_h (tau = log value)
    double ref_background(double a_inv, double c0, double c1,
                          double tau_switch, int has_switch, int which,
                          double tau)
    {
        const Background b = make_bg(a_inv, c0, c1, tau_switch, has_switch);
        switch (which)
        {
        case 0:
            return b.R(tau);
        case 1:
            return b.R_prime(tau);
        case 2:
            return b.curvature_term_coeff(tau);
        case 3:
            return b.lambda(tau);
        case 4:
            return b.H(tau);
        case 5:
            return b.H_over_mr_direct(tau);
        case 6:
            return b.H_over_mr_closed_form(tau);
        case 7:
            return b.tau_from_log_mr_over_h(tau);
        default:
            return 0.0;
        }
    }

    int ref_plaquette_winding(double t1, double t2, double t3, double t4)
    {
        return plaquette_winding(t1, t2, t3, t4);
    }

    double ref_wrap_angle_diff(double d) { return wrap_angle_diff(d); }

    double ref_masking_weight(int scheme, double thr, double psi1, double psi2,
                              double R)
    {
        MaskingParams p{};
        p.scheme = scheme == 0   ? MaskingScheme::None
                   : scheme == 1 ? MaskingScheme::A
                                 : MaskingScheme::B;
        p.threshold = thr;
        return masking_weight(p, psi1, psi2, R);
    }

    double ref_masked_a_dot(int scheme, double thr, double psi1, double psi2,
                            double Pi1, double Pi2, double R)
    {
        MaskingParams p{};
        p.scheme = scheme == 0   ? MaskingScheme::None
                   : scheme == 1 ? MaskingScheme::A
                                 : MaskingScheme::B;
        p.threshold = thr;
        return masked_a_dot(p, psi1, psi2, Pi1, Pi2, R);
    }

    double ref_rho_tot_pointwise(double psi1, double psi2, double Pi1,
                                 double Pi2, double g1, double g2, double R,
                                 double lambda, double RpR)
    {
        return rho_tot_pointwise(psi1, psi2, Pi1, Pi2, g1, g2, R, lambda, RpR);
    }
    double ref_axion_kinetic(double tp, double R)
    {
        return axion_kinetic_energy_pointwise(tp, R);
    }
    double ref_axion_gradient(double g, double R)
    {
        return axion_gradient_energy_pointwise(g, R);
    }
    double ref_radial_kinetic(double psi1, double psi2, double Pi1, double Pi2,
                              double R, double RpR)
    {
        return radial_kinetic_energy_pointwise(psi1, psi2, Pi1, Pi2, R, RpR);
    }
    double ref_radial_gradient(double g1, double g2, double psq, double gt,
                               double R)
    {
        return radial_gradient_energy_pointwise(g1, g2, psq, gt, R);
    }
    double ref_radial_mass(double psi1, double psi2, double R, double lambda)
    {
        return radial_mass_energy_pointwise(psi1, psi2, R, lambda);
    }

    double ref_gamma_sq_v_sq_from_phi(double phi1, double phi2, double pd1,
                                      double pd2, double m_r)
    {
        return gamma_sq_v_sq_from_phi(phi1, phi2, pd1, pd2, m_r);
    }

    double ref_xi_from_comoving_length(double ell, double L, double a_inv,
                                       double tau)
    {
        return xi_from_comoving_length(ell, L, a_inv, tau);
    }
    double ref_xi_from_plaquette_count(double N_p, double dx, double L,
                                       double a_inv, double tau)
    {
        return xi_from_plaquette_count(N_p, dx, L, a_inv, tau);
    }
    double ref_plaquette_count_for_target_xi(double xi, int N, double L,
                                             double a_inv, double tau)
    {
        return plaquette_count_for_target_xi(xi, N, L, a_inv, tau);
    }
}
