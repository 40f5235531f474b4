// AxionStringsOps.hpp -- the reference's operator surface for the hot path,
// as host C++ over libaxs_hip.so's C ABI.
//
// Names, argument meaning and results follow the reference so that code
// written against it ports by replacing `const amrex::MultiFab &state` with
// an `axs_level *` (one uniform periodic level owned by the HIP library):
//   AxionStringsRHS               AxionStringsRHS.hpp:28-54
//   PlaquetteCounts, count_plaquettes          StringFinder.hpp:27-122
//   StringTagger                  StringTagger.hpp:231-274
//   TotalEnergySums/Result, compute_total_energy_sums, compute_total_energy,
//   EnergyLevelInput, compute_composite_total_energy  EnergyKernel.hpp:42-339
//   VelocityResult, compute_velocity_at_pierced_corners VelocityKernel.hpp
//   fill_masked_a_dot_buffer      MaskedFieldBuffer.hpp:15-35
//   Spectrum, compute_spectrum    SpectrumKernel.hpp:42-153
//   xi_from_comoving_length, ...  XiFormula.hpp:32-57
//   FourierIC::generate, wrapped_freq  FourierIC.hpp:179-287
// Every call forwards to one axs_* entry point; errors become AxsError.
#pragma once

#include "AxionStringsParams.hpp"
#include "axs.h"

#include <stdexcept>
#include <string>
#include <vector>

namespace axs_host
{

struct AxsError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

inline void axs_check(int rc, const char *what)
{
    if (rc != AXS_OK)
        throw AxsError(std::string(what) + ": " + axs_last_error());
}

inline axs_masking to_c(const MaskingParams &m)
{
    return axs_masking{m.scheme == MaskingScheme::None ? AXS_MASK_NONE
                       : m.scheme == MaskingScheme::A  ? AXS_MASK_A
                                                       : AXS_MASK_B,
                       m.threshold};
}

// ---------------------------------------------------------------- RHS --
class AxionStringsRHS
{
  public:
    AxionStringsRHS(double a_dx, double a_curvature_coeff, double a_lambda,
                    double a_R_squared)
        : m_dx(a_dx), m_c{a_curvature_coeff, a_lambda, a_R_squared}
    {
    }
    // rhs_state <- RHS(state) for every cell of the level (the ParallelFor
    // of AxionStringsLevel.cpp:915-918).  rhs_state: device array of
    // 4 * cells words of the level's precision.
    void operator()(axs_level *state, void *rhs_state) const
    {
        axs_check(axs_rhs_eval(state, &m_c, rhs_state), "AxionStringsRHS");
    }

  private:
    double m_dx;
    axs_rhs_coeffs m_c;
};

// ----------------------------------------------------------- strings --
struct PlaquetteCounts
{
    long n_p_plain{0};
    long n_p_weighted{0};
};

[[nodiscard]] inline PlaquetteCounts
count_plaquettes(axs_level *state, const uint8_t *mask = nullptr)
{
    int64_t c[2];
    axs_check(axs_count_plaquettes(state, mask, c), "count_plaquettes");
    return {static_cast<long>(c[0]), static_cast<long>(c[1])};
}

class StringTagger
{
  public:
    StringTagger(double a_dx, StringTaggerParams a_params)
        : m_dx(a_dx), m_params(a_params)
    {
    }
    // Sets tags (device uint8 per cell) wherever a criterion holds; never
    // clears.  R is R(tau) for the amplitude extension.
    PlaquetteCounts operator()(axs_level *state, uint8_t *tags,
                               double R = 1.0) const
    {
        axs_tagger_params p{m_params.gradient_threshold,
                            m_params.radial_gradient_threshold,
                            m_params.amplitude_threshold, R};
        int64_t c[2];
        axs_check(axs_tag_cells(state, &p, tags, c), "StringTagger");
        return {static_cast<long>(c[0]), static_cast<long>(c[1])};
    }

  private:
    double m_dx;
    StringTaggerParams m_params;
};

// ------------------------------------------------------------ energy --
struct TotalEnergyResult
{
    double rho_tot_unscreened{0.0};
    double rho_tot_screened{0.0};
    double rho_axion_kinetic_unscreened{0.0};
    double rho_axion_kinetic_screened{0.0};
    double rho_axion_gradient_unscreened{0.0};
    double rho_axion_gradient_screened{0.0};
    double rho_radial_kinetic_unscreened{0.0};
    double rho_radial_kinetic_screened{0.0};
    double rho_radial_gradient_unscreened{0.0};
    double rho_radial_gradient_screened{0.0};
    double rho_radial_mass_unscreened{0.0};
    double rho_radial_mass_screened{0.0};
    double n_total{0.0};
    double n_unmasked{0.0};
};

struct TotalEnergySums
{
    double sum_rho{0.0};
    double sum_rho_w{0.0};
    double sum_rho_a_kin{0.0};
    double sum_rho_a_kin_w{0.0};
    double sum_rho_a_grad{0.0};
    double sum_rho_a_grad_w{0.0};
    double sum_rho_r_kin{0.0};
    double sum_rho_r_kin_w{0.0};
    double sum_rho_r_grad{0.0};
    double sum_rho_r_grad_w{0.0};
    double sum_rho_r_mass{0.0};
    double sum_rho_r_mass_w{0.0};
    double sum_w{0.0};
    double n_uncovered{0.0};
};

[[nodiscard]] inline TotalEnergySums
compute_total_energy_sums(axs_level *state, double /*dx: the level's own*/,
                          double R, double lambda, double R_prime_over_R,
                          const MaskingParams &screening,
                          const uint8_t *mask = nullptr)
{
    const axs_masking m = to_c(screening);
    double s[14];
    axs_check(axs_energy_sums(state, R, lambda, R_prime_over_R, &m, mask, s),
              "compute_total_energy_sums");
    TotalEnergySums o;
    double *f[14] = {&o.sum_rho,        &o.sum_rho_w,        &o.sum_rho_a_kin,
                     &o.sum_rho_a_kin_w, &o.sum_rho_a_grad,   &o.sum_rho_a_grad_w,
                     &o.sum_rho_r_kin,  &o.sum_rho_r_kin_w,  &o.sum_rho_r_grad,
                     &o.sum_rho_r_grad_w, &o.sum_rho_r_mass, &o.sum_rho_r_mass_w,
                     &o.sum_w,          &o.n_uncovered};
    for (int i = 0; i < 14; ++i)
        *f[i] = s[i];
    return o;
}

// compute_total_energy, EnergyKernel.hpp:239-268
[[nodiscard]] inline TotalEnergyResult
compute_total_energy(axs_level *state, double dx, double R, double lambda,
                     double R_prime_over_R, const MaskingParams &screening,
                     long n_cells_total)
{
    const TotalEnergySums s = compute_total_energy_sums(
        state, dx, R, lambda, R_prime_over_R, screening, nullptr);
    const double n = static_cast<double>(n_cells_total);
    TotalEnergyResult o{};
    o.n_total = n;
    o.n_unmasked = s.sum_w;
    o.rho_tot_unscreened = s.sum_rho / n;
    o.rho_axion_kinetic_unscreened = s.sum_rho_a_kin / n;
    o.rho_axion_gradient_unscreened = s.sum_rho_a_grad / n;
    o.rho_radial_kinetic_unscreened = s.sum_rho_r_kin / n;
    o.rho_radial_gradient_unscreened = s.sum_rho_r_grad / n;
    o.rho_radial_mass_unscreened = s.sum_rho_r_mass / n;
    if (s.sum_w > 0.0)
    {
        o.rho_tot_screened = s.sum_rho_w / s.sum_w;
        o.rho_axion_kinetic_screened = s.sum_rho_a_kin_w / s.sum_w;
        o.rho_axion_gradient_screened = s.sum_rho_a_grad_w / s.sum_w;
        o.rho_radial_kinetic_screened = s.sum_rho_r_kin_w / s.sum_w;
        o.rho_radial_gradient_screened = s.sum_rho_r_grad_w / s.sum_w;
        o.rho_radial_mass_screened = s.sum_rho_r_mass_w / s.sum_w;
    }
    return o;
}

struct EnergyLevelInput
{
    axs_level *state{nullptr};
    double dx{0.0};
    const uint8_t *mask{nullptr};
};

// compute_composite_total_energy, EnergyKernel.hpp:287-339: per-level raw
// sums weighted by that level's cell volume dx^3, normalised once.
[[nodiscard]] inline TotalEnergyResult
compute_composite_total_energy(const std::vector<EnergyLevelInput> &levels,
                               double R, double lambda, double R_prime_over_R,
                               const MaskingParams &screening)
{
    double acc[12] = {0.0};
    double n_total = 0.0, n_unmasked = 0.0;
    for (const EnergyLevelInput &l : levels)
    {
        const TotalEnergySums s = compute_total_energy_sums(
            l.state, l.dx, R, lambda, R_prime_over_R, screening, l.mask);
        const double vol = l.dx * l.dx * l.dx;
        const double v[12] = {s.sum_rho,        s.sum_rho_w,
                              s.sum_rho_a_kin,  s.sum_rho_a_kin_w,
                              s.sum_rho_a_grad, s.sum_rho_a_grad_w,
                              s.sum_rho_r_kin,  s.sum_rho_r_kin_w,
                              s.sum_rho_r_grad, s.sum_rho_r_grad_w,
                              s.sum_rho_r_mass, s.sum_rho_r_mass_w};
        for (int i = 0; i < 12; ++i)
            acc[i] += v[i] * vol;
        n_total += s.n_uncovered * vol;
        n_unmasked += s.sum_w * vol;
    }
    TotalEnergyResult o{};
    o.n_total = n_total;
    o.n_unmasked = n_unmasked;
    o.rho_tot_unscreened = acc[0] / n_total;
    o.rho_axion_kinetic_unscreened = acc[2] / n_total;
    o.rho_axion_gradient_unscreened = acc[4] / n_total;
    o.rho_radial_kinetic_unscreened = acc[6] / n_total;
    o.rho_radial_gradient_unscreened = acc[8] / n_total;
    o.rho_radial_mass_unscreened = acc[10] / n_total;
    if (n_unmasked > 0.0)
    {
        o.rho_tot_screened = acc[1] / n_unmasked;
        o.rho_axion_kinetic_screened = acc[3] / n_unmasked;
        o.rho_axion_gradient_screened = acc[5] / n_unmasked;
        o.rho_radial_kinetic_screened = acc[7] / n_unmasked;
        o.rho_radial_gradient_screened = acc[9] / n_unmasked;
        o.rho_radial_mass_screened = acc[11] / n_unmasked;
    }
    return o;
}

// -------------------------------------------------------- velocities --
struct VelocityResult
{
    double sum_gamma_sq_v_sq{0.0};
    long count{0};
};

[[nodiscard]] inline VelocityResult
compute_velocity_at_pierced_corners(axs_level *state, double R,
                                    double R_prime_over_R, double m_r,
                                    const uint8_t *mask = nullptr)
{
    VelocityResult v{};
    int64_t n = 0;
    axs_check(axs_velocity_at_pierced_corners(state, R, R_prime_over_R, m_r,
                                              mask, &v.sum_gamma_sq_v_sq, &n,
                                              nullptr),
              "compute_velocity_at_pierced_corners");
    v.count = static_cast<long>(n);
    return v;
}

// ---------------------------------------------------------- spectrum --
inline void fill_masked_a_dot_buffer(void *buffer, axs_level *state,
                                     const MaskingParams &params, double R)
{
    const axs_masking m = to_c(params);
    axs_check(axs_fill_field(state, AXS_FIELD_MASKED_A_DOT, &m, R, buffer),
              "fill_masked_a_dot_buffer");
}

struct Spectrum
{
    std::vector<double> p;
    std::vector<double> shell_average;
    std::vector<long> shell_count;
    double full_cube_energy{0.0};
    double inscribed_sphere_energy{0.0};
    double real_space_sum_sq{0.0}; // sum of squares of the transformed field
};

// compute_spectrum of the level's `kind` field (AXS_FIELD_*).  For the
// |Pi|^2 spectrum pass kind = -1: Pi1 and Pi2 are transformed and their
// shell sums added (counts are those of one field).
[[nodiscard]] inline Spectrum compute_spectrum(axs_level *state, int kind,
                                               const MaskingParams &masking,
                                               double R, int n_max)
{
    const int n_shells = n_max / 2 + 1;
    std::vector<double> sum(n_shells, 0.0);
    std::vector<int64_t> cnt(n_shells, 0), cnt2(n_shells, 0);
    double tot[2] = {0.0, 0.0};
    double ssq = 0.0, ssq2 = 0.0;
    const axs_masking m = to_c(masking);
    if (kind >= 0)
        axs_check(axs_level_spectrum(state, kind, &m, R, sum.data(), cnt.data(),
                                     tot, &ssq),
                  "compute_spectrum");
    else
    {
        axs_check(axs_level_spectrum(state, AXS_FIELD_PI1, &m, R, sum.data(),
                                     cnt.data(), tot, &ssq),
                  "compute_spectrum");
        axs_check(axs_level_spectrum(state, AXS_FIELD_PI2, &m, R, sum.data(),
                                     cnt2.data(), tot, &ssq2),
                  "compute_spectrum");
        ssq += ssq2;
    }
    Spectrum s{};
    s.full_cube_energy = tot[0];
    s.inscribed_sphere_energy = tot[1];
    s.real_space_sum_sq = ssq;
    for (int i = 0; i < n_shells; ++i)
        if (cnt[i] > 0)
        {
            s.p.push_back(i);
            s.shell_average.push_back(sum[i] / static_cast<double>(cnt[i]));
            s.shell_count.push_back(static_cast<long>(cnt[i]));
        }
    return s;
}

// ---------------------------------------------------------------- xi --
[[nodiscard]] inline double xi_from_comoving_length(double ell, double L_tilde,
                                                    double a_inv, double tau)
{
    const double am1 = a_inv - 1.0;
    return ell / (L_tilde * L_tilde * L_tilde) * (am1 * am1 * am1 / (a_inv * a_inv)) *
           (tau * tau);
}
[[nodiscard]] inline double xi_from_plaquette_count(double N_p, double dx,
                                                    double L_tilde, double a_inv,
                                                    double tau)
{
    return xi_from_comoving_length((2.0 / 3.0) * N_p * dx, L_tilde, a_inv, tau);
}
[[nodiscard]] inline double plaquette_count_for_target_xi(double xi, int N,
                                                          double L_tilde,
                                                          double a_inv,
                                                          double tau)
{
    const double am1 = a_inv - 1.0;
    return 3.0 * N * xi * L_tilde * L_tilde * a_inv * a_inv /
           (2.0 * am1 * am1 * am1 * tau * tau);
}

// --------------------------------------------------------- Fourier IC --
namespace FourierIC
{
inline int wrapped_freq(int idx, int n) { return (idx <= n / 2) ? idx : idx - n; }

inline void generate(axs_level *state, int comp, double k_max_cells,
                     double target_mean_square, std::uint64_t seed,
                     int comp_seed_offset)
{
    axs_check(axs_level_fourier_ic(state, comp, k_max_cells, target_mean_square,
                                   seed, comp_seed_offset),
              "FourierIC::generate");
}
} // namespace FourierIC

} // namespace axs_host
