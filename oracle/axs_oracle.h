/*
 * axs_oracle.h -- CPU restatement of the axion-string per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product (libaxs_hip.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links, loads
 * or falls back to it.
 *
 * Every function restates one function of the reference
 * (ehardy8/axion_strings_adaptive, directory AxionStrings/) formula for
 * formula, in FP64, and cites the file:line it follows.  Where the reference
 * delegates to the GRTeclyn / AMReX submodules (absent from the reference
 * checkout: 4th-order stencil, RK4 driver, FFT), the restated algorithm is
 * the published one and the function is marked "parity unpinned against the
 * reference"; tests/ pin those by closed forms and invariants instead.
 *
 * Field layout (all functions): structure-of-arrays, component-major,
 * x fastest:  f[((c*nz + z)*ny + y)*nx + x],  components {psi1, psi2, Pi1,
 * Pi2} (StateVariables.hpp:9-17).  Grids are periodic in all directions;
 * ghost cells are realised by index wrapping (the reference's
 * FillBoundary(periodicity), AxionStringsLevel.cpp:820,1255).
 */
#ifndef AXS_ORACLE_H_
#define AXS_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Background (Background.hpp:21-134) ---------------- */
typedef struct
{
    double a_inv;      /* Background.hpp:40 */
    double c0;         /* CTauSchedule::c0, Background.hpp:23 */
    double c1;         /* CTauSchedule::c1 */
    double tau_switch; /* CTauSchedule::tau_switch */
    int has_switch;    /* CTauSchedule::has_switch */
} or_background;

double or_bg_R(const or_background *b, double tau);
double or_bg_R_prime(const or_background *b, double tau);
double or_bg_curvature_term_coeff(const or_background *b, double tau);
double or_bg_lambda(const or_background *b, double tau);
double or_bg_H(const or_background *b, double tau);
double or_bg_H_over_mr_direct(const or_background *b, double tau);
double or_bg_H_over_mr_closed_form(const or_background *b, double tau);
double or_bg_tau_from_log_mr_over_h(const or_background *b, double l);

/* -------- plaquette winding (PlaquetteWinding.hpp:24-51) ------------ */
double or_wrap_angle_diff(double delta);
int or_plaquette_winding(double t1, double t2, double t3, double t4);

/* ---------------- xi (XiFormula.hpp:32-57) -------------------------- */
double or_xi_from_comoving_length(double ell, double L, double a_inv,
                                  double tau);
double or_xi_from_plaquette_count(double N_p, double dx, double L,
                                  double a_inv, double tau);
double or_plaquette_count_for_target_xi(double xi, int N, double L,
                                        double a_inv, double tau);

/* ------------- masking (Masking.hpp:39-118) ------------------------- */
enum { OR_MASK_NONE = 0, OR_MASK_A = 1, OR_MASK_B = 2 };
double or_masking_weight(int scheme, double thr, double psi1, double psi2,
                         double R);
double or_masked_a_dot(int scheme, double thr, double psi1, double psi2,
                       double Pi1, double Pi2, double R);

/* ------------- pointwise energies (Energy.hpp:64-189) --------------- */
double or_rho_tot_pointwise(double psi1, double psi2, double Pi1, double Pi2,
                            double g1sq, double g2sq, double R, double lambda,
                            double RpR);
double or_axion_kinetic_energy_pointwise(double theta_prime, double R);
double or_axion_gradient_energy_pointwise(double gtsq, double R);
double or_radial_kinetic_energy_pointwise(double psi1, double psi2, double Pi1,
                                          double Pi2, double R, double RpR);
double or_radial_gradient_energy_pointwise(double g1sq, double g2sq,
                                           double psi_sq, double gtsq,
                                           double R);
double or_radial_mass_energy_pointwise(double psi1, double psi2, double R,
                                       double lambda);

/* ---------------- RHS + RK4 (grid operations) ----------------------- */
void or_rhs(const double *state, double *rhs, int nx, int ny, int nz,
            double dx, double curvature_coeff, double lambda, double R_sq);
/* One classic RK4 step.  coeffs[s*3 + {0,1,2}] = {curvature, lambda, R^2}
 * at the stage times tau, tau+dt/2, tau+dt/2, tau+dt. */
void or_rk4_step(double *state, int nx, int ny, int nz, double dx,
                 const double coeffs[12], double dt);
/* Fills coeffs[12] from the background at tau (stage times above). */
void or_stage_coeffs(const or_background *b, double tau, double dt,
                     double coeffs[12]);

/* ------ pre-evolution background (PreEvolutionBackground.hpp:45-85) --- */
double or_pre_R(double gamma_pre, double tau_pre);
double or_pre_R_prime(double gamma_pre, double tau_pre);
double or_pre_curvature_term_coeff(double gamma_pre, double tau_pre);
double or_pre_lambda(double gamma_pre, double tau_pre);
double or_pre_m_r_over_H_direct(double gamma_pre, double tau_pre);
/* Stage scalars of the relaxing phase (AxionStringsLevel.cpp:869-876) at
 * the stage times tau_pre, +dt/2, +dt/2, +dt. */
void or_pre_stage_coeffs(double gamma_pre, double tau_pre, double dt,
                         double coeffs[12]);

/* -- pre-evolution -> main handoff (AxionStringsLevel.cpp:2027-2125) --- */
/* out = {kappa, common_term, R_pre, R_main} at the end of relaxation. */
void or_handoff_constants(const or_background *b, double gamma_pre,
                          double tau_pre_end, double tau_i, double out[4]);
/* psi <- kappa psi; Pi <- kappa^2 Pi + (psi / R_pre) common_term. */
void or_handoff_rescale(double *state, size_t ncell, double kappa,
                        double common_term, double R_pre);
/* Smoothstep blend of lambda and the curvature coefficient after the
 * handoff (AxionStringsLevel.cpp:889-907); *lambda, *curv hold the main
 * schedule's values on entry. */
void or_handoff_blend(double tau, double tau_start, double dtau,
                      double lambda_start, double curv_start, double *lambda,
                      double *curv);
/* Length of the blend window: n_periods main-schedule core periods
 * 2 pi / sqrt(lambda_main(tau_i)) (AxionStringsLevel.cpp:2070-2081). */
double or_handoff_dtau(double n_periods, double lambda_main_tau_i,
                       double lambda_start);

/* ----------- windings / counts (StringFinder.hpp:45-122) ------------ */
void or_windings_at(const double *state, int nx, int ny, int nz, int i, int j,
                    int k, int w[3]);
/* out[0] = n_p_plain, out[1] = n_p_weighted; mask may be NULL. */
void or_count_plaquettes(const double *state, int nx, int ny, int nz,
                         const uint8_t *mask, int64_t out[2]);
/* per-cell packed windings, for bit-exact comparisons: w[cell*3 + d]. */
void or_windings_field(const double *state, int nx, int ny, int nz,
                       int8_t *w);

/* ------- string velocities (Velocity.hpp:53-78, VelocityKernel.hpp) --- */
double or_gamma_sq_v_sq_from_phi(double phi1, double phi2, double pd1,
                                 double pd2, double m_r);
/* compute_velocity_at_pierced_corners (VelocityKernel.hpp:37-141):
 * *sum = sum of gamma^2 v^2 over the 4 corners of every pierced plaquette,
 * *count = number of corner evaluations.  mask may be NULL. */
void or_velocity(const double *state, int nx, int ny, int nz, double R,
                 double RpR, double m_r, const uint8_t *mask, double *sum,
                 int64_t *count);

/* ---------------- tagger (StringTagger.hpp:83-218) ------------------ */
/* Only sets tags (never clears), like the reference.  amp_thr <= 0 turns
 * the optional amplitude criterion |psi|/R < amp_thr off. */
void or_tag_cells(const double *state, int nx, int ny, int nz, double dx,
                  double gradient_threshold, double radial_gradient_threshold,
                  double amp_thr, double R, uint8_t *tags);

/* ---------------- energy sums (EnergyKernel.hpp:106-233) ------------ */
void or_energy_sums(const double *state, int nx, int ny, int nz, double dx,
                    double R, double lambda, double RpR, int scheme,
                    double thr, const uint8_t *mask, double out[14]);

/* ---------- masked buffer fill (MaskedFieldBuffer.hpp:15-35) -------- */
/* kind 0: masked a_dot (reference); kind 1: Pi1; kind 2: Pi2. */
void or_fill_field(const double *state, int nx, int ny, int nz, int kind,
                   int scheme, double thr, double R, double *buf);

/* ---------------- spectrum (SpectrumKernel.hpp:52-153) -------------- */
/* shell_sum / shell_count have max(nx,ny,nz)/2+1 entries; tot[0] =
 * full_cube_energy, tot[1] = inscribed_sphere_energy.  Accumulates (+=). */
int or_spectrum(const double *field, int nx, int ny, int nz,
                double *shell_sum, int64_t *shell_count, double tot[2]);

/* ---------------- Fourier IC (FourierIC.hpp:28-130) ----------------- */
uint64_t or_hash_u64(uint64_t x);
/* BoxPlan.hpp:16-196: box planning from the resolution targets N1, N2.
 * out = {tau_f, L_tilde, delta_x, delta_tau_leapfrog_bound}. */
void or_general_box_plan(int N, double N1, double N2, double a_inv, double c,
                         double out[4]);
/* out = {L_tilde, dx_finest, dx_base, tau_f}; *N_base (-1 if N_effective is
 * not a multiple of 2^max_level); log_add[max_level]. */
void or_amr_box_plan(int N_effective, double N1, double N2, double a_inv,
                     double c, int max_level, double out[4], int *N_base,
                     double *log_add);
double or_pre_evolution_L_tilde(double L_tilde_main, double a_inv, double c,
                                double tau_i);
double or_moore_max_log_dynamic_range(int N, double N1, double N2,
                                      double gamma);
/* out = {dx, L_tilde, D, tau_end}. */
void or_moore_box_plan(int N, double N1, double N2, double gamma, double a_inv,
                       double b_inv, double R_switch, double m_r_switch,
                       double tau_switch, double out[4]);

double or_uniform_from_hash(uint64_t h);
void or_gaussian_mode(uint64_t seed, int comp, int i, int j, int k,
                      double *re, double *im);
int or_wrapped_freq(int idx, int n);
int or_fourier_ic(double *field, int nx, int ny, int nz, double k_max_cells,
                  double target_mean_square, uint64_t seed,
                  int comp_seed_offset);

/* ---------------- small FFT used by the two above ------------------- */
/* In-place 3D complex DFT (sign -1 forward, +1 backward, unnormalised),
 * interleaved re/im, power-of-two sizes.  Returns 0 on success. */
int or_fft3d(double *data, int nx, int ny, int nz, int sign);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
