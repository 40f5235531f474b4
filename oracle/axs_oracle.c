/*
 * axs_oracle.c -- CPU restatement (FP64, OpenMP) of the reference's per-step
 * hot path.  TEST INFRASTRUCTURE ONLY: loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py as the checker; never by the product.
 *
 * Built with -ffp-contract=off so that every + - * / rounds exactly as
 * written: the FP64 HIP kernels are compiled the same way and follow the
 * same operation order, so RHS/RK4 results can be compared bit for bit.
 *
 * Pinning status (each function's header says which test pins it):
 *   Background, winding, xi, masking, pointwise energies: pinned by the
 *     reference's own doctest KATs (tests/test_*.cpp) restated in
 *     tests/test_oracle_golden.py, and against the reference headers
 *     compiled from source (oracle/_ref, oracle/ref_shim.cpp).
 *   diff1/diff2, RK4: GRTeclyn FourthOrderDerivatives / AMReX RK4 are
 *     absent from the reference checkout -> parity unpinned against the
 *     reference; pinned by closed forms (stencil symbol on plane waves,
 *     4th-order convergence, homogeneous exact solution).
 *   FFT-based spectrum / IC: pinned by Parseval, single-mode and brute-force
 *     DFT checks; the C2R of non-Hermitian input is projected explicitly.
 *   Grid-level energy sums (all 14 slots, every masking scheme) and the
 *     tagger's Laplacian and radial-gradient criteria: pinned by closed
 *     forms on fields whose discrete derivatives are known exactly (a
 *     rotating plane wave; a modulus varying along x with a phase along y),
 *     each with a check that a plausible slip (scaling, sqrt, dropped
 *     term, slot order) gives a different answer.
 */
#include "axs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ===================== Background.hpp:21-134 ======================= */
static double bg_b_inv(const or_background *b) { return b->a_inv - 1.0; }
static double bg_R0(const or_background *b) { return 1.0 / (b->a_inv - 1.0); }
static double bg_c(const or_background *b, double tau) /* :28-31 */
{
    return (b->has_switch && tau >= b->tau_switch) ? b->c1 : b->c0;
}

/* Background::R, Background.hpp:55-58 (tau0 = 1) */
double or_bg_R(const or_background *b, double tau)
{
    return bg_R0(b) * pow(tau / 1.0, 1.0 / bg_b_inv(b));
}

/* Background::R_prime, Background.hpp:60-64 */
double or_bg_R_prime(const or_background *b, double tau)
{
    const double b_inv = bg_b_inv(b);
    return (bg_R0(b) / (b_inv * 1.0)) * pow(tau / 1.0, 1.0 / b_inv - 1.0);
}

/* Background::curvature_term_coeff, Background.hpp:67-70 */
double or_bg_curvature_term_coeff(const or_background *b, double tau)
{
    const double b_inv = bg_b_inv(b);
    return (1.0 - b_inv) / (b_inv * b_inv * tau * tau);
}

/* Background::lambda, Background.hpp:75-86 (lambda0 = 1) */
double or_bg_lambda(const or_background *b, double tau)
{
    const double c_now = bg_c(b, tau);
    const double R0 = bg_R0(b);
    if (!b->has_switch || tau < b->tau_switch)
        return 1.0 * pow(or_bg_R(b, tau) / R0, -2.0 * c_now);
    const double tau_s = b->tau_switch;
    const double R_s = or_bg_R(b, tau_s);
    const double lambda_s = 1.0 * pow(R_s / R0, -2.0 * b->c0);
    return lambda_s * pow(or_bg_R(b, tau) / R_s, -2.0 * c_now);
}

/* Background::H, Background.hpp:104-107 */
double or_bg_H(const or_background *b, double tau)
{
    return or_bg_R_prime(b, tau) / (or_bg_R(b, tau) * or_bg_R(b, tau));
}

/* Background::H_over_mr_direct, Background.hpp:112-116 */
double or_bg_H_over_mr_direct(const or_background *b, double tau)
{
    const double m_r = sqrt(or_bg_lambda(b, tau));
    return or_bg_H(b, tau) / m_r;
}

/* Background::H_over_mr_closed_form, Background.hpp:92-96 */
double or_bg_H_over_mr_closed_form(const or_background *b, double tau)
{
    const double c_now = bg_c(b, tau);
    return pow(tau / 1.0, (c_now - b->a_inv) / (b->a_inv - 1.0));
}

/* Background::tau_from_log_mr_over_h, Background.hpp:129-133 */
double or_bg_tau_from_log_mr_over_h(const or_background *b, double l)
{
    return 1.0 * exp(l * (b->a_inv - 1.0) / (b->a_inv - b->c0));
}

/* Stage scalars of AxionStringsLevel::specific_eval_rhs (evolving phase,
 * AxionStringsLevel.cpp:877-885), at the classic RK4 stage times. */
void or_stage_coeffs(const or_background *b, double tau, double dt,
                     double coeffs[12])
{
    const double t[4] = {tau, tau + 0.5 * dt, tau + 0.5 * dt, tau + dt};
    for (int s = 0; s < 4; ++s)
    {
        const double R = or_bg_R(b, t[s]);
        coeffs[3 * s + 0] = or_bg_curvature_term_coeff(b, t[s]);
        coeffs[3 * s + 1] = or_bg_lambda(b, t[s]);
        coeffs[3 * s + 2] = R * R;
    }
}

/* ============ PreEvolutionBackground.hpp:45-85 ======================= */
/* R0 = alpha = 1: R = exp(tau_pre) = R', constant curvature coefficient 1,
 * lambda = gamma^2 / R^2. */
double or_pre_R(double gamma_pre, double tau_pre)
{
    (void)gamma_pre;
    return exp(tau_pre);
}

double or_pre_R_prime(double gamma_pre, double tau_pre)
{
    (void)gamma_pre;
    return exp(tau_pre);
}

double or_pre_curvature_term_coeff(double gamma_pre, double tau_pre)
{
    (void)gamma_pre;
    (void)tau_pre;
    return 1.0;
}

double or_pre_lambda(double gamma_pre, double tau_pre)
{
    const double r = or_pre_R(gamma_pre, tau_pre);
    return gamma_pre * gamma_pre / (r * r);
}

double or_pre_m_r_over_H_direct(double gamma_pre, double tau_pre)
{
    const double R = or_pre_R(gamma_pre, tau_pre);
    const double hubble = or_pre_R_prime(gamma_pre, tau_pre) / (R * R);
    return sqrt(or_pre_lambda(gamma_pre, tau_pre)) / hubble;
}

/* AxionStringsLevel.cpp:869-876: curvature, lambda and R^2 of the
 * pre-evolution background at the stage time. */
void or_pre_stage_coeffs(double gamma_pre, double tau_pre, double dt,
                         double coeffs[12])
{
    const double t[4] = {tau_pre, tau_pre + 0.5 * dt, tau_pre + 0.5 * dt,
                         tau_pre + dt};
    for (int s = 0; s < 4; ++s)
    {
        const double R = or_pre_R(gamma_pre, t[s]);
        coeffs[3 * s + 0] = or_pre_curvature_term_coeff(gamma_pre, t[s]);
        coeffs[3 * s + 1] = or_pre_lambda(gamma_pre, t[s]);
        coeffs[3 * s + 2] = R * R;
    }
}

/* ============ AxionStringsLevel.cpp:2027-2125 (handoff) ================ */
void or_handoff_constants(const or_background *b, double gamma_pre,
                          double tau_pre_end, double tau_i, double out[4])
{
    const double R_pre = or_pre_R(gamma_pre, tau_pre_end);
    const double R_pre_prime = or_pre_R_prime(gamma_pre, tau_pre_end);
    const double R_main = or_bg_R(b, tau_i);
    const double R_main_prime = or_bg_R_prime(b, tau_i);
    const double kappa = R_main / R_pre;
    out[0] = kappa;
    out[1] = R_main_prime - kappa * kappa * R_pre_prime;
    out[2] = R_pre;
    out[3] = R_main;
}

void or_handoff_rescale(double *state, size_t ncell, double kappa,
                        double common_term, double R_pre)
{
    double *p1 = state, *p2 = state + ncell, *q1 = state + 2 * ncell,
           *q2 = state + 3 * ncell;
    for (size_t i = 0; i < ncell; ++i)
    {
        const double psi1 = p1[i], psi2 = p2[i], Pi1 = q1[i], Pi2 = q2[i];
        p1[i] = kappa * psi1;
        p2[i] = kappa * psi2;
        q1[i] = kappa * kappa * Pi1 + (psi1 / R_pre) * common_term;
        q2[i] = kappa * kappa * Pi2 + (psi2 / R_pre) * common_term;
    }
}

void or_handoff_blend(double tau, double tau_start, double dtau,
                      double lambda_start, double curv_start, double *lambda,
                      double *curv)
{
    if (!(dtau > 0.0))
        return;
    const double t = (tau - tau_start) / dtau;
    if (t < 1.0)
    {
        const double tc = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        const double s = tc * tc * (3.0 - 2.0 * tc);
        *lambda = (1.0 - s) * lambda_start + s * *lambda;
        *curv = (1.0 - s) * curv_start + s * *curv;
    }
}

double or_handoff_dtau(double n_periods, double lambda_main_tau_i,
                       double lambda_start)
{
    const double pi = 3.14159265358979323846;
    const double m_r = lambda_main_tau_i > 0.0 ? sqrt(lambda_main_tau_i)
                                               : sqrt(lambda_start);
    return n_periods * 2.0 * pi / m_r;
}

/* ================= PlaquetteWinding.hpp:24-51 ======================= */
double or_wrap_angle_diff(double delta)
{
    const double pi = 3.14159265358979323846;
    const double two_pi = 2.0 * pi;
    delta = fmod(delta, two_pi);
    if (delta <= -pi)
        delta += two_pi;
    else if (delta > pi)
        delta -= two_pi;
    return delta;
}

int or_plaquette_winding(double t1, double t2, double t3, double t4)
{
    const double two_pi = 2.0 * 3.14159265358979323846;
    const double total = or_wrap_angle_diff(t2 - t1) +
                         or_wrap_angle_diff(t3 - t2) +
                         or_wrap_angle_diff(t4 - t3) +
                         or_wrap_angle_diff(t1 - t4);
    return (int)lround(total / two_pi);
}

/* ===================== BoxPlan.hpp:16-196 =========================== */
/* compute_general_box_plan, BoxPlan.hpp:36-50 */
void or_general_box_plan(int N, double N1, double N2, double a_inv, double c,
                         double out[4])
{
    /* tau_f = (N/(N1 N2))^((a_inv-1)/(a_inv-c)) tau0, L = (N1/R0) * same,
     * R0 = 1/(a_inv-1) (lambda0 = 1, tau0 = 1) */
    const double power = pow((double)N / (N1 * N2), (a_inv - 1.0) / (a_inv - c));
    const double R0 = 1.0 / (a_inv - 1.0);
    out[0] = power;
    out[1] = (N1 / R0) * power;
    out[2] = out[1] / N;
    out[3] = out[2] / 3.0;
}

/* compute_amr_box_plan, BoxPlan.hpp:96-125 */
void or_amr_box_plan(int N_effective, double N1, double N2, double a_inv,
                     double c, int max_level, double out[4], int *N_base,
                     double *log_add)
{
    double g[4];
    or_general_box_plan(N_effective, N1, N2, a_inv, c, g);
    const int ratio = 1 << max_level;
    out[0] = g[1];
    out[1] = g[2];
    out[2] = g[2] * (double)ratio;
    out[3] = g[0];
    *N_base = (N_effective % ratio == 0) ? N_effective / ratio : -1;
    /* level ell comes online max_level - ell + 1 level spacings of
     * ln 2 (a_inv - c)/(1 - c) in log(m_r/H) before tau_f */
    const double b_inv = a_inv - 1.0;
    const double at_f = (a_inv - c) / b_inv * log(g[0]);
    const double step = log(2.0) * (a_inv - c) / (1.0 - c);
    for (int ell = 1; ell <= max_level; ++ell)
        log_add[ell - 1] = at_f - (double)(max_level - ell + 1) * step;
}

/* pre_evolution_L_tilde, BoxPlan.hpp:130-137 */
double or_pre_evolution_L_tilde(double L_tilde_main, double a_inv, double c,
                                double tau_i)
{
    return L_tilde_main * (1.0 / (a_inv - 1.0)) *
           pow(tau_i / 1.0, (1.0 - c) / (a_inv - 1.0));
}

/* moore_max_log_dynamic_range, BoxPlan.hpp:144-149 */
double or_moore_max_log_dynamic_range(int N, double N1, double N2,
                                      double gamma)
{
    return 2.0 * log((double)N / (N2 * N1 * gamma));
}

/* compute_moore_box_plan, BoxPlan.hpp:183-194 */
void or_moore_box_plan(int N, double N1, double N2, double gamma, double a_inv,
                       double b_inv, double R_switch, double m_r_switch,
                       double tau_switch, double out[4])
{
    out[0] = 1.0 / (N2 * R_switch * m_r_switch);
    out[1] = N * out[0];
    out[2] = or_moore_max_log_dynamic_range(N, N1, N2, gamma);
    out[3] = tau_switch * exp(out[2] * b_inv / a_inv);
}

/* ===================== XiFormula.hpp:32-57 ========================== */
double or_xi_from_comoving_length(double ell, double L, double a_inv,
                                  double tau)
{
    const double am1 = a_inv - 1.0;
    return ell / (L * L * L) * (am1 * am1 * am1 / (a_inv * a_inv)) *
           (tau * tau);
}

double or_xi_from_plaquette_count(double N_p, double dx, double L,
                                  double a_inv, double tau)
{
    return or_xi_from_comoving_length((2.0 / 3.0) * N_p * dx, L, a_inv, tau);
}

double or_plaquette_count_for_target_xi(double xi, int N, double L,
                                        double a_inv, double tau)
{
    const double am1 = a_inv - 1.0;
    return 3.0 * N * xi * L * L * a_inv * a_inv /
           (2.0 * am1 * am1 * am1 * tau * tau);
}

/* ===================== Masking.hpp:75-118 =========================== */
double or_masking_weight(int scheme, double thr, double psi1, double psi2,
                         double R)
{
    if (scheme == OR_MASK_NONE)
        return 1.0;
    const double mod = sqrt(psi1 * psi1 + psi2 * psi2) / R;
    if (scheme == OR_MASK_B)
        return (mod >= thr) ? 1.0 : 0.0;
    const double f_a = 1.4142135623730951;
    const double r = mod - 1.0;
    const double f = 1.0 + r / f_a;
    return f * f;
}

double or_masked_a_dot(int scheme, double thr, double psi1, double psi2,
                       double Pi1, double Pi2, double R)
{
    const double raw = psi1 * Pi2 - Pi1 * psi2;
    if (scheme == OR_MASK_A)
        return raw;
    if (scheme == OR_MASK_B &&
        or_masking_weight(scheme, thr, psi1, psi2, R) == 0.0)
        return 0.0;
    const double psi_sq = psi1 * psi1 + psi2 * psi2;
    return raw / psi_sq;
}

/* ====================== Energy.hpp:200-325 ========================== */
double or_rho_tot_pointwise(double psi1, double psi2, double Pi1, double Pi2,
                            double g1sq, double g2sq, double R, double lambda,
                            double RpR)
{
    const double c1 = Pi1 - RpR * psi1;
    const double c2 = Pi2 - RpR * psi2;
    const double kinetic = c1 * c1 + c2 * c2;
    const double gradient = g1sq + g2sq;
    const double d = psi1 * psi1 + psi2 * psi2 - R * R;
    const double potential = 0.25 * lambda * d * d;
    const double R2 = R * R;
    return (kinetic + gradient + potential) / (R2 * R2);
}

double or_axion_kinetic_energy_pointwise(double theta_prime, double R)
{
    const double f_a = 1.4142135623730951;
    const double a_dot = f_a * theta_prime / R;
    return 0.5 * a_dot * a_dot;
}

double or_axion_gradient_energy_pointwise(double gtsq, double R)
{
    const double f_a = 1.4142135623730951;
    return 0.5 * f_a * f_a * gtsq / (R * R);
}

double or_radial_kinetic_energy_pointwise(double psi1, double psi2, double Pi1,
                                          double Pi2, double R, double RpR)
{
    const double c1 = Pi1 - RpR * psi1;
    const double c2 = Pi2 - RpR * psi2;
    const double psi_sq = psi1 * psi1 + psi2 * psi2;
    if (psi_sq <= 0.0)
        return 0.0;
    const double radial_dot = (psi1 * c1 + psi2 * c2) / sqrt(psi_sq);
    const double R2 = R * R;
    return (radial_dot * radial_dot) / (R2 * R2);
}

double or_radial_gradient_energy_pointwise(double g1sq, double g2sq,
                                           double psi_sq, double gtsq,
                                           double R)
{
    const double gm = g1sq + g2sq - psi_sq * gtsq;
    const double R2 = R * R;
    return gm / (R2 * R2);
}

double or_radial_mass_energy_pointwise(double psi1, double psi2, double R,
                                       double lambda)
{
    const double d = psi1 * psi1 + psi2 * psi2 - R * R;
    const double potential = 0.25 * lambda * d * d;
    const double R2 = R * R;
    return potential / (R2 * R2);
}

/* ============== grid helpers (periodic index wrapping) ============== */
static inline int wrap(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}
static inline size_t gidx(int c, int x, int y, int z, int nx, int ny, int nz)
{
    return (((size_t)c * nz + wrap(z, nz)) * ny + wrap(y, ny)) * nx +
           wrap(x, nx);
}

/* GRTeclyn FourthOrderDerivatives::diff2 (published form; the submodule is
 * absent from the reference checkout, AxionStringsRHS.impl.hpp:10-23 calls
 * it): (-f[-2] + 16 f[-1] - 30 f[0] + 16 f[+1] - f[+2]) / (12 dx^2), written
 * with the weights 1/12, 4/3, 5/2 and evaluated left to right.  Parity
 * unpinned against the reference. */
static inline double diff2(double fm2, double fm1, double f0, double fp1,
                           double fp2, double one_over_dx2)
{
    const double w_far = 8.33333333333333333333e-2;
    const double w_near = 1.33333333333333333333e+0;
    const double w_local = 2.50000000000000000000e+0;
    return (-w_far * fm2 + w_near * fm1 - w_local * f0 + w_near * fp1 -
            w_far * fp2) *
           one_over_dx2;
}

/* FourthOrderDerivatives::diff1 (published form): (f[-2] - 8 f[-1] + 8 f[+1]
 * - f[+2]) / (12 dx), weights 1/12 and 2/3.  Parity unpinned. */
static inline double diff1(double fm2, double fm1, double fp1, double fp2,
                           double one_over_dx)
{
    const double w_far = 8.33333333333333333333e-2;
    const double w_near = 6.66666666666666666667e-1;
    return (w_far * fm2 - w_near * fm1 + w_near * fp1 - w_far * fp2) *
           one_over_dx;
}

#define F(c, x, y, z) s[gidx((c), (x), (y), (z), nx, ny, nz)]

static double lap_at(const double *s, int c, int x, int y, int z, int nx,
                     int ny, int nz, double i2)
{
    double r = 0.0;
    r += diff2(F(c, x - 2, y, z), F(c, x - 1, y, z), F(c, x, y, z),
               F(c, x + 1, y, z), F(c, x + 2, y, z), i2);
    r += diff2(F(c, x, y - 2, z), F(c, x, y - 1, z), F(c, x, y, z),
               F(c, x, y + 1, z), F(c, x, y + 2, z), i2);
    r += diff2(F(c, x, y, z - 2), F(c, x, y, z - 1), F(c, x, y, z),
               F(c, x, y, z + 1), F(c, x, y, z + 2), i2);
    return r;
}

static void grad_at(const double *s, int c, int x, int y, int z, int nx,
                    int ny, int nz, double i1, double g[3])
{
    g[0] = diff1(F(c, x - 2, y, z), F(c, x - 1, y, z), F(c, x + 1, y, z),
                 F(c, x + 2, y, z), i1);
    g[1] = diff1(F(c, x, y - 2, z), F(c, x, y - 1, z), F(c, x, y + 1, z),
                 F(c, x, y + 2, z), i1);
    g[2] = diff1(F(c, x, y, z - 2), F(c, x, y, z - 1), F(c, x, y, z + 1),
                 F(c, x, y, z + 2), i1);
}

/* ========== AxionStringsRHS::operator() (impl.hpp:25-60) ============ */
void or_rhs(const double *s, double *rhs, int nx, int ny, int nz, double dx,
            double curvature_coeff, double lambda, double R_sq)
{
    const double i2 = 1.0 / (dx * dx);
    const size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
            {
                const size_t id = ((size_t)z * ny + y) * nx + x;
                const double lap1 = lap_at(s, 0, x, y, z, nx, ny, nz, i2);
                const double lap2 = lap_at(s, 1, x, y, z, nx, ny, nz, i2);
                const double psi1 = s[id];
                const double psi2 = s[n + id];
                const double d = psi1 * psi1 + psi2 * psi2 - R_sq;
                const double pc = 0.5 * lambda * d;
                rhs[id] = s[2 * n + id];
                rhs[n + id] = s[3 * n + id];
                rhs[2 * n + id] = lap1 + curvature_coeff * psi1 - pc * psi1;
                rhs[3 * n + id] = lap2 + curvature_coeff * psi2 - pc * psi2;
            }
}

/* Classic RK4 (the reference's integrator is GRTeclyn/AMReX's, absent from
 * the checkout; docs/conventions.md:239-241 "method of lines with
 * Runge-Kutta"): k1..k4 at tau, tau+dt/2, tau+dt/2, tau+dt,
 * y_{n+1} = y_n + dt/6 (((k1 + 2 k2) + 2 k3) + k4).  Parity unpinned
 * against the reference; pinned by 4th-order convergence and the
 * homogeneous exact solution in tests/. */
void or_rk4_step(double *state, int nx, int ny, int nz, double dx,
                 const double coeffs[12], double dt)
{
    const size_t n4 = (size_t)4 * nx * ny * nz;
    double *k = (double *)malloc(n4 * sizeof(double));
    double *acc = (double *)malloc(n4 * sizeof(double));
    double *ys = (double *)malloc(n4 * sizeof(double));
    const double h2 = 0.5 * dt;
    const double h6 = dt / 6.0;
    const double a[3] = {h2, h2, dt};

    or_rhs(state, k, nx, ny, nz, dx, coeffs[0], coeffs[1], coeffs[2]);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n4; ++i)
    {
        acc[i] = k[i];
        ys[i] = state[i] + a[0] * k[i];
    }
    for (int st = 1; st < 3; ++st)
    {
        or_rhs(ys, k, nx, ny, nz, dx, coeffs[3 * st], coeffs[3 * st + 1],
               coeffs[3 * st + 2]);
        const double as = a[st];
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < n4; ++i)
        {
            acc[i] = acc[i] + 2.0 * k[i];
            ys[i] = state[i] + as * k[i];
        }
    }
    or_rhs(ys, k, nx, ny, nz, dx, coeffs[9], coeffs[10], coeffs[11]);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n4; ++i)
        state[i] = state[i] + h6 * (acc[i] + k[i]);
    free(k);
    free(acc);
    free(ys);
}

/* ============ StringFinder.hpp:45-68 plaquette_windings_at ========== */
static inline double phase_at(const double *s, int x, int y, int z, int nx,
                              int ny, int nz)
{
    return atan2(F(1, x, y, z), F(0, x, y, z));
}

void or_windings_at(const double *s, int nx, int ny, int nz, int i, int j,
                    int k, int w[3])
{
    const double th_000 = phase_at(s, i, j, k, nx, ny, nz);
    const double th_100 = phase_at(s, i + 1, j, k, nx, ny, nz);
    const double th_110 = phase_at(s, i + 1, j + 1, k, nx, ny, nz);
    const double th_010 = phase_at(s, i, j + 1, k, nx, ny, nz);
    const double th_001 = phase_at(s, i, j, k + 1, nx, ny, nz);
    const double th_101 = phase_at(s, i + 1, j, k + 1, nx, ny, nz);
    const double th_011 = phase_at(s, i, j + 1, k + 1, nx, ny, nz);
    w[0] = or_plaquette_winding(th_000, th_100, th_110, th_010); /* xy */
    w[1] = or_plaquette_winding(th_000, th_010, th_011, th_001); /* yz */
    w[2] = or_plaquette_winding(th_000, th_001, th_101, th_100); /* zx */
}

/* count_plaquettes, StringFinder.hpp:76-122 */
void or_count_plaquettes(const double *s, int nx, int ny, int nz,
                         const uint8_t *mask, int64_t out[2])
{
    int64_t plain = 0, weighted = 0;
#pragma omp parallel for collapse(2) reduction(+ : plain, weighted)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
            {
                const size_t id = ((size_t)z * ny + y) * nx + x;
                if (mask && mask[id] == 0)
                    continue;
                int w[3];
                or_windings_at(s, nx, ny, nz, x, y, z, w);
                plain += (w[0] != 0) + (w[1] != 0) + (w[2] != 0);
                weighted += abs(w[0]) + abs(w[1]) + abs(w[2]);
            }
    out[0] = plain;
    out[1] = weighted;
}

void or_windings_field(const double *s, int nx, int ny, int nz, int8_t *wf)
{
#pragma omp parallel for collapse(2)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
            {
                const size_t id = ((size_t)z * ny + y) * nx + x;
                int w[3];
                or_windings_at(s, nx, ny, nz, x, y, z, w);
                wf[3 * id + 0] = (int8_t)w[0];
                wf[3 * id + 1] = (int8_t)w[1];
                wf[3 * id + 2] = (int8_t)w[2];
            }
}

/* ========== gamma_sq_v_sq_from_phi, Velocity.hpp:62-78 ============== */
double or_gamma_sq_v_sq_from_phi(double phi1, double phi2, double pd1,
                                 double pd2, double m_r)
{
    const double c1 = 0.41222;
    const double e1 = -0.025763;
    const double phi_sq = phi1 * phi1 + phi2 * phi2;
    const double phi_dot_sq = pd1 * pd1 + pd2 * pd2;
    const double re = phi1 * pd1 + phi2 * pd2;
    const double m_r_sq = m_r * m_r;
    const double c1_sq = c1 * c1;
    const double term1 = (phi_dot_sq / (2.0 * m_r_sq * c1_sq)) *
                         (1.0 - e1 * phi_sq / (c1 * c1_sq));
    const double term2 = (e1 / (4.0 * m_r_sq * c1_sq * c1_sq * c1)) * re * re;
    return term1 - term2;
}

/* compute_velocity_at_pierced_corners, VelocityKernel.hpp:37-141 */
static double gvv_at(const double *s, int x, int y, int z, int nx, int ny,
                     int nz, double R, double RpR, double m_r)
{
    const double psi1 = F(0, x, y, z), psi2 = F(1, x, y, z);
    const double Pi1 = F(2, x, y, z), Pi2 = F(3, x, y, z);
    const double phi1 = psi1 / R, phi2 = psi2 / R;
    const double pd1 = (Pi1 - RpR * psi1) / (R * R);
    const double pd2 = (Pi2 - RpR * psi2) / (R * R);
    return or_gamma_sq_v_sq_from_phi(phi1, phi2, pd1, pd2, m_r);
}

void or_velocity(const double *s, int nx, int ny, int nz, double R,
                 double RpR, double m_r, const uint8_t *mask, double *sum_out,
                 int64_t *count_out)
{
    double sum = 0.0;
    int64_t count = 0;
#pragma omp parallel for collapse(2) reduction(+ : sum, count)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
            {
                const size_t id = ((size_t)z * ny + y) * nx + x;
                if (mask && mask[id] == 0)
                    continue;
                int w[3];
                or_windings_at(s, nx, ny, nz, x, y, z, w);
#define G(a, b, c) gvv_at(s, x + (a), y + (b), z + (c), nx, ny, nz, R, RpR, m_r)
                if (w[0] != 0)
                {
                    sum += G(0, 0, 0) + G(1, 0, 0) + G(1, 1, 0) + G(0, 1, 0);
                    count += 4;
                }
                if (w[1] != 0)
                {
                    sum += G(0, 0, 0) + G(0, 1, 0) + G(0, 1, 1) + G(0, 0, 1);
                    count += 4;
                }
                if (w[2] != 0)
                {
                    sum += G(0, 0, 0) + G(0, 0, 1) + G(1, 0, 1) + G(1, 0, 0);
                    count += 4;
                }
#undef G
            }
    *sum_out = sum;
    *count_out = count;
}

/* ============ StringTagger::operator(), StringTagger.hpp:240-341 ===== */
void or_tag_cells(const double *s, int nx, int ny, int nz, double dx,
                  double gradient_threshold, double radial_gradient_threshold,
                  double amp_thr, double R, uint8_t *tags)
{
    const double i2 = 1.0 / (dx * dx);
    const double i1 = 1.0 / dx;
#pragma omp parallel for collapse(2)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
            {
                const size_t id = ((size_t)z * ny + y) * nx + x;
                int w[3];
                or_windings_at(s, nx, ny, nz, x, y, z, w);
                if (w[0] != 0 || w[1] != 0 || w[2] != 0)
                {
                    tags[id] = 1;
                    continue;
                }
                /* Extension (off unless amp_thr > 0): the amplitude
                 * criterion |psi|/R < amp_thr of the batched AMR workload. */
                const double psi1 = F(0, x, y, z);
                const double psi2 = F(1, x, y, z);
                if (amp_thr > 0.0 &&
                    sqrt(psi1 * psi1 + psi2 * psi2) / R < amp_thr)
                {
                    tags[id] = 1;
                    continue;
                }
                const double dx2 = dx * dx;
                const double lap1 = lap_at(s, 0, x, y, z, nx, ny, nz, i2);
                const double lap2 = lap_at(s, 1, x, y, z, nx, ny, nz, i2);
                if (dx2 * fabs(lap1) > gradient_threshold ||
                    dx2 * fabs(lap2) > gradient_threshold)
                {
                    tags[id] = 1;
                    continue;
                }
                const double psi_sq = psi1 * psi1 + psi2 * psi2;
                double g1[3], g2[3];
                grad_at(s, 0, x, y, z, nx, ny, nz, i1, g1);
                grad_at(s, 1, x, y, z, nx, ny, nz, i1, g2);
                double g1sq = 0.0, g2sq = 0.0, gtsq = 0.0;
                for (int d = 0; d < 3; ++d)
                {
                    g1sq += g1[d] * g1[d];
                    g2sq += g2[d] * g2[d];
                    const double dt =
                        (psi_sq > 0.0) ? (psi1 * g2[d] - psi2 * g1[d]) / psi_sq
                                       : 0.0;
                    gtsq += dt * dt;
                }
                const double rg = g1sq + g2sq - psi_sq * gtsq;
                if (dx * sqrt(rg) > radial_gradient_threshold)
                    tags[id] = 1;
            }
}

/* ====== compute_total_energy_sums, EnergyKernel.hpp:106-233 ========= */
void or_energy_sums(const double *s, int nx, int ny, int nz, double dx,
                    double R, double lambda, double RpR, int scheme,
                    double thr, const uint8_t *mask, double out[14])
{
    const double i1 = 1.0 / dx;
    const size_t n = (size_t)nx * ny * nz;
    double acc[14] = {0};
    const int nt = or_num_threads();
    double *part = (double *)calloc((size_t)nt * 14, sizeof(double));
#pragma omp parallel
    {
#ifdef _OPENMP
        double *p = part + 14 * (size_t)omp_get_thread_num();
#else
        double *p = part;
#endif
#pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x)
                {
                    const size_t id = ((size_t)z * ny + y) * nx + x;
                    if (mask && mask[id] == 0)
                        continue;
                    const double psi1 = s[id], psi2 = s[n + id];
                    const double Pi1 = s[2 * n + id], Pi2 = s[3 * n + id];
                    const double psi_sq = psi1 * psi1 + psi2 * psi2;
                    double g1[3], g2[3];
                    grad_at(s, 0, x, y, z, nx, ny, nz, i1, g1);
                    grad_at(s, 1, x, y, z, nx, ny, nz, i1, g2);
                    double g1sq = 0.0, g2sq = 0.0, gtsq = 0.0;
                    for (int d = 0; d < 3; ++d)
                    {
                        g1sq += g1[d] * g1[d];
                        g2sq += g2[d] * g2[d];
                        const double dth =
                            (psi_sq > 0.0)
                                ? (psi1 * g2[d] - psi2 * g1[d]) / psi_sq
                                : 0.0;
                        gtsq += dth * dth;
                    }
                    const double rho = or_rho_tot_pointwise(
                        psi1, psi2, Pi1, Pi2, g1sq, g2sq, R, lambda, RpR);
                    const double tp =
                        (psi_sq > 0.0) ? (psi1 * Pi2 - Pi1 * psi2) / psi_sq
                                       : 0.0;
                    const double rak = or_axion_kinetic_energy_pointwise(tp, R);
                    const double rag =
                        or_axion_gradient_energy_pointwise(gtsq, R);
                    const double rrk = or_radial_kinetic_energy_pointwise(
                        psi1, psi2, Pi1, Pi2, R, RpR);
                    const double rrg = or_radial_gradient_energy_pointwise(
                        g1sq, g2sq, psi_sq, gtsq, R);
                    const double rrm =
                        or_radial_mass_energy_pointwise(psi1, psi2, R, lambda);
                    const double w =
                        or_masking_weight(scheme, thr, psi1, psi2, R);
                    const double v[14] = {rho, rho * w, rak, rak * w,
                                          rag, rag * w, rrk, rrk * w,
                                          rrg, rrg * w, rrm, rrm * w,
                                          w,   1.0};
                    for (int q = 0; q < 14; ++q)
                        p[q] += v[q];
                }
    }
    for (int t = 0; t < nt; ++t)
        for (int q = 0; q < 14; ++q)
            acc[q] += part[14 * (size_t)t + q];
    free(part);
    memcpy(out, acc, sizeof(acc));
}

/* ========= fill_masked_a_dot_buffer, MaskedFieldBuffer.hpp:15-35 ===== */
void or_fill_field(const double *s, int nx, int ny, int nz, int kind,
                   int scheme, double thr, double R, double *buf)
{
    const size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (size_t id = 0; id < n; ++id)
    {
        if (kind == 0)
            buf[id] = or_masked_a_dot(scheme, thr, s[id], s[n + id],
                                      s[2 * n + id], s[3 * n + id], R);
        else
            buf[id] = s[(size_t)(kind + 1) * n + id];
    }
}

/* ========================= small FFT ================================ */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft1d(double *re, double *im, int n, int sign, const double *cs,
                  const double *sn)
{
    /* bit reversal */
    for (int i = 1, j = 0; i < n; ++i)
    {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j)
        {
            double t = re[i];
            re[i] = re[j];
            re[j] = t;
            t = im[i];
            im[i] = im[j];
            im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1)
    {
        const int half = len >> 1;
        const int step = n / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < half; ++k)
            {
                const double wr = cs[k * step];
                const double wi = sign * sn[k * step];
                const double ur = re[i + k], ui = im[i + k];
                const double xr = re[i + k + half], xi = im[i + k + half];
                const double vr = xr * wr - xi * wi;
                const double vi = xr * wi + xi * wr;
                re[i + k] = ur + vr;
                im[i + k] = ui + vi;
                re[i + k + half] = ur - vr;
                im[i + k + half] = ui - vi;
            }
    }
}

int or_fft3d(double *data, int nx, int ny, int nz, int sign)
{
    if (!is_pow2(nx) || !is_pow2(ny) || !is_pow2(nz))
        return 1;
    const int dims[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d)
    {
        const int n = dims[d];
        double *cs = (double *)malloc(sizeof(double) * n);
        double *sn = (double *)malloc(sizeof(double) * n);
        for (int k = 0; k < n; ++k)
        {
            cs[k] = cos(2.0 * OR_PI * k / n);
            sn[k] = -sin(2.0 * OR_PI * k / n); /* e^{-i 2 pi k/n} */
        }
        const size_t stride = d == 0 ? 1 : (d == 1 ? (size_t)nx
                                                   : (size_t)nx * ny);
        const size_t nlines = (size_t)nx * ny * nz / n;
#pragma omp parallel
        {
            double *re = (double *)malloc(sizeof(double) * n);
            double *im = (double *)malloc(sizeof(double) * n);
#pragma omp for schedule(static)
            for (size_t l = 0; l < nlines; ++l)
            {
                size_t base;
                if (d == 0)
                    base = l * nx;
                else if (d == 1)
                    base = (l / nx) * nx * ny + (l % nx);
                else
                    base = l;
                for (int i = 0; i < n; ++i)
                {
                    re[i] = data[2 * (base + i * stride)];
                    im[i] = data[2 * (base + i * stride) + 1];
                }
                /* cs/sn hold e^{-i...}; sign=-1 forward uses them as is */
                fft1d(re, im, n, -sign, cs, sn);
                for (int i = 0; i < n; ++i)
                {
                    data[2 * (base + i * stride)] = re[i];
                    data[2 * (base + i * stride) + 1] = im[i];
                }
            }
            free(re);
            free(im);
        }
        free(cs);
        free(sn);
    }
    return 0;
}

/* ============ compute_spectrum, SpectrumKernel.hpp:52-153 ============ */
int or_spectrum(const double *field, int nx, int ny, int nz,
                double *shell_sum, int64_t *shell_count, double tot[2])
{
    const size_t n = (size_t)nx * ny * nz;
    double *c = (double *)malloc(2 * n * sizeof(double));
    if (!c)
        return 2;
    for (size_t i = 0; i < n; ++i)
    {
        c[2 * i] = field[i];
        c[2 * i + 1] = 0.0;
    }
    if (or_fft3d(c, nx, ny, nz, -1))
    {
        free(c);
        return 1;
    }
    int n_max = nx;
    if (ny > n_max)
        n_max = ny;
    if (nz > n_max)
        n_max = nz;
    const int n_shells = n_max / 2 + 1;
    /* serial, like the reference (SpectrumKernel.hpp:75-85) */
    for (int k = 0; k < nz; ++k)
    {
        const int kz = or_wrapped_freq(k, nz);
        for (int j = 0; j < ny; ++j)
        {
            const int ky = or_wrapped_freq(j, ny);
            for (int i = 0; i <= nx / 2; ++i)
            {
                const int kx = i;
                const double p_mod_sq = (double)(kx * kx + ky * ky + kz * kz);
                const double p_mod = sqrt(p_mod_sq);
                const size_t id = ((size_t)k * ny + j) * nx + i;
                const double re = c[2 * id], im = c[2 * id + 1];
                const double mag_sq = re * re + im * im;
                const int weight = (kx == 0 || kx == nx / 2) ? 1 : 2;
                tot[0] += weight * mag_sq;
                if (p_mod < n_max / 2.0)
                {
                    tot[1] += weight * mag_sq;
                    const int shell = (int)floor(p_mod + 0.5);
                    if (shell < n_shells)
                    {
                        shell_sum[shell] +=
                            weight * 4.0 * OR_PI * p_mod_sq * mag_sq;
                        shell_count[shell] += weight;
                    }
                }
            }
        }
    }
    free(c);
    return 0;
}

/* ================== FourierIC.hpp:28-130 =========================== */
uint64_t or_hash_u64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    x = x ^ (x >> 31);
    return x;
}

double or_uniform_from_hash(uint64_t h)
{
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void or_gaussian_mode(uint64_t seed, int comp, int i, int j, int k,
                      double *re, double *im)
{
    const uint64_t key = seed ^ ((uint64_t)comp << 60) ^
                         ((uint64_t)(uint32_t)i << 40) ^
                         ((uint64_t)(uint32_t)j << 20) ^ (uint64_t)(uint32_t)k;
    const uint64_t h1 = or_hash_u64(key);
    const uint64_t h2 = or_hash_u64(key ^ 0x1234567890ABCDEFULL);
    double u1 = or_uniform_from_hash(h1);
    if (u1 < 1.0e-300)
        u1 = 1.0e-300;
    const double u2 = or_uniform_from_hash(h2);
    const double r = sqrt(-2.0 * log(u1));
    const double theta = 2.0 * 3.14159265358979323846 * u2;
    *re = r * cos(theta);
    *im = r * sin(theta);
}

int or_wrapped_freq(int idx, int n) { return (idx <= n / 2) ? idx : idx - n; }

/* FourierIC::generate (FourierIC.hpp:237-285).  The spectral half-cube is
 * filled from the hash exactly as the reference's fill lambda (:253-271);
 * the C2R step is stated explicitly: the self-conjugate planes kx = 0 and
 * kx = nx/2 are projected onto their Hermitian part, (X(k) + conj X(-k))/2,
 * which is what a real-output inverse transform of that half spectrum
 * computes, then the remaining half is completed by conjugate symmetry and
 * a full complex inverse DFT is taken.  Then rescaled in real space to the
 * target mean square (:275-284). */
int or_fourier_ic(double *field, int nx, int ny, int nz, double k_max_cells,
                  double target_mean_square, uint64_t seed,
                  int comp_seed_offset)
{
    const size_t n = (size_t)nx * ny * nz;
    double *c = (double *)calloc(2 * n, sizeof(double));
    if (!c)
        return 2;
    const uint64_t full_seed = seed ^ ((uint64_t)comp_seed_offset << 32);
    const double k_max_sq = k_max_cells * k_max_cells;
    const int hx = nx / 2;
#define CI(i, j, k) (((size_t)(k) * ny + (j)) * nx + (i))
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= hx; ++i)
            {
                const int kx = i, ky = or_wrapped_freq(j, ny),
                          kz = or_wrapped_freq(k, nz);
                const double k_sq = (double)(kx * kx + ky * ky + kz * kz);
                if (k_sq > k_max_sq || k_sq == 0.0)
                    continue;
                double re, im;
                or_gaussian_mode(full_seed, 0, i, j, k, &re, &im);
                c[2 * CI(i, j, k)] = re;
                c[2 * CI(i, j, k) + 1] = im;
            }
    /* Hermitian projection of the self-conjugate planes */
    const int planes[2] = {0, hx};
    for (int p = 0; p < (hx == 0 ? 1 : 2); ++p)
    {
        const int i = planes[p];
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
            {
                const int jm = (ny - j) % ny, km = (nz - k) % nz;
                const size_t a = CI(i, j, k), b = CI(i, jm, km);
                if (b < a)
                    continue;
                const double ar = c[2 * a], ai = c[2 * a + 1];
                const double br = c[2 * b], bi = c[2 * b + 1];
                const double nr = 0.5 * (ar + br), ni = 0.5 * (ai - bi);
                c[2 * a] = nr;
                c[2 * a + 1] = ni;
                c[2 * b] = nr;
                c[2 * b + 1] = -ni;
            }
    }
    /* conjugate completion of kx in (nx/2, nx) */
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = hx + 1; i < nx; ++i)
            {
                const size_t a = CI(i, j, k);
                const size_t b = CI(nx - i, (ny - j) % ny, (nz - k) % nz);
                c[2 * a] = c[2 * b];
                c[2 * a + 1] = -c[2 * b + 1];
            }
#undef CI
    if (or_fft3d(c, nx, ny, nz, +1))
    {
        free(c);
        return 1;
    }
    double sum_sq = 0.0;
    for (size_t i = 0; i < n; ++i)
    {
        field[i] = c[2 * i];
        sum_sq += field[i] * field[i];
    }
    free(c);
    const double current = sum_sq / (double)n;
    if (current > 0.0)
    {
        const double rescale = sqrt(target_mean_square / current);
        for (size_t i = 0; i < n; ++i)
            field[i] *= rescale;
    }
    return 0;
}
