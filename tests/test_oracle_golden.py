"""Pins the CPU oracle (oracle/axs_oracle.c) before anything is checked
against it: the reference's own known-answer tests (AxionStrings/tests/*.cpp)
restated, the reference headers compiled from source (oracle/_ref), published
SplitMix64 vectors, and closed forms / brute force for what the reference's
absent engine (GRTeclyn/AMReX) would have defined."""

import math

import numpy as np
import pytest

PI = math.pi


# ------------------------------------------------------------ winding KATs
# tests/test_plaquette_winding.cpp:12-53
WINDING_KATS = [
    ((0.0, PI / 2, PI, 3 * PI / 2), 1),
    ((0.0, -PI / 2, -PI, -3 * PI / 2), -1),
    ((0.1, 0.2, 0.15, 0.05), 0),
    ((0.0, 0.0, 0.0, 0.0), 0),
    ((1.234, 1.234 + PI / 2, 1.234 + PI, 1.234 + 3 * PI / 2), 1),
    ((3.0, -3.0, 3.0, -3.0), 0),
    ((0.0, PI, 2 * PI, 3 * PI), 2),
    ((PI / 2, PI, 3 * PI / 2, 2 * PI + 0.0), 1),
]


@pytest.mark.parametrize("thetas,w", WINDING_KATS)
def test_winding_kats(oracle, thetas, w):
    assert oracle.load().or_plaquette_winding(*thetas) == w


def test_winding_matches_reference_header(oracle, ref):
    rng = np.random.default_rng(1)
    L = oracle.load()
    for _ in range(20000):
        t = rng.uniform(-3.3, 3.3, 4)
        assert L.or_plaquette_winding(*t) == ref.ref_plaquette_winding(*t)
        d = rng.uniform(-7, 7)
        assert L.or_wrap_angle_diff(d) == ref.ref_wrap_angle_diff(d)
    # exact branch-cut values
    for d in (PI, -PI, 2 * PI, -2 * PI, 0.0, -0.0, 3 * PI):
        assert L.or_wrap_angle_diff(d) == ref.ref_wrap_angle_diff(d)


# ------------------------------------------------------------ masking KATs
# tests/test_masking.cpp:21-110
def test_masking_scheme_b_kat(oracle):
    L = oracle.load()
    R = 2.0
    pts = [(0.1 * R, 0.0, 0.3, -0.2), (0.7999 * R, 0.0, 0.3, -0.2),
           (0.8 * R, 0.0, 0.3, -0.2), (0.8001 * R, 0.0, 0.3, -0.2),
           (0.95 * R, 0.2 * R, -0.4, 0.5), (1.0 * R, 0.0, 0.1, 0.1)]
    for psi1, psi2, Pi1, Pi2 in pts:
        mod = math.sqrt(psi1 ** 2 + psi2 ** 2) / R
        raw = psi1 * Pi2 - Pi1 * psi2
        unmasked = raw / (psi1 ** 2 + psi2 ** 2)
        masked = L.or_masked_a_dot(2, 0.8, psi1, psi2, Pi1, Pi2, R)
        if mod < 0.8:
            assert masked == 0.0
            assert L.or_masking_weight(2, 0.8, psi1, psi2, R) == 0.0
        else:
            assert masked == pytest.approx(unmasked, rel=1e-12)
            assert L.or_masking_weight(2, 0.8, psi1, psi2, R) == 1.0


def test_masking_none_and_a_kat(oracle):
    L = oracle.load()
    R = 1.5
    psi1, psi2, Pi1, Pi2 = 0.05 * R, 0.0, 0.2, -0.3
    raw = psi1 * Pi2 - Pi1 * psi2
    assert L.or_masked_a_dot(0, 0.8, psi1, psi2, Pi1, Pi2, R) == \
        pytest.approx(raw / (psi1 ** 2 + psi2 ** 2), rel=1e-12)
    assert L.or_masking_weight(0, 0.8, psi1, psi2, R) == 1.0
    R = 1.0
    psi1 = 0.05
    raw = psi1 * Pi2 - Pi1 * psi2
    assert L.or_masked_a_dot(1, 0.8, psi1, psi2, Pi1, Pi2, R) == \
        pytest.approx(raw, rel=1e-12)
    f = 1.0 + (psi1 / R - 1.0) / 1.4142135623730951
    assert L.or_masking_weight(1, 0.8, psi1, psi2, R) == \
        pytest.approx(f * f, rel=1e-12)


def test_masking_matches_reference_header(oracle, ref):
    L = oracle.load()
    rng = np.random.default_rng(2)
    for _ in range(5000):
        psi1, psi2, Pi1, Pi2 = rng.normal(size=4)
        R = rng.uniform(0.5, 3)
        thr = rng.uniform(0.05, 0.95)
        for sch in (0, 1, 2):
            assert L.or_masking_weight(sch, thr, psi1, psi2, R) == \
                ref.ref_masking_weight(sch, thr, psi1, psi2, R)
            assert L.or_masked_a_dot(sch, thr, psi1, psi2, Pi1, Pi2, R) == \
                ref.ref_masked_a_dot(sch, thr, psi1, psi2, Pi1, Pi2, R)


# ------------------------------------------------------------- energy KATs
# tests/test_energy.cpp:15-168
def test_energy_vacuum_is_zero(oracle):
    L = oracle.load()
    b = oracle.bg(2.0)
    for tau in (0.5, 1.0, 3.3, 10.0):
        R = L.or_bg_R(b, tau)
        Rp = L.or_bg_R_prime(b, tau)
        lam = L.or_bg_lambda(b, tau)
        rho = L.or_rho_tot_pointwise(R, 0.0, Rp, 0.0, 0.0, 0.0, R, lam,
                                     1.0 / (1.0 * tau))
        assert abs(rho) <= 1e-10


def test_energy_axion_terms(oracle):
    L = oracle.load()
    assert L.or_axion_kinetic_energy_pointwise(0.0, 2.0) == 0.0
    assert L.or_axion_gradient_energy_pointwise(0.0, 2.0) == 0.0
    e1 = L.or_axion_gradient_energy_pointwise(1.3, 1.0)
    e2 = L.or_axion_gradient_energy_pointwise(1.3, 2.0)
    assert e1 / e2 == pytest.approx(4.0, rel=1e-12)


CASES = [(0.6, -0.3, 1.1, 0.4), (2.5, 1.8, -0.7, 2.2), (-1.2, 0.9, 0.05, -1.6)]


@pytest.mark.parametrize("psi1,psi2,Pi1,Pi2", CASES)
def test_energy_radial_plus_tangential_kinetic(oracle, psi1, psi2, Pi1, Pi2):
    L = oracle.load()
    b_inv, tau, R = 1.7, 2.3, 1.9
    c1 = Pi1 - psi1 / (b_inv * tau)
    c2 = Pi2 - psi2 / (b_inv * tau)
    psq = psi1 ** 2 + psi2 ** 2
    full = (c1 * c1 + c2 * c2) / R ** 4
    tp = (psi1 * Pi2 - Pi1 * psi2) / psq
    tang = tp * tp * psq / R ** 4
    rad = L.or_radial_kinetic_energy_pointwise(psi1, psi2, Pi1, Pi2, R,
                                               1.0 / (b_inv * tau))
    assert rad + tang == pytest.approx(full, rel=1e-10)


@pytest.mark.parametrize("psi1,psi2,g1,g2", CASES)
def test_energy_radial_plus_tangential_gradient(oracle, psi1, psi2, g1, g2):
    L = oracle.load()
    R = 1.9
    psq = psi1 ** 2 + psi2 ** 2
    gt = (psi1 * g2 - psi2 * g1) / psq
    full = (g1 * g1 + g2 * g2) / R ** 4
    tang = gt * gt * psq / R ** 4
    rad = L.or_radial_gradient_energy_pointwise(g1 * g1, g2 * g2, psq, gt * gt,
                                                R)
    assert rad + tang == pytest.approx(full, rel=1e-10)


def test_energy_guards_and_mass(oracle):
    L = oracle.load()
    assert L.or_radial_kinetic_energy_pointwise(0.0, 0.0, 1.0, -0.4, 1.5,
                                                1.0 / (1.7 * 2.3)) == 0.0
    assert L.or_radial_gradient_energy_pointwise(0.7, 1.3, 0.0, 0.0, 1.5) == \
        pytest.approx(2.0 / 1.5 ** 4)
    assert L.or_radial_mass_energy_pointwise(1.5, 0.0, 1.5, 0.8) == 0.0
    d = 2.1 ** 2 + 0.4 ** 2 - 1.5 ** 2
    assert L.or_radial_mass_energy_pointwise(2.1, -0.4, 1.5, 0.8) == \
        pytest.approx(0.25 * 0.8 * d * d / 1.5 ** 4)


def test_energy_matches_reference_header(oracle, ref):
    L = oracle.load()
    rng = np.random.default_rng(3)
    for _ in range(3000):
        v = rng.normal(size=9)
        v[6] = abs(v[6]) + 0.1
        assert L.or_rho_tot_pointwise(*v) == ref.ref_rho_tot_pointwise(*v)
        psi1, psi2, Pi1, Pi2, R, RpR = v[0], v[1], v[2], v[3], v[6], v[8]
        assert L.or_radial_kinetic_energy_pointwise(psi1, psi2, Pi1, Pi2, R,
                                                    RpR) == \
            ref.ref_radial_kinetic(psi1, psi2, Pi1, Pi2, R, RpR)
        assert L.or_radial_mass_energy_pointwise(psi1, psi2, R, v[7]) == \
            ref.ref_radial_mass(psi1, psi2, R, v[7])
        a, g = abs(v[4]), abs(v[5])
        assert L.or_axion_kinetic_energy_pointwise(v[4], R) == \
            ref.ref_axion_kinetic(v[4], R)
        assert L.or_axion_gradient_energy_pointwise(a, R) == \
            ref.ref_axion_gradient(a, R)
        assert L.or_radial_gradient_energy_pointwise(a, g, a + g, 0.3, R) == \
            ref.ref_radial_gradient(a, g, a + g, 0.3, R)


# ------------------------------------------------------------------ xi KATs
# tests/test_xi_formula.cpp:7-84
def test_xi_kats(oracle):
    L = oracle.load()
    xi = L.or_xi_from_plaquette_count(12345.0, 0.05, 0.05 * 64, 2.0, 3.7)
    assert L.or_plaquette_count_for_target_xi(xi, 64, 0.05 * 64, 2.0, 3.7) == \
        pytest.approx(12345.0, rel=1e-9)
    a = L.or_plaquette_count_for_target_xi(1.0, 128, 2.0, 2.0, 1.0)
    b = L.or_plaquette_count_for_target_xi(1.0, 128, 2.0, 2.0, 2.0)
    assert a / b == pytest.approx(4.0, rel=1e-9)
    assert L.or_plaquette_count_for_target_xi(2.0, 128, 2.0, 2.0, 1.0) == \
        pytest.approx(2 * a, rel=1e-9)
    assert L.or_xi_from_comoving_length(2.0 / 3.0 * 987.0 * 0.03, 4.2, 2.0,
                                        5.5) == \
        pytest.approx(L.or_xi_from_plaquette_count(987.0, 0.03, 4.2, 2.0, 5.5),
                      rel=1e-12)
    assert L.or_xi_from_comoving_length(15.0, 4.2, 2.0, 5.5) == pytest.approx(
        L.or_xi_from_comoving_length(12.0, 4.2, 2.0, 5.5) +
        L.or_xi_from_comoving_length(3.0, 4.2, 2.0, 5.5), rel=1e-12)


def test_xi_matches_reference_header(oracle, ref):
    L = oracle.load()
    rng = np.random.default_rng(4)
    for _ in range(1000):
        ell, Lt, tau = rng.uniform(0.1, 10, 3)
        a_inv = rng.uniform(1.2, 3)
        assert L.or_xi_from_comoving_length(ell, Lt, a_inv, tau) == \
            ref.ref_xi_from_comoving_length(ell, Lt, a_inv, tau)
        assert L.or_xi_from_plaquette_count(ell * 100, 0.1, Lt, a_inv, tau) == \
            ref.ref_xi_from_plaquette_count(ell * 100, 0.1, Lt, a_inv, tau)
        assert L.or_plaquette_count_for_target_xi(ell, 64, Lt, a_inv, tau) == \
            ref.ref_plaquette_count_for_target_xi(ell, 64, Lt, a_inv, tau)


# ----------------------------------------------------------- background KATs
# tests/test_background.cpp:37-182
@pytest.mark.parametrize("c", [0.0, 1.0, 2.0])
def test_background_closed_form_matches_direct(oracle, c):
    L = oracle.load()
    b = oracle.bg(2.0, c0=c)
    for tau in (0.3, 1.0, 2.7, 15.4):
        assert L.or_bg_H_over_mr_direct(b, tau) == pytest.approx(
            L.or_bg_H_over_mr_closed_form(b, tau), rel=1e-12)


def test_background_derivatives_and_switch(oracle):
    L = oracle.load()
    b = oracle.bg(2.0)
    for tau in (0.5, 1.0, 3.3):
        h = 1e-6
        num = (L.or_bg_R(b, tau + h) - L.or_bg_R(b, tau - h)) / (2 * h)
        assert L.or_bg_R_prime(b, tau) == pytest.approx(num, rel=1e-6)
    for tau in (0.5, 1.0, 3.3, 10.0):
        h = 1e-5
        R2 = (L.or_bg_R(b, tau + h) - 2 * L.or_bg_R(b, tau) +
              L.or_bg_R(b, tau - h)) / (h * h)
        assert R2 == pytest.approx(
            L.or_bg_curvature_term_coeff(b, tau) * L.or_bg_R(b, tau),
            rel=1e-4, abs=1e-4)
    s = oracle.bg(2.0, c0=1.0, c1=2.0, tau_switch=5.0, has_switch=True)
    assert L.or_bg_lambda(s, 5.0 - 1e-9) == pytest.approx(
        L.or_bg_lambda(s, 5.0 + 1e-9), rel=1e-6)
    b15 = oracle.bg(2.0, c0=1.5)
    for tau in (0.5, 1.0, 3.3):
        assert L.or_bg_lambda(b15, tau) == pytest.approx(
            (L.or_bg_R(b15, tau) / 1.0) ** (-3.0), rel=1e-12)
    for a_inv in (1.5, 2.0, 3.0):
        for c0 in (0.0, 1.0):
            bb = oracle.bg(a_inv, c0=c0)
            for l in (-2.0, -0.5, 0.0, 0.7, 3.1):
                tau = L.or_bg_tau_from_log_mr_over_h(bb, l)
                assert tau > 0
                assert -math.log(L.or_bg_H_over_mr_closed_form(bb, tau)) == \
                    pytest.approx(l, rel=1e-12, abs=1e-12)
    assert L.or_bg_tau_from_log_mr_over_h(oracle.bg(2.0, c0=0.7), 0.0) == \
        pytest.approx(1.0, rel=1e-12)
    for sw in (False, True):
        bb = oracle.bg(2.0, c0=1.0, c1=2.0, tau_switch=5.0, has_switch=sw)
        for tau in (1.0, 5.0, 5.001, 12.3):
            mr = math.sqrt(L.or_bg_lambda(bb, tau))
            assert L.or_bg_H(bb, tau) == pytest.approx(
                L.or_bg_H_over_mr_direct(bb, tau) * mr, rel=1e-12)


def test_background_matches_reference_header(oracle, ref):
    L = oracle.load()
    fns = [L.or_bg_R, L.or_bg_R_prime, L.or_bg_curvature_term_coeff,
           L.or_bg_lambda, L.or_bg_H, L.or_bg_H_over_mr_direct,
           L.or_bg_H_over_mr_closed_form, L.or_bg_tau_from_log_mr_over_h]
    for (a_inv, c0, c1, ts, sw) in [(2.0, 0.0, 0.0, 0.0, 0), (2.0, 1.0, 2.0, 5.0, 1),
                                    (3.0, 0.5, 3.0, 2.0, 1), (1.5, 0.0, 0.0, 0, 0),
                                    (2.0, 0.0, 2.0, 1.0, 1)]:
        b = oracle.bg(a_inv, c0, c1, ts, bool(sw))
        for tau in (0.3, 1.0, 1.05, 2.0, 4.99, 5.0, 7.7, 151.0):
            for w, f in enumerate(fns):
                assert f(b, tau) == ref.ref_background(a_inv, c0, c1, ts, sw, w,
                                                       tau)


# ------------------------------------------------------------ box planning
# tests/test_box_plan.cpp:8-228, checked against the background formulas (not
# against the plan's own algebra)
def _phys(L, b, tau):
    R = L.or_bg_R(b, tau)
    H = L.or_bg_R_prime(b, tau) / (R * R)
    return R, H, math.sqrt(L.or_bg_lambda(b, tau))


@pytest.mark.parametrize("N,N1,N2,c", [(512, 1.0, 4.0, 0.0), (800, 1.5, 6.0, 1.0),
                                       (1024, 1.0, 4.0, 1.5), (300, 0.7, 3.0, 0.5)])
def test_general_box_plan_delivers_N1_N2_at_tau_f(oracle, N, N1, N2, c):
    L = oracle.load()
    tau_f, Lt, dx, dt_bound = oracle.general_box_plan(N, N1, N2, 2.0, c)
    R, H, mr = _phys(L, oracle.bg(2.0, c0=c), tau_f)
    assert H * Lt * R == pytest.approx(N1, rel=1e-9)
    assert 1.0 / (dx * R * mr) == pytest.approx(N2, rel=1e-9)
    assert dx * N == pytest.approx(Lt, rel=1e-15)
    assert dt_bound == pytest.approx(dx / 3.0, rel=1e-15)


def test_pre_evolution_L_tilde_kats(oracle):
    assert oracle.pre_evolution_L_tilde(10.0, 2.0, 1.0, 4.0) == \
        pytest.approx(10.0, rel=1e-12)
    assert oracle.pre_evolution_L_tilde(10.0, 2.0, 0.0, 4.0) == \
        pytest.approx(40.0, rel=1e-12)
    # a_inv = 3: L (1/2) tau_i^((1-c)/2)
    assert oracle.pre_evolution_L_tilde(8.0, 3.0, 0.0, 9.0) == \
        pytest.approx(12.0, rel=1e-12)


def test_moore_dynamic_range_inverse(oracle):
    for (N1, N2, g, D) in [(1.0, 5.0, 3.0, 6.0), (1.2, 8.0, 2.0, 4.0)]:
        N = int(round(N2 * N1 * g * math.exp(D / 2.0)))
        assert oracle.moore_max_log_dynamic_range(N, N1, N2, g) == \
            pytest.approx(D, rel=1e-3)


def test_amr_box_plan_base_grid_and_divisibility(oracle):
    Lt, dxf, dxb, tau_f, nb, la = oracle.amr_box_plan(4096, 1.2, 8.0, 2.0, 0.0, 4)
    assert nb == 4096 // 16
    assert dxb == pytest.approx(dxf * 16, rel=1e-15)
    assert len(la) == 4
    assert oracle.amr_box_plan(4097, 1.2, 8.0, 2.0, 0.0, 4)[4] == -1
    g = oracle.general_box_plan(4096, 1.2, 8.0, 2.0, 0.0)
    assert (tau_f, Lt, dxf) == (g[0], g[1], g[2])


def test_amr_box_plan_level_spacing(oracle):
    L = oracle.load()
    Lt, dxf, dxb, tau_f, nb, la = oracle.amr_box_plan(4096, 1.0, 4.0, 2.0, 0.0, 5)
    d = math.log(4.0)
    for i in range(1, len(la)):
        assert la[i] - la[i - 1] == pytest.approx(d, rel=1e-9)
    at_f = -math.log(L.or_bg_H_over_mr_closed_form(oracle.bg(2.0), tau_f))
    assert la[-1] == pytest.approx(at_f - d, rel=1e-9)


@pytest.mark.parametrize("N,lev,N2,a_inv", [(8192, 3, 6.0, 2.5), (2048, 2, 4.0, 2.0),
                                             (4096, 4, 5.0, 3.0)])
def test_amr_box_plan_level_thresholds_hit_N2(oracle, N, lev, N2, a_inv):
    """At each level's switch-on time, the level below it alone gives exactly
    N2 (level 1: the base grid)."""
    L = oracle.load()
    Lt, dxf, dxb, tau_f, nb, la = oracle.amr_box_plan(N, 1.0, N2, a_inv, 0.0, lev)
    b = oracle.bg(a_inv)
    for ell in range(1, lev + 1):
        tau = L.or_bg_tau_from_log_mr_over_h(b, la[ell - 1])
        R, _, mr = _phys(L, b, tau)
        dx_below = dxb / 2 ** (ell - 1)
        assert 1.0 / (R * dx_below * mr) == pytest.approx(N2, rel=1e-6)


def test_moore_box_plan_N2_at_switch_and_dynamic_range(oracle):
    L = oracle.load()
    b = oracle.bg(2.0, c0=1.0)
    tau_sw = 30.0
    R_sw, _, mr_sw = _phys(L, b, tau_sw)
    gamma = 1.0 / L.or_bg_H_over_mr_direct(b, tau_sw)
    dx, Lt, D, tau_end = oracle.moore_box_plan(4000, 1.2, 8.0, gamma, 2.0, 1.0,
                                              R_sw, mr_sw, tau_sw)
    assert 1.0 / (8.0 * dx * R_sw * mr_sw) == pytest.approx(1.0, rel=1e-12)
    R12, _, mr12 = _phys(L, b, 12.0)
    assert R12 * mr12 == pytest.approx(R_sw * mr_sw, rel=1e-9)
    assert math.log(L.or_bg_H(b, tau_sw) / L.or_bg_H(b, tau_end)) == \
        pytest.approx(D, rel=1e-9)
    assert Lt == pytest.approx(4000 * dx, rel=1e-15)


def test_box_plan_matches_reference_header(oracle, ref):
    for args in [(512, 1.0, 4.0, 2.0, 0.0), (800, 1.5, 6.0, 2.0, 1.0),
                 (1024, 1.0, 4.0, 3.0, 2.0), (4097, 0.9, 5.0, 2.5, 0.3)]:
        assert oracle.general_box_plan(*args) == \
            oracle.general_box_plan(*args, lib=ref)
        for lev in (0, 1, 3, 5):
            assert oracle.amr_box_plan(*args, lev) == \
                oracle.amr_box_plan(*args, lev, lib=ref)
    for a in [(10.0, 2.0, 1.0, 4.0), (10.0, 2.0, 0.0, 4.0), (7.5, 3.0, 0.5, 2.5)]:
        assert oracle.pre_evolution_L_tilde(*a) == \
            oracle.pre_evolution_L_tilde(*a, lib=ref)
    for a in [(1024, 1.0, 4.0, 3.0), (4000, 1.2, 8.0, 17.3)]:
        assert oracle.moore_max_log_dynamic_range(*a) == \
            oracle.moore_max_log_dynamic_range(*a, lib=ref)
    a = (4000, 1.2, 8.0, 17.3, 2.0, 1.0, 0.5, 2.0, 30.0)
    assert oracle.moore_box_plan(*a) == oracle.moore_box_plan(*a, lib=ref)


# ------------------------------- pre-evolution background and the handoff
PRE_FNS = ["R", "R_prime", "curvature_term_coeff", "lambda",
           "m_r_over_H_direct"]


def test_pre_evolution_background_kats(oracle):
    """tests/test_pre_evolution_background.cpp:12-78: m_r/H = gamma_pre for
    every tau_pre, R m_r constant, R' the derivative of R, R'' = c R."""
    for g in (0.5, 1.0, 3.0):
        for t in (-2.0, -0.3, 0.0, 0.7, 4.0):
            assert oracle.pre_bg(g, "m_r_over_H_direct", t) == \
                pytest.approx(g, rel=1e-9)
    rm0 = oracle.pre_bg(2.0, "R", 0.0) * math.sqrt(oracle.pre_bg(2.0, "lambda", 0.0))
    for t in (-1.5, 0.0, 0.6, 2.2):
        rm = oracle.pre_bg(2.0, "R", t) * math.sqrt(oracle.pre_bg(2.0, "lambda", t))
        assert rm == pytest.approx(rm0, rel=1e-9)
    h = 1e-6
    for t in (-1.0, 0.0, 1.3):
        num = (oracle.pre_bg(1.7, "R", t + h) - oracle.pre_bg(1.7, "R", t - h)) / (2 * h)
        assert oracle.pre_bg(1.7, "R_prime", t) == pytest.approx(num, rel=1e-6)
    h = 1e-5
    for t in (-1.0, 0.0, 1.3):
        R2 = (oracle.pre_bg(1.0, "R", t + h) - 2 * oracle.pre_bg(1.0, "R", t) +
              oracle.pre_bg(1.0, "R", t - h)) / (h * h)
        assert R2 == pytest.approx(oracle.pre_bg(1.0, "curvature_term_coeff", t) *
                                   oracle.pre_bg(1.0, "R", t), rel=1e-4)
    # stage scalars: {curvature, lambda, R^2} at tau, +dt/2, +dt/2, +dt
    c = oracle.pre_stage_coeffs(1.3, 0.4, 0.1)
    for s, t in enumerate((0.4, 0.45, 0.45, 0.5)):
        assert c[3 * s] == 1.0
        assert c[3 * s + 1] == pytest.approx(1.69 * math.exp(-2 * t), rel=1e-15)
        assert c[3 * s + 2] == pytest.approx(math.exp(2 * t), rel=1e-15)


def test_pre_evolution_background_matches_reference_header(oracle, ref):
    for g in (0.5, 1.0, 1.7, 3.0):
        for t in (-2.0, -0.3, 0.0, 0.7, 4.0, 21.6):
            for w, name in enumerate(PRE_FNS):
                assert oracle.pre_bg(g, name, t) == ref.ref_pre_background(g, w, t)


def test_pre_evolution_homogeneous_solution(oracle):
    """psi = (R_pre, 0), Pi = (R_pre', 0) solves the relaxing-phase EOM
    (R'' = 1 * R, |psi| = R): RK4 tracks R = exp(tau_pre) with 4th-order
    error in dt."""
    errs = []
    for dt, nsteps in ((0.1, 20), (0.05, 40)):
        s = np.zeros((4, 4, 4, 4))
        s[0] = 1.0
        s[2] = 1.0
        t = 0.0
        for _ in range(nsteps):
            oracle.rk4_step(s, 0.5, oracle.pre_stage_coeffs(1.0, t, dt), dt)
            t += dt
        assert np.max(np.abs(s[1])) == 0.0
        errs.append(abs(s[0].max() - math.exp(t)) / math.exp(t))
    assert errs[0] < 2e-6
    assert 12.0 < errs[0] / errs[1] < 20.0


def test_handoff_rescale_maps_vacuum_to_vacuum(oracle):
    """The handoff (AxionStringsLevel.cpp:2027-2125) sends the pre-evolution
    vacuum psi = R_pre e^{i theta}, Pi = R_pre' e^{i theta} to the main
    one, R_main e^{i theta}, R_main' e^{i theta}: kappa matches the moduli
    and common_term the velocities."""
    rng = np.random.default_rng(3)
    for (a_inv, c0, g, tp, ti) in [(2.0, 0.0, 1.0, 3.7, 1.0),
                                   (3.0, 1.0, 2.5, 0.4, 2.2),
                                   (2.0, 0.0, 0.5, 12.0, 0.7)]:
        b = oracle.bg(a_inv, c0=c0)
        kappa, common, R_pre, R_main = oracle.handoff_constants(b, g, tp, ti)
        L = oracle.load()
        assert R_pre == math.exp(tp)
        assert R_main == L.or_bg_R(b, ti)
        th = rng.uniform(-PI, PI, (2, 2, 3))
        s = np.stack([R_pre * np.cos(th), R_pre * np.sin(th),
                      math.exp(tp) * np.cos(th), math.exp(tp) * np.sin(th)])
        s = np.ascontiguousarray(s)
        oracle.handoff_rescale(s, kappa, common, R_pre)
        Rp = L.or_bg_R_prime(b, ti)
        np.testing.assert_allclose(s[0], R_main * np.cos(th), rtol=1e-14, atol=1e-14 * R_main)
        np.testing.assert_allclose(s[1], R_main * np.sin(th), rtol=1e-14, atol=1e-14 * R_main)
        np.testing.assert_allclose(s[2], Rp * np.cos(th), rtol=1e-12, atol=1e-12 * abs(Rp))
        np.testing.assert_allclose(s[3], Rp * np.sin(th), rtol=1e-12, atol=1e-12 * abs(Rp))


def test_handoff_blend_smoothstep(oracle):
    """AxionStringsLevel.cpp:889-907: the start values at the handoff, the
    main schedule once t >= 1, the midpoint at t = 1/2, zero slope at both
    ends; n_periods main-schedule core periods long (:2070-2081)."""
    lam_s, cur_s, lam_m, cur_m = 0.3, 1.0, 2.0, -0.25
    t0, d = 1.0, 4.0
    assert oracle.handoff_blend(t0, t0, d, lam_s, cur_s, lam_m, cur_m) == (lam_s, cur_s)
    assert oracle.handoff_blend(t0 - 1, t0, d, lam_s, cur_s, lam_m, cur_m) == (lam_s, cur_s)
    assert oracle.handoff_blend(t0 + d, t0, d, lam_s, cur_s, lam_m, cur_m) == (lam_m, cur_m)
    assert oracle.handoff_blend(t0 + 9, t0, d, lam_s, cur_s, lam_m, cur_m) == (lam_m, cur_m)
    assert oracle.handoff_blend(t0 + 5, t0, 0.0, lam_s, cur_s, lam_m, cur_m) == (lam_m, cur_m)
    mid = oracle.handoff_blend(t0 + d / 2, t0, d, lam_s, cur_s, lam_m, cur_m)
    assert mid == pytest.approx(((lam_s + lam_m) / 2, (cur_s + cur_m) / 2), rel=1e-15)
    h = 1e-4
    for te in (t0, t0 + d):
        a = oracle.handoff_blend(te - h, t0, d, lam_s, cur_s, lam_m, cur_m)[0]
        b = oracle.handoff_blend(te + h, t0, d, lam_s, cur_s, lam_m, cur_m)[0]
        assert abs(b - a) / (2 * h) < 1e-3
    assert oracle.handoff_dtau(3.0, 4.0, 9.0) == pytest.approx(3 * PI, rel=1e-15)
    assert oracle.handoff_dtau(3.0, 0.0, 9.0) == pytest.approx(2 * PI, rel=1e-15)


# ------------------------------------------- stencil / RHS / RK4 closed forms
def test_laplacian_symbol_on_plane_waves(oracle):
    """diff2's symbol on e^{ikx}: (-2cos2t + 32 cos t - 30)/(12 dx^2), t=k dx
    (pins the 4th-order weights and their signs per direction)."""
    n, dx = 16, 0.5
    x = np.arange(n)
    for axis, m in [(0, 1), (1, 3), (2, 5)]:
        t = 2 * np.pi * m / n
        s = np.zeros((4, n, n, n))
        wave = np.cos(t * x)
        shape = [1, 1, 1]
        shape[2 - axis] = n
        s[0] = np.broadcast_to(wave.reshape(shape), (n, n, n))
        r = oracle.rhs(s, dx, 0.0, 0.0, 0.0)
        sym = (-2 * np.cos(2 * t) + 32 * np.cos(t) - 30) / (12 * dx * dx)
        np.testing.assert_allclose(r[2], sym * s[0], atol=1e-12)
        assert np.all(r[3] == 0.0)
        # 4th order: relative error of the symbol vs -k^2 at small k
    k = 2 * np.pi / 64 / dx
    t = k * dx
    sym = (-2 * np.cos(2 * t) + 32 * np.cos(t) - 30) / (12 * dx * dx)
    assert abs(sym + k * k) / (k * k) < 1e-4


def test_rhs_pointwise_terms(oracle):
    """Uniform fields: lap = 0, so rhs_Pi = c psi - (lambda/2)(|psi|^2-R^2) psi
    and rhs_psi = Pi exactly (AxionStringsRHS.impl.hpp:105-117)."""
    n = 8
    s = np.zeros((4, n, n, n))
    s[0], s[1], s[2], s[3] = 1.3, -0.4, 0.25, -0.75
    cc, lam, R2 = 0.1, 0.7, 2.0
    r = oracle.rhs(s, 0.5, cc, lam, R2)
    d = 1.3 ** 2 + 0.4 ** 2 - R2
    assert np.all(r[0] == 0.25) and np.all(r[1] == -0.75)
    # the weight form of diff2 leaves O(1e-16) of a constant field
    np.testing.assert_allclose(r[2], cc * 1.3 - 0.5 * lam * d * 1.3,
                               rtol=0, atol=1e-14)
    np.testing.assert_allclose(r[3], cc * -0.4 - 0.5 * lam * d * -0.4,
                               rtol=0, atol=1e-14)


def test_rk4_homogeneous_exact_solution(oracle):
    """psi = (R, 0), Pi = (R', 0) solves the full EOM when R'' = c_curv R
    (AxionStringsLevel.cpp:824-847, tests/test_background.cpp:69-89).  In
    radiation domination c_curv = 0 and R = tau, so RK4 is exact."""
    n, dx, dt = 8, 0.5, 0.1
    for a_inv, tol in [(2.0, 1e-13), (3.0, 1e-7)]:
        b = oracle.bg(a_inv, c0=0.0)
        L = oracle.load()
        tau = 1.0
        s = np.zeros((4, n, n, n))
        s[0] = L.or_bg_R(b, tau)
        s[2] = L.or_bg_R_prime(b, tau)
        tau = oracle.advance(s, dx, b, tau, dt, 20)
        R = L.or_bg_R(b, tau)
        assert np.max(np.abs(s[0] - R)) / R < tol
        assert np.max(np.abs(s[1])) == 0.0


def test_rk4_fourth_order_convergence(oracle):
    """Global error of RK4 scales as dt^4 on a smooth nonlinear field."""
    n, dx = 8, 0.5
    rng = np.random.default_rng(5)
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    s0 = np.zeros((4, n, n, n))
    s0[0] = 1.0 + 0.3 * np.cos(X.T) + 0.1 * np.sin(Y.T + Z.T)
    s0[1] = 0.2 * np.sin(X.T - Z.T)
    b = oracle.bg(2.0, c0=1.0)
    T = 0.8

    def run(nsteps):
        s = s0.copy()
        oracle.advance(s, dx, b, 1.0, T / nsteps, nsteps)
        return s

    ref = run(256)
    e1 = np.max(np.abs(run(16) - ref))
    e2 = np.max(np.abs(run(32) - ref))
    assert 12.0 < e1 / e2 < 20.0


# ------------------------------------------------------------------ windings
def _string_lattice(n, dx=1.0):
    """psi = (sin(2 pi (x+0.3)/n), sin(2 pi (y+0.3)/n)): four straight
    strings along z (at x, y in {-0.3, n/2-0.3}), alternating winding."""
    i = np.arange(n)
    sx = np.sin(2 * np.pi * (i + 0.3) / n)
    s = np.zeros((4, n, n, n))
    s[0] = sx[None, None, :]
    s[1] = sx[None, :, None]
    return s


def test_plaquette_count_string_lattice(oracle):
    for n in (8, 12, 16):
        s = _string_lattice(n)
        plain, weighted = oracle.count_plaquettes(s)
        assert plain == 4 * n and weighted == 4 * n  # 4 strings x n layers
        w = oracle.windings_field(s)
        assert np.all(w[..., 1] == 0) and np.all(w[..., 2] == 0)
        assert sorted(np.unique(w[..., 0])) == [-1, 0, 1]
        assert w[..., 0].sum() == 0


def test_plaquette_count_brute_force(oracle):
    """Brute-force enumeration with the reference formula on 6^3."""
    L = oracle.load()
    rng = np.random.default_rng(7)
    n = 6
    s = np.zeros((4, n, n, n))
    s[0], s[1] = rng.normal(size=(2, n, n, n))
    th = np.arctan2(s[1], s[0])  # [z, y, x]
    plain = weighted = 0
    for k in range(n):
        for j in range(n):
            for i in range(n):
                t = lambda a, b, c: th[(k + c) % n, (j + b) % n, (i + a) % n]
                ws = [L.or_plaquette_winding(t(0, 0, 0), t(1, 0, 0), t(1, 1, 0), t(0, 1, 0)),
                      L.or_plaquette_winding(t(0, 0, 0), t(0, 1, 0), t(0, 1, 1), t(0, 0, 1)),
                      L.or_plaquette_winding(t(0, 0, 0), t(0, 0, 1), t(1, 0, 1), t(1, 0, 0))]
                plain += sum(w != 0 for w in ws)
                weighted += sum(abs(w) for w in ws)
    assert oracle.count_plaquettes(s) == (plain, weighted)
    # invariances: global phase rotation, conjugation flips the sign
    rot = s.copy()
    a = 0.7
    rot[0] = np.cos(a) * s[0] - np.sin(a) * s[1]
    rot[1] = np.sin(a) * s[0] + np.cos(a) * s[1]
    assert oracle.count_plaquettes(rot) == (plain, weighted)
    conj = s.copy()
    conj[1] = -s[1]
    np.testing.assert_array_equal(oracle.windings_field(conj),
                                  -oracle.windings_field(s))


def test_tagger_semantics(oracle):
    """Thresholds off: tags == cells with any nonzero low-corner winding."""
    s = _string_lattice(12)
    t = oracle.tag_cells(s, 1.0, np.finfo(float).max, np.finfo(float).max)
    w = oracle.windings_field(s)
    np.testing.assert_array_equal(t, np.any(w != 0, axis=-1).astype(np.uint8))
    # Laplacian branch on a closed-form quadratic-like field: a huge
    # gradient threshold tags nothing extra, a tiny one tags every cell
    t2 = oracle.tag_cells(s, 1.0, 1e-12, np.finfo(float).max)
    assert t2.all()
    # amplitude extension: |psi|/R < thr tags the cores only
    t3 = oracle.tag_cells(s, 1.0, np.finfo(float).max, np.finfo(float).max,
                          amp_thr=0.5, R=1.0)
    amp = np.sqrt(s[0] ** 2 + s[1] ** 2)
    assert np.all(t3[amp < 0.5] == 1)


# ------------------------------------------------------------------- energy
def test_energy_sums_homogeneous_vacuum(oracle):
    L = oracle.load()
    b = oracle.bg(2.0)
    tau = 2.0
    R, Rp, lam = L.or_bg_R(b, tau), L.or_bg_R_prime(b, tau), L.or_bg_lambda(b, tau)
    n = 8
    s = np.zeros((4, n, n, n))
    s[0], s[2] = R, Rp
    e = oracle.energy_sums(s, 0.5, R, lam, 1.0 / tau, scheme=2, thr=0.8)
    assert abs(e[0]) < 1e-14 and e[12] == n ** 3 and e[13] == n ** 3


def test_energy_sums_plane_wave_gradient(oracle):
    """psi1 = A cos(k x): sum |grad psi1|^2 = n^3 A^2 g(k)^2 / 2 with the
    4th-order first-derivative symbol g(k) = (8 sin t - sin 2t)/(6 dx)."""
    n, dx, A = 16, 0.5, 0.3
    m = 2
    t = 2 * np.pi * m / n
    x = np.arange(n)
    s = np.zeros((4, n, n, n))
    s[0] = A * np.cos(t * x)[None, None, :]
    R, lam = 1.0, 0.0
    e = oracle.energy_sums(s, dx, R, lam, 0.0)
    g = (8 * np.sin(t) - np.sin(2 * t)) / (6 * dx)
    pot = 0.25 * lam
    kin_grad = n ** 3 * A * A * g * g / 2
    # rho = (|grad psi|^2 + 0.25 lam (|psi|^2 - R^2)^2) / R^4 with lam = 0
    assert e[0] == pytest.approx(kin_grad, rel=1e-12)
    assert pot == 0.0


def _rotating_wave(n, A, m, omega):
    """psi = A (cos kx, sin kx), Pi = i omega psi = omega A (-sin, cos), k =
    2 pi m / n per cell, constant along y and z."""
    t = 2 * np.pi * m / n
    ph = t * np.arange(n)
    s = np.zeros((4, n, n, n))
    s[0] = (A * np.cos(ph))[None, None, :]
    s[1] = (A * np.sin(ph))[None, None, :]
    s[2] = -omega * s[1]
    s[3] = omega * s[0]
    return s, t


def test_energy_sums_rotating_plane_wave_every_slot(oracle):
    """All 14 sums in TotalEnergySums order (EnergyKernel.hpp:85-101), in
    closed form (Energy.hpp:64-189) on psi = A e^{ikx}, Pi = i omega psi:
    theta' = omega, grad theta = (g, 0, 0) with the 4th-order first-
    derivative symbol g = (8 sin t - sin 2t)/(6 dx), |grad psi|^2 = A^2 g^2,
    psi . Pi = 0.  Per cell, with f_a^2 / 2 = 1:
      rho       = (omega^2 A^2 + (R'/R)^2 A^2 + A^2 g^2
                   + lambda/4 (A^2 - R^2)^2) / R^4
      axion kin = omega^2 / R^2,       axion grad = g^2 / R^2
      radial kin = (R'/R)^2 A^2 / R^4, radial grad = 0,
      radial mass = lambda/4 (A^2 - R^2)^2 / R^4;
    the masked slot of each is the unmasked one times the masking weight
    (scheme B top-hat on |psi|/R = A/R, scheme A (1 + (A/R - 1)/sqrt 2)^2),
    slot 12 sums the weight and slot 13 counts the cells."""
    n, dx, A, m, omega = 16, 0.5, 1.3, 2, 0.7
    R, lam, RpR = 1.1, 0.9, 0.35
    s, t = _rotating_wave(n, A, m, omega)
    g = (8 * np.sin(t) - np.sin(2 * t)) / (6 * dx)
    N = n ** 3
    R4 = R ** 4
    pot = 0.25 * lam * (A * A - R * R) ** 2
    vals = [(omega ** 2 * A ** 2 + RpR ** 2 * A ** 2 + A * A * g * g + pot) / R4,
            omega ** 2 / R ** 2,
            g * g / R ** 2,
            RpR ** 2 * A * A / R4,
            0.0,
            pot / R4]
    wA = (1.0 + (A / R - 1.0) / math.sqrt(2.0)) ** 2
    for scheme, thr, w in [(0, 0.8, 1.0), (2, 0.8, 1.0), (2, 1.5, 0.0),
                           (1, 0.8, wA)]:
        e = oracle.energy_sums(s, dx, R, lam, RpR, scheme=scheme, thr=thr)
        for k, v in enumerate(vals):
            scale = N * max(abs(x) for x in vals)
            assert e[2 * k] == pytest.approx(N * v, rel=1e-12,
                                             abs=1e-12 * scale), k
            assert e[2 * k + 1] == pytest.approx(N * v * w, rel=1e-12,
                                                 abs=1e-12 * scale), k
        assert e[12] == pytest.approx(N * w, rel=1e-13)
        assert e[13] == N


def _distinct_midpoint(v, rank=1):
    """A threshold halfway between the rank-th and (rank+1)-th distinct
    values of v (so that no cell sits on it)."""
    u = np.unique(np.round(v, 12))
    return 0.5 * (u[rank] + u[rank + 1])


def test_tagger_laplacian_threshold_closed_form(oracle):
    """StringTagger.hpp:130-141: a cell is tagged where dx^2 |lap psi_i| >
    threshold.  psi1 = B + A cos(kx) > 0, psi2 = 0 (no windings): dx^2 |lap
    psi1| = dx^2 A |s(k)| |cos(k x)| with the diff2 symbol s(k) = (-2 cos 2t
    + 32 cos t - 30)/(12 dx^2); the tags follow that closed form at
    thresholds between its distinct values, and a dx (not dx^2) scaling or a
    missing |.| would tag other cells."""
    n, dx, A, B, m = 16, 0.5, 0.4, 1.0, 3
    t = 2 * np.pi * m / n
    x = np.arange(n)
    s = np.zeros((4, n, n, n))
    s[0] = (B + A * np.cos(t * x))[None, None, :]
    sym = (-2 * np.cos(2 * t) + 32 * np.cos(t) - 30) / (12 * dx * dx)
    v = dx * dx * A * abs(sym) * np.abs(np.cos(t * x))
    big = np.finfo(float).max
    caught = {"dx": False, "signed": False}
    for rank in (0, 1, 2):
        thr = _distinct_midpoint(v, rank)
        want = np.broadcast_to((v > thr)[None, None, :], (n, n, n))
        got = oracle.tag_cells(s, dx, thr, big)
        np.testing.assert_array_equal(got, want.astype(np.uint8))
        assert 0 < want.sum() < n ** 3
        wrong = dx * A * abs(sym) * np.abs(np.cos(t * x)) > thr
        caught["dx"] |= not np.array_equal(wrong, v > thr)
        signed = dx * dx * A * sym * np.cos(t * x) > thr
        caught["signed"] |= not np.array_equal(signed, v > thr)
    assert all(caught.values()), caught


def test_tagger_radial_gradient_closed_form(oracle):
    """StringTagger.hpp:143-150, 182-216: a cell is tagged where
    dx sqrt(|grad psi1|^2 + |grad psi2|^2 - |psi|^2 |grad theta|^2) >
    threshold.  psi = rho(x) (cos qy, sin qy), rho = B + A cos(kx) > 0: the
    modulus varies along x, the phase along y (no windings).  With the
    discrete first-derivative symbols g(k), g(q) the gradient terms are
    A^2 g(k)^2 sin^2(kx) + rho^2 g(q)^2 and rho^2 |grad theta|^2 =
    rho^2 g(q)^2, so the criterion is dx A g(k) |sin(kx)| > threshold; a
    missing sqrt, a dx^2 scaling or a dropped phase term would tag other
    cells."""
    n, dx, A, B, m, mq = 16, 0.5, 0.4, 1.0, 3, 1
    t, tq = 2 * np.pi * m / n, 2 * np.pi * mq / n
    i = np.arange(n)
    rho = B + A * np.cos(t * i)
    s = np.zeros((4, n, n, n))
    s[0] = rho[None, None, :] * np.cos(tq * i)[None, :, None]
    s[1] = rho[None, None, :] * np.sin(tq * i)[None, :, None]
    assert oracle.count_plaquettes(s) == (0, 0)
    gk = (8 * np.sin(t) - np.sin(2 * t)) / (6 * dx)
    gq = (8 * np.sin(tq) - np.sin(2 * tq)) / (6 * dx)
    v = dx * A * abs(gk) * np.abs(np.sin(t * i))
    big = np.finfo(float).max
    caught = {"sqrt": False, "dx2": False, "phase": False}
    for rank in (0, 1, 2):
        thr = _distinct_midpoint(v, rank)
        want = np.broadcast_to((v > thr)[None, None, :], (n, n, n))
        got = oracle.tag_cells(s, dx, big, thr)
        np.testing.assert_array_equal(got, want.astype(np.uint8))
        assert 0 < want.sum() < n ** 3
        no_sqrt = dx * (A * gk * np.sin(t * i)) ** 2 > thr
        caught["sqrt"] |= not np.array_equal(no_sqrt, v > thr)
        caught["dx2"] |= not np.array_equal(dx * v > thr, v > thr)
        no_phase = dx * np.sqrt((A * gk * np.sin(t * i)) ** 2
                                + (rho * gq) ** 2) > thr
        caught["phase"] |= not np.array_equal(no_phase, v > thr)
    assert all(caught.values()), caught


# ----------------------------------------------------------------- spectrum
def test_fft_matches_brute_force_dft(oracle):
    rng = np.random.default_rng(8)
    n = 4
    x = rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n))
    buf = np.zeros(2 * n ** 3)
    buf[0::2] = x.real.ravel()
    buf[1::2] = x.imag.ravel()
    assert oracle.load().or_fft3d(buf.ctypes.data_as(
        __import__("ctypes").POINTER(__import__("ctypes").c_double)), n, n, n,
        -1) == 0
    got = buf[0::2] + 1j * buf[1::2]
    np.testing.assert_allclose(got.reshape(n, n, n), np.fft.fftn(x), atol=1e-12)


def test_spectrum_parseval_and_single_mode(oracle):
    n = 16
    rng = np.random.default_rng(9)
    f = rng.normal(size=(n, n, n))
    ss, sc, tot = oracle.spectrum(f)
    assert tot[0] == pytest.approx(n ** 3 * np.sum(f * f), rel=1e-12)
    assert sc.sum() == pytest.approx(
        sum(1 for a in range(-n // 2, n // 2) for b in range(-n // 2, n // 2)
            for c in range(-n // 2, n // 2)
            if math.sqrt(a * a + b * b + c * c) < n / 2))
    assert ss[0] == 0.0  # DC shell carries 4 pi p^2 = 0
    # single mode p0 = (3, 0, 0): all power in shell 3
    x = np.arange(n)
    g = np.cos(2 * np.pi * 3 * x / n)[None, None, :] * np.ones((n, n, n))
    ss, sc, tot = oracle.spectrum(g)
    nz = np.nonzero(ss > 1e-9 * ss.max())[0]
    assert list(nz) == [3]
    # |X|^2 at (3,0,0) = (n^3/2)^2, weight 2 (its conjugate is not stored)
    assert ss[3] == pytest.approx(2 * 4 * np.pi * 9 * (n ** 3 / 2) ** 2,
                                  rel=1e-12)


# ----------------------------------------------------------------------- IC
def test_splitmix64_published_vectors(oracle):
    L = oracle.load()
    assert L.or_hash_u64(0) == 0xE220A8397B1DCDAF
    g = 0x9E3779B97F4A7C15
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    got = [L.or_hash_u64((1234567 + i * g) % 2 ** 64) for i in range(5)]
    assert got == want


def test_gaussian_mode_box_muller(oracle):
    """gaussian_mode = Box-Muller on uniform_from_hash of the two hashes of
    the key (FourierIC.hpp:202-221), restated independently here."""
    L = oracle.load()
    seed = 42 ^ (1 << 32)
    for (i, j, k) in [(0, 1, 2), (5, 0, 63), (31, 17, 4)]:
        key = seed ^ (i << 40) ^ (j << 20) ^ k
        h1 = L.or_hash_u64(key)
        h2 = L.or_hash_u64(key ^ 0x1234567890ABCDEF)
        u1 = max(((h1 >> 11) + 0.5) / 2 ** 53, 1e-300)
        u2 = ((h2 >> 11) + 0.5) / 2 ** 53
        r = math.sqrt(-2 * math.log(u1))
        re, im = oracle.gaussian_mode(seed, 0, i, j, k)
        assert re == pytest.approx(r * math.cos(2 * math.pi * u2), rel=1e-14)
        assert im == pytest.approx(r * math.sin(2 * math.pi * u2), rel=1e-14)


def test_fourier_ic_invariants_and_modes(oracle):
    n, kmax, ms, seed = 16, 2.0, 0.5, 42
    f = oracle.fourier_ic(n, kmax, ms, seed, 1)
    assert abs(f.mean()) < 1e-14
    assert np.mean(f * f) == pytest.approx(ms, rel=1e-12)
    X = np.fft.rfftn(f)  # [kz][ky][kx<=n/2]
    full_seed = seed ^ (1 << 32)
    scale = None
    for k in range(n):
        for j in range(n):
            for i in range(n // 2 + 1):
                kx, ky, kz = i, (j if j <= n // 2 else j - n), \
                    (k if k <= n // 2 else k - n)
                ksq = kx * kx + ky * ky + kz * kz
                if ksq == 0 or ksq > kmax * kmax:
                    assert abs(X[k, j, i]) < 1e-10
                    continue
                re, im = oracle.gaussian_mode(full_seed, 0, i, j, k)
                want = complex(re, im)
                if i in (0, n // 2):  # Hermitian projection of the plane
                    r2, i2 = oracle.gaussian_mode(full_seed, 0, i, (-j) % n,
                                                  (-k) % n)
                    mk = (-kx) ** 2 + ky ** 2 + kz ** 2
                    other = complex(r2, -i2) if 0 < mk <= kmax * kmax else 0
                    want = 0.5 * (want + other)
                if scale is None:
                    scale = X[k, j, i] / want
                assert X[k, j, i] == pytest.approx(scale * want, rel=1e-10,
                                                   abs=1e-10)
    assert abs(scale.imag) < 1e-10 and scale.real > 0


# ---------------------------------------------------------------- velocity
def test_velocity_formula_matches_reference_header(oracle, ref):
    L = oracle.load()
    rng = np.random.default_rng(10)
    for _ in range(3000):
        v = rng.normal(size=5)
        v[4] = abs(v[4]) + 0.05
        assert L.or_gamma_sq_v_sq_from_phi(*v) == \
            ref.ref_gamma_sq_v_sq_from_phi(*v)


def test_velocity_static_field_and_corner_count(oracle):
    """Pi = 0 and R'/R = 0 give phi_dot = 0, so gamma^2 v^2 = 0 at every
    corner; the string lattice pierces 4n xy plaquettes -> 16n corners."""
    for n in (8, 12):
        i = np.arange(n)
        sx = np.sin(2 * np.pi * (i + 0.3) / n)
        s = np.zeros((4, n, n, n))
        s[0] = sx[None, None, :]
        s[1] = sx[None, :, None]
        total, count = oracle.velocity(s, 1.0, 0.0, 1.0)
        assert count == 16 * n and total == 0.0
        # a uniform rotation psi -> psi e^{i w tau}: Pi = i w psi, so at every
        # corner gamma^2 v^2 = term1 - term2 with re(phi* phi_dot) = 0
        w = 0.3
        s[2], s[3] = -w * s[1], w * s[0]
        total, count = oracle.velocity(s, 1.0, 0.0, 1.0)
        L = oracle.load()
        wf = oracle.windings_field(s)
        acc = 0.0
        for z, yy, x in zip(*np.nonzero(wf[..., 0])):
            for (a, b) in [(0, 0), (1, 0), (1, 1), (0, 1)]:
                xx, yb = (x + a) % n, (yy + b) % n
                p1, p2 = s[0, z, yb, xx], s[1, z, yb, xx]
                acc += L.or_gamma_sq_v_sq_from_phi(p1, p2, -w * p2, w * p1,
                                                   1.0)
        assert total == pytest.approx(acc, rel=1e-12)
