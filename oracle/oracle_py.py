"""ctypes wrapper of the CPU oracle (oracle/_build/libaxs_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "_build", "libaxs_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libaxs_ref.so")

_PD = C.POINTER(C.c_double)
_P = C.c_void_p
_D = C.c_double
_I = C.c_int


class OrBackground(C.Structure):
    _fields_ = [("a_inv", C.c_double), ("c0", C.c_double), ("c1", C.c_double),
                ("tau_switch", C.c_double), ("has_switch", C.c_int)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build()
    lib = C.CDLL(ORACLE_LIB)
    pb = C.POINTER(OrBackground)
    for name in ["or_bg_R", "or_bg_R_prime", "or_bg_curvature_term_coeff",
                 "or_bg_lambda", "or_bg_H", "or_bg_H_over_mr_direct",
                 "or_bg_H_over_mr_closed_form", "or_bg_tau_from_log_mr_over_h"]:
        getattr(lib, name).restype = _D
        getattr(lib, name).argtypes = [pb, _D]
    lib.or_wrap_angle_diff.restype = _D
    lib.or_wrap_angle_diff.argtypes = [_D]
    lib.or_plaquette_winding.restype = _I
    lib.or_plaquette_winding.argtypes = [_D] * 4
    lib.or_xi_from_comoving_length.restype = _D
    lib.or_xi_from_comoving_length.argtypes = [_D] * 4
    lib.or_xi_from_plaquette_count.restype = _D
    lib.or_xi_from_plaquette_count.argtypes = [_D] * 5
    lib.or_plaquette_count_for_target_xi.restype = _D
    lib.or_plaquette_count_for_target_xi.argtypes = [_D, _I, _D, _D, _D]
    lib.or_masking_weight.restype = _D
    lib.or_masking_weight.argtypes = [_I, _D, _D, _D, _D]
    lib.or_masked_a_dot.restype = _D
    lib.or_masked_a_dot.argtypes = [_I, _D, _D, _D, _D, _D, _D]
    lib.or_rho_tot_pointwise.restype = _D
    lib.or_rho_tot_pointwise.argtypes = [_D] * 9
    lib.or_axion_kinetic_energy_pointwise.restype = _D
    lib.or_axion_kinetic_energy_pointwise.argtypes = [_D] * 2
    lib.or_axion_gradient_energy_pointwise.restype = _D
    lib.or_axion_gradient_energy_pointwise.argtypes = [_D] * 2
    lib.or_radial_kinetic_energy_pointwise.restype = _D
    lib.or_radial_kinetic_energy_pointwise.argtypes = [_D] * 6
    lib.or_radial_gradient_energy_pointwise.restype = _D
    lib.or_radial_gradient_energy_pointwise.argtypes = [_D] * 5
    lib.or_radial_mass_energy_pointwise.restype = _D
    lib.or_radial_mass_energy_pointwise.argtypes = [_D] * 4
    lib.or_rhs.argtypes = [_PD, _PD, _I, _I, _I, _D, _D, _D, _D]
    lib.or_rk4_step.argtypes = [_PD, _I, _I, _I, _D, _PD, _D]
    lib.or_stage_coeffs.argtypes = [pb, _D, _D, _PD]
    lib.or_count_plaquettes.argtypes = [_PD, _I, _I, _I, _P,
                                        C.POINTER(C.c_int64)]
    lib.or_windings_field.argtypes = [_PD, _I, _I, _I, _P]
    lib.or_tag_cells.argtypes = [_PD, _I, _I, _I, _D, _D, _D, _D, _D, _P]
    lib.or_energy_sums.argtypes = [_PD, _I, _I, _I, _D, _D, _D, _D, _I, _D,
                                   _P, _PD]
    lib.or_fill_field.argtypes = [_PD, _I, _I, _I, _I, _I, _D, _D, _PD]
    lib.or_spectrum.restype = _I
    lib.or_spectrum.argtypes = [_PD, _I, _I, _I, _PD, C.POINTER(C.c_int64),
                                _PD]
    lib.or_gamma_sq_v_sq_from_phi.restype = _D
    lib.or_gamma_sq_v_sq_from_phi.argtypes = [_D] * 5
    lib.or_velocity.argtypes = [_PD, _I, _I, _I, _D, _D, _D, _P, _PD,
                                C.POINTER(C.c_int64)]
    lib.or_hash_u64.restype = C.c_uint64
    lib.or_hash_u64.argtypes = [C.c_uint64]
    lib.or_uniform_from_hash.restype = _D
    lib.or_uniform_from_hash.argtypes = [C.c_uint64]
    lib.or_gaussian_mode.argtypes = [C.c_uint64, _I, _I, _I, _I, _PD, _PD]
    lib.or_wrapped_freq.restype = _I
    lib.or_wrapped_freq.argtypes = [_I, _I]
    lib.or_fourier_ic.restype = _I
    lib.or_fourier_ic.argtypes = [_PD, _I, _I, _I, _D, _D, C.c_uint64, _I]
    lib.or_fft3d.restype = _I
    lib.or_fft3d.argtypes = [_PD, _I, _I, _I, _I]
    lib.or_num_threads.restype = _I
    for name in ["or_pre_R", "or_pre_R_prime", "or_pre_curvature_term_coeff",
                 "or_pre_lambda", "or_pre_m_r_over_H_direct"]:
        getattr(lib, name).restype = _D
        getattr(lib, name).argtypes = [_D, _D]
    lib.or_pre_stage_coeffs.argtypes = [_D, _D, _D, _PD]
    lib.or_handoff_constants.argtypes = [pb, _D, _D, _D, _PD]
    lib.or_handoff_rescale.argtypes = [_PD, C.c_size_t, _D, _D, _D]
    lib.or_handoff_blend.argtypes = [_D, _D, _D, _D, _D, _PD, _PD]
    lib.or_handoff_dtau.restype = _D
    lib.or_handoff_dtau.argtypes = [_D, _D, _D]
    _lib = lib
    _bind_box_plan(lib, "or_")
    return lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_PD)


def bg(a_inv=2.0, c0=0.0, c1=0.0, tau_switch=0.0, has_switch=False):
    return OrBackground(a_inv, c0, c1, tau_switch, int(has_switch))


def shape_of(state):
    _, nz, ny, nx = state.shape
    return nx, ny, nz


def rhs(state, dx, cc, lam, R2):
    s = np.ascontiguousarray(state, dtype=np.float64)
    out = np.empty_like(s)
    nx, ny, nz = shape_of(s)
    load().or_rhs(_d(s), _d(out), nx, ny, nz, dx, cc, lam, R2)
    return out


def stage_coeffs(b, tau, dt):
    c = np.zeros(12)
    load().or_stage_coeffs(C.byref(b), tau, dt, _d(c))
    return c


def pre_stage_coeffs(gamma_pre, tau_pre, dt):
    c = np.zeros(12)
    load().or_pre_stage_coeffs(gamma_pre, tau_pre, dt, _d(c))
    return c


def pre_bg(gamma_pre, which, tau_pre):
    """which: R, R_prime, curvature_term_coeff, lambda, m_r_over_H_direct."""
    return getattr(load(), "or_pre_" + which)(gamma_pre, tau_pre)


def handoff_constants(b, gamma_pre, tau_pre_end, tau_i):
    out = np.zeros(4)
    load().or_handoff_constants(C.byref(b), gamma_pre, tau_pre_end, tau_i,
                                _d(out))
    return tuple(out)  # kappa, common_term, R_pre, R_main


def handoff_rescale(state, kappa, common_term, R_pre):
    """In place on a (4, nz, ny, nx) float64 state."""
    assert state.dtype == np.float64 and state.flags.c_contiguous
    load().or_handoff_rescale(_d(state), state[0].size, kappa, common_term,
                              R_pre)


def handoff_blend(tau, tau_start, dtau, lambda_start, curv_start, lam, curv):
    lv = C.c_double(lam)
    cv = C.c_double(curv)
    load().or_handoff_blend(tau, tau_start, dtau, lambda_start, curv_start,
                            C.byref(lv), C.byref(cv))
    return lv.value, cv.value


def handoff_dtau(n_periods, lambda_main_tau_i, lambda_start):
    return load().or_handoff_dtau(n_periods, lambda_main_tau_i, lambda_start)


def rk4_step(state, dx, coeffs12, dt):
    """In place."""
    nx, ny, nz = shape_of(state)
    c = np.ascontiguousarray(coeffs12, dtype=np.float64)
    load().or_rk4_step(_d(state), nx, ny, nz, dx, _d(c), dt)


def advance(state, dx, b, tau, dt, nsteps):
    """nsteps RK4 steps in place; tau advanced by repeated += dt."""
    for _ in range(nsteps):
        rk4_step(state, dx, stage_coeffs(b, tau, dt), dt)
        tau += dt
    return tau


def count_plaquettes(state, mask=None):
    s = np.ascontiguousarray(state, dtype=np.float64)
    out = (C.c_int64 * 2)()
    nx, ny, nz = shape_of(s)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    load().or_count_plaquettes(_d(s), nx, ny, nz,
                               m.ctypes.data_as(_P) if m is not None else None,
                               out)
    return int(out[0]), int(out[1])


def velocity(state, R, RpR, m_r, mask=None):
    s = np.ascontiguousarray(state, dtype=np.float64)
    nx, ny, nz = shape_of(s)
    out = C.c_double()
    cnt = C.c_int64()
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    load().or_velocity(_d(s), nx, ny, nz, R, RpR, m_r,
                       m.ctypes.data_as(_P) if m is not None else None,
                       C.byref(out), C.byref(cnt))
    return out.value, cnt.value


def windings_field(state):
    s = np.ascontiguousarray(state, dtype=np.float64)
    nx, ny, nz = shape_of(s)
    w = np.zeros((nz, ny, nx, 3), dtype=np.int8)
    load().or_windings_field(_d(s), nx, ny, nz, w.ctypes.data_as(_P))
    return w


def tag_cells(state, dx, grad_thr, rgrad_thr, amp_thr=0.0, R=1.0):
    s = np.ascontiguousarray(state, dtype=np.float64)
    nx, ny, nz = shape_of(s)
    t = np.zeros((nz, ny, nx), dtype=np.uint8)
    load().or_tag_cells(_d(s), nx, ny, nz, dx, grad_thr, rgrad_thr, amp_thr,
                        R, t.ctypes.data_as(_P))
    return t


def energy_sums(state, dx, R, lam, RpR, scheme=0, thr=0.8, mask=None):
    s = np.ascontiguousarray(state, dtype=np.float64)
    nx, ny, nz = shape_of(s)
    out = np.zeros(14)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    load().or_energy_sums(_d(s), nx, ny, nz, dx, R, lam, RpR, scheme, thr,
                          m.ctypes.data_as(_P) if m is not None else None,
                          _d(out))
    return out


def fill_field(state, kind, scheme=0, thr=0.8, R=1.0):
    s = np.ascontiguousarray(state, dtype=np.float64)
    nx, ny, nz = shape_of(s)
    buf = np.zeros((nz, ny, nx))
    load().or_fill_field(_d(s), nx, ny, nz, kind, scheme, thr, R, _d(buf))
    return buf


def spectrum(field):
    f = np.ascontiguousarray(field, dtype=np.float64)
    nz, ny, nx = f.shape
    n_shells = max(nx, ny, nz) // 2 + 1
    ss = np.zeros(n_shells)
    sc = np.zeros(n_shells, dtype=np.int64)
    tot = np.zeros(2)
    rc = load().or_spectrum(_d(f), nx, ny, nz, _d(ss),
                            sc.ctypes.data_as(C.POINTER(C.c_int64)), _d(tot))
    assert rc == 0
    return ss, sc, tot


def fourier_ic(n, k_max_cells, target_ms, seed, comp_seed_offset):
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    f = np.zeros((nz, ny, nx))
    rc = load().or_fourier_ic(_d(f), nx, ny, nz, k_max_cells, target_ms,
                              seed, comp_seed_offset)
    assert rc == 0
    return f


def gaussian_mode(seed, comp, i, j, k):
    re, im = C.c_double(), C.c_double()
    load().or_gaussian_mode(seed, comp, i, j, k, C.byref(re), C.byref(im))
    return re.value, im.value


def ref_lib():
    """The reference's own standalone headers compiled from source, if
    oracle/_ref was built (needs the reference checkout); else None."""
    if not os.path.exists(REF_LIB):
        return None
    lib = C.CDLL(REF_LIB)
    lib.ref_background.restype = _D
    lib.ref_background.argtypes = [_D, _D, _D, _D, _I, _I, _D]
    lib.ref_plaquette_winding.restype = _I
    lib.ref_plaquette_winding.argtypes = [_D] * 4
    lib.ref_wrap_angle_diff.restype = _D
    lib.ref_wrap_angle_diff.argtypes = [_D]
    lib.ref_masking_weight.restype = _D
    lib.ref_masking_weight.argtypes = [_I, _D, _D, _D, _D]
    lib.ref_masked_a_dot.restype = _D
    lib.ref_masked_a_dot.argtypes = [_I, _D, _D, _D, _D, _D, _D]
    lib.ref_rho_tot_pointwise.restype = _D
    lib.ref_rho_tot_pointwise.argtypes = [_D] * 9
    for n, k in [("ref_axion_kinetic", 2), ("ref_axion_gradient", 2),
                 ("ref_radial_kinetic", 6), ("ref_radial_gradient", 5),
                 ("ref_radial_mass", 4)]:
        getattr(lib, n).restype = _D
        getattr(lib, n).argtypes = [_D] * k
    lib.ref_pre_background.restype = _D
    lib.ref_pre_background.argtypes = [_D, _I, _D]
    lib.ref_gamma_sq_v_sq_from_phi.restype = _D
    lib.ref_gamma_sq_v_sq_from_phi.argtypes = [_D] * 5
    lib.ref_xi_from_comoving_length.restype = _D
    lib.ref_xi_from_comoving_length.argtypes = [_D] * 4
    lib.ref_xi_from_plaquette_count.restype = _D
    lib.ref_xi_from_plaquette_count.argtypes = [_D] * 5
    lib.ref_plaquette_count_for_target_xi.restype = _D
    lib.ref_plaquette_count_for_target_xi.argtypes = [_D, _I, _D, _D, _D]
    _bind_box_plan(lib, "ref_")
    return lib


def _bind_box_plan(lib, pre):
    _PD = C.POINTER(C.c_double)
    getattr(lib, pre + "general_box_plan").restype = None
    getattr(lib, pre + "general_box_plan").argtypes = [_I, _D, _D, _D, _D, _PD]
    getattr(lib, pre + "amr_box_plan").restype = None
    getattr(lib, pre + "amr_box_plan").argtypes = [
        _I, _D, _D, _D, _D, _I, _PD, C.POINTER(C.c_int), _PD]
    getattr(lib, pre + "pre_evolution_L_tilde").restype = _D
    getattr(lib, pre + "pre_evolution_L_tilde").argtypes = [_D] * 4
    getattr(lib, pre + "moore_max_log_dynamic_range").restype = _D
    getattr(lib, pre + "moore_max_log_dynamic_range").argtypes = [_I, _D, _D, _D]
    getattr(lib, pre + "moore_box_plan").restype = None
    getattr(lib, pre + "moore_box_plan").argtypes = [_I] + [_D] * 8 + [_PD]


def general_box_plan(N, N1, N2, a_inv, c, lib=None):
    """(tau_f, L_tilde, delta_x, delta_tau_leapfrog_bound), BoxPlan.hpp."""
    lib, pre = (load(), "or_") if lib is None else (lib, "ref_")
    out = (C.c_double * 4)()
    getattr(lib, pre + "general_box_plan")(N, N1, N2, a_inv, c, out)
    return tuple(out)


def amr_box_plan(N_eff, N1, N2, a_inv, c, max_level, lib=None):
    """(L_tilde, dx_finest, dx_base, tau_f, N_base, [log_add per level])."""
    lib, pre = (load(), "or_") if lib is None else (lib, "ref_")
    out = (C.c_double * 4)()
    nb = C.c_int()
    la = (C.c_double * max(1, max_level))()
    getattr(lib, pre + "amr_box_plan")(N_eff, N1, N2, a_inv, c, max_level,
                                       out, C.byref(nb), la)
    return tuple(out) + (nb.value, list(la)[:max_level])


def pre_evolution_L_tilde(L, a_inv, c, tau_i, lib=None):
    lib, pre = (load(), "or_") if lib is None else (lib, "ref_")
    return getattr(lib, pre + "pre_evolution_L_tilde")(L, a_inv, c, tau_i)


def moore_max_log_dynamic_range(N, N1, N2, gamma, lib=None):
    lib, pre = (load(), "or_") if lib is None else (lib, "ref_")
    return getattr(lib, pre + "moore_max_log_dynamic_range")(N, N1, N2, gamma)


def moore_box_plan(N, N1, N2, gamma, a_inv, b_inv, R_sw, mr_sw, tau_sw,
                   lib=None):
    """(dx, L_tilde, D, tau_end), BoxPlan.hpp."""
    lib, pre = (load(), "or_") if lib is None else (lib, "ref_")
    out = (C.c_double * 4)()
    getattr(lib, pre + "moore_box_plan")(N, N1, N2, gamma, a_inv, b_inv, R_sw,
                                         mr_sw, tau_sw, out)
    return tuple(out)


def seeded_ic(n, seed=42, dx=0.5, target_ms=0.5, k_max_cells=None):
    """The benchmark IC: psi1, psi2 hash-seeded Gaussian fields (seed offsets
    0 and 1), modes 0 < |p| <= n/8, <psi_i^2> = target_ms, Pi = 0."""
    nx, ny, nz = (n, n, n) if isinstance(n, int) else n
    if k_max_cells is None:
        k_max_cells = max(nx, ny, nz) / 8.0
    s = np.zeros((4, nz, ny, nx))
    s[0] = fourier_ic((nx, ny, nz), k_max_cells, target_ms, seed, 0)
    s[1] = fourier_ic((nx, ny, nz), k_max_cells, target_ms, seed, 1)
    return s
