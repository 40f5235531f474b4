"""The pre-evolution pieces of the HIP path against the oracle: RK4 steps
with caller-supplied stage scalars (the relaxing phase,
AxionStringsLevel.cpp:869-876) and the in-place handoff rescale
(AxionStringsLevel.cpp:2027-2125).

Bars: FP64 bit-exact (same operation order, contraction off); FP32C within
the rounding of its hi + lo storage."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capi():
    from axion_strings_adaptive_amd import capi
    return capi


def _pre_rows(capi, oracle, gamma, t, dt, nsteps):
    rows = []
    for _ in range(nsteps):
        c = oracle.pre_stage_coeffs(gamma, t, dt)
        rows += [capi.RhsCoeffs(*c[3 * k:3 * k + 3]) for k in range(4)]
        t += dt
    return rows, t


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("n", [32, 40])
def test_relaxing_steps_fp64_bitexact(ctx, oracle, use_graph, n):
    capi = _capi()
    dx, dt, gamma = 0.5, 0.1, 1.3
    s = oracle.seeded_ic(32, seed=5) if n == 32 else \
        np.random.default_rng(n).normal(size=(4, n, n, n))
    lvl = capi.Level(ctx, n, dx, capi.AXS_FP64)
    lvl.upload(s)
    rows, _ = _pre_rows(capi, oracle, gamma, 0.25, dt, 7)
    lvl.advance_coeffs(rows, dt, use_graph=use_graph)
    t = 0.25
    for _ in range(7):
        oracle.rk4_step(s, dx, oracle.pre_stage_coeffs(gamma, t, dt), dt)
        t += dt
    assert np.array_equal(lvl.download(), s)
    lvl.close()


def test_advance_coeffs_equals_advance_with_background(ctx, oracle):
    """The same stage scalars through both entry points give the same bits."""
    capi = _capi()
    s = oracle.seeded_ic(32, seed=9)
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.2,
                              has_switch=True)
    a = capi.Level(ctx, 32, 0.5, capi.AXS_FP32C)
    b = capi.Level(ctx, 32, 0.5, capi.AXS_FP32C)
    a.upload(s)
    b.upload(s)
    a.advance(bg, 1.0, 0.1, 5)
    rows, tau, dt = [], 1.0, 0.1
    for _ in range(5):
        rows += [capi.rhs_coeffs(bg, t) for t in
                 (tau, tau + 0.5 * dt, tau + 0.5 * dt, tau + dt)]
        tau += dt
    b.advance_coeffs(rows, dt, use_graph=True)
    ha, la = a.download_raw()
    hb, lb = b.download_raw()
    assert np.array_equal(ha, hb) and np.array_equal(la, lb)
    a.close()
    b.close()


def test_handoff_rescale_fp64_bitexact(ctx, oracle):
    capi = _capi()
    s = oracle.seeded_ic(32, seed=3)
    s[2:] = np.random.default_rng(1).normal(size=s[2:].shape)
    s = np.ascontiguousarray(s)
    ob = oracle.bg(2.0)
    kappa, common, R_pre, _ = oracle.handoff_constants(ob, 1.0, 4.7, 1.0)
    lvl = capi.Level(ctx, 32, 0.5, capi.AXS_FP64)
    lvl.upload(s)
    lvl.handoff_rescale(kappa, common, R_pre)
    oracle.handoff_rescale(s, kappa, common, R_pre)
    assert np.array_equal(lvl.download(), s)
    lvl.close()


@pytest.mark.parametrize("precision", [1, 2])
def test_handoff_rescale_fp32(ctx, oracle, precision):
    """FP32C keeps hi + lo (rescaled in FP64, re-split): ~1e-14 relative;
    plain FP32 rounds to one float."""
    capi = _capi()
    s = oracle.seeded_ic(32, seed=4)
    s[2:] = 0.3 * np.random.default_rng(2).normal(size=s[2:].shape)
    s = np.ascontiguousarray(s)
    ob = oracle.bg(2.0)
    kappa, common, R_pre, _ = oracle.handoff_constants(ob, 1.0, 2.2, 1.0)
    lvl = capi.Level(ctx, 32, 0.5, precision)
    lvl.upload(s)
    start = lvl.download()
    lvl.handoff_rescale(kappa, common, R_pre)
    oracle.handoff_rescale(start, kappa, common, R_pre)
    got = lvl.download()
    scale = np.max(np.abs(start), axis=(1, 2, 3), keepdims=True)
    err = np.max(np.abs(got - start) / scale)
    assert err < (1e-13 if precision == 1 else 1.2e-7)
    lvl.close()


def test_handoff_rejects_zero_R_pre(ctx):
    capi = _capi()
    lvl = capi.Level(ctx, 8, 0.5, capi.AXS_FP64)
    with pytest.raises(capi.AxsError):
        lvl.handoff_rescale(1.0, 0.0, 0.0)
    lvl.close()


def test_placement_trials_keep_a_working_level(ctx, oracle):
    """axs_level_place re-allocates the arrays (state reset to zero); the
    level it keeps computes the same bits as one that was never moved."""
    capi = _capi()
    s = oracle.seeded_ic(32, seed=8)
    bg = capi.make_background(2.0)
    a = capi.Level(ctx, 32, 0.5, capi.AXS_FP32C)
    ms = a.place(3)
    assert 1 <= len(ms) <= 3 and all(t > 0 for t in ms)
    assert not np.any(a.download())
    b = capi.Level(ctx, 32, 0.5, capi.AXS_FP32C)
    for lvl in (a, b):
        lvl.upload(s)
        lvl.advance(bg, 1.0, 0.1, 3, use_graph=True)
    ha, la = a.download_raw()
    hb, lb = b.download_raw()
    assert np.array_equal(ha, hb) and np.array_equal(la, lb)
    with pytest.raises(capi.AxsError):
        a.place(0)
    a.close()
    b.close()
