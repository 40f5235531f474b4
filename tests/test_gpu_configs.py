"""The plain-at-scale, batching and differentiation workloads at their own
sizes (the suite's other tests run them on small grids):

* plain method at scale: 512^3 FP64, c = 2 (Moore), dx = 0.5, dtau = 0.1,
  seeded IC -- the device IC against the oracle's, two RK4 steps bit for
  bit against the oracle, and the |Pi|^2 spectra in 257 shells
  (SpectrumKernel.hpp:52-153) against the oracle's;
* batching form: a 256^3 base with two ratio-2 levels in FP32C from the
  amplitude / winding tagger on the seeded IC, two coarse steps and a
  regrid -- composite volume equal to L^3, level-0 tags equal to the
  oracle's tagger on the same field and a superset of the pierced cells,
  graph replays equal to direct launches;
* differentiation form: 128^3, 400 RK4 steps, checkpoints every 20 -- the
  FP64 adjoint against central differences (eps = 1e-3) on 8 seeded
  directions and lambda, the FP32C-forward gradient against the FP64 one.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capi():
    from axion_strings_adaptive_amd import capi
    return capi


def test_plain_method_512_fp64_steps_and_spectrum(ctx, oracle):
    capi = _capi()
    n, dx, dt = 512, 0.5, 0.1
    lvl = capi.Level(ctx, n, dx, capi.AXS_FP64)
    lvl.fourier_ic(0, n / 8.0, 0.5, 42, 0)
    lvl.fourier_ic(1, n / 8.0, 0.5, 42, 1)
    lvl.zero_components(2, 2)
    s = oracle.seeded_ic(n)
    got = lvl.download()
    assert np.max(np.abs(got - s)) < 1e-12
    del got
    lvl.upload(s)  # identical inputs for the bit-for-bit comparison
    bg = capi.make_background(2.0, c0=2.0)
    lvl.advance(bg, 1.0, dt, 2, use_graph=True)
    tau = oracle.advance(s, dx, oracle.bg(2.0, c0=2.0), 1.0, dt, 2)
    assert tau == pytest.approx(1.2, rel=1e-15)
    assert np.array_equal(lvl.download(), s)
    for kind in (capi.AXS_FIELD_PI1, capi.AXS_FIELD_PI2):
        ss, sc, tot, ssq = lvl.spectrum(kind)
        wss, wsc, wtot = oracle.spectrum(oracle.fill_field(s, kind))
        assert len(sc) == n // 2 + 1 == 257
        assert np.array_equal(sc, wsc)
        np.testing.assert_allclose(ss, wss, rtol=1e-10,
                                   atol=1e-12 * wss.max())
        # hipFFT and the oracle's radix-2 FFT round differently: the totals
        # over 6.7e7 modes agree to ~1e-10 (7.6e-11 measured)
        np.testing.assert_allclose(tot, wtot, rtol=2e-10)
        assert tot[0] == pytest.approx(n ** 3 * ssq, rel=2e-10)  # Parseval
    lvl.close()


def _hierarchy_run(ctx, use_graph, oracle=None):
    from axion_strings_adaptive_amd import amr, capi
    h = amr.Hierarchy(ctx, n0=256, levels=2, precision=capi.AXS_FP32C,
                      verbose=False, use_graph=use_graph, regrid_int=2)
    L3 = (256 * 0.5) ** 3
    assert len(h.levels) == 3
    assert h.composite_energy()["n_total"] == pytest.approx(L3, rel=1e-12)
    tag_check = None
    if oracle is not None:
        # level-0 tags: the amplitude / winding tagger on the level's field
        L0 = h.levels[0]
        R = capi.background_eval(h.bg, 0, h.tau)
        tags, counts = L0.tag(amp_thr=0.5, R=R)
        t = tags.to_numpy(np.uint8).reshape(256, 256, 256)
        s = L0.download_dense()
        want = oracle.tag_cells(s, 0.5, capi.DBL_MAX, capi.DBL_MAX,
                                amp_thr=0.5, R=R)
        w = oracle.windings_field(s)
        tag_check = (np.array_equal(t, want),
                     bool(np.all(t[np.any(w != 0, axis=-1)] == 1)),
                     counts == oracle.count_plaquettes(s), int(t.sum()))
    sums = []
    for _ in range(2):
        h.step()
        e = h.composite_energy()
        assert e["n_total"] == pytest.approx(L3, rel=1e-12)
        sums.append((e["rho_tot_unscreened"], e["rho_tot_screened"]))
    assert h.nregrids == 1 and len(h.levels) == 3
    cells = list(h.ncells)
    h.close()
    ctx.trim()
    return sums, cells, tag_check


def test_batching_form_256_two_levels(ctx, oracle):
    g_sums, g_cells, tags = _hierarchy_run(ctx, True, oracle)
    same, superset, counts_ok, ntag = tags
    assert same and superset and counts_ok and ntag > 0
    assert g_cells[1] > 0 and g_cells[2] > 0
    d_sums, d_cells, _ = _hierarchy_run(ctx, False)
    assert d_cells == g_cells
    assert d_sums == g_sums  # graph replays give the direct launches' bits
    assert all(np.isfinite(v) and v > 0 for pair in g_sums for v in pair)


def test_differentiation_form_128_400_steps(ctx):
    from axion_strings_adaptive_amd import capi
    from axion_strings_adaptive_amd.adjoint_check import seeded_state
    from axion_strings_adaptive_amd.torch_adjoint import GradEnergyProblem
    n, nsteps, K = 128, 400, 20
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    s0 = seeded_state(ctx, n, 42)
    p64 = GradEnergyProblem(ctx, n, precision=capi.AXS_FP64, bg=bg,
                            nsteps=nsteps, checkpoint_every=K)
    p32 = GradEnergyProblem(ctx, n, precision=capi.AXS_FP32C, bg=bg,
                            nsteps=nsteps, checkpoint_every=K)
    J64, g64, l64 = p64.value_and_grad(s0, 1.0)
    J32, g32, l32 = p32.value_and_grad(s0, 1.0)
    assert J32 == pytest.approx(J64, rel=1e-5)
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-5
    assert l32 == pytest.approx(l64, rel=1e-4)
    rng = np.random.default_rng(43)
    eps = 1e-3
    for _ in range(8):
        v = rng.standard_normal(s0.shape)
        fd = (p64.forward_value(s0 + eps * v, 1.0) -
              p64.forward_value(s0 - eps * v, 1.0)) / (2 * eps)
        assert float(np.sum(g64 * v)) == pytest.approx(fd, rel=1e-4)
    fdl = (p64.forward_value(s0, 1.0 + eps) -
           p64.forward_value(s0, 1.0 - eps)) / (2 * eps)
    assert l64 == pytest.approx(fdl, rel=1e-4)
    for p in (p32, p64):
        p.level.close()
