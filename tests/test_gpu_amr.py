"""The static AMR hierarchy (amr.Hierarchy) built from the tagger on the
seeded IC: composite volume bookkeeping, nesting, and subcycled stepping."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_hierarchy_volume_and_nesting(ctx):
    from axion_strings_adaptive_amd import amr, capi
    h = amr.Hierarchy(ctx, n0=32, levels=2, precision=capi.AXS_FP64,
                      verbose=False)
    assert len(h.levels) == 3
    for l in range(1, 3):
        for (lo, hi) in h.levels[l].boxes:
            n = [b - a + 1 for a, b in zip(lo, hi)]
            assert all(v % 16 == 0 and 16 <= v <= 64 for v in n)
            assert all(a % 16 == 0 for a in lo)
    # every comoving cell counted exactly once across the hierarchy
    e = h.composite_energy()
    assert e["n_total"] == pytest.approx((32 * 0.5) ** 3, rel=1e-12)
    e0 = e["rho_tot_unscreened"]
    for _ in range(3):
        h.step()
    e1 = h.composite_energy()
    assert e1["n_total"] == pytest.approx((32 * 0.5) ** 3, rel=1e-12)
    assert np.isfinite(e1["rho_tot_unscreened"]) and e1["rho_tot_unscreened"] > 0
    assert h.tau == pytest.approx(1.3, rel=1e-12)
    assert e0 > 0
    h.close()


def test_graph_replay_matches_direct_launches(ctx):
    from axion_strings_adaptive_amd import amr, capi
    out = []
    for use_graph in (False, True):
        h = amr.Hierarchy(ctx, n0=32, levels=2, precision=capi.AXS_FP32C,
                          verbose=False, use_graph=use_graph)
        for _ in range(3):
            h.step()
        out.append([L.download_dense(fill=0.0) for L in h.levels])
        h.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_fully_refined_hierarchy_tracks_uniform_fine_grid(ctx, oracle):
    """With force_full_refinement, level 1 covers the whole domain: its own
    periodic ghosts make it independent of level 0, so after one coarse
    step (two fine steps) it equals a uniform fine grid started from the
    prolonged IC."""
    from axion_strings_adaptive_amd import amr, capi
    h = amr.Hierarchy(ctx, n0=32, levels=1, precision=capi.AXS_FP64,
                      verbose=False, force_full_refinement=True)
    fine = h.levels[1]
    assert fine.num_cells == 64 ** 3
    s0 = fine.download_dense()
    ref = capi.Level(ctx, 64, 0.25, capi.AXS_FP64)
    ref.upload(s0)
    h.step()
    bg = h.bg
    tau, dtf = 1.0, 0.05
    for _ in range(2):
        ref.rk4_step([capi.rhs_coeffs(bg, t) for t in
                      (tau, tau + dtf / 2, tau + dtf / 2, tau + dtf)], dtf)
        tau += dtf
    assert np.array_equal(fine.download_dense(), ref.download())
    h.close()
