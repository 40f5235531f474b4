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


@pytest.mark.parametrize("precision", [0, 1])
def test_copy_valid_between_box_levels(ctx, precision):
    """axs_boxlevel_copy_valid copies exactly the valid cells both levels
    cover and leaves the rest of the destination alone."""
    from axion_strings_adaptive_amd import capi
    n = 64
    src_boxes = [((0, 0, 0), (31, 63, 63))]
    dst_boxes = [((16, 0, 0), (47, 31, 63)), ((48, 32, 0), (63, 63, 31))]
    src = capi.BoxLevel(ctx, src_boxes, (n, n, n), 0.5, precision)
    dst = capi.BoxLevel(ctx, dst_boxes, (n, n, n), 0.5, precision)
    rng = np.random.default_rng(3)
    a = rng.normal(size=(4, n, n, n))
    b = rng.normal(size=(4, n, n, n))
    src.upload_dense(a)
    dst.upload_dense(b)
    a = src.download_dense(fill=0.0)   # as stored (FP32C: hi + lo)
    b = dst.download_dense(fill=0.0)
    copied = dst.copy_valid_from(src)
    got = dst.download_dense(fill=0.0)
    assert copied == 16 * 32 * 64
    want = b.copy()
    want[:, :, 0:32, 16:32] = a[:, :, 0:32, 16:32]
    assert np.array_equal(got, want)
    src.close()
    dst.close()


def test_regrid_with_unchanged_tags_keeps_levels(ctx):
    from axion_strings_adaptive_amd import amr, capi
    h = amr.Hierarchy(ctx, n0=32, levels=1, precision=capi.AXS_FP64,
                      verbose=False, force_full_refinement=True)
    before = h.levels[1]
    s0 = before.download_dense(fill=0.0)
    h.regrid()
    assert h.levels[1] is before and h.nregrids == 1
    assert np.array_equal(h.levels[1].download_dense(fill=0.0), s0)
    h.close()


def test_regridding_hierarchy_steps_and_keeps_volume(ctx):
    """Regrid after coarse steps: cells the new level shares with the old
    one keep their values (not re-interpolated), the composite volume stays
    the domain's and the energies stay finite."""
    from axion_strings_adaptive_amd import amr, capi
    h = amr.Hierarchy(ctx, n0=32, levels=2, precision=capi.AXS_FP32C,
                      verbose=False, amp_thr=0.45)
    for _ in range(3):
        h.step()
        d_old = h.levels[1].download_dense(fill=np.nan)
        h.regrid()
        d_new = h.levels[1].download_dense(fill=np.nan)
        both = ~np.isnan(d_old) & ~np.isnan(d_new)
        assert both.any()
        assert np.array_equal(d_old[both], d_new[both])
        e = h.composite_energy()
        assert e["n_total"] == pytest.approx((32 * 0.5) ** 3, rel=1e-12)
        assert np.isfinite(e["rho_tot_unscreened"])
    assert h.nregrids == 3
    h2 = amr.Hierarchy(ctx, n0=32, levels=2, precision=capi.AXS_FP32C,
                       verbose=False, amp_thr=0.45, regrid_int=1)
    for _ in range(2):
        h2.step()
    assert h2.nregrids == 2
    h2.close()
    h.close()


def test_composite_level0_collapses_by_average_down(ctx):
    """collapse_to_level0: level 0 where nothing refines it, the 2x2x2
    average of the (recursively collapsed) finer level where something
    does; the hierarchy itself unchanged."""
    from axion_strings_adaptive_amd import amr, capi
    h = amr.Hierarchy(ctx, n0=32, levels=2, precision=capi.AXS_FP64,
                      verbose=False, regrid_int=1)
    h.step()  # regrid: fine data partly interpolated, not yet averaged down
    dense = [L.download_dense() for L in h.levels]
    want = dense[-1]
    for l in range(len(dense) - 2, -1, -1):
        d = dense[l].copy()
        nz, ny, nx = d.shape[1:]
        avg = want.reshape(4, nz, 2, ny, 2, nx, 2).mean(axis=(2, 4, 6))
        covered = ~np.isnan(avg[0])
        d[:, covered] = avg[:, covered]
        want = d
    got = h.composite_level0()
    g = got.download()
    got.close()
    np.testing.assert_allclose(g, want, rtol=1e-13, atol=1e-13)
    for L, d in zip(h.levels, dense):
        assert np.array_equal(L.download_dense(), d, equal_nan=True)
    ss, sc, tot, sumsq = h.composite_spectrum()
    assert (sc > 0).sum() > 0 and np.all(np.isfinite(ss))
    # the spectrum of the collapsed field, whose level is kept between calls
    lvl = capi.Level(ctx, 32, h.dx[0], capi.AXS_FP64)
    lvl.upload(g)
    R = capi.background_eval(h.bg, 0, h.tau)
    want_spec = lvl.spectrum(capi.AXS_FIELD_MASKED_A_DOT, capi.AXS_MASK_B,
                             0.8, R)
    lvl.close()
    np.testing.assert_allclose(ss, want_spec[0], rtol=1e-12,
                               atol=1e-14 * np.abs(want_spec[0]).max())
    assert np.array_equal(sc, want_spec[1])
    again = h.composite_spectrum()
    assert np.array_equal(again[0], ss) and np.array_equal(again[1], sc)
    h.step()  # a later call collapses the advanced hierarchy
    later = h.composite_spectrum()
    assert not np.array_equal(later[0], ss)
    h.close()


@pytest.mark.gpu
def test_schedule_gates_level_creation(ctx):
    """With a level schedule (the box plan's log(m_r/H) per level), a level
    appears at the first regrid after log(m_r/H) = 2 ln tau (c = 0) reaches
    its entry, not before; force_full_refinement bypasses the gate
    (AxionStringsLevel.cpp:1939-1983)."""
    from axion_strings_adaptive_amd import amr, capi
    bg = capi.make_background(2.0, c0=0.0)
    sched = [2 * np.log(1.15), 2 * np.log(1.35)]
    h = amr.Hierarchy(ctx, n0=32, levels=2, precision=capi.AXS_FP64,
                      verbose=False, bg=bg, tau0=1.0, regrid_int=1,
                      level_add_log_mr_over_h=sched)
    seen = [len(h.levels)]
    for _ in range(5):
        h.step()
        seen.append(len(h.levels))
    assert seen == [1, 1, 2, 2, 3, 3], seen
    e = h.composite_energy()
    assert np.isfinite(e["rho_tot_unscreened"])
    h.close()
    h = amr.Hierarchy(ctx, n0=32, levels=2, precision=capi.AXS_FP64,
                      verbose=False, bg=bg, tau0=1.0, regrid_int=1,
                      level_add_log_mr_over_h=sched[:1])
    for _ in range(5):
        h.step()
    assert len(h.levels) == 2  # no schedule entry for level 2
    h.close()
    h = amr.Hierarchy(ctx, n0=32, levels=1, precision=capi.AXS_FP64,
                      verbose=False, bg=bg, tau0=1.0,
                      level_add_log_mr_over_h=sched,
                      force_full_refinement=True)
    assert len(h.levels) == 2
    h.close()


@pytest.mark.gpu
@pytest.mark.parametrize("block,r", [(8, 2), (8, 0), (8, 8), (4, 3)])
def test_tagged_blocks_device_equals_host(ctx, block, r):
    """The regrid's block mask, formed on the device from device tags, is
    the host reduction's (tags grown by r cells, periodic, any per block)
    for sparse, medium and dense tags, including tags next to the seam."""
    import ctypes as C
    from axion_strings_adaptive_amd import capi
    lib = capi.load()
    u8 = C.POINTER(C.c_uint8)
    rng = np.random.default_rng(10 * block + r)
    nx, ny, nz = 64, 32, 48
    for p in (0.0005, 0.01, 0.2):
        tags = (rng.random((nz, ny, nx)) < p).astype(np.uint8)
        tags[0, 0, nx - 1] = 1  # grows across the periodic seams
        want = np.zeros((nz // block, ny // block, nx // block), np.uint8)
        capi.check(lib.axs_amr_tagged_blocks(
            tags.ctypes.data_as(u8), (C.c_int * 3)(nx, ny, nz), block, r,
            want.ctypes.data_as(u8)))
        tb = capi.DeviceBuffer(tags.size)
        tb.from_numpy(tags)
        ob = capi.DeviceBuffer(want.size)
        capi.check(lib.axs_amr_tagged_blocks_device(
            ctx.h, C.c_void_p(tb.ptr), (C.c_int * 3)(nx, ny, nz), block, r,
            C.c_void_p(ob.ptr)))
        got = ob.to_numpy(np.uint8).reshape(want.shape)
        assert np.array_equal(got, want)
        assert 0 < want.sum()
