"""AMR box levels (axs_boxlevel_*): the work-list stage kernel and ghost
fills against the uniform level and closed forms, tagging / energies /
coverage / average-down against the oracle and numpy."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capi():
    from axion_strings_adaptive_amd import capi
    return capi


def tiling(n, edges):
    """Boxes tiling a periodic n^3 domain, x/y/z edges from `edges` cyclic."""
    boxes = []
    cuts = []
    pos = 0
    i = 0
    while pos < n:
        e = min(edges[i % len(edges)], n - pos)
        cuts.append((pos, pos + e - 1))
        pos += e
        i += 1
    for (z0, z1) in cuts:
        for (y0, y1) in cuts:
            for (x0, x1) in cuts:
                boxes.append(((x0, y0, z0), (x1, y1, z1)))
    return boxes


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("edges", [[32], [16, 48], [24, 8, 64]])
def test_tiled_boxes_match_uniform_level(ctx, oracle, precision, edges):
    """Boxes tiling the periodic domain see the same neighbours as the
    uniform level, so the same arithmetic gives the same bits."""
    capi = _capi()
    n = 64 if edges != [24, 8, 64] else 96
    if n & (n - 1) == 0:
        s0 = oracle.seeded_ic(n, seed=3)
    else:  # the oracle's FFT is power-of-two only
        s0 = 1.0 + 0.3 * np.random.default_rng(3).normal(size=(4, n, n, n))
    bg = capi.make_background(2.0, c0=1.0)
    lvl = capi.Level(ctx, n, 0.5, precision)
    lvl.upload(s0)
    bl = capi.BoxLevel(ctx, tiling(n, edges), (n, n, n), 0.5, precision)
    bl.upload_dense(s0)
    tau, dt = 1.0, 0.1
    for _ in range(3):
        c = [capi.rhs_coeffs(bg, t) for t in (tau, tau + dt / 2, tau + dt / 2,
                                              tau + dt)]
        lvl.rk4_step(c, dt)
        bl.rk4_step(c, dt)
        tau += dt
    assert np.array_equal(bl.download_dense(), lvl.download())


def test_copy_between_uniform_and_box_levels(ctx, oracle):
    capi = _capi()
    n = 32
    s0 = oracle.seeded_ic(n, seed=4)
    lvl = capi.Level(ctx, n, 0.5, capi.AXS_FP32C)
    lvl.upload(s0)
    bl = capi.BoxLevel(ctx, tiling(n, [16]), (n, n, n), 0.5, capi.AXS_FP32C)
    bl.copy_from_level(lvl)
    out = capi.Level(ctx, n, 0.5, capi.AXS_FP32C)
    bl.copy_to_level(out)
    assert np.array_equal(out.download(), lvl.download())
    assert np.array_equal(bl.download_dense(), lvl.download())


def test_quartic_prolongation_is_exact_on_polynomials(ctx):
    """Tensor-product 5-point Lagrange interpolation reproduces polynomials
    of degree <= 4 per direction exactly (away from the periodic seam)."""
    capi = _capi()
    nc = 32
    coarse = capi.BoxLevel(ctx, tiling(nc, [16]), (nc,) * 3, 1.0)
    xc = np.arange(nc) + 0.5  # coarse cell centres, coarse units
    X, Y, Z = xc[None, None, :], xc[None, :, None], xc[:, None, None]

    def poly(X, Y, Z, c):
        return (1.0 + 0.1 * c) * (0.3 + 0.02 * X ** 4 - 0.1 * X * Y ** 2 +
                                  0.001 * (Y ** 3) * Z + 0.05 * Z ** 2 * X ** 2)

    dense = np.stack([np.broadcast_to(poly(X, Y, Z, c), (nc,) * 3)
                      for c in range(4)])
    coarse.upload_dense(dense)
    coarse.fill_ghosts()
    fine_boxes = [((16, 16, 16), (31, 31, 47)), ((32, 16, 16), (47, 31, 31))]
    fine = capi.BoxLevel(ctx, fine_boxes, (2 * nc,) * 3, 0.5)
    fine.fill_ghosts(coarse, 1.0, interior=True)
    got = fine.download_dense()
    xf = (np.arange(2 * nc) + 0.5) / 2.0
    XF, YF, ZF = xf[None, None, :], xf[None, :, None], xf[:, None, None]
    for (lo, hi) in fine_boxes:
        sl = (slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1),
              slice(lo[0], hi[0] + 1))
        for c in range(4):
            want = np.broadcast_to(poly(XF, YF, ZF, c), (2 * nc,) * 3)[sl]
            np.testing.assert_allclose(got[c][sl], want, rtol=1e-12, atol=1e-12)


def test_unnested_fill_is_reported(ctx):
    capi = _capi()
    fine = capi.BoxLevel(ctx, [((0, 0, 0), (15, 15, 15))], (32, 32, 32), 0.5)
    with pytest.raises(capi.AxsError, match="no owning box"):
        fine.fill_ghosts()


@pytest.mark.parametrize("use_graph", [False, True])
def test_unnested_step_is_reported(ctx, use_graph):
    """A box level that does not tile its domain, stepped without a coarse
    level: the ghost fills find cells with no owning box, and both the
    direct and the graph-replayed step return AXS_ERR_ARG (replayed twice:
    the count is cleared before every replay), while a properly tiled level
    steps cleanly on the same path."""
    capi = _capi()
    coeffs = [capi.rhs_coeffs(capi.make_background(2.0), t)
              for t in (1.0, 1.05, 1.05, 1.1)]
    fine = capi.BoxLevel(ctx, [((0, 0, 0), (15, 15, 15))], (32, 32, 32), 0.5)
    for _ in range(2):
        with pytest.raises(capi.AxsError, match="no owning box"):
            fine.rk4_step(coeffs, 0.1, use_graph=use_graph)
    fine.close()
    ok = capi.BoxLevel(ctx, tiling(32, [16]), (32, 32, 32), 0.5)
    for _ in range(2):
        ok.rk4_step(coeffs, 0.1, use_graph=use_graph)
    ok.close()


def test_box_tag_and_energy_with_coverage(ctx, oracle):
    capi = _capi()
    n = 32
    lvl = capi.Level(ctx, n, 0.5, capi.AXS_FP64)
    lvl.upload(oracle.seeded_ic(n, seed=9))
    lvl.advance(capi.make_background(2.0), 1.0, 0.1, 20)
    s = lvl.download()
    bl = capi.BoxLevel(ctx, tiling(n, [16]), (n, n, n), 0.5)
    bl.copy_from_level(lvl)
    bl.fill_ghosts()
    tags, counts = bl.tag()
    assert counts == oracle.count_plaquettes(s)
    t = tags.to_numpy(np.uint8).reshape(n, n, n)
    want = oracle.tag_cells(s, 0.5, capi.DBL_MAX, capi.DBL_MAX)
    assert np.array_equal(t, want)
    # a finer level covering part of the domain -> coverage mask
    fine = capi.BoxLevel(ctx, [((16, 0, 8), (47, 31, 23))], (2 * n,) * 3, 0.25)
    cover = fine.coverage_mask((n, n, n))
    m = cover.to_numpy(np.uint8).reshape(n, n, n)
    want_m = np.ones((n, n, n), np.uint8)
    want_m[4:12, 0:16, 8:24] = 0
    assert np.array_equal(m, want_m)
    e = bl.energy_sums(2.0, 1.0, 0.5, capi.AXS_MASK_B, 0.8, cover=cover)
    np.testing.assert_allclose(
        e, oracle.energy_sums(s, 0.5, 2.0, 1.0, 0.5, 2, 0.8, mask=m),
        rtol=1e-11)
    _, c2 = bl.tag(cover=cover)
    assert c2 == oracle.count_plaquettes(s, m)


def test_average_down(ctx):
    capi = _capi()
    nc = 16
    rng = np.random.default_rng(0)
    fine_boxes = [((8, 8, 8), (23, 15, 31))]
    fine = capi.BoxLevel(ctx, fine_boxes, (2 * nc,) * 3, 0.25)
    dense_f = rng.normal(size=(4, 2 * nc, 2 * nc, 2 * nc))
    fine.upload_dense(dense_f)
    coarse = capi.BoxLevel(ctx, tiling(nc, [8]), (nc,) * 3, 0.5)
    dense_c = rng.normal(size=(4, nc, nc, nc))
    coarse.upload_dense(dense_c)
    fine.average_down_to(coarse)
    got = coarse.download_dense()
    want = dense_c.copy()
    blk = dense_f[:, 8:32, 8:16, 8:24].reshape(4, 12, 2, 4, 2, 8, 2)
    want[:, 4:16, 4:8, 4:12] = blk.mean(axis=(2, 4, 6))
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-14)


def test_two_level_subcycled_step_tracks_uniform_fine(ctx, oracle):
    """A fine level covering the whole domain, stepped twice per coarse step
    with coarse-fine ghosts, equals the uniform fine evolution (every ghost
    has a same-level owner, so the coarse level only supplies nothing)."""
    capi = _capi()
    nf = 32
    s0 = oracle.seeded_ic(nf, seed=12)
    fine = capi.BoxLevel(ctx, tiling(nf, [16]), (nf,) * 3, 0.25)
    fine.upload_dense(s0)
    coarse = capi.BoxLevel(ctx, tiling(nf // 2, [16]), (nf // 2,) * 3, 0.5)
    ref = capi.Level(ctx, nf, 0.25)
    ref.upload(s0)
    bg = capi.make_background(2.0)
    tau, dtf = 1.0, 0.05
    for half in range(2):
        c = [capi.rhs_coeffs(bg, t) for t in (tau, tau + dtf / 2,
                                              tau + dtf / 2, tau + dtf)]
        fine.rk4_step(c, dtf, coarse, 0.5 * half, 0.5)
        ref.rk4_step(c, dtf)
        tau += dtf
    assert np.array_equal(fine.download_dense(), ref.download())


def test_recycled_arrays_start_clean(ctx, oracle):
    """A box level built on arrays the context kept from a destroyed one
    starts from zero state and steps exactly like one on fresh arrays; trim
    releases the kept arrays.  192^3 in 64^3 boxes: arrays of ~150 MB,
    above the size the context keeps."""
    capi = _capi()
    n = 192
    boxes = tiling(n, [64])
    s0 = np.random.default_rng(7).normal(size=(4, n, n, n))
    coeffs = [capi.RhsCoeffs(0.0, 1.0, 1.0)] * 4
    first = capi.BoxLevel(ctx, boxes, (n, n, n), 0.5, capi.AXS_FP32C)
    first.upload_dense(s0)
    first.fill_ghosts()
    first.rk4_step(coeffs, 0.1)
    want = first.download_dense()
    first.close()
    for trim in (False, True):
        if trim:
            ctx.trim()
        b = capi.BoxLevel(ctx, boxes, (n, n, n), 0.5, capi.AXS_FP32C)
        assert np.array_equal(b.download_dense(), np.zeros_like(want))
        b.upload_dense(s0)
        b.fill_ghosts()
        b.rk4_step(coeffs, 0.1)
        assert np.array_equal(b.download_dense(), want)
        b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_box_tag_derivative_criteria(ctx, oracle, precision):
    """StringTagger's Laplacian and radial-gradient criteria on a box level
    (StringTagger.hpp:83-218) equal the oracle's tags on the same state."""
    capi = _capi()
    n = 32
    lvl = capi.Level(ctx, n, 0.5, precision)
    lvl.upload(oracle.seeded_ic(n, seed=11))
    lvl.advance(capi.make_background(2.0), 1.0, 0.1, 10)
    s = lvl.download()
    bl = capi.BoxLevel(ctx, tiling(n, [16, 8, 8]), (n, n, n), 0.5, precision)
    bl.copy_from_level(lvl)
    bl.fill_ghosts()
    for g, r in ((0.05, capi.DBL_MAX), (capi.DBL_MAX, 0.3), (0.1, 0.4)):
        tags, counts = bl.tag(grad_thr=g, rgrad_thr=r)
        t = tags.to_numpy(np.uint8).reshape(n, n, n)
        want = oracle.tag_cells(s, 0.5, g, r)
        assert np.array_equal(t, want), (g, r)
        assert 0 < t.sum() < t.size
        assert counts == oracle.count_plaquettes(s)
