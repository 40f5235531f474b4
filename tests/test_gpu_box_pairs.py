"""Fused stage pairs on box levels (axs_boxlevel_set_stage_fusion(bl, 1)):
the two launches per RK4 step give the bits of one launch per stage --
every precision, tilings with boxes of 8 to 64 cells, coarse-fine ghosts
of a partly refined level stepped in two subcycles per coarse step,
direct launches and graph replays (whose state alternates between two
arrays) -- and every entry point reads the state where the step left
it."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capi():
    from axion_strings_adaptive_amd import capi
    return capi


def tiling(n, edges):
    cuts, pos, i = [], 0, 0
    while pos < n:
        e = min(edges[i % len(edges)], n - pos)
        cuts.append((pos, pos + e - 1))
        pos += e
        i += 1
    return [((x0, y0, z0), (x1, y1, z1)) for (z0, z1) in cuts
            for (y0, y1) in cuts for (x0, x1) in cuts]


def _coeffs(capi, bg, tau, dt):
    return [capi.rhs_coeffs(bg, t) for t in (tau, tau + dt / 2, tau + dt / 2,
                                             tau + dt)]


def _state(n, seed):
    rng = np.random.default_rng(seed)
    s = rng.normal(size=(4, n, n, n)) * 0.3
    s[0] += 1.0
    return s


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("edges", [[32], [16, 48], [24, 8, 64]])
@pytest.mark.parametrize("use_graph", [False, True])
def test_pairs_equal_per_stage_tiled(ctx, precision, edges, use_graph):
    capi = _capi()
    n = 64 if edges != [24, 8, 64] else 96
    s0 = _state(n, 17 + precision)
    bg = capi.make_background(2.0, c0=1.0)
    out = {}
    for mode in (0, 1):
        bl = capi.BoxLevel(ctx, tiling(n, edges), (n, n, n), 0.5, precision)
        bl.set_fusion(mode)
        assert bl.fusion() == mode
        bl.upload_dense(s0)
        tau, dt = 1.0, 0.1
        for _ in range(3):  # odd: the pairs' state ends in the other array
            bl.rk4_step(_coeffs(capi, bg, tau, dt), dt, use_graph=use_graph)
            tau += dt
        bl.fill_ghosts()  # the energy sums read the ghosts
        out[mode] = (bl.download_dense(fill=0.0),
                     bl.energy_sums(1.3, 1.0, 0.1))
        bl.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def _two_levels(capi, ctx, precision, mode, use_graph, steps=2):
    nc = 32
    sc = _state(nc, 5)
    bg = capi.make_background(2.0, c0=1.0)
    coarse = capi.BoxLevel(ctx, tiling(nc, [16]), (nc,) * 3, 0.5, precision)
    coarse.upload_dense(sc)
    # partial refinement: boxes at the periodic seam, a pair that touch,
    # and one inside, each with coarse-fine faces
    fine_boxes = [((0, 0, 0), (15, 15, 15)),
                  ((16, 16, 16), (31, 31, 47)), ((32, 16, 16), (47, 31, 31)),
                  ((40, 40, 40), (63, 55, 63))]
    fine = capi.BoxLevel(ctx, fine_boxes, (2 * nc,) * 3, 0.25, precision)
    fine.fill_ghosts(coarse, 1.0, interior=True)
    coarse.set_fusion(mode)
    fine.set_fusion(mode)
    tau, dtc = 1.0, 0.1
    for _ in range(steps):
        fine.snapshot_coarse(coarse)
        coarse.rk4_step(_coeffs(capi, bg, tau, dtc), dtc, use_graph=use_graph)
        for half in (0, 1):
            tf = tau + 0.5 * half * dtc
            fine.rk4_step(_coeffs(capi, bg, tf, 0.5 * dtc), 0.5 * dtc, coarse,
                          0.5 * half, 0.5, use_graph=use_graph)
        tau += dtc
    res = (coarse.download_dense(fill=0.0), fine.download_dense(fill=0.0))
    fine.close()
    coarse.close()
    return res


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("use_graph", [False, True])
def test_pairs_equal_per_stage_coarse_fine(ctx, precision, use_graph):
    """Ghost cells on coarse-fine faces take the coarse interpolation at
    each stage's time on both paths; three coarse steps (six fine steps)
    alternate both levels' state arrays."""
    capi = _capi()
    ref = _two_levels(capi, ctx, precision, 0, use_graph, steps=3)
    got = _two_levels(capi, ctx, precision, 1, use_graph, steps=3)
    assert np.array_equal(ref[0], got[0])
    assert np.array_equal(ref[1], got[1])


def test_pairs_direct_equal_graph(ctx):
    capi = _capi()
    a = _two_levels(capi, ctx, capi.AXS_FP32C, 1, False)
    b = _two_levels(capi, ctx, capi.AXS_FP32C, 1, True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_pairs_state_followed_by_other_entry_points(ctx):
    """After an odd number of fused-pair steps: average-down, copies to a
    uniform level, ghost fills and tags read the advanced state."""
    capi = _capi()
    n = 64
    s0 = _state(n, 23)
    bg = capi.make_background(2.0, c0=1.0)
    res = {}
    for mode in (0, 1):
        bl = capi.BoxLevel(ctx, tiling(n, [32]), (n,) * 3, 0.5, capi.AXS_FP32C)
        bl.set_fusion(mode)
        bl.upload_dense(s0)
        bl.rk4_step(_coeffs(capi, bg, 1.0, 0.1), 0.1)
        lvl = capi.Level(ctx, n, 0.5, capi.AXS_FP32C)
        bl.copy_to_level(lvl)
        coarse = capi.BoxLevel(ctx, tiling(n // 2, [16]), (n // 2,) * 3, 1.0,
                               capi.AXS_FP32C)
        bl.average_down_to(coarse)
        bl.fill_ghosts()
        _, counts = bl.tag(amp_thr=0.5)
        res[mode] = (lvl.download(), coarse.download_dense(fill=0.0),
                     tuple(counts))
        coarse.close()
        lvl.close()
        bl.close()
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]


def test_fusion_setter_rejects_bad_modes(ctx):
    capi = _capi()
    bl = capi.BoxLevel(ctx, tiling(32, [16]), (32,) * 3, 0.5)
    with pytest.raises(capi.AxsError, match="fused stage pairs"):
        bl.set_fusion(2)
    bl.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_pairs_after_state_writes(ctx, precision):
    """The fills of the fused pairs read their neighbours' x-face cells from
    packed x-edge arrays that the second pair writes for the next step; every
    other write of the state (an upload, an average-down onto the level, a
    copy from a uniform level, a regrid copy, a step with one launch per
    stage, an interior fill from a coarse level) makes the next step rebuild
    them.  Each such write followed by steps gives the per-stage bits."""
    capi = _capi()
    n = 64
    bg = capi.make_background(2.0, c0=1.0)
    boxes = tiling(n, [32, 16])

    def run(mode):
        bl = capi.BoxLevel(ctx, boxes, (n,) * 3, 0.5, precision)
        bl.set_fusion(mode)
        bl.upload_dense(_state(n, 31))
        clock = [1.0]

        def step(k=1):
            for _ in range(k):
                dt = 0.1
                bl.rk4_step(_coeffs(capi, bg, clock[0], dt), dt)
                clock[0] += dt

        step(2)
        bl.upload_dense(_state(n, 32))
        step(2)
        fine = capi.BoxLevel(ctx, [((0, 0, 0), (31, 31, 31)),
                                   ((64, 32, 0), (95, 63, 31))],
                             (2 * n,) * 3, 0.25, precision)
        fine.upload_dense(_state(2 * n, 33))
        fine.average_down_to(bl)
        fine.close()
        step()
        lvl = capi.Level(ctx, n, 0.5, precision)
        lvl.upload(_state(n, 34))
        bl.copy_from_level(lvl)
        lvl.close()
        step()
        other = capi.BoxLevel(ctx, tiling(n, [16]), (n,) * 3, 0.5, precision)
        other.upload_dense(_state(n, 35))
        bl.copy_valid_from(other)
        other.close()
        step(2)
        bl.set_fusion(0)
        step()
        bl.set_fusion(mode)
        step()
        coarse = capi.BoxLevel(ctx, tiling(n // 2, [16]), (n // 2,) * 3, 1.0,
                               precision)
        coarse.upload_dense(_state(n // 2, 36))
        bl.fill_ghosts(coarse, 1.0, interior=True)
        coarse.close()
        step(2)
        out = bl.download_dense(fill=0.0)
        bl.close()
        return out

    assert np.array_equal(run(0), run(1))


@pytest.mark.parametrize("use_graph", [False, True])
def test_hierarchy_pairs_equal_per_stage(ctx, use_graph):
    """A whole AMR hierarchy (two refined levels, subcycling, regrids every
    coarse step, average-down) stepped with the fused pairs gives the bits
    of one launch per stage on every level."""
    from axion_strings_adaptive_amd import amr
    capi = _capi()
    out = {}
    for per_stage in (True, False):
        h = amr.Hierarchy(ctx, n0=32, levels=2, precision=capi.AXS_FP32C,
                          verbose=False, regrid_int=1, use_graph=use_graph,
                          per_stage_launches=per_stage)
        for _ in range(3):
            h.step()
        out[per_stage] = ([L.boxes for L in h.levels],
                          [L.download_dense(fill=0.0) for L in h.levels])
        h.close()
    assert out[True][0] == out[False][0]
    for a, b in zip(out[True][1], out[False][1]):
        assert np.array_equal(a, b)
