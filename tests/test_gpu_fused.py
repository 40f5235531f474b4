"""Stage fusion: the two fused launches per RK4 step (stages 1+2 and 3+4,
the second stage's psi formed at every stencil point from the first stage's
inputs) give the same bits as one launch per stage -- every precision, pair
tile variant and z chunk, ragged shapes, odd and even step counts, direct
launches and graph replays, stage scalars that change between steps and
stages -- and FP64 matches the oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _capi():
    from axion_strings_adaptive_amd import capi
    return capi


def random_state(shape, seed=0):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    return rng.normal(size=(4, nz, ny, nx))


def _bg(capi):
    # c0 = 0 -> c1 = 2 at tau = 1.15: lambda and the curvature coefficient
    # change between steps and across the stages of a step
    return capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.15,
                                has_switch=True)


def _run(ctx, shape, precision, s0, nsteps, graph, fused, variant=-1, zc=0):
    capi = _capi()
    lvl = capi.Level(ctx, shape, 0.5, precision)
    lvl.set_fusion(fused)
    lvl.set_pair_tile(variant, zc)
    lvl.upload(s0)
    lvl.advance(_bg(capi), 1.0, 0.1, nsteps, use_graph=graph)
    hi, lo = lvl.download_raw()
    lvl.close()
    return hi, lo


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    assert (a[1] is None and b[1] is None) or np.array_equal(a[1], b[1])


@pytest.mark.parametrize("shape", [(64, 64, 64), (40, 24, 36), (4, 8, 12),
                                   (128, 16, 8), (36, 8, 8), (130, 10, 7),
                                   (9, 4, 4), (67, 13, 5)])
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_fused_equals_per_stage_shapes(ctx, shape, precision):
    s0 = random_state(shape, seed=sum(shape) + precision)
    for nsteps in (1, 2, 5):
        ref = _run(ctx, shape, precision, s0, nsteps, False, 0)
        for graph in (False, True):
            for mode in (1, 2):
                got = _run(ctx, shape, precision, s0, nsteps, graph, mode)
                _same(got, ref)


@pytest.mark.parametrize("variant", list(range(48)))
@pytest.mark.parametrize("zc", [0, 4, 16])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("mode", [1, 2])
def test_pair_variants_identical(ctx, variant, zc, precision, mode):
    shape = (256, 24, 20)  # several x tiles at every cells-per-lane width
    s0 = random_state(shape, seed=5)
    if mode == 2 and precision != 0 and variant % 6 == 4:
        # FP32 / FP32C pairs without an accumulator in 4 x 4 tiles are
        # refused: a GPU test of that tile gave wrong results, cause not
        # shown, so it is not launched
        with pytest.raises(_capi().AxsError, match="4 x 4"):
            _run(ctx, shape, precision, s0, 4, True, mode, variant, zc)
        return
    ref = _run(ctx, shape, precision, s0, 4, False, 0)
    got = _run(ctx, shape, precision, s0, 4, True, mode, variant, zc)
    _same(got, ref)


def test_fused_fp64_matches_oracle(ctx, oracle):
    """7 steps (three fused pairs and a per-stage last step) across the
    fat -> Moore switch, graph replay: bit-exact with the oracle."""
    capi = _capi()
    n, dx, dt = 32, 0.5, 0.1
    s0 = oracle.seeded_ic(n, seed=3)
    ob = oracle.bg(2.0, c0=0.0, c1=2.0, tau_switch=1.15, has_switch=True)
    want = s0.copy()
    oracle.advance(want, dx, ob, 1.0, dt, 7)
    lvl = capi.Level(ctx, n, dx, capi.AXS_FP64)
    lvl.upload(s0)
    lvl.advance(_bg(capi), 1.0, dt, 7, use_graph=True)
    assert np.array_equal(lvl.download(), want)
    lvl.set_fusion(2)
    lvl.upload(s0)
    lvl.advance(_bg(capi), 1.0, dt, 7, use_graph=True)
    assert np.array_equal(lvl.download(), want)


@pytest.mark.parametrize("shape", [(64, 64, 64), (40, 24, 36), (4, 8, 12),
                                   (128, 16, 8), (130, 10, 7), (61, 9, 5),
                                   (9, 4, 4), (67, 13, 5), (200, 5, 4),
                                   (256, 90, 9), (256, 30, 6)])
@pytest.mark.parametrize("precision", [1, 2])
def test_one_launch_step_equals_per_stage(ctx, shape, precision):
    """Fusion mode 3 (the whole RK4 step in one launch, FP32 / FP32C):
    the same bits as one launch per stage, odd and even step counts, graph
    replay, z chunks down to 4 planes."""
    s0 = random_state(shape, seed=sum(shape) + 7 * precision)
    for nsteps in (1, 2, 5):
        ref = _run(ctx, shape, precision, s0, nsteps, False, 0)
        for graph in (False, True):
            for zc in (0, 4, 16):
                # -1, 4: ring kernel, loads one / two planes ahead (6, 5:
                # the same with tiles in 4 x 8 patches; 8, 7: XCDs on
                # adjacent bands of tiles); 3: register-queue kernel (all
                # 64 x 11 aligned); 0, 1, 2: 60 x 8, 4, 12
                for v in (-1, 0, 1, 2, 3, 4, 5, 6, 7, 8):
                    got = _run(ctx, shape, precision, s0, nsteps, graph, 3, v,
                               zc)
                    _same(got, ref)


@pytest.mark.parametrize("precision", [1, 2])
def test_one_launch_step_long_z_chunks(ctx, precision):
    """The production z chunks (128 and 64 planes, several chunks along z,
    explicit and automatic) give the per-stage bits."""
    shape = (64, 22, 256)
    s0 = random_state(shape, seed=911 + precision)
    for nsteps in (1, 3):
        ref = _run(ctx, shape, precision, s0, nsteps, False, 0)
        for zc in (0, 64, 128, 96):
            for v in (-1, 3, 4, 5, 6, 7, 8):
                got = _run(ctx, shape, precision, s0, nsteps, True, 3, v, zc)
                _same(got, ref)


@pytest.mark.parametrize("precision", [1, 2])
def test_one_launch_step_xcd_bands(ctx, precision):
    """Variants 7 and 8 deal the tiles to the XCDs in super-bands of 8 x 32
    tiles; a grid of 4 x 10 tiles in x, y and 8 z chunks has 320 tiles:
    one full super-band of 256 and 64 tiles in launch order.  Both parts,
    and the other tile orders, give the per-stage bits."""
    shape = (256, 100, 64)
    s0 = random_state(shape, seed=313 + precision)
    ref = _run(ctx, shape, precision, s0, 2, False, 0)
    for v in (7, 8, 4, 5):
        got = _run(ctx, shape, precision, s0, 2, False, 3, v, 8)
        _same(got, ref)


def test_fused_timing_reports_pairs(ctx):
    capi = _capi()
    lvl = capi.Level(ctx, 64, 0.5, capi.AXS_FP32C)
    lvl.upload(random_state((64, 64, 64), seed=2))
    lvl.enable_timing(True)
    lvl.stage_times()
    for k in (4, 3):  # every step, an odd last one too, is one launch
        lvl.advance(_bg(capi), 1.0, 0.1, k)  # default: one launch per step
        ms, nsteps = lvl.stage_times()
        assert nsteps == k
        assert ms[0] > 0 and ms[1] == 0 and ms[2] == 0 and ms[3] == 0
    lvl.set_fusion(2)
    for k in (4, 3):
        lvl.advance(_bg(capi), 1.0, 0.1, k)
        ms, nsteps = lvl.stage_times()
        assert nsteps == k
        assert ms[0] > 0 and ms[1] > 0 and ms[2] == 0 and ms[3] == 0
    lvl.set_fusion(False)
    lvl.advance(_bg(capi), 1.0, 0.1, 2)
    ms, nsteps = lvl.stage_times()
    assert nsteps == 2 and min(ms) > 0
    lvl.close()


def test_fusion_setters_reject_bad_arguments(ctx):
    capi = _capi()
    lvl = capi.Level(ctx, 16, 0.5, capi.AXS_FP64)
    with pytest.raises(capi.AxsError):
        capi.check(capi.load().axs_level_set_stage_fusion(lvl.h, 4))
    with pytest.raises(capi.AxsError):  # one launch per step: FP32 only
        lvl.set_fusion(3)
    with pytest.raises(capi.AxsError):
        lvl.set_pair_tile(capi.load().axs_num_pair_variants(), 0)
    with pytest.raises(capi.AxsError):
        lvl.set_pair_tile(0, -1)
    lvl.close()


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("graph", [False, True])
def test_single_step_calls_equal_one_call(ctx, precision, graph):
    """N calls of one step each (advance and rk4_step) give the bits of one
    N-step call in every precision's default fusion mode; each single step
    is one fused launch (FP32 / FP32C: the step kernel alone in the timing
    slots; FP64: the two pair launches), and the state pointer follows the
    array the step wrote."""
    capi = _capi()
    shape = (40, 24, 36)
    s0 = random_state(shape, seed=55 + precision)
    bg = _bg(capi)
    dt, nsteps = 0.1, 5
    ref = capi.Level(ctx, shape, 0.5, precision)
    ref.upload(s0)
    ref.advance(bg, 1.0, dt, nsteps, use_graph=graph)
    want = ref.download_raw()
    ref.close()
    one = capi.Level(ctx, shape, 0.5, precision)
    one.upload(s0)
    ptrs = set()
    tau = 1.0  # accumulated as axs_rk4_advance accumulates it
    for i in range(nsteps):
        one.advance(bg, tau, dt, 1, use_graph=graph)
        ptrs.add(one.device_ptrs()[0])
        tau += dt
    _same(one.download_raw(), want)
    assert len(ptrs) == 2  # the state alternates between two arrays
    # rk4_step with the same stage scalars
    stp = capi.Level(ctx, shape, 0.5, precision)
    stp.upload(s0)
    tau = 1.0
    for i in range(nsteps):
        coeffs = [capi.rhs_coeffs(bg, t)
                  for t in (tau, tau + 0.5 * dt, tau + 0.5 * dt, tau + dt)]
        stp.rk4_step(coeffs, dt)
        tau += dt
    _same(stp.download_raw(), want)
    stp.enable_timing(True)
    stp.stage_times()
    stp.advance(bg, 1.5, dt, 1)
    ms, n = stp.stage_times()
    assert n == 1 and ms[0] > 0 and ms[2] == 0 and ms[3] == 0
    assert (ms[1] > 0) == (precision == capi.AXS_FP64)
    stp.close()
    one.close()
