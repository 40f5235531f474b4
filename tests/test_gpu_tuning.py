"""The tuning table (axs_tuning_load) on the device: levels and box levels
created after a load take the table's launch shapes, a level switched to
another stage fusion mode takes that mode's line, every tuned shape gives
the built-in shapes' bits, and the box levels' work shape
(axs_boxlevel_set_work_shape) changes the item list without changing a
bit."""

import numpy as np
import pytest

from test_gpu_boxes import tiling

pytestmark = pytest.mark.gpu


@pytest.fixture
def table(tmp_path):
    """Writes a table, loads it, and puts the packaged one back after."""
    from axion_strings_adaptive_amd import capi
    import os

    def load(text):
        p = tmp_path / "t.txt"
        p.write_text(text)
        capi.load_tuning(p)

    yield load
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    capi.load_tuning(capi.TUNING_PATH if os.path.exists(capi.TUNING_PATH)
                     else empty)


def _steps(capi, lvl, s0, nsteps=3):
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    lvl.upload(s0)
    lvl.advance(bg, 1.0, 0.1, nsteps)
    return lvl.download()


@pytest.mark.parametrize("precision", [0, 1])
def test_levels_take_the_table_shapes(ctx, oracle, table, precision):
    from axion_strings_adaptive_amd import capi
    n = 64
    s0 = oracle.seeded_ic(n, seed=5)
    table("level 64 64 64 1 3 -1 0 0 16\n"      # FP32C one launch: 60 x 8
          "level 64 64 64 1 1 5 16 9 16\n"      # FP32C pairs: 8 x 2 + prefetch
          "level 64 64 64 0 1 -1 0 6 32\n"      # FP64 pairs: 4 x 1 + prefetch
          "level 64 64 64 0 0 3 16 -1 0\n"      # FP64 per stage
          "level 32 32 32 0 1 -1 0 2 8\n")      # another shape
    lvl = capi.Level(ctx, n, 0.5, precision)
    sh = lvl.launch_shape()
    if precision == capi.AXS_FP32C:
        assert sh == {"fusion": 3, "tile_variant": -1, "tile_z_chunk": 0,
                      "pair_variant": 0, "pair_z_chunk": 16}
    else:
        assert sh == {"fusion": 1, "tile_variant": -1, "tile_z_chunk": 0,
                      "pair_variant": 6, "pair_z_chunk": 32}
    tuned = _steps(capi, lvl, s0)
    # another fusion mode: that mode's line, or the shapes kept
    lvl.set_fusion(1 if precision == capi.AXS_FP32C else 0)
    sh2 = lvl.launch_shape()
    if precision == capi.AXS_FP32C:
        assert (sh2["pair_variant"], sh2["pair_z_chunk"]) == (9, 16)
        assert (sh2["tile_variant"], sh2["tile_z_chunk"]) == (5, 16)
    else:
        assert (sh2["tile_variant"], sh2["tile_z_chunk"]) == (3, 16)
    other = _steps(capi, lvl, s0)
    lvl.close()
    table("")  # built-in shapes
    ref = capi.Level(ctx, n, 0.5, precision)
    assert ref.launch_shape()["pair_variant"] == -1
    builtin = _steps(capi, ref, s0)
    ref.close()
    assert np.array_equal(tuned, builtin)
    assert np.array_equal(other, builtin)


@pytest.mark.parametrize("precision", [0, 1])
def test_box_work_shapes_give_the_same_bits(ctx, oracle, table, precision):
    from axion_strings_adaptive_amd import capi
    n = 64
    s0 = oracle.seeded_ic(n, seed=6)
    boxes = tiling(n, [16, 8, 40])
    bg = capi.make_background(2.0, c0=1.0)
    c = [capi.rhs_coeffs(bg, t) for t in (1.0, 1.05, 1.05, 1.1)]
    table("")  # built-in shapes first

    def run(shape=None):
        bl = capi.BoxLevel(ctx, boxes, (n, n, n), 0.5, precision)
        if shape:
            bl.set_work_shape(*shape)
        got = bl.work_shape()
        bl.upload_dense(s0)
        for _ in range(2):
            bl.rk4_step(c, 0.1)
        bl.rk4_step(c, 0.1, use_graph=True)
        out = bl.download_dense()
        bl.close()
        return got, out

    (r0, g0, n0), base = run()
    assert (r0, g0) == (512, 2048)
    # items of whole-width rows: W x H <= 512 with H >= 4
    def items(b, rows):
        w, ny = b[1][0] - b[0][0] + 1, b[1][1] - b[0][1] + 1
        h = max(4, min(rows, 512 // w))
        return -(-ny // h)
    want = sum(items(b, 512) for b in boxes)
    assert n0 == want
    for shape in [(4, 7), (8, 2048), (16, 100), (1, 65536)]:
        (r, g, ni), out = run(shape)
        assert (r, g) == shape
        assert ni == sum(items(b, r) for b in boxes) >= n0
        assert np.array_equal(out, base), shape
    table(f"box {precision} 4 33\n")
    (r, g, ni), out = run()
    assert (r, g) == (4, 33) and ni > n0
    assert np.array_equal(out, base)
    bl = capi.BoxLevel(ctx, boxes, (n, n, n), 0.5, precision)
    with pytest.raises(capi.AxsError, match="item_rows"):
        bl.set_work_shape(0, 64)
    with pytest.raises(capi.AxsError, match="grid_blocks"):
        bl.set_work_shape(8, 0)
    bl.close()
