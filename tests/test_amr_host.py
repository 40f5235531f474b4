"""Host-side box building of the static AMR hierarchy (amr.py): the boxes
partition exactly the tagged blocks, respect the blocking factor and the
maximum grid size, and a fully tagged region becomes max-size boxes; the
hierarchy's box plan equals the oracle's."""

import numpy as np
import pytest

from axion_strings_adaptive_amd import amr


def _check_partition(blk, boxes, fb, max_blocks):
    cov = np.zeros(tuple(s * fb for s in blk.shape), bool)
    for lo, hi in boxes:
        for d in range(3):
            n = hi[d] - lo[d] + 1
            assert lo[d] % fb == 0 and n % fb == 0 and fb <= n <= fb * max_blocks
        sl = (slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1),
              slice(lo[0], hi[0] + 1))
        assert not cov[sl].any()
        cov[sl] = True
    assert np.array_equal(cov, np.kron(blk, np.ones((fb,) * 3, bool)))


def test_boxes_partition_random_tag_patterns():
    rng = np.random.default_rng(0)
    for _ in range(200):
        shape = tuple(int(v) for v in rng.integers(1, 9, 3))
        blk = rng.random(shape) < rng.random()
        _check_partition(blk, amr.boxes_from_blocks(blk, 16, 4), 16, 4)


def test_fully_tagged_region_gives_max_size_boxes():
    blk = np.ones((8, 8, 8), bool)
    boxes = amr.boxes_from_blocks(blk, 16, 4)
    assert len(boxes) == 8
    assert all(hi[d] - lo[d] + 1 == 64 for lo, hi in boxes for d in range(3))


def test_dilation_and_erosion_are_periodic():
    m = np.zeros((8, 8, 8), bool)
    m[0, 0, 0] = True
    d = amr.dilate_periodic(m, 1)
    assert d.sum() == 27 and d[7, 7, 7] and d[1, 1, 1]
    assert np.array_equal(amr.erode_periodic(d, 1), m)


def test_tagged_blocks_equals_dilate_then_block():
    rng = np.random.default_rng(1)
    for _ in range(60):
        block = int(rng.choice([2, 4, 8, 16]))
        r = int(rng.integers(0, block + 1))
        shape = tuple(block * int(v) for v in rng.integers(1, 5, 3))
        tags = rng.random(shape) < rng.choice([0.001, 0.01, 0.1])
        want = amr.blocks_from_tags(amr.dilate_periodic(tags, r), block)
        assert np.array_equal(amr.tagged_blocks(tags, block, r), want)


@pytest.mark.parametrize("args", [(1024, 1.0, 4.0, 2.0, 0.0, 2),
                                  (4096, 1.2, 8.0, 2.0, 0.0, 4),
                                  (4097, 1.2, 8.0, 2.0, 0.0, 4),
                                  (8192, 1.0, 6.0, 2.5, 0.0, 3),
                                  (512, 1.0, 4.0, 2.0, 0.5, 1)])
def test_amr_box_plan_equals_oracle(oracle, args):
    """The hierarchy's box plan (base grid, spacings, tau_f, level schedule)
    is the oracle's compute_amr_box_plan exactly."""
    p = amr.amr_box_plan(*args)
    L, dxf, dxb, tau_f, nb, la = oracle.amr_box_plan(*args)
    assert (p["L_tilde"], p["dx_finest"], p["dx_base"], p["tau_f"],
            p["N_base"], p["log_add"]) == (L, dxf, dxb, tau_f, nb, la)


# ---- the native host steps (axs_amr_*) against the numpy restatements ----
def _native():
    import ctypes as C
    from axion_strings_adaptive_amd import capi
    lib = capi.load()
    u8 = C.POINTER(C.c_uint8)

    def tagged_blocks(tags, block, r):
        t = np.ascontiguousarray(tags, dtype=np.uint8)
        nz, ny, nx = t.shape
        out = np.zeros((nz // block, ny // block, nx // block), np.uint8)
        capi.check(lib.axs_amr_tagged_blocks(
            t.ctypes.data_as(u8), (C.c_int * 3)(nx, ny, nz), block, r,
            out.ctypes.data_as(u8)))
        return out.astype(bool)

    def erode(mask, r):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        bz, by, bx = m.shape
        out = np.zeros_like(m)
        capi.check(lib.axs_amr_erode(m.ctypes.data_as(u8),
                                     (C.c_int * 3)(bx, by, bz), r,
                                     out.ctypes.data_as(u8)))
        return out.astype(bool)

    def boxes(blk, fb, max_blocks):
        b = np.ascontiguousarray(blk, dtype=np.uint8)
        bz, by, bx = b.shape
        cap = b.size
        lo, hi = (C.c_int * (3 * cap + 3))(), (C.c_int * (3 * cap + 3))()
        n = C.c_int()
        capi.check(lib.axs_amr_boxes_from_blocks(
            b.ctypes.data_as(u8), (C.c_int * 3)(bx, by, bz), fb, max_blocks,
            cap, C.byref(n), lo, hi))
        return [(tuple(lo[3 * i:3 * i + 3]), tuple(hi[3 * i:3 * i + 3]))
                for i in range(n.value)]

    return tagged_blocks, erode, boxes


def test_native_tagged_blocks_equal_numpy():
    tagged_blocks, _, _ = _native()
    rng = np.random.default_rng(11)
    for _ in range(60):
        block = int(rng.choice([2, 4, 8, 16]))
        r = int(rng.integers(0, block + 1))
        shape = tuple(block * int(v) for v in rng.integers(1, 5, 3))
        tags = rng.random(shape) < rng.choice([0.001, 0.01, 0.1])
        assert np.array_equal(tagged_blocks(tags, block, r),
                              amr.tagged_blocks(tags, block, r))


def test_native_erode_equals_numpy():
    _, erode, _ = _native()
    rng = np.random.default_rng(12)
    for _ in range(40):
        shape = tuple(int(v) for v in rng.integers(1, 9, 3))
        m = rng.random(shape) < rng.random()
        for r in (0, 1, 2):
            assert np.array_equal(erode(m, r), amr.erode_periodic(m, r))


def test_native_boxes_equal_numpy():
    _, _, boxes = _native()
    rng = np.random.default_rng(13)
    for _ in range(200):
        shape = tuple(int(v) for v in rng.integers(1, 9, 3))
        blk = rng.random(shape) < rng.random()
        fb, mb = int(rng.choice([8, 16, 32])), int(rng.integers(1, 5))
        got = boxes(blk, fb, mb)
        assert got == amr.boxes_from_blocks(blk, fb, mb)
        _check_partition(blk, got, fb, mb)
