"""The entry points over caller-owned box data (axs_rhs_array4,
axs_tag_array4): on views laid out as AMReX's Array4 -- grown boxes with
periodic ghosts filled by the caller, sub-boxes, component strides with
padding -- they give the bits of the level entry points (axs_rhs_eval,
axs_tag_cells), write nothing outside the box, and reject views that do not
hold the stencil."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _view(capi, t, begin, shape, ncomp, pad=0):
    """Array4 of a device tensor holding [ncomp][nz][ny][nx] (+ pad elements
    between components)."""
    nx, ny, nz = shape
    a = capi.Array4()
    a.p = t.data_ptr()
    a.jstride, a.kstride = nx, nx * ny
    a.nstride = nx * ny * nz + pad
    a.begin = (C.c_int * 3)(*begin)
    a.end = (C.c_int * 3)(*(b + n for b, n in zip(begin, shape)))
    a.ncomp = ncomp
    return a


def _grown(s, g, pad, dtype):
    """[4][n+2g]^3 periodic copy of s with `pad` elements after each
    component, flattened."""
    w = np.pad(s, ((0, 0), (g, g), (g, g), (g, g)), mode="wrap").astype(dtype)
    flat = np.zeros(4 * (w[0].size + pad), dtype)
    for c in range(4):
        flat[c * (w[0].size + pad):c * (w[0].size + pad) + w[0].size] = \
            w[c].ravel()
    return flat


@pytest.mark.parametrize("precision", [0, 2])
def test_rhs_array4_equals_level_rhs(ctx, precision):
    from axion_strings_adaptive_amd import capi
    import torch
    n, g, pad = 40, 3, 17
    dt = np.float64 if precision == capi.AXS_FP64 else np.float32
    rng = np.random.default_rng(5)
    s = 1.0 + 0.3 * rng.standard_normal((4, n, n, n))
    lvl = capi.Level(ctx, n, 0.5, precision)
    lvl.upload(s)
    stored = s.astype(dt)  # what the level holds (FP32: rounded)
    bg = capi.make_background(2.0, c0=1.0)
    c = capi.rhs_coeffs(bg, 1.7)
    want = lvl.rhs(c)
    st = _dev(_grown(stored, g, pad, dt))
    sv = _view(capi, st, (-g, -g, -g), (n + 2 * g,) * 3, 4, pad)
    lib = capi.load()
    for lo, hi in [((0, 0, 0), (n - 1,) * 3), ((8, 5, 3), (30, 39, 20)),
                   ((0, 33, 38), (39, 39, 39))]:
        # the rhs view: the grown box, NaN outside lo..hi afterwards
        rt = torch.full(((n + 2 * g) ** 3 * 4,), float("nan"),
                        dtype=torch.float64 if dt is np.float64
                        else torch.float32, device="cuda")
        rv = _view(capi, rt, (-g, -g, -g), (n + 2 * g,) * 3, 4)
        capi.check(lib.axs_rhs_array4(
            C.byref(sv), C.byref(rv), (C.c_int * 3)(*lo), (C.c_int * 3)(*hi),
            C.byref(c), 0.5, precision, None))
        torch.cuda.synchronize()
        got = rt.cpu().numpy().reshape(4, n + 2 * g, n + 2 * g, n + 2 * g)
        box = (slice(None), slice(lo[2] + g, hi[2] + g + 1),
               slice(lo[1] + g, hi[1] + g + 1), slice(lo[0] + g, hi[0] + g + 1))
        ref = want[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        assert np.array_equal(got[box], ref), (lo, hi)
        outside = np.ones(got.shape, bool)
        outside[box] = False
        assert np.all(np.isnan(got[outside]))
    # a view one ghost short of the star is refused
    sv2 = _view(capi, st, (-1, -1, -1), (n + 2,) * 3, 4, pad)
    with pytest.raises(capi.AxsError, match="lo-2"):
        capi.check(lib.axs_rhs_array4(
            C.byref(sv2), C.byref(rv), (C.c_int * 3)(0, 0, 0),
            (C.c_int * 3)(n - 1, n - 1, n - 1), C.byref(c), 0.5, precision,
            None))
    lvl.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("criteria", [False, True])
def test_tag_array4_equals_level_tagger(ctx, oracle, precision, criteria):
    from axion_strings_adaptive_amd import capi
    import torch
    n, g = 32, 3
    lvl = capi.Level(ctx, n, 0.5, precision)
    s0 = oracle.seeded_ic(n, seed=9)
    lvl.upload(s0)
    thr = dict(grad_thr=0.02, rgrad_thr=0.3, amp_thr=0.5) if criteria else {}
    R = 1.3
    want_tags, want_counts = lvl.tag(R=R, **thr)
    if precision == capi.AXS_FP64:
        hi = lvl.download()
        lo_arr = None
        dt = np.float64
    else:
        hi_raw = np.zeros(4 * n ** 3, np.float32)
        lo_raw = np.zeros(4 * n ** 3, np.float32)
        capi.check(capi.load().axs_level_download_raw(
            lvl.h, hi_raw.ctypes.data_as(C.c_void_p),
            lo_raw.ctypes.data_as(C.c_void_p)))
        hi = hi_raw.reshape(4, n, n, n)
        lo_arr = lo_raw.reshape(4, n, n, n)
        dt = np.float32
    st = _dev(_grown(hi, g, 0, dt))
    sv = _view(capi, st, (-g,) * 3, (n + 2 * g,) * 3, 4)
    lv = None
    if lo_arr is not None:
        lt = _dev(_grown(lo_arr, g, 0, np.float32))
        lv = _view(capi, lt, (-g,) * 3, (n + 2 * g,) * 3, 4)
    tags = torch.zeros(n ** 3, dtype=torch.uint8, device="cuda")
    tv = _view(capi, tags, (0, 0, 0), (n, n, n), 1)
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    p = capi.TaggerParams(thr.get("grad_thr", capi.DBL_MAX),
                          thr.get("rgrad_thr", capi.DBL_MAX),
                          thr.get("amp_thr", 0.0), R)
    capi.check(capi.load().axs_tag_array4(
        C.byref(sv), C.byref(lv) if lv is not None else None,
        (C.c_int * 3)(0, 0, 0), (C.c_int * 3)(n - 1, n - 1, n - 1),
        C.byref(p), 0.5, capi.AXS_FP64 if precision == 0 else capi.AXS_FP32,
        C.byref(tv), C.c_void_p(counts.data_ptr()), None))
    torch.cuda.synchronize()
    assert tuple(counts.cpu().numpy()) == want_counts
    assert np.array_equal(tags.cpu().numpy().reshape(n, n, n), want_tags)
    assert want_counts[0] > 0
    # counts only, on a sub-box: the counts of that box's cells
    counts.zero_()
    lo_b, hi_b = (4, 0, 10), (20, 31, 17)
    capi.check(capi.load().axs_tag_array4(
        C.byref(sv), C.byref(lv) if lv is not None else None,
        (C.c_int * 3)(*lo_b), (C.c_int * 3)(*hi_b), None, 0.5,
        capi.AXS_FP64 if precision == 0 else capi.AXS_FP32, None,
        C.c_void_p(counts.data_ptr()), None))
    torch.cuda.synchronize()
    w = oracle.windings_field(s0 if precision == 0 else
                              (hi.astype(np.float64) + lo_arr))
    sub = w[lo_b[2]:hi_b[2] + 1, lo_b[1]:hi_b[1] + 1, lo_b[0]:hi_b[0] + 1]
    assert counts.cpu().numpy()[0] == int(np.count_nonzero(sub))
    assert counts.cpu().numpy()[1] == int(np.abs(sub).sum())
    lvl.close()
