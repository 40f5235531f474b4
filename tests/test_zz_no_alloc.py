"""axs_energy_sums_noalloc: compute_total_energy_sums over caller-owned
device state with its reduction partials in the library's fixed scratch
array (no device allocation).  Through the C ABI it must give the bits of
axs_energy_sums on the same state -- on two calls in a row, the second on a
smaller grid (the scratch keeps the first call's partials) -- and refuse a
grid beyond its scratch before any GPU work."""

import ctypes as C

import numpy as np
import pytest


def _capi():
    from axion_strings_adaptive_amd import capi
    return capi


def _noalloc(capi, lvl, R, lam, RpR, masking, out, stream):
    lib = capi.load()
    state, comp = lvl.device_ptrs()
    return lib.axs_energy_sums_noalloc(
        C.c_void_p(state), C.c_void_p(comp or None), lvl.nx, lvl.ny, lvl.nz,
        lvl.precision, lvl.dx, R, lam, RpR, C.byref(masking), None,
        C.c_void_p(out.ptr), C.c_void_p(stream))


@pytest.mark.gpu
def test_noalloc_energy_sums_equal_allocating_entry_point(ctx):
    capi = _capi()
    lib = capi.load()
    stream = lib.axs_ctx_stream(ctx.h)
    out = capi.DeviceBuffer(14 * 8)
    R, lam, RpR = 1.3, 0.8, 0.2
    # larger grid first, then smaller ones (stale partials must not leak)
    cases = [((128, 96, 80), capi.AXS_FP32C, 3, capi.AXS_MASK_B),
             ((40, 24, 36), capi.AXS_FP32C, 4, capi.AXS_MASK_B),
             ((64, 64, 64), capi.AXS_FP64, 5, capi.AXS_MASK_A),
             ((36, 20, 12), capi.AXS_FP32, 6, capi.AXS_MASK_NONE)]
    for shape, precision, seed, scheme in cases:
        nx, ny, nz = shape
        lvl = capi.Level(ctx, shape, 0.5, precision)
        rng = np.random.default_rng(seed)
        s = rng.normal(size=(4, nz, ny, nx))
        s[0] += 1.0  # cells on both sides of the scheme-B threshold
        lvl.upload(s)
        want = lvl.energy_sums(R, lam, RpR, scheme, 0.8)
        m = capi.Masking(scheme, 0.8)
        assert _noalloc(capi, lvl, R, lam, RpR, m, out, stream) == capi.AXS_OK
        ctx.synchronize()
        got = out.to_numpy(np.float64)[:14]
        assert np.array_equal(got, want), (shape, precision)
        lvl.close()


def test_noalloc_refuses_grids_beyond_its_scratch():
    """1025 x 1023 x 1023 needs 34816 energy blocks (> 32768): refused on
    the host before any GPU work, so this runs without a device."""
    capi = _capi()
    lib = capi.load()
    m = capi.Masking(capi.AXS_MASK_NONE, 0.8)
    rc = lib.axs_energy_sums_noalloc(None, None, 1025, 1023, 1023,
                                     capi.AXS_FP32, 0.5, 1.0, 1.0, 0.0,
                                     C.byref(m), None, None, None)
    assert rc == capi.AXS_ERR_ARG
    assert b"32768" in lib.axs_last_error()
