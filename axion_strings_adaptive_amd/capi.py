"""Thin ctypes binding over libaxs_hip.so's C ABI (include/axs.h).

This is test/bench plumbing: it loads the in-tree HIP library, declares the
C signatures and wraps a level in a small Python class.  It never computes
anything itself and raises if the library is missing -- there is no CPU
fallback.

Loading the library also loads the autotuner's launch-shape table
(tune_table.txt next to this file, or the file AXS_TUNING_TABLE names; set
it to the empty string for the built-in shapes), so that levels created
through this binding run the tuned tiles.
"""

from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libaxs_hip.so")
TUNING_PATH = os.path.join(HERE, "tune_table.txt")

AXS_OK = 0
AXS_ERR_ARG, AXS_ERR_HIP, AXS_ERR_NOMEM, AXS_ERR_UNSUPPORTED = 1, 2, 3, 4
AXS_FP64, AXS_FP32C, AXS_FP32 = 0, 1, 2
AXS_MASK_NONE, AXS_MASK_A, AXS_MASK_B = 0, 1, 2
AXS_FIELD_MASKED_A_DOT, AXS_FIELD_PI1, AXS_FIELD_PI2 = 0, 1, 2
DBL_MAX = np.finfo(np.float64).max


class AxsError(RuntimeError):
    pass


class Background(C.Structure):
    _fields_ = [("a_inv", C.c_double), ("c0", C.c_double), ("c1", C.c_double),
                ("tau_switch", C.c_double), ("has_switch", C.c_int),
                ("flat_space", C.c_int), ("flat_m_r", C.c_double)]


class RhsCoeffs(C.Structure):
    _fields_ = [("curvature_coeff", C.c_double), ("lambda_", C.c_double),
                ("R_squared", C.c_double)]


class TaggerParams(C.Structure):
    _fields_ = [("gradient_threshold", C.c_double),
                ("radial_gradient_threshold", C.c_double),
                ("amplitude_threshold", C.c_double), ("R", C.c_double)]


class Array4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("jstride", C.c_int64),
                ("kstride", C.c_int64), ("nstride", C.c_int64),
                ("begin", C.c_int * 3), ("end", C.c_int * 3),
                ("ncomp", C.c_int)]


AXS_MAX_LEVELS = 8


class AmrParams(C.Structure):
    _fields_ = [("max_level", C.c_int), ("n_error_buf", C.c_int),
                ("blocking_factor", C.c_int), ("max_grid_size", C.c_int),
                ("regrid_int", C.c_int), ("force_full_refinement", C.c_int),
                ("gradient_threshold", C.c_double),
                ("radial_gradient_threshold", C.c_double),
                ("amplitude_threshold", C.c_double),
                ("n_schedule", C.c_int),
                ("level_add_log_mr_over_h", C.c_double * AXS_MAX_LEVELS),
                ("use_graph", C.c_int), ("per_stage_launches", C.c_int)]


class CoeffBlend(C.Structure):
    _fields_ = [("tau_start", C.c_double), ("dtau", C.c_double),
                ("lambda_start", C.c_double),
                ("curvature_start", C.c_double)]


class Masking(C.Structure):
    _fields_ = [("scheme", C.c_int), ("threshold", C.c_double)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
_SIGS = {
    "axs_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "axs_ctx_destroy": (_I, [_P]),
    "axs_ctx_synchronize": (_I, [_P]),
    "axs_ctx_trim": (_I, [_P]),
    "axs_ctx_stream": (_P, [_P]),
    "axs_device_alloc": (_I, [_P, C.c_size_t, C.POINTER(_P)]),
    "axs_device_free": (_I, [_P, _P]),
    "axs_device_memset": (_I, [_P, _P, _I, C.c_size_t]),
    "axs_device_download": (_I, [_P, _P, _P, C.c_size_t]),
    "axs_last_error": (C.c_char_p, []),
    "axs_num_tile_variants": (_I, []),
    "axs_tuning_load": (_I, [C.c_char_p]),
    "axs_background_eval": (_I, [C.POINTER(Background), _I, _D, _PD]),
    "axs_background_rhs_coeffs": (_I, [C.POINTER(Background), _D,
                                       C.POINTER(RhsCoeffs)]),
    "axs_level_create": (_I, [_P, _I, _I, _I, _D, _I, C.POINTER(_P)]),
    "axs_level_destroy": (_I, [_P]),
    "axs_level_upload": (_I, [_P, _PD]),
    "axs_level_download": (_I, [_P, _PD]),
    "axs_level_download_raw": (_I, [_P, _P, _P]),
    "axs_level_upload_raw": (_I, [_P, _P, _P]),
    "axs_level_device_ptrs": (_I, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "axs_level_set_homogeneous": (_I, [_P, _D, _D]),
    "axs_level_handoff_rescale": (_I, [_P, _D, _D, _D]),
    "axs_level_set_tile": (_I, [_P, _I, _I]),
    "axs_num_pair_variants": (_I, []),
    "axs_level_set_stage_fusion": (_I, [_P, _I]),
    "axs_level_set_pair_tile": (_I, [_P, _I, _I]),
    "axs_level_get_tile": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "axs_level_get_launch": (_I, [_P, C.POINTER(_I), C.POINTER(_I),
                                  C.POINTER(_I)]),
    "axs_rhs_eval": (_I, [_P, C.POINTER(RhsCoeffs), _P]),
    "axs_rk4_step": (_I, [_P, C.POINTER(RhsCoeffs), _D]),
    "axs_rk4_advance": (_I, [_P, C.POINTER(Background), _D, _D, _I, _I]),
    "axs_rk4_advance_coeffs": (_I, [_P, C.POINTER(RhsCoeffs), _D, _I, _I]),
    "axs_rk4_advance_windings": (_I, [_P, C.POINTER(Background), _D, _D, _I,
                                      _I, C.POINTER(TaggerParams), _P, _PI64,
                                      _I]),
    "axs_level_place": (_I, [_P, _I, C.POINTER(_D), C.POINTER(_I)]),
    "axs_level_enable_timing": (_I, [_P, _I]),
    "axs_level_stage_times": (_I, [_P, _PD, _PI64]),
    "axs_count_plaquettes": (_I, [_P, _P, _PI64]),
    "axs_velocity_at_pierced_corners": (_I, [_P, _D, _D, _D, _P, _PD, _PI64,
                                             _PI64]),
    "axs_windings_field": (_I, [_P, _P]),
    "axs_tag_cells": (_I, [_P, C.POINTER(TaggerParams), _P, _PI64]),
    "axs_energy_sums": (_I, [_P, _D, _D, _D, C.POINTER(Masking), _P, _PD]),
    "axs_energy_sums_noalloc": (_I, [_P, _P, _I, _I, _I, _I, _D, _D, _D, _D,
                                     C.POINTER(Masking), _P, _P, _P]),
    "axs_fill_field": (_I, [_P, _I, C.POINTER(Masking), _D, _P]),
    "axs_level_spectrum": (_I, [_P, _I, C.POINTER(Masking), _D, _PD, _PI64,
                                _PD, _PD]),
    "axs_level_fourier_ic": (_I, [_P, _I, _D, _D, C.c_uint64, _I]),
    "axs_level_zero_components": (_I, [_P, _I, _I]),
    "axs_level_compare": (_I, [_P, _P, _PD, _PD]),
    "axs_boxlevel_create": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I),
                                 C.POINTER(_I), _I, _D, _I, C.POINTER(_P)]),
    "axs_boxlevel_destroy": (_I, [_P]),
    "axs_boxlevel_num_cells": (C.c_int64, [_P]),
    "axs_boxlevel_set_work_shape": (_I, [_P, _I, _I]),
    "axs_boxlevel_set_stage_fusion": (_I, [_P, _I]),
    "axs_boxlevel_get_stage_fusion": (_I, [_P, C.POINTER(_I)]),
    "axs_boxlevel_get_work_shape": (_I, [_P, C.POINTER(_I), C.POINTER(_I),
                                         C.POINTER(_I)]),
    "axs_boxlevel_upload_dense": (_I, [_P, _PD]),
    "axs_boxlevel_download_dense": (_I, [_P, _PD]),
    "axs_boxlevel_copy_from_level": (_I, [_P, _P]),
    "axs_boxlevel_copy_to_level": (_I, [_P, _P]),
    "axs_boxlevel_copy_valid": (_I, [_P, _P, C.POINTER(C.c_int64)]),
    "axs_boxlevel_snapshot_coarse": (_I, [_P, _P]),
    "axs_boxlevel_fill_ghosts": (_I, [_P, _P, _D, _I]),
    "axs_boxlevel_rk4_step": (_I, [_P, C.POINTER(RhsCoeffs), _D, _P, _D, _D]),
    "axs_boxlevel_rk4_step_graph": (_I, [_P, C.POINTER(RhsCoeffs), _D, _P, _D,
                                         _D]),
    "axs_boxlevel_tag": (_I, [_P, C.POINTER(TaggerParams), _P, _P, _PI64]),
    "axs_boxlevel_energy_sums": (_I, [_P, _D, _D, _D, C.POINTER(Masking), _P,
                                      _PD]),
    "axs_rhs_array4": (_I, [C.POINTER(Array4), C.POINTER(Array4),
                            C.POINTER(_I), C.POINTER(_I),
                            C.POINTER(RhsCoeffs), _D, _I, _P]),
    "axs_tag_array4": (_I, [C.POINTER(Array4), C.POINTER(Array4),
                            C.POINTER(_I), C.POINTER(_I),
                            C.POINTER(TaggerParams), _D, _I, C.POINTER(Array4),
                            _P, _P]),
    "axs_hierarchy_create": (_I, [_P, _P, C.POINTER(AmrParams),
                                  C.POINTER(Background), C.POINTER(CoeffBlend),
                                  _D, _D, C.POINTER(_P)]),
    "axs_hierarchy_destroy": (_I, [_P]),
    "axs_hierarchy_step": (_I, [_P]),
    "axs_hierarchy_regrid": (_I, [_P]),
    "axs_hierarchy_info": (_I, [_P, C.POINTER(_I), _PD, _PI64, _PI64, _PD]),
    "axs_hierarchy_level": (_I, [_P, _I, C.POINTER(_P), C.POINTER(_I), _PD,
                                 C.POINTER(_I)]),
    "axs_hierarchy_boxes": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I)]),
    "axs_hierarchy_energy_sums": (_I, [_P, _D, _D, _D, C.POINTER(Masking),
                                       _PD]),
    "axs_hierarchy_plaquette_counts": (_I, [_P, _PI64]),
    "axs_hierarchy_collapse": (_I, [_P, _P]),
    "axs_amr_tagged_blocks": (_I, [_P, C.POINTER(_I), _I, _I, _P]),
    "axs_amr_tagged_blocks_device": (_I, [_P, _P, C.POINTER(_I), _I, _I, _P]),
    "axs_amr_erode": (_I, [_P, C.POINTER(_I), _I, _P]),
    "axs_amr_boxes_from_blocks": (_I, [_P, C.POINTER(_I), _I, _I, _I,
                                       C.POINTER(_I), C.POINTER(_I),
                                       C.POINTER(_I)]),
    "axs_boxlevel_coverage_mask": (_I, [_P, C.POINTER(_I), _P]),
    "axs_boxlevel_average_down": (_I, [_P, _P]),
    "axs_adjoint_grad_energy": (_I, [_P, C.POINTER(Background), _D, _D, _D,
                                     _I, _I, _PD, _PD, _PD]),
    "axs_forward_grad_energy": (_I, [_P, C.POINTER(Background), _D, _D, _D,
                                     _I, _PD]),
}

_lib: Optional[C.CDLL] = None


def declared_symbols() -> list[str]:
    return list(_SIGS)


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads libaxs_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise AxsError(f"{path} not built: run __graft_entry__.build() or "
                       "python -m axion_strings_adaptive_amd.build")
    # torch (which provides the device buffers of this binding) must bring
    # up its HIP runtime before the library is loaded, so that both share
    # one runtime instance in the process
    import torch
    torch.cuda.is_available()
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    table = os.environ.get("AXS_TUNING_TABLE", TUNING_PATH)
    if table and (table != TUNING_PATH or os.path.exists(table)):
        check(lib.axs_tuning_load(table.encode()))
    _lib = lib
    return lib


def load_tuning(path: str, lib: C.CDLL | None = None) -> None:
    """Loads a tuning table (autotune.py's output) into the library; levels
    and box levels created afterwards take their launch shapes from it."""
    lib = lib or load()
    check(lib.axs_tuning_load(str(path).encode()))


def check(rc: int) -> None:
    if rc != AXS_OK:
        msg = load().axs_last_error().decode(errors="replace")
        raise AxsError(f"axs error {rc}: {msg}")


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_PD)


def make_background(a_inv=2.0, c0=0.0, c1=0.0, tau_switch=0.0,
                    has_switch=False, flat_space=False, flat_m_r=1.0):
    return Background(a_inv, c0, c1, tau_switch, int(has_switch),
                      int(flat_space), flat_m_r)


def background_eval(bg: Background, which: int, x: float) -> float:
    out = C.c_double()
    check(load().axs_background_eval(C.byref(bg), which, x, C.byref(out)))
    return out.value


def rhs_coeffs(bg: Background, tau: float) -> RhsCoeffs:
    c = RhsCoeffs()
    check(load().axs_background_rhs_coeffs(C.byref(bg), tau, C.byref(c)))
    return c


_live_contexts: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _close_all_contexts():
    # Before interpreter teardown: levels first, then their contexts, while
    # the HIP runtime is still loaded.
    for c in list(_live_contexts):
        c.close()


class Context:
    def __init__(self, device: int = 0):
        self.lib = load()
        h = _P()
        check(self.lib.axs_ctx_create(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()
        _live_contexts.add(self)

    def adopt(self, obj):
        self._children.add(obj)

    def synchronize(self):
        check(self.lib.axs_ctx_synchronize(self.h))

    def trim(self):
        """Frees the device arrays kept from destroyed box levels."""
        check(self.lib.axs_ctx_trim(self.h))

    def close(self):
        """Destroys every level still alive on this context, then itself."""
        if self.h:
            for child in list(self._children):
                child.close()
            self.lib.axs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Raw device memory owned through torch (plumbing only)."""

    def __init__(self, nbytes: int):
        import torch
        self.t = torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")
        self.ptr = self.t.data_ptr()
        self.nbytes = nbytes

    def to_numpy(self, dtype) -> np.ndarray:
        import torch
        torch.cuda.synchronize()
        return self.t.cpu().numpy()[: self.nbytes].view(dtype).copy()

    def zero(self):
        import torch
        self.t.zero_()
        torch.cuda.synchronize()

    def from_numpy(self, a: np.ndarray):
        import torch
        b = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
        self.t[: b.size].copy_(torch.from_numpy(b))
        torch.cuda.synchronize()


class Level:
    """A uniform periodic level (axs_level) owned by the HIP library."""

    def __init__(self, ctx: Context, n, dx: float, precision: int = AXS_FP64):
        if isinstance(n, int):
            n = (n, n, n)
        self.ctx = ctx
        self.lib = ctx.lib
        self.nx, self.ny, self.nz = n
        self.dx = dx
        self.precision = precision
        h = _P()
        check(self.lib.axs_level_create(ctx.h, self.nx, self.ny, self.nz, dx,
                                        precision, C.byref(h)))
        self.h = h
        ctx.adopt(self)

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def word(self) -> int:
        return 8 if self.precision == AXS_FP64 else 4

    def upload(self, state: np.ndarray):
        s = np.ascontiguousarray(state, dtype=np.float64).reshape(-1)
        assert s.size == 4 * self.ncell
        check(self.lib.axs_level_upload(self.h, dptr(s)))

    def download(self) -> np.ndarray:
        out = np.empty(4 * self.ncell, dtype=np.float64)
        check(self.lib.axs_level_download(self.h, dptr(out)))
        return out.reshape(4, self.nz, self.ny, self.nx)

    def download_raw(self):
        dt = np.float64 if self.precision == AXS_FP64 else np.float32
        hi = np.empty(4 * self.ncell, dtype=dt)
        lo = np.empty(4 * self.ncell, dtype=np.float32) \
            if self.precision == AXS_FP32C else None
        check(self.lib.axs_level_download_raw(
            self.h, hi.ctypes.data_as(_P),
            lo.ctypes.data_as(_P) if lo is not None else None))
        return hi, lo

    def set_homogeneous(self, R: float, Rp: float):
        check(self.lib.axs_level_set_homogeneous(self.h, R, Rp))

    def place(self, trials: int = 4):
        """Placement trials (state reset to zero): list of ms per step of
        each placement tried; the level keeps the last one."""
        ms = (_D * trials)()
        used = C.c_int()
        check(self.lib.axs_level_place(self.h, trials, ms, C.byref(used)))
        return [ms[i] for i in range(used.value)]

    def device_ptrs(self):
        """(state, comp) device addresses; comp is 0 unless FP32C."""
        st, cp = C.c_void_p(), C.c_void_p()
        check(self.lib.axs_level_device_ptrs(self.h, C.byref(st), C.byref(cp)))
        return st.value or 0, cp.value or 0

    def handoff_rescale(self, kappa: float, common_term: float, R_pre: float):
        check(self.lib.axs_level_handoff_rescale(self.h, kappa, common_term,
                                                 R_pre))

    def set_tile(self, variant: int, z_chunk: int = 0):
        check(self.lib.axs_level_set_tile(self.h, variant, z_chunk))

    def set_fusion(self, on):
        """RK4 stages in two fused launches (1 / True, the default), the same
        without a stored accumulator (2), or one launch per stage (0 /
        False); the same bits every way."""
        check(self.lib.axs_level_set_stage_fusion(self.h, int(on)))

    def set_pair_tile(self, variant: int, z_chunk: int = 0):
        check(self.lib.axs_level_set_pair_tile(self.h, variant, z_chunk))

    def launch_shape(self) -> dict:
        """The level's launch settings: stage fusion mode, stage tile variant
        and z chunk, pair variant and pair z chunk."""
        v, z, f, pv, pz = (C.c_int() for _ in range(5))
        check(self.lib.axs_level_get_tile(self.h, C.byref(v), C.byref(z)))
        check(self.lib.axs_level_get_launch(self.h, C.byref(f), C.byref(pv),
                                            C.byref(pz)))
        return {"fusion": f.value, "tile_variant": v.value,
                "tile_z_chunk": z.value, "pair_variant": pv.value,
                "pair_z_chunk": pz.value}

    def rhs(self, coeffs: RhsCoeffs) -> np.ndarray:
        buf = DeviceBuffer(4 * self.ncell * self.word)
        check(self.lib.axs_rhs_eval(self.h, C.byref(coeffs), _P(buf.ptr)))
        self.ctx.synchronize()
        dt = np.float64 if self.precision == AXS_FP64 else np.float32
        return buf.to_numpy(dt).reshape(4, self.nz, self.ny, self.nx)

    def rk4_step(self, coeffs, dt: float):
        arr = (RhsCoeffs * 4)(*coeffs)
        check(self.lib.axs_rk4_step(self.h, arr, dt))

    def advance(self, bg: Background, tau: float, dt: float, nsteps: int,
                use_graph: bool = False):
        check(self.lib.axs_rk4_advance(self.h, C.byref(bg), tau, dt, nsteps,
                                       int(use_graph)))

    def advance_windings(self, bg: Background, tau: float, dt: float,
                         nsteps: int, every: int, tags=None, grad_thr=DBL_MAX,
                         rgrad_thr=DBL_MAX, amp_thr=0.0,
                         use_graph: bool = False) -> np.ndarray:
        """nsteps steps with the windings (and, given a tags DeviceBuffer,
        the tagger) after every `every` steps; returns the [segment,
        (plain, weighted)] counts."""
        nseg = nsteps // every
        out = np.zeros((max(nseg, 1), 2), np.int64)
        p = TaggerParams(grad_thr, rgrad_thr, amp_thr, 1.0)
        check(self.lib.axs_rk4_advance_windings(
            self.h, C.byref(bg), tau, dt, nsteps, every, C.byref(p),
            _P(tags.ptr) if tags is not None else None,
            out.ctypes.data_as(C.POINTER(C.c_int64)), int(use_graph)))
        return out[:nseg]

    def advance_coeffs(self, coeffs, dt: float, use_graph: bool = False):
        """len(coeffs) // 4 RK4 steps; coeffs[4 i + s] = stage s of step i."""
        assert len(coeffs) % 4 == 0
        arr = (RhsCoeffs * len(coeffs))(*coeffs)
        check(self.lib.axs_rk4_advance_coeffs(self.h, arr, dt,
                                              len(coeffs) // 4,
                                              int(use_graph)))

    def enable_timing(self, on: bool = True):
        check(self.lib.axs_level_enable_timing(self.h, int(on)))

    def stage_times(self):
        """(ms per stage summed over the timed steps, number of steps)."""
        ms = np.zeros(4)
        n = C.c_int64()
        check(self.lib.axs_level_stage_times(self.h, dptr(ms), C.byref(n)))
        return ms, n.value

    def count_plaquettes(self, mask: Optional[DeviceBuffer] = None):
        out = (C.c_int64 * 2)()
        check(self.lib.axs_count_plaquettes(
            self.h, _P(mask.ptr) if mask is not None else None, out))
        return int(out[0]), int(out[1])

    def velocity(self, R, RpR, m_r, mask: Optional[DeviceBuffer] = None):
        """(sum gamma^2 v^2, corner count, (n_p_plain, n_p_weighted))."""
        s = C.c_double()
        n = C.c_int64()
        counts = (C.c_int64 * 2)()
        check(self.lib.axs_velocity_at_pierced_corners(
            self.h, R, RpR, m_r, _P(mask.ptr) if mask is not None else None,
            C.byref(s), C.byref(n), counts))
        return s.value, n.value, (int(counts[0]), int(counts[1]))

    def windings_field(self) -> np.ndarray:
        buf = DeviceBuffer(3 * self.ncell)
        check(self.lib.axs_windings_field(self.h, _P(buf.ptr)))
        self.ctx.synchronize()
        return buf.to_numpy(np.int8).reshape(self.nz, self.ny, self.nx, 3)

    def tag(self, grad_thr=DBL_MAX, rgrad_thr=DBL_MAX, amp_thr=0.0, R=1.0,
            tags: Optional[DeviceBuffer] = None):
        if tags is None:
            tags = DeviceBuffer(self.ncell)
            tags.zero()
        p = TaggerParams(grad_thr, rgrad_thr, amp_thr, R)
        counts = (C.c_int64 * 2)()
        check(self.lib.axs_tag_cells(self.h, C.byref(p), _P(tags.ptr), counts))
        self.ctx.synchronize()
        return (tags.to_numpy(np.uint8).reshape(self.nz, self.ny, self.nx),
                (int(counts[0]), int(counts[1])))

    def energy_sums(self, R, lam, RpR, scheme=AXS_MASK_NONE, thr=0.8,
                    mask: Optional[DeviceBuffer] = None) -> np.ndarray:
        out = np.zeros(14)
        m = Masking(scheme, thr)
        check(self.lib.axs_energy_sums(
            self.h, R, lam, RpR, C.byref(m),
            _P(mask.ptr) if mask is not None else None, dptr(out)))
        return out

    def fill_field(self, kind, scheme=AXS_MASK_NONE, thr=0.8, R=1.0):
        w = 8 if self.precision == AXS_FP64 else 4
        buf = DeviceBuffer(self.ncell * w)
        m = Masking(scheme, thr)
        check(self.lib.axs_fill_field(self.h, kind, C.byref(m), R,
                                      _P(buf.ptr)))
        self.ctx.synchronize()
        dt = np.float64 if w == 8 else np.float32
        return buf.to_numpy(dt).reshape(self.nz, self.ny, self.nx)

    def spectrum(self, kind, scheme=AXS_MASK_NONE, thr=0.8, R=1.0):
        n_shells = max(self.nx, self.ny, self.nz) // 2 + 1
        ss = np.zeros(n_shells)
        sc = np.zeros(n_shells, dtype=np.int64)
        tot = np.zeros(2)
        sumsq = C.c_double()
        m = Masking(scheme, thr)
        check(self.lib.axs_level_spectrum(
            self.h, kind, C.byref(m), R, dptr(ss),
            sc.ctypes.data_as(_PI64), dptr(tot), C.byref(sumsq)))
        return ss, sc, tot, sumsq.value

    def fourier_ic(self, comp, k_max_cells, target_ms, seed, comp_seed_offset):
        check(self.lib.axs_level_fourier_ic(self.h, comp, k_max_cells,
                                            target_ms, seed, comp_seed_offset))

    def zero_components(self, c0, ncomp):
        check(self.lib.axs_level_zero_components(self.h, c0, ncomp))

    def compare(self, other: "Level"):
        a, b = C.c_double(), C.c_double()
        check(self.lib.axs_level_compare(self.h, other.h, C.byref(a),
                                         C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            self.lib.axs_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BoxLevel:
    """An AMR level stored as boxes with 3 ghost layers (axs_boxlevel)."""

    def __init__(self, ctx: Context, boxes, domain, dx: float,
                 precision: int = AXS_FP64):
        """boxes: list of ((lo_x, lo_y, lo_z), (hi_x, hi_y, hi_z)), inclusive."""
        self.ctx = ctx
        self.lib = ctx.lib
        self.domain = tuple(domain)
        self.dx = dx
        self.precision = precision
        self.boxes = [tuple(map(tuple, b)) for b in boxes]
        lo = (C.c_int * (3 * len(boxes)))(*[v for b in boxes for v in b[0]])
        hi = (C.c_int * (3 * len(boxes)))(*[v for b in boxes for v in b[1]])
        dom = (C.c_int * 3)(*domain)
        h = _P()
        check(self.lib.axs_boxlevel_create(ctx.h, len(boxes), lo, hi, dom, 3,
                                           dx, precision, C.byref(h)))
        self.h = h
        self.owned = True
        ctx.adopt(self)

    @classmethod
    def borrowed(cls, ctx: Context, handle: int, boxes, domain, dx: float,
                 precision: int) -> "BoxLevel":
        """A view of a box level that another object (an AMR hierarchy)
        owns: close() forgets the handle without destroying the level."""
        self = cls.__new__(cls)
        self.ctx, self.lib = ctx, ctx.lib
        self.domain, self.dx, self.precision = tuple(domain), dx, precision
        self.boxes = [tuple(map(tuple, b)) for b in boxes]
        self.h = _P(handle)
        self.owned = False
        return self

    @property
    def num_cells(self) -> int:
        return int(self.lib.axs_boxlevel_num_cells(self.h))

    def set_work_shape(self, item_rows: int, grid_blocks: int):
        check(self.lib.axs_boxlevel_set_work_shape(self.h, item_rows,
                                                   grid_blocks))

    def work_shape(self) -> tuple:
        """(item_rows, grid_blocks, number of work items)."""
        r, g, n = C.c_int(), C.c_int(), C.c_int()
        check(self.lib.axs_boxlevel_get_work_shape(self.h, C.byref(r),
                                                   C.byref(g), C.byref(n)))
        return r.value, g.value, n.value

    def set_fusion(self, mode: int):
        """RK4 launches: 0 one per stage, 1 fused stage pairs."""
        check(self.lib.axs_boxlevel_set_stage_fusion(self.h, mode))

    def fusion(self) -> int:
        m = C.c_int()
        check(self.lib.axs_boxlevel_get_stage_fusion(self.h, C.byref(m)))
        return m.value

    def upload_dense(self, dense: np.ndarray):
        d = np.ascontiguousarray(dense, dtype=np.float64)
        check(self.lib.axs_boxlevel_upload_dense(self.h, dptr(d)))

    def download_dense(self, fill=np.nan) -> np.ndarray:
        nx, ny, nz = self.domain
        d = np.full((4, nz, ny, nx), fill)
        check(self.lib.axs_boxlevel_download_dense(self.h, dptr(d)))
        return d

    def copy_from_level(self, lvl: Level):
        check(self.lib.axs_boxlevel_copy_from_level(self.h, lvl.h))

    def copy_to_level(self, lvl: Level):
        check(self.lib.axs_boxlevel_copy_to_level(self.h, lvl.h))

    def copy_valid_from(self, src: "BoxLevel") -> int:
        """Copies the valid cells `src` also covers; returns their count."""
        n = C.c_int64()
        check(self.lib.axs_boxlevel_copy_valid(self.h, src.h, C.byref(n)))
        return n.value

    def snapshot_coarse(self, coarse: "BoxLevel"):
        check(self.lib.axs_boxlevel_snapshot_coarse(self.h, coarse.h))

    def fill_ghosts(self, coarse: Optional["BoxLevel"] = None, t_frac=1.0,
                    interior=False):
        check(self.lib.axs_boxlevel_fill_ghosts(
            self.h, coarse.h if coarse is not None else None, t_frac,
            int(interior)))

    def rk4_step(self, coeffs, dt, coarse: Optional["BoxLevel"] = None,
                 t0_frac=0.0, dt_frac=1.0, use_graph=False):
        arr = (RhsCoeffs * 4)(*coeffs)
        fn = (self.lib.axs_boxlevel_rk4_step_graph if use_graph
              else self.lib.axs_boxlevel_rk4_step)
        check(fn(self.h, arr, dt, coarse.h if coarse is not None else None,
                 t0_frac, dt_frac))

    def tag(self, amp_thr=0.0, R=1.0, tags: Optional[DeviceBuffer] = None,
            cover: Optional[DeviceBuffer] = None, grad_thr=DBL_MAX,
            rgrad_thr=DBL_MAX):
        nx, ny, nz = self.domain
        if tags is None:
            tags = DeviceBuffer(nx * ny * nz)
            tags.zero()
        p = TaggerParams(grad_thr, rgrad_thr, amp_thr, R)
        counts = (C.c_int64 * 2)()
        check(self.lib.axs_boxlevel_tag(
            self.h, C.byref(p), _P(tags.ptr),
            _P(cover.ptr) if cover is not None else None, counts))
        self.ctx.synchronize()
        return tags, (int(counts[0]), int(counts[1]))

    def energy_sums(self, R, lam, RpR, scheme=AXS_MASK_NONE, thr=0.8,
                    cover: Optional[DeviceBuffer] = None) -> np.ndarray:
        out = np.zeros(14)
        m = Masking(scheme, thr)
        check(self.lib.axs_boxlevel_energy_sums(
            self.h, R, lam, RpR, C.byref(m),
            _P(cover.ptr) if cover is not None else None, dptr(out)))
        return out

    def coverage_mask(self, coarse_domain) -> DeviceBuffer:
        cd = (C.c_int * 3)(*coarse_domain)
        buf = DeviceBuffer(int(np.prod(coarse_domain)))
        check(self.lib.axs_boxlevel_coverage_mask(self.h, cd, _P(buf.ptr)))
        self.ctx.synchronize()
        return buf

    def average_down_to(self, coarse: "BoxLevel"):
        check(self.lib.axs_boxlevel_average_down(self.h, coarse.h))

    def close(self):
        if self.h:
            if self.owned:
                self.lib.axs_boxlevel_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
