"""AMR hierarchy on box levels: the batching form of the hot path.

Level 0 is a periodic base grid (tiled into boxes); each finer level
(ratio 2) is built from the string tagger on the level below -- tag where
the plaquette winding is nonzero, where dx^2 |laplacian psi_i| exceeds
gradient_threshold or dx |grad |psi|| exceeds radial_gradient_threshold
(StringTagger.hpp:83-274; both off by default, as in the reference) or, as
an extension, where |psi|/R < amplitude_threshold -- grown by n_error_buf cells
(the reference's amr.n_error_buf), cut into blocking-factor blocks and
merged along x, then y, then z into boxes of at most max_grid_size cells
(amr.blocking_factor / amr.max_grid_size), and kept properly nested one
block inside the coarser level's coverage.

With regrid_int > 0 the finer levels are rebuilt every regrid_int coarse
steps from fresh tags (the reference's amr.regrid_int): a new level takes
the old level's data where the two overlap and quartic interpolation from
the level below elsewhere.

The hierarchy's host logic is native (axs_amr.hip: the axs_hierarchy_*
entry points that the C++ driver uses too); Hierarchy below drives it and
views its levels, and the numpy block / box functions restate the native
host steps for the CPU tests.

One coarse step advances level l once and level l+1 twice (subcycling,
docs/conventions.md:239-241): the coarse state is snapshotted, the coarse
level takes its RK4 step, the fine level takes two half steps whose
coarse-fine ghosts interpolate the coarse state in time, then the fine
valid cells are averaged down onto the coarse level.  Each RK stage of each
level is one persistent launch over that level's box work-list.

With a box plan (the reference's axion_strings.N, N1, N2 with
amr.max_level > 0) the base grid, its spacing and the level schedule come
from amr_box_plan, and a level is only created once log(m_r/H) has reached
its entry of the schedule (the reference's schedule-gated tagger).

    python -m axion_strings_adaptive_amd.amr --n0 256 --levels 2 --steps 4
    python -m axion_strings_adaptive_amd.amr --plan 1024 1 4 --levels 2 \
        --log-mr-over-h-i 1.0 --regrid-int 4 --steps 100
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import capi


def amr_box_plan(N_effective: int, N1: float, N2: float, a_inv: float,
                 c: float, max_level: int) -> dict:
    """compute_amr_box_plan (BoxPlan.hpp:96-125; the same arithmetic as the
    driver's csrc/host/BoxPlan.hpp): N is the finest level's effective
    resolution, reached at tau_f with H L = N1 and 1/(dx m_r) = N2 (tau0 = 1,
    lambda0 = 1, R0 = 1/(a_inv - 1)); level l (entry l-1 of log_add) is
    needed from log(m_r/H) = log_f - (max_level - l + 1) ln2 (a_inv-c)/(1-c).
    N_base is -1 when 2^max_level does not divide N."""
    grow = (N_effective / (N1 * N2)) ** ((a_inv - 1.0) / (a_inv - c))
    R0 = 1.0 / (a_inv - 1.0)
    L_tilde = (N1 / R0) * grow
    dx_finest = L_tilde / N_effective
    refine = 1 << max_level
    log_f = (a_inv - c) / (a_inv - 1.0) * np.log(grow)
    spacing = np.log(2.0) * (a_inv - c) / (1.0 - c)
    return {"L_tilde": L_tilde, "dx_finest": dx_finest,
            "dx_base": dx_finest * refine, "tau_f": grow,
            "N_base": N_effective // refine if N_effective % refine == 0 else -1,
            "log_add": [float(log_f - (max_level - l + 1) * spacing)
                        for l in range(1, max_level + 1)]}


def dilate_periodic(mask: np.ndarray, r: int) -> np.ndarray:
    """Periodic dilation of a boolean [z, y, x] mask by an r-cell cube."""
    out = mask.copy()
    for ax in range(3):
        acc = out.copy()
        for s in range(1, r + 1):
            acc |= np.roll(out, s, axis=ax) | np.roll(out, -s, axis=ax)
        out = acc
    return out


def blocks_from_tags(tags: np.ndarray, block: int) -> np.ndarray:
    """Boolean [bz, by, bx] of blocks containing any tagged cell."""
    nz, ny, nx = tags.shape
    t = tags.reshape(nz // block, block, ny // block, block, nx // block, block)
    return t.any(axis=(1, 3, 5))


def tagged_blocks(tags: np.ndarray, block: int, r: int) -> np.ndarray:
    """blocks_from_tags(dilate_periodic(tags, r), block) without forming the
    dilated cell mask: dilation by a cube and "any over a block" both act
    axis by axis, so each axis is reduced to blocks at once, a block taking
    its own cells plus the r cells on either side of it (periodic, r <=
    block).  One pass over the cell mask, then passes over arrays that are
    block times smaller."""
    assert 0 <= r <= block
    out = tags
    axes = (2, 1, 0)  # x, then y, then z ([z, y, x] arrays)
    if (block % 8 == 0 and r <= 8 and sys.byteorder == "little"
            and tags.dtype == np.bool_ and tags.shape[2] % block == 0):
        # x first, 8 cells per 64-bit word: a block is block / 8 words, its
        # first r cells the low r bytes of its first word, its last r cells
        # the high r bytes of its last word
        nz, ny, nx = tags.shape
        w = np.ascontiguousarray(tags).view(np.uint64).reshape(
            nz, ny, nx // block, block // 8)
        red = (w != 0).any(axis=-1)
        if r > 0:
            head = (w[..., 0] & np.uint64((1 << (8 * r)) - 1)) != 0
            tail = (w[..., -1] >> np.uint64(8 * (8 - r))) != 0
            red = red | np.roll(head, -1, axis=-1) | np.roll(tail, 1, axis=-1)
        out = red
        axes = (1, 0)
    for ax in axes:
        a = np.moveaxis(out, ax, -1)
        nb = a.shape[-1] // block
        a = a.reshape(a.shape[:-1] + (nb, block))
        red = a.any(axis=-1)
        if r > 0:
            head = a[..., :r].any(axis=-1)   # first r cells of each block
            tail = a[..., -r:].any(axis=-1)  # last r cells of each block
            red = red | np.roll(head, -1, axis=-1) | np.roll(tail, 1, axis=-1)
        out = np.moveaxis(red, -1, ax)
    return np.ascontiguousarray(out)


def boxes_from_blocks(blk: np.ndarray, fine_block: int, max_blocks: int):
    """Fine-level boxes (ratio 2) covering the tagged coarse blocks: runs of
    up to max_blocks consecutive blocks along x, runs with the same x extent
    merged along y, then along z (each again up to max_blocks), so that a
    fully tagged region becomes boxes of max_blocks^3 blocks."""
    bz, by, bx = blk.shape
    runs = {}  # (k, j) -> list of (i0, i1) block runs, i1 exclusive
    for k in range(bz):
        for j in range(by):
            row = blk[k, j]
            i = 0
            while i < bx:
                if not row[i]:
                    i += 1
                    continue
                i0 = i
                while i < bx and row[i] and i - i0 < max_blocks:
                    i += 1
                runs.setdefault((k, j), []).append((i0, i))
    # merge along y: same (k, i0, i1) in consecutive rows j
    slabs = {}  # k -> list of (i0, i1, j0, j1)
    for k in range(bz):
        open_ = {}
        out = []
        for j in range(by + 1):
            cur = set(runs.get((k, j), [])) if j < by else set()
            for key in list(open_):
                j0 = open_[key]
                if key not in cur or j - j0 == max_blocks:
                    out.append((key[0], key[1], j0, j))
                    del open_[key]
            for key in cur:
                if key not in open_:
                    open_[key] = j
        slabs[k] = out
    # merge along z: same (i0, i1, j0, j1) in consecutive planes k
    boxes = []
    open_ = {}
    for k in range(bz + 1):
        cur = set(slabs.get(k, [])) if k < bz else set()
        for key in list(open_):
            k0 = open_[key]
            if key not in cur or k - k0 == max_blocks:
                i0, i1, j0, j1 = key
                boxes.append(((i0 * fine_block, j0 * fine_block, k0 * fine_block),
                              (i1 * fine_block - 1, j1 * fine_block - 1,
                               k * fine_block - 1)))
                del open_[key]
        for key in cur:
            if key not in open_:
                open_[key] = k
    boxes.sort(key=lambda b: (b[0][2], b[0][1], b[0][0]))
    return boxes


def erode_periodic(mask: np.ndarray, r: int) -> np.ndarray:
    return ~dilate_periodic(~mask, r)


class Hierarchy:
    """The AMR hierarchy (the native axs_hierarchy_* of libaxs_hip.so: tag
    -> blocks -> boxes, subcycled steps, regrids, composite sums) built on a
    seeded-IC base level; its levels are BoxLevel views of the native ones.
    The block and box functions above restate the native host steps
    (axs_amr_tagged_blocks / axs_amr_erode / axs_amr_boxes_from_blocks) for
    testing."""

    def __init__(self, ctx: capi.Context, n0: int = 256, levels: int = 2,
                 dx0: float = 0.5, precision: int = capi.AXS_FP32C,
                 amp_thr: float = 0.5, n_error_buf: int = 2,
                 blocking_factor: int = 16, max_grid_size: int = 64,
                 seed: int = 42, bg=None, tau0: float = 1.0,
                 dt_multiplier: float = 0.2, verbose: bool = True,
                 force_full_refinement: bool = False, use_graph: bool = True,
                 regrid_int: int = 0, gradient_threshold: float = capi.DBL_MAX,
                 radial_gradient_threshold: float = capi.DBL_MAX,
                 level_add_log_mr_over_h=None,
                 per_stage_launches: bool = False):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = None
        # the box plan's level schedule (entry l-1: log(m_r/H) from which
        # level l may exist), or None for no gating
        self.level_add = (None if level_add_log_mr_over_h is None
                          else [float(v) for v in level_add_log_mr_over_h])
        self.use_graph = use_graph
        self.max_levels = levels
        self.regrid_int = regrid_int
        self.force_full_refinement = force_full_refinement
        self.n_error_buf = n_error_buf
        self.blocking_factor = blocking_factor
        self.max_grid_size = max_grid_size
        self.bg = bg if bg is not None else capi.make_background(
            2.0, c0=0.0, c1=2.0, tau_switch=1.0, has_switch=True)
        self.precision = precision
        self.dx = [dx0 / 2 ** l for l in range(levels + 1)]
        self.dt0 = dt_multiplier * dx0
        self.amp_thr = amp_thr
        # axion_strings.tagging.{gradient,radial_gradient}_threshold
        self.grad_thr = gradient_threshold
        self.rgrad_thr = radial_gradient_threshold
        self.verbose = verbose
        self._views = {}
        self._composite = None  # collapsed level-0 field of the spectra
        self.levels: list[capi.BoxLevel] = []
        p = capi.AmrParams()
        p.max_level = levels
        p.n_error_buf = n_error_buf
        p.blocking_factor = blocking_factor
        p.max_grid_size = max_grid_size
        p.regrid_int = regrid_int
        p.force_full_refinement = int(force_full_refinement)
        p.gradient_threshold = gradient_threshold
        p.radial_gradient_threshold = radial_gradient_threshold
        p.amplitude_threshold = amp_thr
        sched = self.level_add or []
        if len(sched) > capi.AXS_MAX_LEVELS:
            raise ValueError("at most 8 schedule entries")
        p.n_schedule = len(sched)
        for i, v in enumerate(sched):
            p.level_add_log_mr_over_h[i] = v
        p.use_graph = int(use_graph)
        p.per_stage_launches = int(per_stage_launches)
        self.params = p
        # ---- level 0: seeded IC on a uniform level ----
        base = capi.Level(ctx, n0, dx0, precision)
        try:
            base.fourier_ic(0, n0 / 8.0, 0.5, seed, 0)
            base.fourier_ic(1, n0 / 8.0, 0.5, seed, 1)
            base.zero_components(2, 2)
            h = capi._P()
            capi.check(self.lib.axs_hierarchy_create(
                ctx.h, base.h, C.byref(p), C.byref(self.bg), None, tau0,
                self.dt0, C.byref(h)))
        finally:
            base.close()
        self.h = h
        ctx.adopt(self)
        self._sync()
        if verbose:
            for l in range(1, len(self.levels)):
                print(f"[amr] level {l}: {len(self.levels[l].boxes)} boxes, "
                      f"{self.levels[l].num_cells} cells", file=sys.stderr,
                      flush=True)

    def _sync(self):
        """Refreshes the level views, their boxes and the counters."""
        n, tau = C.c_int(), C.c_double()
        ns, nr = C.c_int64(), C.c_int64()
        secs = (C.c_double * 6)()
        capi.check(self.lib.axs_hierarchy_info(self.h, C.byref(n),
                                               C.byref(tau), C.byref(ns),
                                               C.byref(nr), secs))
        self.tau, self.nsteps, self.nregrids = tau.value, ns.value, nr.value
        self.regrid_s = secs[0]
        self.regrid_parts = {k: secs[i + 1] for i, k in enumerate(
            ("tag_and_boxes", "create", "interpolate", "copy_valid", "free"))
            if secs[i + 1] > 0.0}
        levels, domains, views = [], [], {}
        for l in range(n.value):
            hl, dom = capi._P(), (C.c_int * 3)()
            dxl, nb = C.c_double(), C.c_int()
            capi.check(self.lib.axs_hierarchy_level(
                self.h, l, C.byref(hl), dom, C.byref(dxl), C.byref(nb)))
            lo = (C.c_int * (3 * nb.value))()
            hi = (C.c_int * (3 * nb.value))()
            capi.check(self.lib.axs_hierarchy_boxes(self.h, l, lo, hi))
            boxes = [(tuple(lo[3 * b:3 * b + 3]), tuple(hi[3 * b:3 * b + 3]))
                     for b in range(nb.value)]
            v = self._views.get(hl.value)
            if v is None or v.boxes != boxes or v.h is None:
                v = capi.BoxLevel.borrowed(self.ctx, hl.value, boxes,
                                           tuple(dom), dxl.value,
                                           self.precision)
            views[hl.value] = v
            levels.append(v)
            domains.append(tuple(dom))
        for k, v in self._views.items():
            if views.get(k) is not v:
                v.h = None  # its native level is gone
        self._views = views
        self.levels = levels
        self.domains = domains
        self.ncells = [L.num_cells for L in levels]

    def level_allowed(self, l: int) -> bool:
        """Schedule gating of the tagger (AxionStringsLevel.cpp:1957-1983):
        level l may be created once log(m_r/H) at the current tau has
        reached the schedule's entry for it; with force_full_refinement,
        or without a schedule, always."""
        if self.force_full_refinement or self.level_add is None:
            return True
        if l > len(self.level_add):
            return False
        log_mr_over_h = -np.log(capi.background_eval(self.bg, 5, self.tau))
        return bool(log_mr_over_h >= self.level_add[l - 1])

    def regrid(self):
        """Rebuilds levels 1.. from fresh tags at a synchronised time."""
        capi.check(self.lib.axs_hierarchy_regrid(self.h))
        self._sync()

    def coeffs(self, tau, dt):
        return [capi.rhs_coeffs(self.bg, t) for t in
                (tau, tau + 0.5 * dt, tau + 0.5 * dt, tau + dt)]

    def step(self):
        """One coarse step of the whole hierarchy (then a regrid every
        regrid_int coarse steps)."""
        capi.check(self.lib.axs_hierarchy_step(self.h))
        self._sync()

    def cell_updates_per_coarse_step(self) -> int:
        return sum(n * 2 ** l for l, n in enumerate(self.ncells))

    def composite_energy(self, scheme=capi.AXS_MASK_B, thr=0.8):
        """compute_composite_total_energy (EnergyKernel.hpp:287-339) over the
        hierarchy, each level masked by the coverage of the next finer."""
        R = capi.background_eval(self.bg, 0, self.tau)
        lam = capi.background_eval(self.bg, 3, self.tau)
        RpR = capi.background_eval(self.bg, 8, self.tau)
        acc = np.zeros(14)
        m = capi.Masking(scheme, thr)
        capi.check(self.lib.axs_hierarchy_energy_sums(
            self.h, R, lam, RpR, C.byref(m), capi.dptr(acc)))
        n_total, n_unmasked = acc[13], acc[12]
        return {"rho_tot_unscreened": acc[0] / n_total,
                "rho_tot_screened": acc[1] / n_unmasked if n_unmasked else 0.0,
                "n_total": n_total, "n_unmasked": n_unmasked}

    def plaquette_counts(self) -> np.ndarray:
        """[level, (plain, weighted)] pierced plaquettes over the cells no
        finer level covers."""
        out = np.zeros((len(self.levels), 2), np.int64)
        capi.check(self.lib.axs_hierarchy_plaquette_counts(
            self.h, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def composite_level0(self) -> capi.Level:
        """collapse_to_level0 (AxionStringsLevel.cpp:110-134): level 0's own
        data where nothing refines it and the finer data's recursive average
        where something does, as a uniform level-0 resolution field (the
        caller closes it).  The hierarchy itself is left unchanged."""
        out = capi.Level(self.ctx, self.domains[0], self.dx[0], self.precision)
        try:
            capi.check(self.lib.axs_hierarchy_collapse(self.h, out.h))
        except Exception:
            out.close()
            raise
        return out

    def composite_spectrum(self, kind=capi.AXS_FIELD_MASKED_A_DOT,
                           scheme=capi.AXS_MASK_B, thr=0.8):
        """The spectrum of the collapsed level-0 field
        (AxionStringsLevel.cpp:1617-1619 on collapse_to_level0's result):
        (shell sums, shell counts, [full cube, inscribed sphere], sum of
        squares)."""
        # the collapsed level is kept between calls, so its FFT plans
        # (built on a level's first spectrum) are built once
        lvl = self._composite
        if lvl is None:
            lvl = self._composite = capi.Level(
                self.ctx, self.domains[0], self.dx[0], self.precision)
        capi.check(self.lib.axs_hierarchy_collapse(self.h, lvl.h))
        R = capi.background_eval(self.bg, 0, self.tau)
        return lvl.spectrum(kind, scheme, thr, R)

    def close(self):
        if self._composite is not None:
            self._composite.close()
            self._composite = None
        if self.h:
            for v in self._views.values():
                v.h = None
            self.levels = []
            self.lib.axs_hierarchy_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n0", type=int, default=256)
    ap.add_argument("--levels", type=int, default=2)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--precision", type=int, default=capi.AXS_FP32C)
    ap.add_argument("--gradient-threshold", type=float, default=capi.DBL_MAX,
                    help="tag where dx^2 |laplacian psi_i| exceeds this")
    ap.add_argument("--radial-gradient-threshold", type=float,
                    default=capi.DBL_MAX,
                    help="tag where dx |grad |psi|| exceeds this")
    ap.add_argument("--regrid-int", type=int, default=0,
                    help="rebuild the finer levels every N coarse steps")
    ap.add_argument("--plan", type=float, nargs=3, metavar=("N", "N1", "N2"),
                    help="box plan (axion_strings.N, N1, N2): the base grid "
                         "N/2^levels, its spacing, and the level schedule; "
                         "constant c = --c0 from log(m_r/H) = "
                         "--log-mr-over-h-i, stopping at tau_f")
    ap.add_argument("--c0", type=float, default=0.0)
    ap.add_argument("--log-mr-over-h-i", type=float, default=0.0)
    ap.add_argument("--per-stage-launches", action="store_true",
                    help="step the box levels with one launch per RK stage "
                         "instead of the fused stage pairs (same bits)")
    args = ap.parse_args(argv)
    ctx = capi.Context(0)
    kw = {}
    n0 = args.n0
    tau_f = None
    if args.plan:
        N, N1, N2 = int(args.plan[0]), args.plan[1], args.plan[2]
        plan = amr_box_plan(N, N1, N2, 2.0, args.c0, args.levels)
        if plan["N_base"] <= 0:
            ap.error("--plan: N is not divisible by 2^levels")
        bg = capi.make_background(2.0, c0=args.c0)
        tau0 = capi.background_eval(bg, 7, args.log_mr_over_h_i)
        if plan["tau_f"] <= tau0:
            ap.error("--plan: tau_f <= tau_i")
        n0, tau_f = plan["N_base"], plan["tau_f"]
        kw = dict(bg=bg, tau0=tau0, dx0=plan["dx_base"],
                  level_add_log_mr_over_h=plan["log_add"])
    h = Hierarchy(ctx, n0, args.levels, precision=args.precision,
                  regrid_int=args.regrid_int,
                  gradient_threshold=args.gradient_threshold,
                  radial_gradient_threshold=args.radial_gradient_threshold,
                  per_stage_launches=args.per_stage_launches, **kw)
    h.step()  # warm-up
    ctx.synchronize()
    rs0, n_regrid0, parts0 = h.regrid_s, h.nregrids, dict(h.regrid_parts)
    t0 = time.perf_counter()
    steps = 0
    for _ in range(args.steps):
        if tau_f is not None and h.tau >= tau_f:
            break
        h.step()
        steps += 1
    ctx.synchronize()
    dt = time.perf_counter() - t0
    args.steps = max(1, steps)
    e = h.composite_energy()
    ctx.synchronize()
    t0 = time.perf_counter()
    ss, sc, tot, _ = h.composite_spectrum()
    ctx.synchronize()
    t_spec_first = time.perf_counter() - t0  # builds the FFT plans
    t0 = time.perf_counter()
    ss, sc, tot, _ = h.composite_spectrum()
    ctx.synchronize()
    t_spec = time.perf_counter() - t0
    print(json.dumps({
        "levels": len(h.levels), "cells": h.ncells,
        "boxes": [len(L.boxes) for L in h.levels], "regrids": h.nregrids,
        # ghost cells the RK stage fills write per level (2-cell shells)
        "stage_fill_ghost_cells": [
            sum((b[1][0] - b[0][0] + 5) * (b[1][1] - b[0][1] + 5) *
                (b[1][2] - b[0][2] + 5) - (b[1][0] - b[0][0] + 1) *
                (b[1][1] - b[0][1] + 1) * (b[1][2] - b[0][2] + 1)
                for b in L.boxes) for L in h.levels],
        "cell_updates_per_coarse_step": h.cell_updates_per_coarse_step(),
        "cell_updates_per_s": h.cell_updates_per_coarse_step() * args.steps / dt,
        "ms_per_coarse_step": 1e3 * dt / args.steps,
        "regrid_ms_mean": (1e3 * (h.regrid_s - rs0) / (h.nregrids - n_regrid0)
                           if h.nregrids > n_regrid0 else None),
        "ms_per_coarse_step_without_regrids":
            1e3 * (dt - (h.regrid_s - rs0)) / args.steps,
        "regrid_ms_by_phase": {k: 1e3 * (v - parts0.get(k, 0.0)) /
                               max(1, h.nregrids - n_regrid0)
                               for k, v in h.regrid_parts.items()},
        "tau": h.tau, "tau_f": tau_f,
        "n_total": e["n_total"],
        "L_tilde^3": (n0 * h.dx[0]) ** 3,
        "rho_tot_unscreened": e["rho_tot_unscreened"],
        "composite_spectrum_ms": 1e3 * t_spec,
        "composite_spectrum_first_call_ms": 1e3 * t_spec_first,
        "composite_spectrum_nonempty_shells": int((sc > 0).sum())}))
    h.close()
    ctx.close()


if __name__ == "__main__":
    main()
