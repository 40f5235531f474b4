"""Autotuner for every launch shape of the RK4 step.

What it tunes, per grid size and precision (every candidate gives the same
bits; only the time differs):

* stage fusion 3 (FP32 / FP32C, the whole step in one launch): the step
  kernel's variant (-1 = the 64 x 11 ring kernel loading one plane ahead,
  4 = the same loading two planes ahead, 6 / 5 = those two with the tiles
  walked in 4 x 8 patches, 8 / 7 = those two with the 8 XCDs on adjacent
  bands of 32 tiles, 3 = 64 x 11 with rows aligned to cache lines,
  2 = 60 x 12, 0 = 60 x 8, 1 = 60 x 4) and its z chunk;
* stage fusions 1 and 2 (the fused pairs, with and without a stored
  accumulator): the pair tile (TY x VX cells per lane, centre-load prefetch,
  non-temporal stores) and z chunk;
* stage fusion 0 (one launch per stage): the stage kernel's tile variant
  and z chunk;
* box levels: the rows per work item and the persistent grid of the box
  stage launches, timed on box levels of 16^3, 32^3 and 64^3 boxes tiling a
  256^3 domain.

Each candidate runs one untimed RK4 step, then `steps` steps bracketed by
HIP events on the level's stream (the library's step timing); box levels
are timed by the wall clock over `steps` steps between stream syncs.  The
fastest candidate of each (size, precision, fusion) becomes a `level` line
of the tuning table, the fastest box shape of each precision a `box` line
(the format axs_tuning_load reads), written to tune_table.txt next to this
file; every candidate's time goes to tune_results.json.  bench.py and the
C++ driver (axion_strings.tuning_table) load the table; levels created
after the load take their launch shapes from it.

    python -m axion_strings_adaptive_amd.autotune --n 1024 --precision 1 0
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tune_table.txt")
RESULTS = os.path.join(HERE, "tune_results.json")

STEP_VARIANTS = (-1, 4, 6, 5, 8, 7, 3, 2, 0, 1)
STEP_Z_CHUNKS = (32, 64, 128, 256)
PAIR_Z_CHUNKS = (16, 32, 64)
STAGE_Z_CHUNKS = (16, 32, 64)
BOX_ROWS = (4, 8, 16, 32, 64, 512)
BOX_GRIDS = (512, 1024, 2048, 4096, 8192)
BOX_EDGES = (16, 32, 64)
BOX_DOMAIN = 256


def _log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def _pair_candidates(capi, precision, fusion):
    """Pair variants worth timing: base tiles x {prefetch, non-temporal
    stores} (the y-major order bit only permutes tiles).  FP64 runs tiles
    that would spill in smaller shapes, so only its distinct bases are
    timed; the accumulator-free pairs do not run in 4 x 4 tiles."""
    nbase = capi.load().axs_num_pair_variants() // 8
    bases = (0, 1, 2) if precision == capi.AXS_FP64 else tuple(range(nbase))
    out = []
    for g in (0, 1, 2, 3):
        for b in bases:
            if fusion == 2 and b == 4 and precision != capi.AXS_FP64:
                continue
            out.append(b + nbase * g)
    return out


def _candidates(capi, precision, fusion):
    if fusion == 3:
        return [(v, z) for v in STEP_VARIANTS for z in STEP_Z_CHUNKS]
    if fusion in (1, 2):
        return [(v, z) for v in _pair_candidates(capi, precision, fusion)
                for z in PAIR_Z_CHUNKS]
    return [(v, z) for v in range(capi.load().axs_num_tile_variants())
            for z in STAGE_Z_CHUNKS]


def _time_level(capi, lvl, bg, steps):
    lvl.advance(bg, 1.0, 0.1, 1)
    lvl.enable_timing(True)
    lvl.stage_times()
    lvl.advance(bg, 1.1, 0.1, steps)
    ms, ns = lvl.stage_times()
    lvl.enable_timing(False)
    return float(sum(ms) / ns)


def tune_level(ctx, n: int, precision: int, fusions=None, steps: int = 3):
    """Times every candidate of each stage fusion mode on an n^3 level of
    the seeded IC; returns {fusion: (ms, variant, z_chunk, candidates)}."""
    from . import capi
    if fusions is None:
        fusions = (0, 1, 2) if precision == capi.AXS_FP64 else (3, 1, 2, 0)
    lvl = capi.Level(ctx, n, 0.5, precision)
    lvl.fourier_ic(0, n / 8.0, 0.5, 42, 0)
    lvl.fourier_ic(1, n / 8.0, 0.5, 42, 1)
    lvl.zero_components(2, 2)
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    best = {}
    for fusion in fusions:
        lvl.set_fusion(fusion)
        res = []
        for v, z in _candidates(capi, precision, fusion):
            if fusion == 0:
                lvl.set_tile(v, z)
            else:
                lvl.set_pair_tile(v, z)
            try:
                t = _time_level(capi, lvl, bg, steps)
            except capi.AxsError as e:
                _log(f"n={n} prec={precision} fusion={fusion} v={v} z={z}: "
                     f"skipped ({e})")
                continue
            res.append((t, v, z))
            _log(f"n={n} prec={precision} fusion={fusion} v={v} z={z}: "
                 f"{t:.3f} ms/step")
        # back to the defaults before the next mode
        lvl.set_tile(-1, 0)
        lvl.set_pair_tile(-1, 0)
        t, v, z = min(res)
        best[fusion] = (t, v, z, sorted(res))
    lvl.close()
    return best


def _box_list(edge):
    m = BOX_DOMAIN // edge
    return [((i * edge, j * edge, k * edge),
             ((i + 1) * edge - 1, (j + 1) * edge - 1, (k + 1) * edge - 1))
            for k in range(m) for j in range(m) for i in range(m)]


def tune_boxes(ctx, precision: int, steps: int = 4):
    """Times every (item_rows, grid_blocks) on box levels of 16^3, 32^3 and
    64^3 boxes over a 256^3 domain; returns (ms, rows, grid, candidates)
    with ms the sum of the three levels' ms per step."""
    import numpy as np
    from . import capi
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    coeffs = [capi.rhs_coeffs(bg, 1.0 + f * 0.1) for f in (0, 0.5, 0.5, 1)]
    rng = np.random.default_rng(7)
    n = BOX_DOMAIN
    dense = rng.standard_normal((4, n, n, n)) * 0.1
    dense[0] += 1.0
    totals = {}
    for edge in BOX_EDGES:
        bl = capi.BoxLevel(ctx, _box_list(edge), (n, n, n), 0.5, precision)
        bl.upload_dense(dense)
        for rows in BOX_ROWS:
            for grid in BOX_GRIDS:
                bl.set_work_shape(rows, grid)
                bl.rk4_step(coeffs, 0.1)
                ctx.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    bl.rk4_step(coeffs, 0.1)
                ctx.synchronize()
                ms = 1e3 * (time.perf_counter() - t0) / steps
                totals.setdefault((rows, grid), []).append(ms)
                _log(f"boxes {edge}^3 prec={precision} rows={rows} "
                     f"grid={grid}: {ms:.3f} ms/step")
        bl.close()
    ctx.trim()
    res = sorted((sum(v), r, g) for (r, g), v in totals.items())
    t, r, g = res[0]
    return t, r, g, res


def results_of(level_best: dict, box_best: dict) -> dict:
    """level_best: {(n, precision): {fusion: (ms, v, z, cands)}};
    box_best: {precision: (ms, rows, grid, cands)} -> the results record,
    keyed "level <n> <precision> <fusion>" / "box <precision>"."""
    results = {}
    for (n, p), modes in level_best.items():
        for fusion, (t, v, z, cands) in modes.items():
            results[f"level {n} {p} {fusion}"] = {
                "best": [v, z], "ms_per_step": t,
                "candidates": [[c[1], c[2], round(c[0], 4)] for c in cands]}
    for p, (t, r, g, cands) in box_best.items():
        results[f"box {p}"] = {
            "best": [r, g], "ms_sum": t,
            "candidates": [[c[1], c[2], round(c[0], 4)] for c in cands]}
    return results


def write_table(results: dict, path: str = TABLE,
                results_path: str = RESULTS):
    """The tuning table (axs_tuning_load format) and the results record."""
    lines = ["# Launch shapes timed by axion_strings_adaptive_amd.autotune "
             "(axs_tuning_load format).",
             "# level <nx> <ny> <nz> <precision> <fusion> <tile_variant> "
             "<tile_z_chunk> <pair_variant> <pair_z_chunk>",
             "# box <precision> <item_rows> <grid_blocks>"]

    def order(k):
        f = k.split()
        return (0, int(f[1]), int(f[2]), int(f[3])) if f[0] == "level" \
            else (1, int(f[1]), 0, 0)
    for k in sorted(results, key=order):
        f, e = k.split(), results[k]
        if f[0] == "level":
            n, p, fusion = int(f[1]), int(f[2]), int(f[3])
            v, z = e["best"]
            tv, tz, pv, pz = (v, z, -1, 0) if fusion == 0 else (-1, 0, v, z)
            lines.append(f"level {n} {n} {n} {p} {fusion} {tv} {tz} {pv} {pz}"
                         f"  # {e['ms_per_step']:.3f} ms/step")
        else:
            r, g = e["best"]
            lines.append(f"box {f[1]} {r} {g}  # {e['ms_sum']:.3f} ms over 3 "
                         "box sizes")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    with open(results_path, "w") as fh:
        json.dump(results, fh, indent=1, sort_keys=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs="*", default=[1024],
                    help="level sizes to tune (none: box shapes only)")
    ap.add_argument("--precision", type=int, nargs="+", default=[1, 0])
    ap.add_argument("--fusion", type=int, nargs="*", default=None,
                    help="stage fusion modes to tune (default: all the "
                         "precision runs)")
    ap.add_argument("--box-precision", type=int, nargs="*", default=[1, 0])
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--table", default=TABLE)
    ap.add_argument("--results", default=RESULTS)
    ap.add_argument("--merge", default=None,
                    help="a results record (tune_results.json) whose entries "
                         "this run does not retune are kept")
    args = ap.parse_args(argv)
    from . import capi
    ctx = capi.Context(0)
    level_best, box_best = {}, {}
    for n in args.n:
        for p in args.precision:
            fus = args.fusion
            if fus is not None and p == capi.AXS_FP64:
                fus = [f for f in fus if f != 3]
            level_best[(n, p)] = tune_level(ctx, n, p, fus, args.steps)
            ctx.trim()
            for f, (t, v, z, _) in sorted(level_best[(n, p)].items()):
                _log(f"best n={n} prec={p} fusion={f}: v={v} z={z} "
                     f"{t:.3f} ms/step")
    for p in args.box_precision:
        box_best[p] = tune_boxes(ctx, p)
        _log(f"best boxes prec={p}: rows={box_best[p][1]} "
             f"grid={box_best[p][2]} {box_best[p][0]:.3f} ms")
    results = {}
    if args.merge:
        with open(args.merge) as fh:
            results = json.load(fh)
    results.update(results_of(level_best, box_best))
    write_table(results, args.table, args.results)
    print(json.dumps({"table": args.table, "results": args.results}),
          flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
