"""Autotuner for the fused RK-stage kernel's launch shape.

The stage kernel is compiled in several tile variants ((64 VX) x TY column
tiles with TY in {4, 8, 16} and VX in {1, 2, 4} cells per lane; each with or
without non-temporal output stores and with or without the next plane's
centre loads issued one plane ahead) and takes the z-chunk (planes streamed
per workgroup) at run time.  tune() times every (variant, z_chunk) candidate on a real level
with HIP events on the level's stream (two RK4 steps per candidate after one
untimed step) and returns the fastest; the result is cached per
(grid size, precision) in tune_cache.json next to this file so later runs
(bench, driver) reuse it without re-tuning.

    python -m axion_strings_adaptive_amd.autotune --n 1024 --precision 1
"""

from __future__ import annotations

import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tune_cache.json")

Z_CHUNKS = (16, 32, 64)


def _key(n: int, precision: int) -> str:
    return f"{n}:{precision}"


def load_cache() -> dict:
    if not os.path.exists(CACHE):
        return {}
    with open(CACHE) as f:
        return json.load(f)


def cached_choice(n: int, precision: int):
    """(variant, z_chunk) tuned for this size and precision, or None."""
    e = load_cache().get(_key(n, precision))
    return (e["variant"], e["z_chunk"]) if e else None


def tune(ctx, n: int, precision: int, steps: int = 2, verbose=True,
         cache_path: str = CACHE):
    from . import capi
    lvl = capi.Level(ctx, n, 0.5, precision)
    lvl.fourier_ic(0, n / 8.0, 0.5, 42, 0)
    lvl.fourier_ic(1, n / 8.0, 0.5, 42, 1)
    lvl.zero_components(2, 2)
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    lvl.enable_timing(True)
    results = []
    for v in range(capi.load().axs_num_tile_variants()):
        for zc in Z_CHUNKS:
            if zc > n:
                continue
            lvl.set_tile(v, zc)
            lvl.advance(bg, 1.0, 0.1, 1)
            lvl.stage_times()
            lvl.advance(bg, 1.1, 0.1, steps)
            ms, ns = lvl.stage_times()
            t = float(sum(ms) / ns)
            results.append((t, v, zc))
            if verbose:
                print(f"[autotune] n={n} prec={precision} variant={v} "
                      f"z_chunk={zc}: {t:.3f} ms/step", file=sys.stderr,
                      flush=True)
    lvl.close()
    t, v, zc = min(results)
    cache = load_cache()
    if cache_path != CACHE and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache.update(json.load(f))
    cache[_key(n, precision)] = {"variant": v, "z_chunk": zc, "ms_per_step": t,
                                 "candidates": [[a, b, c] for a, b, c in
                                                sorted(results)]}
    for path in {CACHE, cache_path}:
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return v, zc, t


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs="+", default=[1024])
    ap.add_argument("--precision", type=int, nargs="+", default=[1, 0])
    ap.add_argument("--out", default=CACHE,
                    help="also write the cache to this path")
    args = ap.parse_args(argv)
    from . import capi
    ctx = capi.Context(0)
    for n in args.n:
        for p in args.precision:
            print(json.dumps({"n": n, "precision": p,
                              "best": tune(ctx, n, p, cache_path=args.out)}),
                  flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
