#!/usr/bin/env python3
"""Benchmark: cell-updates/s of classic RK4 on a 1024^3 periodic level,
FP32 + Kahan compensation versus FP64, on one MI355X.

Workload (the precision-form configuration): 1024^3 cells, dx = 0.5,
radiation era a = tau from tau = 1, lambda = eta = 1 with the c = 2 (Moore)
schedule reached through a switch at tau = 1 (c0 = 0, c1 = 2), dtau = 0.1;
psi1, psi2 hash-seeded Gaussian fields (seed 42, modes 0 < |p| <= N/8,
<psi_i^2> = 0.5), Pi = 0, generated on the device.  One step = one RK4 step
of every cell: for FP32C one launch that forms every stage input from y_n
and stores only y_{n+1}; for FP64 two launches, stages 1+2 and 3+4, each
fusing the RHS of two stages with their RK updates (--fusion picks a mode,
--no-fusion one launch per stage; the same bits every way).  The FP32C run is the
headline `value`; the FP64 run on the same IC is timed the same way and
reported beside it together with the precision metrics (max |dphi|/eta,
relative energy error, delta xi) at the end of the timed steps.

The timed region holds exactly K steps, bracketed by a device
synchronisation on both sides, with all data resident in HBM.  Stage
kernels are launched directly with HIP events recorded on the level's own
stream between them, which gives the per-launch durations used for the
roofline figure.  The CPU baseline is the FP64 OpenMP oracle timed on a
bounded sample (a 512^3 level, 5 RK4 steps, about 10-20 s on 16 host
threads) of the same per-cell work.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("cell-updates/s 1024^3 RK4: FP32-compensated vs FP64 at <=1e-5 "
          "rel energy err")
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell of each stage launch (4 components):
#   FP32C (S = 16 B): stage1 3S, stage2 5S, stage3 5S, stage4 6S (y, lo rd/wr)
#   FP64  (S = 32 B): stage1 3S, stage2 5S, stage3 5S, stage4 4S
STAGE_BYTES = {1: [48, 80, 80, 96], 0: [96, 160, 160, 128]}
# ... of each fused-pair launch: pair 1 reads y_n, writes acc and y_2 (3S);
# pair 2 reads y_2, y_n, acc, writes y_{n+1} (4S) + FP32C lo read/write (2S)
PAIR_BYTES = {1: [48, 96], 0: [96, 128]}
# ... without a stored accumulator (fusion mode 2): pair 1 reads y_n and
# writes y_2 (2S); pair 2 reads y_2, y_n, writes y_{n+1} (3S) (+ lo 2S)
PAIR_BYTES_NOACC = {1: [32, 80], 0: [64, 96]}
# ... of the one launch per step (fusion mode 3, FP32C): read y_n and lo,
# write y_{n+1} and lo (4S)
STEP_BYTES_ONE = 64
# HBM bytes per cell of the dominant FP32C launch at 1024^3, from
# rocprofv3 --pmc FETCH_SIZE (x 2.0) and WRITE_SIZE (x 1.0), the scales
# measured on device copies and reads of known size; keyed by the launch
# shape the level reports ((tile variant, z chunk) for one launch per
# stage, (pair variant, z chunk) otherwise).  Each table holds the digest
# of the kernel sources it was measured on: when the sources differ the
# stored figure describes another kernel, and `traffic` is reported as null
# (source_digest() of the files named).
MEASURED_TRAFFIC = {
    # one launch per step (stage fusion 3): k_step_r / k_step
    "step": {"sources": ["axs_step.hip", "axs_stage_math.hpp"],
             "digest": "e642f85f2cbfd8b0",
             "profile": "profiles/k_step_1024cubed_trace_and_traffic.md",
             "bytes": {(7, 128): 77.32}},
    # fused pairs, stages 3+4 (modes 1 and 2)
    "pair": {"sources": ["axs_pair.hip", "axs_stage_math.hpp"],
             "digest": None, "profile": None, "bytes": {}},
    # one launch per stage, stages 2/3
    "stage": {"sources": ["axs_rk.hip", "axs_stage_math.hpp"],
              "digest": None, "profile": None, "bytes": {}},
}
CSRC = os.path.join(ROOT, "axion_strings_adaptive_amd", "csrc")


def source_digest(names):
    """sha256 (first 16 hex digits) of the named kernel sources, in order."""
    import hashlib
    h = hashlib.sha256()
    for name in names:
        with open(os.path.join(CSRC, name), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(family, key):
    """Stored HBM bytes per cell for a launch shape, or (None, reason)."""
    t = MEASURED_TRAFFIC[family]
    if key not in t["bytes"]:
        return None, "no stored measurement for this launch shape"
    if t["digest"] != source_digest(t["sources"]):
        return None, "stored measurement is of other kernel sources"
    return t["bytes"][key], t["profile"]


def step_kernel_name(pair_variant):
    """The one-launch step kernel a pair variant selects."""
    if pair_variant in (-1, 4, 5, 6, 7, 8):
        d = 2 if pair_variant in (4, 5, 7) else 1
        order = {5: ", 4 x 8 tile patches", 6: ", 4 x 8 tile patches",
                 7: ", XCDs on adjacent tile bands",
                 8: ", XCDs on adjacent tile bands"}.get(pair_variant, "")
        return f"k_step_r<FP32C, 11, {d}>{order}"
    ty = {0: 8, 1: 4, 3: 11}.get(pair_variant, 12)
    return f"k_step<FP32C, {ty}, {'aligned' if ty == 11 else 'unaligned'}>"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_level(capi, ctx, n, precision, bg, dt, warmup, steps, tile=None,
              place_trials=0, fused=True, pair_tile=None):
    lvl = capi.Level(ctx, n, 0.5, precision)
    if tile is not None:
        lvl.set_tile(*tile)
    lvl.set_fusion(fused)
    if pair_tile is not None:
        lvl.set_pair_tile(*pair_tile)
    placement = lvl.place(place_trials) if place_trials > 0 else []
    kmax = n / 8.0
    lvl.fourier_ic(0, kmax, 0.5, 42, 0)
    lvl.fourier_ic(1, kmax, 0.5, 42, 1)
    lvl.zero_components(2, 2)
    tau = 1.0
    lvl.advance(bg, tau, dt, warmup, use_graph=True)
    tau += warmup * dt
    ctx.synchronize()
    lvl.enable_timing(True)
    lvl.stage_times()  # reset
    ctx.synchronize()
    t0 = time.perf_counter()
    lvl.advance(bg, tau, dt, steps, use_graph=False)
    ctx.synchronize()
    t1 = time.perf_counter()
    ms, nsteps = lvl.stage_times()
    lvl.enable_timing(False)
    tau_end = 1.0 + (warmup + steps) * dt
    return lvl, (t1 - t0), ms, nsteps, tau_end, placement


def timing_check(capi, ctx, n, precision, bg, dt, steps, tile):
    """The same K steps timed three ways on one level: graph replay and
    direct launches without per-stage events (wall clock), and direct
    launches with HIP events between the stages."""
    lvl = capi.Level(ctx, n, 0.5, precision)
    lvl.fourier_ic(0, n / 8.0, 0.5, 42, 0)
    lvl.fourier_ic(1, n / 8.0, 0.5, 42, 1)
    lvl.zero_components(2, 2)
    if tile is not None:
        lvl.set_tile(*tile)
    tau = 1.0
    lvl.advance(bg, tau, dt, 2, use_graph=True)
    tau += 2 * dt
    out = {}
    for mode in ("graph", "direct", "events", "graph", "events"):
        lvl.enable_timing(mode == "events")
        lvl.stage_times()
        ctx.synchronize()
        t0 = time.perf_counter()
        lvl.advance(bg, tau, dt, steps, use_graph=(mode == "graph"))
        ctx.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            tau += dt
        rec = {"ms_per_step": 1e3 * (t1 - t0) / steps}
        if mode == "events":
            ms, ns = lvl.stage_times()
            rec["stage_ms"] = [m / ns for m in ms]
        out.setdefault(mode, []).append(rec)
        log(f"[timing] {mode}: {rec}")
    lvl.enable_timing(False)
    lvl.close()
    return out


def alloc_check(capi, ctx, n, precision, bg, dt, steps, tile, rounds=4):
    """K graph-replayed steps on `rounds` successive allocations of the
    same level (each freed before the next): placement sensitivity."""
    out = []
    for r in range(rounds):
        lvl = capi.Level(ctx, n, 0.5, precision)
        lvl.fourier_ic(0, n / 8.0, 0.5, 42, 0)
        lvl.fourier_ic(1, n / 8.0, 0.5, 42, 1)
        lvl.zero_components(2, 2)
        if tile is not None:
            lvl.set_tile(*tile)
        lvl.advance(bg, 1.0, dt, 2, use_graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        lvl.advance(bg, 1.2, dt, steps, use_graph=True)
        ctx.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        ptrs = lvl.device_ptrs() if hasattr(lvl, "device_ptrs") else None
        out.append({"round": r, "ms_per_step": ms, "ptrs": ptrs})
        log(f"[alloc] round {r}: {ms:.2f} ms/step ptrs {ptrs}")
        lvl.close()
    return out


def diagnostics(capi, lvl, bg, tau, n):
    R = capi.background_eval(bg, 0, tau)
    lam = capi.background_eval(bg, 3, tau)
    RpR = capi.background_eval(bg, 8, tau)
    e = lvl.energy_sums(R, lam, RpR, capi.AXS_MASK_B, 0.8)
    plain, weighted = lvl.count_plaquettes()
    dx = 0.5
    L = n * dx
    ell = (2.0 / 3.0) * plain * dx
    a_inv = bg.a_inv
    xi = ell / L ** 3 * ((a_inv - 1) ** 3 / a_inv ** 2) * tau * tau
    return {"rho_tot_unscreened": e[0] / e[13], "n_p": plain, "xi": xi, "R": R}


def diag_timing(capi, ctx, lvl, bg, tau, n, reps=3):
    """Wall time (best of `reps`, device-synchronised) of the per-output
    diagnostics on the FP32C level after the timed steps, with their
    algorithmic HBM bytes per cell (FP32C words are read as hi + lo):
    plaquette counts 16 B (psi), velocities 32 B (psi, Pi), energy sums
    32 B (psi, Pi), masked a_dot spectrum (fill 32 B + R2C FFT + binning)."""
    R = capi.background_eval(bg, 0, tau)
    lam = capi.background_eval(bg, 3, tau)
    RpR = capi.background_eval(bg, 8, tau)
    m_r = math.sqrt(lam)
    calls = {
        "count_plaquettes": (16, lambda: lvl.count_plaquettes()),
        "velocity": (32, lambda: lvl.velocity(R, RpR, m_r)),
        "energy_sums": (32, lambda: lvl.energy_sums(R, lam, RpR,
                                                     capi.AXS_MASK_B, 0.8)),
        "spectrum_masked_a_dot": (None, lambda: lvl.spectrum(
            capi.AXS_FIELD_MASKED_A_DOT, capi.AXS_MASK_B, 0.8, R)),
    }
    out = {}
    for name, (bpc, f) in calls.items():
        f()
        ctx.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            ctx.synchronize()
            best = min(best, time.perf_counter() - t0)
        out[name] = {"ms": 1e3 * best}
        if bpc:
            out[name]["alg_gbs"] = bpc * n ** 3 / best / 1e9
    return out


def pcie_check(capi, ctx, n, bg, dt, steps, tile):
    """The host-buffer boundary: wall time of axs_level_upload and
    axs_level_download (FP64 host state, 32 B per cell, across PCIe) of an
    n^3 FP32C level, and the step rate a caller would see if the state went
    down and up around every `steps` RK4 steps."""
    import numpy as np
    lvl = capi.Level(ctx, n, 0.5, capi.AXS_FP32C)
    if tile is not None:
        lvl.set_tile(*tile)
    host = np.random.default_rng(1).normal(size=(4, n, n, n))
    lvl.upload(host)  # first touch of the pages
    ctx.synchronize()
    t0 = time.perf_counter()
    lvl.upload(host)
    ctx.synchronize()
    t_up = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = lvl.download()
    t_down = time.perf_counter() - t0
    lvl.advance(bg, 1.0, dt, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    lvl.advance(bg, 1.1, dt, steps)
    ctx.synchronize()
    t_steps = time.perf_counter() - t0
    lvl.close()
    nbytes = host.nbytes
    del host, out
    ncell = n ** 3
    return {"n": n, "state_bytes": nbytes, "upload_s": t_up,
            "download_s": t_down, "upload_gbs": nbytes / t_up / 1e9,
            "download_gbs": nbytes / t_down / 1e9, "steps": steps,
            "device_cell_updates_per_s": ncell * steps / t_steps,
            "with_round_trip_cell_updates_per_s":
                ncell * steps / (t_steps + t_up + t_down)}


def precision_run(capi, ctx, n, bg, dt, steps, every, tile32=None, tile64=None):
    """FP32C and FP64 levels side by side from the same IC for `steps` RK4
    steps, with the precision metrics every `every` steps: max |dphi|/eta,
    relative energy error and delta xi."""
    levels = []
    for prec, tile in ((capi.AXS_FP32C, tile32), (capi.AXS_FP64, tile64)):
        lvl = capi.Level(ctx, n, 0.5, prec)
        lvl.fourier_ic(0, n / 8.0, 0.5, 42, 0)
        lvl.fourier_ic(1, n / 8.0, 0.5, 42, 1)
        lvl.zero_components(2, 2)
        if tile is not None:
            lvl.set_tile(*tile)
        levels.append(lvl)
    a, b = levels
    tau, done, series = 1.0, 0, []
    while done < steps:
        m = min(every, steps - done)
        for lvl in levels:
            lvl.advance(bg, tau, dt, m, use_graph=True)
        for _ in range(m):
            tau += dt
        done += m
        d32 = diagnostics(capi, a, bg, tau, n)
        d64 = diagnostics(capi, b, bg, tau, n)
        dpsi, _ = a.compare(b)
        row = {"step": done, "tau": tau,
               "max_dphi_over_eta": dpsi / d64["R"],
               "rel_energy_err": abs(d32["rho_tot_unscreened"] -
                                     d64["rho_tot_unscreened"]) /
               abs(d64["rho_tot_unscreened"]),
               "delta_xi": d32["xi"] - d64["xi"], "xi_fp64": d64["xi"],
               "n_p_fp32c": d32["n_p"], "n_p_fp64": d64["n_p"]}
        series.append(row)
        log(f"[precision] step {done} tau {tau:.2f}: max|dphi|/eta "
            f"{row['max_dphi_over_eta']:.3e} rel energy err "
            f"{row['rel_energy_err']:.3e} dxi {row['delta_xi']:.3e}")
    for lvl in levels:
        lvl.close()
    worst = {k: max(abs(r[k]) for r in series)
             for k in ("max_dphi_over_eta", "rel_energy_err", "delta_xi")}
    return {"workload": f"{n}^3, {steps} RK4 steps, dtau={dt}, metrics every "
                        f"{every} steps", "worst": worst, "series": series}


def cpu_baseline(n_sample=512, steps=5):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle_py as o
    L = o.load()
    s = np.zeros((4, n_sample, n_sample, n_sample))
    rng = np.random.default_rng(0)
    s[0] = 1.0 + 0.1 * rng.standard_normal(s[0].shape)
    s[1] = 0.1 * rng.standard_normal(s[1].shape)
    b = o.bg(2.0, c0=0.0, c1=2.0, tau_switch=1.0, has_switch=True)
    t0 = time.perf_counter()
    o.advance(s, 0.5, b, 1.0, 0.1, steps)
    dt = time.perf_counter() - t0
    return {"value": n_sample ** 3 * steps / dt, "unit": "cell-updates/s",
            "cores": int(L.or_num_threads()), "kind": "port",
            "sample": f"FP64 OpenMP oracle, {n_sample}^3 periodic level, "
                      f"{steps} RK4 steps ({dt:.1f} s)"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pcie-check", action="store_true",
                    help="only time the host-buffer upload / download of a "
                         "--pcie-n^3 FP32C level around --steps RK4 steps")
    ap.add_argument("--pcie-n", type=int, default=512)
    ap.add_argument("--no-diag", action="store_true",
                    help="skip the diagnostics timing after the FP32C run")
    ap.add_argument("--place-trials", type=int, default=8,
                    help="placement trials per level before the IC "
                         "(axs_level_place); 0 = keep the first placement")
    ap.add_argument("--cpu-n", type=int, default=512,
                    help="CPU baseline sample: level size")
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="CPU baseline sample: RK4 steps")
    ap.add_argument("--tile", type=str, default="",
                    help="stage tile variant,z_chunk (default: the tuning "
                         "table's, else built-in)")
    ap.add_argument("--no-fusion", action="store_true",
                    help="one launch per RK stage instead of two fused pairs")
    ap.add_argument("--fusion", type=int, default=0, choices=(0, 1, 2, 3),
                    help="fused pairs (1), fused pairs without a stored "
                         "accumulator (2), the whole step in one launch (3, "
                         "FP32C only); 0 = the library's default per "
                         "precision (FP32C 3, FP64 1)")
    ap.add_argument("--pair-tile", type=str, default="",
                    help="fused-pair / step variant,z_chunk (default: the "
                         "tuning table's, else built-in)")
    ap.add_argument("--alloc-check", action="store_true",
                    help="only time the FP32C level on successive "
                         "allocations")
    ap.add_argument("--timing-check", action="store_true",
                    help="only compare timing modes (graph / direct / "
                         "per-stage events) on the FP32C level")
    ap.add_argument("--precision-steps", type=int, default=0,
                    help="only run the FP32C-vs-FP64 precision comparison "
                         "for this many steps")
    ap.add_argument("--precision-every", type=int, default=100)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 device copies of this many MiB (counter "
                         "calibration: read N + write N bytes per copy) and 3 "
                         "sums over it (read N bytes)")
    args = ap.parse_args(argv)
    if args.copy_probe_mib:
        import torch
        nbytes = args.copy_probe_mib << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        for _ in range(3):  # read-only: a reduction over N bytes
            t = a.sum()
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nbytes, "sum": float(t)}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if world != 1:
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        return 2

    from axion_strings_adaptive_amd import capi
    import torch
    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        return 2
    ctx = capi.Context(0)
    n, dt = args.n, 0.1
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    # launch shapes: the tuning table's (capi.load applies it) unless given
    tile = None
    if args.tile:
        v, z = (int(t) for t in args.tile.split(","))
        tile = (v, z)

    # stage fusion mode per precision (0 = one launch per stage)
    fused = 0 if args.no_fusion else (args.fusion or 3)
    fused64 = 0 if args.no_fusion else (min(args.fusion, 2) or 1)
    pair_tile = (tuple(int(t) for t in args.pair_tile.split(","))
                 if args.pair_tile else None)

    if args.alloc_check:
        res = alloc_check(capi, ctx, n, capi.AXS_FP32C, bg, dt, args.steps,
                          tile)
        print(json.dumps({"alloc_check": res}), flush=True)
        ctx.close()
        return 0
    if args.pcie_check:
        res = pcie_check(capi, ctx, args.pcie_n, bg, dt, args.steps, tile)
        print(json.dumps({"pcie_check": res}), flush=True)
        ctx.close()
        return 0
    if args.timing_check:
        res = timing_check(capi, ctx, n, capi.AXS_FP32C, bg, dt, args.steps,
                           tile)
        print(json.dumps({"timing_check": res}), flush=True)
        ctx.close()
        return 0
    if args.precision_steps:
        res = precision_run(capi, ctx, n, bg, dt, args.precision_steps,
                            args.precision_every, tile, tile)
        print(json.dumps({"precision_run": res}), flush=True)
        ctx.close()
        return 0

    log(f"[bench] FP32C {n}^3: IC + {args.warmup} warmup + {args.steps} timed steps")
    a, wall32, ms32, ns32, tau_end, place32 = run_level(
        capi, ctx, n, capi.AXS_FP32C, bg, dt, args.warmup, args.steps, tile,
        args.place_trials, fused, pair_tile)
    if place32:
        log(f"[bench] FP32C placement trials (ms/step): "
            f"{', '.join(f'{t:.2f}' for t in place32)}")
    ncell = n ** 3
    value = ncell * args.steps / wall32
    d32 = diagnostics(capi, a, bg, tau_end, n)
    log(f"[bench] FP32C {value:.4e} cell-updates/s, {1e3 * wall32 / args.steps:.2f} ms/step")
    diag_ms = None
    if not args.no_diag:
        diag_ms = diag_timing(capi, ctx, a, bg, tau_end, n)
        log("[bench] FP32C diagnostics (ms): " + ", ".join(
            f"{k} {v['ms']:.2f}" for k, v in diag_ms.items()))

    res64 = None
    if not args.no_fp64:
        log(f"[bench] FP64 {n}^3 same IC and steps")
        b, wall64, ms64, ns64, _, place64 = run_level(
            capi, ctx, n, capi.AXS_FP64, bg, dt, args.warmup, args.steps,
            tile, args.place_trials, fused64, pair_tile)
        if place64:
            log(f"[bench] FP64 placement trials (ms/step): "
                f"{', '.join(f'{t:.2f}' for t in place64)}")
        d64 = diagnostics(capi, b, bg, tau_end, n)
        dpsi, _ = a.compare(b)
        value64 = ncell * args.steps / wall64
        res64 = {
            "value": value64, "ms_per_step": 1e3 * wall64 / args.steps,
            "stage_ms": [m / ns64 for m in ms64], "stage_fusion": fused64,
            "launch_shape": b.launch_shape(),
            "speedup_fp32c_over_fp64": value / value64,
            "max_dphi_over_eta": dpsi / d64["R"],
            "rel_energy_err": abs(d32["rho_tot_unscreened"] -
                                  d64["rho_tot_unscreened"]) /
            abs(d64["rho_tot_unscreened"]),
            "delta_xi": d32["xi"] - d64["xi"], "xi_fp64": d64["xi"],
            "tau": tau_end, "placement_trials_ms": place64,
        }
        log(f"[bench] FP64 {value64:.4e} cell-updates/s; ratio {value / value64:.3f}")
        b.close()

    st = [m / ns32 for m in ms32]  # ms per launch of each stage (pair)
    shape32 = a.launch_shape()  # the tuning table's or the given shapes
    pkey = (shape32["pair_variant"], shape32["pair_z_chunk"])
    if fused == 3:
        # one launch per step: the step kernel is the only one
        achieved = STEP_BYTES_ONE * ncell / (st[0] * 1e-3) / 1e9
        per_cell, tsrc = measured_traffic("step", pkey)
        roofline = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": per_cell * ncell if per_cell and n == 1024 else None,
            "traffic_unit": "bytes per launch",
            "traffic_source": tsrc if n == 1024 else "1024^3 only",
            "kernel": step_kernel_name(pkey[0]) + " (the whole RK4 step)",
            "algorithmic_bytes_per_cell": STEP_BYTES_ONE,
            "step_launch_ms": st[0], "step_gbs": achieved}
    elif fused:
        # roofline of the dominant kernel: the longer of the two pair
        # launches (one each per step)
        k = 0 if st[0] > st[1] else 1
        pb = PAIR_BYTES if fused == 1 else PAIR_BYTES_NOACC
        achieved = pb[1][k] * ncell / (st[k] * 1e-3) / 1e9
        per_cell, tsrc = (measured_traffic("pair", (fused,) + pkey)
                          if k == 1 else (None, "not measured"))
        roofline = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": per_cell * ncell if per_cell and n == 1024 else None,
            "traffic_unit": "bytes per launch",
            "traffic_source": tsrc if n == 1024 else "1024^3 only",
            "kernel": f"k_pair<float, FP32C, {k + 1 + 2 * (fused - 1)}> "
                      f"(RK stages {'1+2' if k == 0 else '3+4'})",
            "algorithmic_bytes_per_cell": pb[1][k],
            "pair_ms": st[:2],
            "pair_gbs": [pb[1][i] * ncell / (st[i] * 1e-3) / 1e9
                         for i in range(2)],
            "step_gbs": sum(pb[1]) * ncell / (sum(st[:2]) * 1e-3) / 1e9}
    else:
        # roofline of the dominant kernel: the stage-2/3 launch (two per step)
        bytes23 = STAGE_BYTES[1][1] * ncell
        achieved = bytes23 / (0.5 * (st[1] + st[2]) * 1e-3) / 1e9
        per_cell, tsrc = measured_traffic(
            "stage", (shape32["tile_variant"], shape32["tile_z_chunk"]))
        roofline = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": per_cell * ncell if per_cell and n == 1024 else None,
            "traffic_unit": "bytes per launch",
            "traffic_source": tsrc if n == 1024 else "1024^3 only",
            "kernel": "k_stage<float, FP32C, MODE_S23> (RK stages 2/3)",
            "algorithmic_bytes_per_cell": STAGE_BYTES[1][1],
            "stage_ms": st,
            "stage_gbs": [STAGE_BYTES[1][i] * ncell / (st[i] * 1e-3) / 1e9
                          for i in range(4)],
            "step_gbs": sum(STAGE_BYTES[1]) * ncell / (sum(st) * 1e-3) / 1e9}
    a.close()

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, bounded sample)")
        cpu = cpu_baseline(args.cpu_n, args.cpu_steps)

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall32 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 periodic RK4, dx=0.5, dtau=0.1, a=tau "
                               "from tau=1, c0=0 -> c1=2 at tau=1, seeded "
                               "Gaussian IC (seed 42, |p|<=N/8)",
                   "precision": "fp32 storage + Kahan-compensated y_n "
                                "(fp32 hi + fp32 lo)",
                   "stage_fusion": fused,
                   "launch_shape": shape32,
                   "placement_trials_ms": place32},
        "fp64": res64, "roofline": roofline, "cpu_baseline": cpu,
        "diagnostics_fp32c": diag_ms,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
