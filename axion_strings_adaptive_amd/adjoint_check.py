"""The differentiation form at full size: gradient of the final-time total
gradient energy with respect to the initial field and lambda, on a 128^3
level, c = 2 (Moore, reached by a switch at tau = 1), 400 RK4 steps,
checkpoints every 20 steps; FP32C forward with the FP64 adjoint, checked
against the FP64-forward gradient and against central finite differences
(eps = 1e-3) of the FP64 forward on 64 seeded random directions.

    python -m axion_strings_adaptive_amd.adjoint_check [--n 128 --steps 400]

Prints one JSON line with the timings and the error figures.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import capi
from .torch_adjoint import GradEnergyProblem


def seeded_state(ctx, n, seed):
    """The seeded Gaussian IC (|p| <= N/8, <psi_i^2> = 0.5) generated on the
    device, plus small seeded momenta so every component has a gradient."""
    lvl = capi.Level(ctx, n, 0.5, capi.AXS_FP64)
    lvl.fourier_ic(0, n / 8.0, 0.5, seed, 0)
    lvl.fourier_ic(1, n / 8.0, 0.5, seed, 1)
    lvl.zero_components(2, 2)
    s = lvl.download()
    lvl.close()
    rng = np.random.default_rng(seed)
    s[2:] = 0.05 * rng.standard_normal(s[2:].shape)
    return np.ascontiguousarray(s)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--checkpoint-every", type=int, default=20)
    ap.add_argument("--directions", type=int, default=64)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=42)
    args = ap.parse_args(argv)
    ctx = capi.Context(0)
    bg = capi.make_background(2.0, c0=0.0, c1=2.0, tau_switch=1.0,
                              has_switch=True)
    s0 = seeded_state(ctx, args.n, args.seed)
    lam0 = 1.0
    p32 = GradEnergyProblem(ctx, args.n, precision=capi.AXS_FP32C, bg=bg,
                            nsteps=args.steps,
                            checkpoint_every=args.checkpoint_every)
    p64 = GradEnergyProblem(ctx, args.n, precision=capi.AXS_FP64, bg=bg,
                            nsteps=args.steps,
                            checkpoint_every=args.checkpoint_every)
    ctx.synchronize()
    t0 = time.perf_counter()
    J32 = p32.forward_value(s0, lam0)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J32b, g32, l32 = p32.value_and_grad(s0, lam0)
    t_grad = time.perf_counter() - t0
    J64, g64, l64 = p64.value_and_grad(s0, lam0)
    print(f"[adjoint] J fp32c {J32:.12e} fp64 {J64:.12e}; forward "
          f"{t_fwd:.2f} s, forward + adjoint {t_grad:.2f} s",
          file=sys.stderr, flush=True)
    rel_g = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    rel_l = abs(l32 - l64) / abs(l64)
    rng = np.random.default_rng(args.seed + 1)
    errs, errs32 = [], []
    for d in range(args.directions):
        v = rng.standard_normal(s0.shape)
        fd = (p64.forward_value(s0 + args.eps * v, lam0) -
              p64.forward_value(s0 - args.eps * v, lam0)) / (2 * args.eps)
        ad = float(np.sum(g64 * v))
        ad32 = float(np.sum(g32 * v))
        errs.append(abs(ad - fd) / abs(fd))
        errs32.append(abs(ad32 - fd) / abs(fd))
        if (d + 1) % 16 == 0:
            print(f"[adjoint] {d + 1} directions: max rel err fp64 "
                  f"{max(errs):.3e}, fp32c {max(errs32):.3e}",
                  file=sys.stderr, flush=True)
    fdl = (p64.forward_value(s0, lam0 + args.eps) -
           p64.forward_value(s0, lam0 - args.eps)) / (2 * args.eps)
    out = {
        "workload": f"{args.n}^3, {args.steps} RK4 steps, checkpoints every "
                    f"{args.checkpoint_every}, c0=0 -> c1=2 at tau=1, "
                    "seeded IC",
        "J_fp32c": J32, "J_fp64": J64,
        "forward_s": t_fwd, "forward_plus_adjoint_s": t_grad,
        "cell_updates_per_s_forward": args.n ** 3 * args.steps / t_fwd,
        "grad_rel_err_fp32c_vs_fp64": rel_g,
        "lambda_grad_rel_err_fp32c_vs_fp64": rel_l,
        "fd_directions": args.directions, "eps": args.eps,
        "fd_max_rel_err_fp64_adjoint": max(errs),
        "fd_max_rel_err_fp32c_adjoint": max(errs32),
        "fd_lambda_rel_err_fp64": abs(l64 - fdl) / abs(fdl),
    }
    print(json.dumps(out), flush=True)
    for p in (p32, p64):
        p.level.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
