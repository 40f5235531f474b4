"""The C++ driver on a refined run (amr.max_level = 2): the parameter file
of the batching form, run from the driver, gives the same hierarchy as
amr.Hierarchy built with the same keys -- the same cells per level after
every regrid, the same composite pierced-plaquette counts and the same
composite energies in network_scalars.dat -- and the regrid buffer policy
(AxionStringsParams.hpp:403-494) derives and checks amr.n_error_buf and
amr.regrid_int."""

import os
import re

import numpy as np
import pytest

from test_driver import PARAMS, run

AMR = os.path.join(PARAMS, "params_amr_256_two_levels.txt")


def test_amr_params_validate():
    r = run([AMR, "--check"])
    assert r.returncode == 0, r.stderr


def test_regrid_buffer_policy_derives_and_checks():
    base = [AMR, "--check", "axion_strings.tagging.regrid_interval_steps=4",
            "axion_strings.tagging.buffer_safety_factor=1.5"]
    # ceil(1.5 * 4 * 0.2) = 2 cells; regrid_int must equal 4
    r = run(base + ["amr.n_error_buf=2", "amr.regrid_int=4"])
    assert r.returncode == 0, r.stderr
    r = run(base + ["amr.n_error_buf=1", "amr.regrid_int=4"])
    assert r.returncode == 3 and "n_error_buf" in r.stderr
    r = run(base + ["amr.n_error_buf=2"])
    assert r.returncode == 3 and "regrid_int" in r.stderr
    r = run(base + ["amr.n_error_buf=2", "amr.regrid_int=4",
                    "axion_strings.tagging.buffer_safety_factor=0.5"])
    assert r.returncode == 3 and "buffer_safety_factor" in r.stderr


@pytest.mark.parametrize("extra,key", [
    (["amr.blocking_factor=12"], "blocking_factor"),
    (["amr.max_grid_size=128"], "max_grid_size"),
    (["amr.max_grid_size=24"], "max_grid_size"),
    (["amr.n_error_buf=9"], "n_error_buf"),
    (["amr.max_level=9"], "max_level"),
])
def test_amr_key_errors(extra, key):
    r = run([AMR, "--check", *extra])
    assert r.returncode == 3 and key in r.stderr, (r.stdout, r.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("n0,steps", [(64, 4), (256, 2)])
def test_driver_hierarchy_matches_python(tmp_path, ctx, n0, steps):
    from axion_strings_adaptive_amd import amr, capi
    L = n0 * 0.5
    print(f"[driver amr] n0={n0} steps={steps}", flush=True)
    r = run([AMR, f"amr.n_cell={n0} {n0} {n0}",
             f"geometry.prob_extent={L} {L} {L}",
             f"axion_strings.k_max_over_mr={n0 // 4}",
             f"evolution.max_steps={steps}", "--out", str(tmp_path)])
    assert r.returncode == 0, r.stdout + r.stderr
    rows = np.atleast_2d(np.loadtxt(tmp_path / "network_scalars.dat"))
    assert rows.shape[0] == steps // 2
    cells_lines = [list(map(int, re.findall(r" (\d+) \(\d+ boxes\)", ln)))
                   for ln in r.stdout.splitlines()
                   if ln.startswith("  [AxionStrings AMR] levels")]
    assert len(cells_lines) == rows.shape[0]
    spec = np.loadtxt(tmp_path / "axion_spectrum.dat")
    assert spec.shape[1] == 15 and np.all(np.isfinite(spec))

    h = amr.Hierarchy(ctx, n0=n0, levels=2, precision=capi.AXS_FP32C,
                      amp_thr=0.5, n_error_buf=2, blocking_factor=16,
                      max_grid_size=64, regrid_int=2, verbose=False)
    for k in range(rows.shape[0]):
        h.step()
        h.step()
        tau = h.tau
        assert rows[k, 0] == pytest.approx(tau, rel=1e-12)
        assert cells_lines[k] == h.ncells
        c = h.plaquette_counts().sum(axis=0)
        assert rows[k, 1] == c[0] and rows[k, 2] == c[1]
        R = capi.background_eval(h.bg, 0, tau)
        lam = capi.background_eval(h.bg, 3, tau)
        RpR = capi.background_eval(h.bg, 8, tau)
        acc = np.zeros(14)
        capi.check(h.lib.axs_hierarchy_energy_sums(
            h.h, R, lam, RpR, capi.C.byref(capi.Masking(capi.AXS_MASK_B, 0.8)),
            capi.dptr(acc)))
        assert acc[13] == pytest.approx(L ** 3, rel=1e-12)
        assert rows[k, 6] == pytest.approx(acc[0] / acc[13], rel=1e-9)
        assert rows[k, 7] == pytest.approx(acc[1] / acc[12], rel=1e-9)
        dxs = [0.5 / 2 ** l for l in range(len(h.levels))]
        ell = sum((2.0 / 3.0) * float(n) * d for n, d in
                  zip(h.plaquette_counts()[:, 0], dxs))
        xi = ell / L ** 3 * 0.25 * tau * tau
        assert rows[k, 3] == pytest.approx(xi, rel=1e-9)
    assert h.nregrids == rows.shape[0]
    h.close()
    ctx.trim()
