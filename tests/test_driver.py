"""The host C++ driver (axion_strings): parameter validation without a GPU
(the reference's just_check_params path, AxionStringsParams.hpp:30-1208,
with its error messages), and on the GPU an end-to-end run whose
network_scalars.dat rows agree with the same quantities computed through the
C ABI directly."""

import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "axion_strings_adaptive_amd", "_build",
                      "axion_strings")
PARAMS = os.path.join(ROOT, "axion_strings_adaptive_amd", "params")


def run(args, cwd=None):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    return subprocess.run([DRIVER, *args], capture_output=True, text=True,
                          cwd=cwd, timeout=300)


@pytest.mark.parametrize("name", ["params_plain_64.txt", "params_scale_512.txt",
                                  "params_precision_1024.txt"])
def test_shipped_params_validate(name):
    r = run([os.path.join(PARAMS, name), "--check"])
    assert r.returncode == 0, r.stderr
    assert "parameters ok" in r.stdout


BAD = [
    (["axion_strings.c0=2.0"], "log_mr_over_h_i"),           # Moore from start
    (["axion_strings.compute_spectrum=1", "axion_strings.masking.scheme=A"],
     "compute_spectrum"),
    (["axion_strings.masking.scheme=C"], "masking.scheme"),
    (["axion_strings.c1=2.0"], "c1"),                         # no tau_switch
    (["axion_strings.log_mr_over_h_switch=1.0", "axion_strings.c1=2",
      "axion_strings.tau_switch=3"], "log_mr_over_h_switch"),
    (["axion_strings.a_inv=1.0"], "a_inv"),
    (["axion_strings.tagging.gradient_threshold=0"], "gradient_threshold"),
    (["axion_strings.masking.scheme=B", "axion_strings.masking.threshold=1.5"],
     "masking.threshold"),
    (["axion_strings.k_max_over_mr=-1"], "k_max_over_mr"),
    (["axion_strings.output_cadence_steps=0"], "output_cadence_steps"),
    (["axion_strings.precision=fp16"], "precision"),
]


@pytest.mark.parametrize("extra,key", BAD)
def test_parameter_errors_name_the_key(extra, key):
    p = os.path.join(PARAMS, "params_plain_64.txt")
    r = run([p, "--check", *extra])
    if key == "precision":  # read when the level is set up, not in --check
        r = run([p, *extra, "evolution.max_steps=0"])
    assert r.returncode == 3, (r.stdout, r.stderr)
    assert key in r.stderr


@pytest.mark.gpu
def test_driver_end_to_end_matches_capi(tmp_path, ctx, oracle):
    from axion_strings_adaptive_amd import capi
    p = os.path.join(PARAMS, "params_plain_64.txt")
    r = run([p, "amr.n_cell=32 32 32", "geometry.prob_extent=16 16 16",
             "axion_strings.k_max_over_mr=8", "evolution.max_steps=40",
             "axion_strings.output_cadence_steps=20",
             "axion_strings.compute_spectrum=1", "--out", str(tmp_path)])
    assert r.returncode == 0, r.stderr
    rows = np.loadtxt(tmp_path / "network_scalars.dat")
    assert rows.shape == (2, 25)
    # same run through the C ABI
    lvl = capi.Level(ctx, 32, 0.5, capi.AXS_FP64)
    lvl.fourier_ic(0, 4.0, 0.5, 42, 0)
    lvl.fourier_ic(1, 4.0, 0.5, 42, 1)
    lvl.zero_components(2, 2)
    bg = capi.make_background(2.0)
    lvl.advance(bg, 1.0, 0.1, 40)
    tau = rows[1, 0]
    assert tau == pytest.approx(5.0, rel=1e-12)
    R, lam = capi.background_eval(bg, 0, tau), capi.background_eval(bg, 3, tau)
    e = lvl.energy_sums(R, lam, 1.0 / tau, capi.AXS_MASK_B, 0.8)
    plain, weighted = lvl.count_plaquettes()
    assert rows[1, 1] == plain and rows[1, 2] == weighted
    assert rows[1, 6] == pytest.approx(e[0] / e[13], rel=1e-9)
    assert rows[1, 7] == pytest.approx(e[1] / e[12], rel=1e-9)
    xi = (2.0 / 3.0) * plain * 0.5 / 16.0 ** 3 * 0.25 * tau * tau
    assert rows[1, 3] == pytest.approx(xi, rel=1e-9)
    spec = np.loadtxt(tmp_path / "axion_spectrum.dat")
    assert spec.shape[1] == 15
