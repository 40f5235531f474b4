"""Box planning in the host driver (AxionStringsParams.hpp:594-1130
apply_box_plan, BoxPlan.hpp): the keys `axion_strings --check` derives from
N, N1, N2 against the oracle's box plan, the cross-checks of keys already
set and the error cases, each naming its key; on the GPU, a planned run
that stops at the first step with tau >= tau_f (AxionStringsLevel.cpp:
2007-2025)."""

import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "axion_strings_adaptive_amd", "_build",
                      "axion_strings")

BASE = {
    "amr.max_level": "0",
    "geometry.is_periodic": "1 1 1",
    "evolution.dt_multiplier": "0.2",
    "axion_strings.a_inv": "2.0",
    "axion_strings.c0": "0.0",
    "axion_strings.log_mr_over_h_i": "0.0",
    "axion_strings.N": "512",
    "axion_strings.N1": "1.0",
    "axion_strings.N2": "4.0",
    "axion_strings.ic_mode": "fourier",
    "axion_strings.k_max_over_mr": "8",
    "axion_strings.mean_square_variance": "0.5",
    "axion_strings.seed": "42",
    "axion_strings.masking.scheme": "B",
    "axion_strings.output_cadence_steps": "50",
}


def check(tmp_path, over=(), drop=(), args=()):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    d = dict(BASE)
    for k in drop:
        d.pop(k)
    d.update(dict(over))
    p = tmp_path / "params.txt"
    p.write_text("".join(f"{k} = {v}\n" for k, v in d.items()))
    r = subprocess.run([DRIVER, str(p), *args], capture_output=True,
                       text=True, timeout=300, cwd=str(tmp_path))
    derived = {}
    for line in r.stdout.splitlines():
        if line.startswith("derived "):
            k, v = line[len("derived "):].split(" = ")
            derived[k] = [float(x) for x in v.split()]
    return r, derived


@pytest.mark.parametrize("N,N1,N2,c", [(512, 1.0, 4.0, 0.0), (800, 1.5, 6.0, 1.0),
                                       (300, 0.7, 3.0, 0.5)])
def test_general_plan_derives_grid_box_and_stop(tmp_path, oracle, N, N1, N2, c):
    r, d = check(tmp_path, {"axion_strings.N": N, "axion_strings.N1": N1,
                            "axion_strings.N2": N2, "axion_strings.c0": c},
                 args=["--check"])
    assert r.returncode == 0, r.stderr
    tau_f, L, dx, _ = oracle.general_box_plan(N, N1, N2, 2.0, c)
    assert d["amr.n_cell"] == [N] * 3
    assert d["geometry.prob_extent"] == [L] * 3
    assert d["axion_strings.derived_tau_f"] == [tau_f]
    assert d["evolution.stop_time"] == [-1.0]
    assert "axion_strings.pre_evolution.gamma" not in d  # not relaxing


def test_relaxed_run_derives_gamma_as_one_over_dx(tmp_path, oracle):
    over = {"axion_strings.ic_mode": "fourier_relaxed",
            "axion_strings.pre_evolution.xi_target": "1.0",
            "axion_strings.pre_evolution.k_max_over_mr": "8",
            "axion_strings.pre_evolution.mean_square_variance": "0.5",
            "axion_strings.pre_evolution.seed": "1"}
    r, d = check(tmp_path, over, args=["--check"])
    assert r.returncode == 0, r.stderr
    dx = oracle.general_box_plan(512, 1.0, 4.0, 2.0, 0.0)[2]
    assert d["axion_strings.pre_evolution.gamma"] == [1.0 / dx]
    # a value that differs only warns, and is kept
    over["axion_strings.pre_evolution.gamma"] = "3.0"
    r, d = check(tmp_path, over, args=["--check"])
    assert r.returncode == 0, r.stderr
    assert "WARNING" in r.stdout
    assert "axion_strings.pre_evolution.gamma" not in d


def test_keys_already_set_are_checked(tmp_path, oracle):
    tau_f, L, dx, _ = oracle.general_box_plan(512, 1.0, 4.0, 2.0, 0.0)
    ok = {"amr.n_cell": "512 512 512",
          "geometry.prob_extent": " ".join([repr(L * (1 + 1e-8))] * 3),
          "evolution.stop_time": "7.5"}
    r, d = check(tmp_path, ok, args=["--check"])
    assert r.returncode == 0, r.stderr
    assert set(d) == {"axion_strings.derived_tau_f"}
    for key, val in [("amr.n_cell", "256 256 256"),
                     ("geometry.prob_extent", " ".join([repr(L * 1.001)] * 3))]:
        r, _ = check(tmp_path, {key: val}, args=["--check"])
        assert r.returncode == 3, r.stdout
        assert key in r.stderr


@pytest.mark.parametrize("over,key", [
    ({"axion_strings.N": "0"}, "axion_strings.N"),
    ({"axion_strings.N2": "-1"}, "axion_strings.N"),
    ({"axion_strings.log_mr_over_h_i": "40"}, "axion_strings.N"),  # tau_f<=tau_i
    ({"amr.max_level": "2", "axion_strings.N": "514"}, "axion_strings.N"),
    ({"axion_strings.log_mr_over_h_switch": "3.0", "amr.max_level": "1",
      "axion_strings.c0": "1.0"}, "axion_strings.N"),
    ({"axion_strings.log_mr_over_h_switch": "3.0", "axion_strings.c0": "1.0",
      "axion_strings.target_log_range": "30"}, "target_log_range"),
])
def test_plan_errors_name_the_key(tmp_path, over, key):
    r, _ = check(tmp_path, over, args=["--check"])
    assert r.returncode == 3, r.stdout
    assert key in r.stderr


def test_missing_N1_is_an_error(tmp_path):
    r, _ = check(tmp_path, drop=["axion_strings.N1"], args=["--check"])
    assert r.returncode == 3
    assert "axion_strings.N1" in r.stderr


def test_no_N_means_no_plan(tmp_path):
    r, d = check(tmp_path, {"amr.n_cell": "64 64 64",
                            "geometry.prob_extent": "32 32 32"},
                 drop=["axion_strings.N"], args=["--check"])
    assert r.returncode == 0, r.stderr
    assert d == {}


@pytest.mark.parametrize("lev", [1, 2, 3])
def test_refined_plan_derives_base_grid_and_schedule(tmp_path, oracle, lev):
    r, d = check(tmp_path, {"amr.max_level": lev, "axion_strings.N": 1024},
                 args=["--check"])
    assert r.returncode == 0, r.stderr
    L, dxf, dxb, tau_f, nb, la = oracle.amr_box_plan(1024, 1.0, 4.0, 2.0, 0.0,
                                                     lev)
    assert d["amr.n_cell"] == [nb] * 3
    assert d["geometry.prob_extent"] == [L] * 3
    assert d["axion_strings.derived_tau_f"] == [tau_f]
    assert d["axion_strings.derived_level_add_log_mr_over_h"] == la


def test_moore_plan_through_a_switch(tmp_path, oracle):
    """c0 = 1 switching to c = a_inv where log(m_r/H) = 4: dx from N2 at the
    switch, tau_f = the Moore plan's tau_end."""
    N, N1, N2 = 384, 1.0, 2.0
    r, d = check(tmp_path, {"axion_strings.c0": "1.0",
                            "axion_strings.log_mr_over_h_i": "2.0",
                            "axion_strings.log_mr_over_h_switch": "4.0",
                            "axion_strings.N": N, "axion_strings.N1": N1,
                            "axion_strings.N2": N2,
                            "amr.n_cell": f"{N} {N} {N}"},
                 args=["--check"])
    assert r.returncode == 0, r.stderr
    L = oracle.load()
    tau_sw = L.or_bg_tau_from_log_mr_over_h(oracle.bg(2.0, c0=1.0), 4.0)
    b = oracle.bg(2.0, c0=1.0, c1=2.0, tau_switch=tau_sw, has_switch=True)
    gamma = 1.0 / L.or_bg_H_over_mr_direct(b, tau_sw)
    dx, Lt, D, tau_end = oracle.moore_box_plan(
        N, N1, N2, gamma, 2.0, 1.0, L.or_bg_R(b, tau_sw),
        math.sqrt(L.or_bg_lambda(b, tau_sw)), tau_sw)
    assert d["geometry.prob_extent"][0] == pytest.approx(Lt, rel=1e-14)
    assert d["axion_strings.derived_tau_f"][0] == pytest.approx(tau_end,
                                                               rel=1e-14)
    assert "amr.n_cell" not in d  # set by hand in Moore runs


def test_shipped_planned_params(tmp_path):
    p = os.path.join(ROOT, "axion_strings_adaptive_amd", "params",
                     "params_planned_256.txt")
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    r = subprocess.run([DRIVER, p, "--check"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "derived axion_strings.derived_tau_f = 256" in r.stdout
    assert "derived geometry.prob_extent = 256 256 256" in r.stdout
    assert "derived axion_strings.pre_evolution.gamma = 1" in r.stdout


@pytest.mark.gpu
def test_planned_run_stops_at_tau_f(tmp_path, oracle):
    """N = 64, N1 = 1, N2 = 1, c = 1 from log(m_r/H) = 3: tau_f = 64, L = 64,
    dx = 1, dt = 0.2; the run takes steps while tau < tau_f and no more."""
    r, _ = check(tmp_path, {"axion_strings.c0": "1.0",
                            "axion_strings.log_mr_over_h_i": "3.0",
                            "axion_strings.N": "64", "axion_strings.N1": "1.0",
                            "axion_strings.N2": "1.0",
                            "axion_strings.output_cadence_steps": "40"},
                 args=["--out", str(tmp_path)])
    assert r.returncode == 0, (r.stdout, r.stderr)
    tau_i = math.exp(3.0)
    n = 0
    t = tau_i
    while t < 64.0:
        t += 0.2
        n += 1
    done = [l for l in r.stdout.splitlines() if l.startswith("done:")][0]
    assert f"steps = {n}" in done
    tau_end = float(done.split("tau = ")[1].split()[0])
    assert tau_end >= 64.0 > tau_end - 0.2 - 1e-9
    assert "tau/tau_f" in r.stdout
    rows = np.loadtxt(tmp_path / "network_scalars.dat", ndmin=2)
    assert rows.shape[0] >= 1 and rows[-1, 0] <= tau_end + 1e-9
