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
                                  "params_precision_1024.txt",
                                  "params_relaxed_64.txt",
                                  "params_planned_256.txt"])
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


# AxionStringsParams.hpp:135-261 (the fourier_relaxed readers)
BAD_RELAXED = [
    (["axion_strings.pre_evolution.gamma=0"], "gamma"),
    (["axion_strings.pre_evolution.xi_target=-1"], "xi_target"),
    (["axion_strings.pre_evolution.check_interval_fine=0"],
     "check_interval_coarse"),
    (["axion_strings.pre_evolution.check_threshold_fine=2.0"],
     "check_threshold_medium"),
    (["axion_strings.pre_evolution.handoff_transition_n_periods=-1"],
     "handoff_transition_n_periods"),
    (["axion_strings.pre_evolution.k_max_over_mr=0"], "k_max_over_mr"),
    (["axion_strings.ic_mode=plane"], "ic_mode"),
]


@pytest.mark.parametrize("extra,key", BAD_RELAXED)
def test_relaxed_parameter_errors_name_the_key(extra, key):
    r = run([os.path.join(PARAMS, "params_relaxed_64.txt"), "--check", *extra])
    assert r.returncode == 3, (r.stdout, r.stderr)
    assert key in r.stderr


def test_relaxed_requires_xi_target():
    p = os.path.join(PARAMS, "params_relaxed_64.txt")
    with open(p) as f:
        text = "".join(l for l in f if "xi_target" not in l)
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        f.write(text)
    try:
        r = run([f.name, "--check"])
    finally:
        os.unlink(f.name)
    assert r.returncode == 3 and "xi_target" in r.stderr


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


def _pre_table(capi, oracle, gamma, t, dt, nsteps):
    rows = []
    for _ in range(nsteps):
        c = oracle.pre_stage_coeffs(gamma, t, dt)
        rows += [capi.RhsCoeffs(*c[3 * k:3 * k + 3]) for k in range(4)]
        t += dt
    return rows, t


@pytest.mark.gpu
def test_driver_relaxed_ic_matches_capi(tmp_path, ctx, oracle):
    """ic_mode = fourier_relaxed end to end (AxionStringsLevel.cpp:784-822,
    :869-876, :1044-1191, :2027-2125): the driver's xi checks, handoff step,
    and first main-phase row agree with the same run assembled from the C
    ABI and the oracle's pre-evolution, handoff and blend functions."""
    from axion_strings_adaptive_amd import capi
    n, dx, dt, gamma, kmax = 32, 0.5, 0.1, 1.0, 6.0
    L = n * dx
    coarse, medium, fine = 10, 5, 2
    thr_med, thr_fine = 1.5, 1.1

    def fresh():
        lvl = capi.Level(ctx, n, dx, capi.AXS_FP64)
        lvl.fourier_ic(0, kmax, 0.5, 1, 0)
        lvl.fourier_ic(1, kmax, 0.5, 1, 1)
        lvl.zero_components(2, 2)
        return lvl

    def xi_of(n_p):  # tau_i = 1, a_inv = 2
        return oracle.load().or_xi_from_plaquette_count(n_p, dx, L, 2.0, 1.0)

    # N_p after every relaxing step (the checks do not change the state)
    lvl = fresh()
    t, n_p = 0.0, [None]
    for _ in range(150):
        rows, t = _pre_table(capi, oracle, gamma, t, dt, 1)
        lvl.advance_coeffs(rows, dt)
        n_p.append(lvl.count_plaquettes()[0])
    lvl.close()
    # a target the network reaches after a few coarse checks
    checks = [s for s in range(coarse, 151, coarse)]
    m = next(i for i in range(3, len(checks))
             if n_p[checks[i]] < min(n_p[c] for c in checks[:i]))
    xi_target = xi_of(n_p[checks[m]])
    # the reference's adaptive cadence on the recorded counts
    step, interval, check_steps = 0, coarse, []
    while True:
        step += interval
        check_steps.append(step)
        ratio = xi_of(n_p[step]) / xi_target
        if ratio <= 1.0:
            break
        interval = fine if ratio <= thr_fine else (
            medium if ratio <= thr_med else coarse)
    h = step
    first_row = (h // 20 + 1) * 20

    r = run([os.path.join(PARAMS, "params_relaxed_64.txt"),
             f"amr.n_cell={n} {n} {n}", f"geometry.prob_extent={L} {L} {L}",
             f"axion_strings.pre_evolution.xi_target={xi_target!r}",
             f"axion_strings.pre_evolution.check_interval_medium={medium}",
             f"evolution.max_steps={first_row}", "--out", str(tmp_path)])
    assert r.returncode == 0, r.stderr
    pre = np.atleast_2d(np.loadtxt(tmp_path / "pre_evolution_scalars.dat"))
    assert pre.shape == (len(check_steps), 13)
    assert [int(v) for v in pre[:, 4]] == [n_p[s] for s in check_steps]
    tp = [0.0]
    for _ in range(h):
        tp.append(tp[-1] + dt)
    assert list(pre[:, 0]) == pytest.approx([tp[s] for s in check_steps],
                                            rel=1e-10)
    assert list(pre[:, 5]) == pytest.approx([xi_of(n_p[s]) for s in check_steps],
                                            rel=1e-9)
    assert "transitioning in place" in r.stdout

    # the same run through the C ABI: relax h steps, hand off, blend
    lvl = fresh()
    rows, t = _pre_table(capi, oracle, gamma, 0.0, dt, h)
    lvl.advance_coeffs(rows, dt)
    bg = capi.make_background(2.0)
    ob = oracle.bg(2.0)
    kappa, common, R_pre, _ = oracle.handoff_constants(ob, gamma, t, 1.0)
    lvl.handoff_rescale(kappa, common, R_pre)
    lam_s = oracle.pre_bg(gamma, "lambda", t)
    dtau = oracle.handoff_dtau(3.0, capi.background_eval(bg, 3, 1.0), lam_s)
    rows, tau = [], 1.0
    for _ in range(first_row - h):
        for ts in (tau, tau + 0.5 * dt, tau + 0.5 * dt, tau + dt):
            c = capi.rhs_coeffs(bg, ts)
            lam, cur = oracle.handoff_blend(ts, 1.0, dtau, lam_s, 1.0,
                                            c.lambda_, c.curvature_coeff)
            rows.append(capi.RhsCoeffs(cur, lam, c.R_squared))
        tau += dt
    lvl.advance_coeffs(rows, dt)
    net = np.atleast_2d(np.loadtxt(tmp_path / "network_scalars.dat"))
    assert net.shape == (1, 25)
    assert net[0, 0] == pytest.approx(tau, rel=1e-12)
    plain, weighted = lvl.count_plaquettes()
    assert net[0, 1] == plain and net[0, 2] == weighted
    R = capi.background_eval(bg, 0, tau)
    lam = capi.background_eval(bg, 3, tau)
    e = lvl.energy_sums(R, lam, 1.0 / tau, capi.AXS_MASK_B, 0.8)
    assert net[0, 6] == pytest.approx(e[0] / e[13], rel=1e-9)
    lvl.close()


def test_restart_missing_file_names_the_key(tmp_path):
    p = os.path.join(PARAMS, "params_plain_64.txt")
    r = run([p, "amr.restart=" + str(tmp_path / "nope.axs"), "--out",
             str(tmp_path)])
    assert r.returncode == 3, (r.stdout, r.stderr)
    assert "amr.restart" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_driver_restart_continues_bit_for_bit(tmp_path, precision):
    """A run checkpointed at step 20 and restarted from that checkpoint into
    the same output folder leaves the same .dat files as one uninterrupted
    run: the state and clocks round-trip exactly, and the rows the killed
    segment wrote after the checkpoint are dropped, not duplicated."""
    p = os.path.join(PARAMS, "params_plain_64.txt")
    common = [p, "amr.n_cell=32 32 32", "geometry.prob_extent=16 16 16",
              "axion_strings.k_max_over_mr=8", "evolution.max_steps=60",
              "axion_strings.output_cadence_steps=10",
              "axion_strings.compute_spectrum=1",
              "axion_strings.precision=" + precision]
    full, seg = tmp_path / "full", tmp_path / "seg"
    full.mkdir()
    seg.mkdir()
    r = run([*common, "--out", str(full)])
    assert r.returncode == 0, r.stderr
    # the interrupted run: checkpoints every 20 steps, stopped at 40
    r = run([*common, "amr.check_int=20", "evolution.max_steps=40",
             "--out", str(seg)])
    assert r.returncode == 0, r.stderr
    assert (seg / "chk00020.axs").exists() and (seg / "chk00040.axs").exists()
    r = run([*common, "amr.restart=" + str(seg / "chk00020.axs"),
             "--out", str(seg)])
    assert r.returncode == 0, r.stderr
    for name in ("network_scalars.dat", "axion_spectrum.dat"):
        a = (full / name).read_text()
        b = (seg / name).read_text()
        assert a == b, name
    # a fresh run into a folder that holds files moves them aside
    r = run([*common, "evolution.max_steps=10", "--out", str(seg)])
    assert r.returncode == 0, r.stderr
    assert any(f.name.startswith("network_scalars.dat.old.")
               for f in seg.iterdir())
    assert len(np.atleast_2d(np.loadtxt(seg / "network_scalars.dat"))) == 1


@pytest.mark.gpu
def test_driver_flat_space_loop_scalars(tmp_path, ctx):
    """Flat space (axion_strings.flat_space = 1): R = 1, lambda = m_r^2, a
    step-count cadence, and loop_scalars.dat rows (AxionStringsLevel.cpp:
    932-1040) equal to the same quantities through the C ABI."""
    from axion_strings_adaptive_amd import capi
    p = os.path.join(PARAMS, "params_plain_64.txt")
    r = run([p, "amr.n_cell=32 32 32", "geometry.prob_extent=16 16 16",
             "axion_strings.flat_space=1", "axion_strings.flat.m_r=1.0",
             "axion_strings.flat.output_cadence_steps=10",
             "axion_strings.k_max_over_mr=8", "evolution.max_steps=20",
             "--out", str(tmp_path)])
    assert r.returncode == 0, r.stderr
    assert not (tmp_path / "network_scalars.dat").exists()
    rows = np.loadtxt(tmp_path / "loop_scalars.dat")
    assert rows.shape == (2, 8)
    lvl = capi.Level(ctx, 32, 0.5, capi.AXS_FP64)
    lvl.fourier_ic(0, 4.0, 0.5, 42, 0)
    lvl.fourier_ic(1, 4.0, 0.5, 42, 1)
    lvl.zero_components(2, 2)
    bg = capi.make_background(flat_space=True, flat_m_r=1.0)
    lvl.advance(bg, 0.0, 0.1, 20)
    assert rows[1, 0] == pytest.approx(2.0, rel=1e-12)
    plain, _ = lvl.count_plaquettes()
    assert rows[1, 1] == plain
    assert rows[1, 2] == pytest.approx((2.0 / 3.0) * plain * 0.5, rel=1e-12)
    e = lvl.energy_sums(1.0, 1.0, 0.0, capi.AXS_MASK_B, 0.8)
    assert rows[1, 3] == pytest.approx(e[0] / e[13], rel=1e-9)
    assert rows[1, 4] == pytest.approx(e[0] * 0.5 ** 3, rel=1e-9)
    total, count, _ = lvl.velocity(1.0, 0.0, 1.0)
    assert rows[1, 7] == count
    g2v2 = total / count if count else 0.0
    assert rows[1, 5] == pytest.approx(g2v2, rel=1e-9)
    assert rows[1, 6] == pytest.approx(np.sqrt(1.0 + g2v2), rel=1e-9)


@pytest.mark.gpu
def test_driver_relaxed_restart_continues_bit_for_bit(tmp_path):
    """Checkpoint / restart across the pre-evolution relaxation: the relaxing
    clock, the adaptive xi-check cadence and (after the handoff) the blend
    window round-trip through the checkpoint, so the .dat files of a run
    restarted from an early checkpoint equal an uninterrupted run's."""
    p = os.path.join(PARAMS, "params_relaxed_64.txt")
    common = [p, "amr.n_cell=32 32 32", "geometry.prob_extent=16 16 16",
              "evolution.max_steps=120",
              "axion_strings.output_cadence_steps=20"]
    full, seg = tmp_path / "full", tmp_path / "seg"
    full.mkdir()
    seg.mkdir()
    r = run([*common, "--out", str(full)])
    assert r.returncode == 0, r.stderr
    r = run([*common, "amr.check_int=10", "evolution.max_steps=40",
             "--out", str(seg)])
    assert r.returncode == 0, r.stderr
    r = run([*common, "amr.restart=" + str(seg / "chk00010.axs"),
             "--out", str(seg)])
    assert r.returncode == 0, r.stderr
    names = sorted(f.name for f in full.iterdir() if f.suffix == ".dat")
    assert "pre_evolution_scalars.dat" in names
    for name in names:
        assert (full / name).read_text() == (seg / name).read_text(), name
