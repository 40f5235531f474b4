"""CPU-side checks of the C-ABI library: it loads without a GPU, exports
every function include/axs.h declares, and its host-only entry points (the
FP64 background that feeds the kernels) agree bit for bit with the oracle and
with the reference's Background.hpp compiled from source."""

import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "axs.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axs_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from axion_strings_adaptive_amd import capi
    return capi.load()


def test_library_exports_every_declared_symbol(lib):
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_binding_declares_every_header_function():
    from axion_strings_adaptive_amd import capi
    assert sorted(capi.declared_symbols()) == header_functions()


def test_background_matches_oracle_bitwise(lib, oracle):
    from axion_strings_adaptive_amd import capi
    L = oracle.load()
    fns = [L.or_bg_R, L.or_bg_R_prime, L.or_bg_curvature_term_coeff,
           L.or_bg_lambda, L.or_bg_H, L.or_bg_H_over_mr_direct,
           L.or_bg_H_over_mr_closed_form, L.or_bg_tau_from_log_mr_over_h]
    for (a_inv, c0, c1, ts, sw) in [(2.0, 0.0, 0.0, 0.0, 0),
                                    (2.0, 0.0, 2.0, 1.0, 1),
                                    (2.0, 1.0, 2.0, 5.0, 1), (3.0, 0.5, 0, 0, 0)]:
        b = capi.make_background(a_inv, c0, c1, ts, bool(sw))
        ob = oracle.bg(a_inv, c0, c1, ts, bool(sw))
        for tau in (0.7, 1.0, 1.05, 4.99, 5.0, 33.3, 151.0):
            for w, f in enumerate(fns):
                assert capi.background_eval(b, w, tau) == f(ob, tau)
            c = capi.rhs_coeffs(b, tau)
            R = L.or_bg_R(ob, tau)
            assert c.curvature_coeff == L.or_bg_curvature_term_coeff(ob, tau)
            assert c.lambda_ == L.or_bg_lambda(ob, tau)
            assert c.R_squared == R * R


def test_background_matches_reference_header(lib, ref):
    from axion_strings_adaptive_amd import capi
    for (a_inv, c0, c1, ts, sw) in [(2.0, 0.0, 2.0, 1.0, 1), (3.0, 1.0, 3.0, 2.0, 1)]:
        b = capi.make_background(a_inv, c0, c1, ts, bool(sw))
        for tau in (1.0, 1.5, 2.0, 10.0):
            for w in range(8):
                assert capi.background_eval(b, w, tau) == \
                    ref.ref_background(a_inv, c0, c1, ts, sw, w, tau)


def test_flat_background_and_errors(lib):
    from axion_strings_adaptive_amd import capi
    f = capi.make_background(flat_space=True, flat_m_r=1.5)
    c = capi.rhs_coeffs(f, 3.0)
    assert (c.curvature_coeff, c.lambda_, c.R_squared) == (0.0, 2.25, 1.0)
    bad = capi.make_background(a_inv=1.0)
    with pytest.raises(capi.AxsError, match="a_inv"):
        capi.background_eval(bad, 0, 1.0)
    with pytest.raises(capi.AxsError):
        capi.background_eval(capi.make_background(), 99, 1.0)


def test_null_arguments_are_rejected_without_a_gpu(lib):
    from axion_strings_adaptive_amd import capi
    assert lib.axs_level_create(None, 8, 8, 8, 0.5, 0, None) == capi.AXS_OK + 1
    assert b"NULL" in lib.axs_last_error()
    assert lib.axs_rk4_step(None, None, 0.1) == 1
    assert lib.axs_energy_sums(None, 1.0, 1.0, 0.0, None, None, None) == 1


def test_tuning_table_parser(lib, tmp_path):
    """axs_tuning_load accepts the autotuner's format (comments, blank
    lines, both entry kinds) and names the file and line of a malformed
    entry: wrong field count, out-of-range variant, fusion, precision or
    box shape, or an unknown kind."""
    from axion_strings_adaptive_amd import capi
    good = tmp_path / "good.txt"
    good.write_text("# launch shapes\n\n"
                    "level 1024 1024 1024 1 3 -1 0 4 128  # one launch\n"
                    "level 256 256 256 0 1 2 32 5 64\n"
                    "box 1 8 1024\n")
    capi.load_tuning(good, lib)
    nt, npv = lib.axs_num_tile_variants(), lib.axs_num_pair_variants()
    bad_lines = [
        "level 1024 1024 1024 1 3 -1 0 4",          # too few fields
        "level 1024 1024 1024 1 3 -1 0 4 128 7",    # too many
        f"level 64 64 64 1 1 {nt} 0 -1 0",           # tile variant
        f"level 64 64 64 1 1 -1 0 {npv} 0",          # pair variant
        "level 64 64 64 1 4 -1 0 -1 0",              # fusion mode
        "level 64 64 64 3 1 -1 0 -1 0",              # precision
        "level 2 64 64 1 1 -1 0 -1 0",               # grid edge
        "level 64 64 64 1 1 -1 -5 -1 0",             # z chunk
        "box 1 0 1024", "box 1 8 0", "box 1 257 64", "box 1 8",
        "tile 1 2 3",
    ]
    for i, line in enumerate(bad_lines):
        f = tmp_path / f"bad{i}.txt"
        f.write_text("# ok\nbox 0 16 512\n" + line + "\n")
        with pytest.raises(capi.AxsError, match=rf"bad{i}\.txt:3"):
            capi.load_tuning(f, lib)
    with pytest.raises(capi.AxsError, match="cannot read"):
        capi.load_tuning(tmp_path / "missing.txt", lib)
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    capi.load_tuning(empty, lib)  # an empty table is valid
    if os.path.exists(capi.TUNING_PATH):  # the table load() applied
        capi.load_tuning(capi.TUNING_PATH, lib)
