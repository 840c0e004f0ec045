"""CPU checks of the product library: it loads, exports every entry point the
headers declare, and its host-side arithmetic matches the oracle. No kernel
is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, bits, golden

LIB = os.path.join(ROOT, "form_factors_amd", "libff_amd.so")


def declared_symbols():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-zA-Z_][\w\s\*]*?\b((?:ffk|ffp|ffio)_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libff_amd.so is not built (run __graft_entry__.build())")
    return C.CDLL(LIB)


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "ffk_cast" in names and "ffk_ff_fused" in names and len(names) > 20


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    from form_factors_amd import ffk

    ffk_names = [n for n in declared_symbols() if n.startswith("ffk_")]
    assert sorted(ffk.SIGNATURES) == ffk_names
    from form_factors_amd import formats

    assert sorted(formats.SIGNATURES) == [n for n in declared_symbols() if n.startswith("ffio_")]


def test_host_area_weights_and_counts_match_oracle(oracle):
    from form_factors_amd import ffk

    g = golden("emitted.npz")
    c = golden("cast_sphere3.npz")
    w, total = ffk.area_weights(c["faces"])
    assert np.array_equal(bits(w), bits(g["sphere3_weights"])) and total == g["sphere3_total"]
    for n in (2_000_000, 40_000, 999):
        fr, rr = ffk.emitted_counts(w, n, total)
        assert np.array_equal(fr, g[f"sphere3_{n}_front"]) and np.array_equal(rr, g[f"sphere3_{n}_rear"])


def test_errors_are_codes_not_aborts(lib):
    from form_factors_amd import ffk

    L = ffk.lib()
    assert L.ffk_scene_create(None) == -1
    assert b"NULL" in L.ffk_last_error()
    # host-only argument checks that never reach the device
    assert L.ffk_emitted_counts(None, 1, 1, 1.0, None, None) == -1


def test_device_step_entry_points_check_their_arguments(lib):
    """The device-resident step's entry points refuse missing arrays and bad
    sizes with FFK_ERR_ARG before any device call (host checks only)."""
    from form_factors_amd import ffk

    L = ffk.lib()
    p = C.c_void_p(0x1000)  # never dereferenced: each call fails on another argument first
    f3 = (C.c_float * 3)(0, 0, 1)
    calls = [
        ("ffk_radiance_step_dev", L.ffk_radiance_step_dev(None, f3, f3, 10, 0, 1, p, p, None)),
        ("ffk_radiance_status", L.ffk_radiance_status(None)),
        ("ffk_sun_count_dev", L.ffk_sun_count_dev(None, f3, f3, 1.0, 1.0, 0, 10, 1, p, None)),
        ("ffk_sun_deposit_dev", L.ffk_sun_deposit_dev(4, None, p, 1.0, p, None)),
        ("ffk_conduction_dev", L.ffk_conduction_dev(0, p, p, p, p, p, p, None)),
        ("ffk_sbff_dev", L.ffk_sbff_dev(4, p, 0, p, None, p, p, p, p, p, p, p, p, p, None)),
        ("ffk_heat_sources_dev", L.ffk_heat_sources_dev(-1, p, p, p, p, None)),
        ("ffk_step_begin_dev", L.ffk_step_begin_dev(0, p, p, p, p, None)),
        ("ffk_adams_step_f", L.ffk_adams_step_f(4, 0, 0.1, p, None, p, p, p, p, None, None)),
    ]
    for name, r in calls:
        assert r == -1, name  # FFK_ERR_ARG
        assert L.ffk_last_error(), name


def test_solution_device_loop_setter_checks_its_system(lib):
    from form_factors_amd import plugin

    L = plugin.lib()
    assert L.ffp_thermal_solution_set_device_loop(None, 1) < 0
