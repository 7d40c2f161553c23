"""CPU checks of the C ABI library: it loads and exports every entry point
declared in include/gm.h (no compute calls without a GPU)."""
import ctypes as C
import os
import re

import pytest

import gm

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "gm.h")).read()
    return sorted(set(re.findall(r"\b(gm_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_the_solver_api():
    names = declared_symbols()
    for required in ("gm_create", "gm_destroy", "gm_initialize", "gm_step", "gm_get_state",
                     "gm_set_initial_state_full", "gm_apply_H", "gm_apply_sink", "gm_schrodinger_rhs"):
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(gm.LIB_PATH):
        pytest.fail("libgm.so is not built (run boson-cloud-binary_amd/build.py)")
    lib = C.CDLL(gm.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_device_reports_an_error():
    """With no GPU the constructor must fail loudly, never fall back to the CPU."""
    lib = gm.lib()
    if lib.gm_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(gm.GMError) as info:
        gm.CrankNicolsonModes(64, 10.0, 0.1, 0.0, 2, 0.2, 0.01, 0.1, 0.0, 30.0)
    assert info.value.code == gm.capi.GM_E_DEVICE


def test_invalid_arguments_are_rejected_before_the_device():
    with pytest.raises(gm.InvalidArgument):
        gm.CrankNicolsonModes(2, 10.0, 0.1, 0.0, 2, 0.2, 0.01, 0.1, 0.0, 30.0)
    with pytest.raises(gm.InvalidArgument):
        gm.CrankNicolsonModes(64, 10.0, -0.1, 0.0, 2, 0.2, 0.01, 0.1, 0.0, 30.0)
    # the kernels' fixed tables hold l_max <= 39
    with pytest.raises(gm.InvalidArgument, match="largest supported"):
        gm.CrankNicolsonModes(64, 10.0, 0.1, 0.0, 40, 0.2, 0.01, 0.1, 0.0, 30.0)
