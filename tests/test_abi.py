"""The C-ABI library loads and exports every symbol include/atjup_hip.h declares (no GPU needed)."""
import ctypes
import os

import pytest

from atjup_amd import abi, lib


def test_header_and_library_agree():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    so = ctypes.CDLL(lib.LIB_PATH)
    names = lib.exported_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_struct_layout_matches_header():
    # atjup_grid: 3 ints, then 5 doubles, 8-byte aligned
    assert ctypes.sizeof(abi.Grid) == 16 + 5 * 8
    assert ctypes.sizeof(abi.Consts) == 23 * 8
    assert abi.FIELD["p_dyn"] == 20 and abi.FIELD["fluxlim_nh4sh"] == 46
    assert abi.FIELD["tn"] == 64 and abi.FIELD["disn"] == 80


def test_default_knobs_symbol():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    L = lib.load()
    k = abi.Knobs()
    L.atjup_hip_default_knobs(ctypes.byref(k))
    assert k.press_omega == 1.0 and k.sinthe_min == 0.55 and k.rk_scalar_sync == 2


def test_default_radiation_params_match_mirror():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    L = lib.load()
    p = abi.RadParams()
    L.atjup_hip_default_rad_params(ctypes.byref(p))
    q = abi.rad_params()
    for name, _ in abi.RadParams._fields_:
        assert getattr(p, name) == getattr(q, name), name
    assert ctypes.sizeof(abi.RadParams) == 9 * 8 + 2 * 4
    assert abi.FIELD["epsilon"] == 47 and abi.FIELD["radiation"] == 48
    assert abi.FIELD["P_rain"] == 49 and abi.FIELD["P_nh4sh"] == 58
    assert abi.FIELD["thermalmassflux"] == 61 and abi.FIELD["difflux_nh4sh"] == 62 and abi.F_END == 63 < abi.NLEVEL


def test_default_chemistry_params_match_mirror():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    L = lib.load()
    p = abi.ChemParams()
    L.atjup_hip_default_chem_params(ctypes.byref(p))
    q = abi.chem_params()
    for name, _ in abi.ChemParams._fields_:
        assert getattr(p, name) == getattr(q, name), name


def test_default_mca_params_match_mirror():
    L = lib.load()
    p = abi.MCAParams()
    L.atjup_hip_default_mca_params(ctypes.byref(p))
    q = abi.mca_params()
    for name, _ in abi.MCAParams._fields_:
        assert getattr(p, name) == getattr(q, name), name


def test_default_precipitation_params_match_mirror():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    L = lib.load()
    p = abi.PrecipParams()
    L.atjup_hip_default_precip_params(ctypes.byref(p))
    q = abi.precip_params()
    for name, _ in abi.PrecipParams._fields_:
        assert getattr(p, name) == getattr(q, name), name
