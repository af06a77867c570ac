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


def _create(G, Cn, K, precision=abi.PRECISION_FP64):
    L = lib.load()
    h = L.atjup_hip_ctx_create(ctypes.byref(G), ctypes.byref(Cn), ctypes.byref(K), precision)
    msg = L.atjup_hip_last_error_string().decode()
    if h:
        L.atjup_hip_ctx_destroy(h)
    return h, msg


@pytest.mark.parametrize("change,expect", [
    (dict(turb_model=3), "turb_model"),
    (dict(rk_scalar_sync=0), "RK_SCALAR_SYNC"),
    (dict(press_omega=2.0), "press_omega"),
    (dict(radiation=2), "radiation"),
    (dict(mca=-1), "mca"),
])
def test_ctx_create_refuses_profiles_it_does_not_implement(change, expect):
    """A knob profile outside the implemented set is refused before any device work, with the knob
    named in atjup_hip_last_error_string() (ATJUP_ERR_PROFILE: atjup_hip_ctx_create returns NULL)."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    G, Cn, K, _ = abi.make_profile("full", sweeps=5)
    for k, v in change.items():
        setattr(K, k, v)
    h, msg = _create(G, Cn, K)
    assert not h and expect in msg, msg


def test_ctx_create_refuses_the_dry_profile_with_chemistry_or_a_closure():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=5)
    K.chemistry = 1
    h, msg = _create(G, Cn, K)
    assert not h and "moist profile" in msg, msg
    G, Cn, K, _ = abi.make_profile("dyn", sweeps=5)
    K.turb_model = 2
    h, msg = _create(G, Cn, K)
    assert not h and "closure off" in msg, msg
    G, Cn, K, _ = abi.make_profile("full", sweeps=5)
    h, msg = _create(G, Cn, K, abi.PRECISION_TANGENT)
    assert not h and "tangent" in msg, msg


@pytest.mark.parametrize("model", ["k_omega_SST", "k_epsilon", "k_omega"])
def test_every_closure_passes_validation(model):
    """The three closures of the reference are accepted: validation passes, and without a GPU the
    creation stops at the device check (with one, it succeeds)."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    G, Cn, K, _ = abi.make_profile("full", sweeps=5, turb=model)
    h, msg = _create(G, Cn, K)
    assert h or "no HIP device" in msg, msg


def test_default_precipitation_params_match_mirror():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libatjup_hip.so not built")
    L = lib.load()
    p = abi.PrecipParams()
    L.atjup_hip_default_precip_params(ctypes.byref(p))
    q = abi.precip_params()
    for name, _ in abi.PrecipParams._fields_:
        assert getattr(p, name) == getattr(q, name), name


REF_DIR = os.environ.get("ATJUP_REFERENCE", "/root/reference")


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF_DIR, "planet")),
                    reason="the reference sources are not present (the check compiles against them)")
def test_reference_side_adapter_compiles_and_links():
    """include/atjup_hip_step.hpp, HipStep<cJupiterModel> with every member function instantiated,
    against the reference's own headers (plus the friend line the adapter asks for) and objects,
    linked with libatjup_hip.so (oracle/Makefile, target `adapter`)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["make", "-C", os.path.join(root, "oracle"), "adapter", f"REF={REF_DIR}", "-j8"],
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    out = subprocess.run([os.path.join(root, "oracle", "_ref", "hip_step_check")], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 0 and "linked with libatjup_hip.so" in out.stdout
