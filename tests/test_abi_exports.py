"""The C-ABI boundary without a GPU: libkm_hydro.so loads and exports every function that
include/km_hydro.h declares, and the host-only entry points answer (no device work)."""
import ctypes as C
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "km_hydro.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)  # comments
    text = re.sub(r"//[^\n]*", " ", text)
    names = re.findall(r"\b(km_\w+)\s*\([^;{]*\)\s*;", text)
    # function declarations only (typedef'd structs and macros do not match "name(...);")
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = _declared()
    assert len(names) > 40, names
    for must in ("km_calculate_fluxes", "km_calculate_emf", "km_fluxes_to_dudt", "km_apply_dudt",
                 "km_prepare_primitive", "km_estimate_dt", "km_fill_divb", "km_ghost_fill",
                 "km_flux_correct", "km_last_error", "km_sim_set_fused_kernel",
                 "km_sim_fused_kernel"):
        assert must in names, must


def test_library_exports_every_declared_function():
    import kamayan_amd as km
    lib = C.CDLL(str(km.LIB_PATH))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_host_entry_points_without_a_gpu():
    import kamayan_amd as km
    lib = km.lib()
    cfg = km.Config(**{"hydro/reconstruction": "wenoz"})
    assert cfg.c.recon == 3
    # an unknown option value is refused with the reference's dispatcher-style message
    rc = lib.km_cfg_set(C.byref(cfg.c), b"hydro/reconstruction", b"weno7")
    assert rc != km.KM_OK
    lib.km_last_error.restype = C.c_char_p
    msg = lib.km_last_error().decode()
    assert "weno7" in msg or "reconstruction" in msg, msg
    with pytest.raises(km.KmError):
        cfg.set("hydro/riemann", "roe")
