"""The C-ABI library loads and exports every symbol include/gprx_hip.h declares."""
import re

from conftest import ROOT


def declared_symbols():
    text = (ROOT / "include" / "gprx_hip.h").read_text()
    return sorted(set(re.findall(r"\b(gprx_hip_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ["gprx_hip_exact_create", "gprx_hip_exact_value_and_grad", "gprx_hip_exact_predict",
                 "gprx_hip_sgpr_create", "gprx_hip_sgpr_value_and_grad", "gprx_hip_svgp_value_and_grad",
                 "gprx_hip_batch_create", "gprx_hip_batch_value_and_grad", "gprx_hip_destroy"]:
        assert must in syms


def test_library_exports_every_declared_symbol():
    import ctypes

    from gprx_amd import _lib

    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_version_and_device_count_without_gpu():
    from gprx_amd import _lib

    assert _lib.lib().gprx_hip_version().startswith(b"gprx-hip")
    assert _lib.lib().gprx_hip_device_count() >= 0
