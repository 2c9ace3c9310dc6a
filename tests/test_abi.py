"""The C-ABI library loads and exports every entry point include/cmcx.h declares (no GPU calls)."""
import ctypes
import os
import re

from cmcx import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "cmcx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cmcx2?d?_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(_declared()) == bound


def test_host_only_calls():
    L = _lib.lib()
    assert L.cmcx_version().startswith(b"cmcx")
    p = _lib.cmcx_params()
    L.cmcx_params_normalized(400.0, 0.72, 1.4, ctypes.byref(p))
    assert p.v_T == 1.0 and abs(p.v_vis - 1 / 400) < 1e-18
    assert abs(p.t_vis - 1 / (400 * 0.72)) < 1e-18 and abs(p.t_phi - 0.4 / (1.4 * 400)) < 1e-18
    L.cmcx_params_dimensional(1.8e-5, 1.2, 287.0, 0.025, 718.0, ctypes.byref(p))
    assert p.v_T == 287.0 and abs(p.v_vis - 1.8e-5 / 1.2) < 1e-20
    # argument errors are reported without touching the device
    assert L.cmcx_create(None, None, 0, None, None) == _lib.CMCX_ERR_ARG
    assert L.cmcx_step(None, 1.0, 1, 1, 0, None) == _lib.CMCX_ERR_ARG
    assert L.cmcx_destroy(None) == _lib.CMCX_ERR_ARG
