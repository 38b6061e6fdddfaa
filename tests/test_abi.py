"""The C-ABI library loads and exports every entry point include/*.h declares (no GPU calls)."""
from __future__ import annotations

import ctypes
import re
import subprocess

import pytest

from conftest import ROOT
from owe_hip import _lib

HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions() -> list[str]:
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(owe_hip_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "owe_hip_trace_batch" in names and "owe_hip_last_error" in names
    assert "owe_hip_lens_batch_run" in names and "owe_hip_renderer_particles" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\b(owe_hip_[a-z0-9_]+)\b", out))
    assert set(declared_functions()) <= exported


def test_library_is_built_for_gfx950_only():
    # The offload bundle ids embedded in our own build name every device target it carries.
    blob = _lib.LIB_PATH.read_bytes()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", blob))
    assert targets == {b"gfx950"}


def test_identity_and_unavailable_reason_without_gpu():
    lib = _lib.load()
    assert lib.owe_hip_backend_name() == b"hip"
    assert b"gfx950" in lib.owe_hip_backend_summary()
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    ok, why = _lib.available()
    assert not ok and why


def test_invalid_arguments_report_status_and_message():
    lib = _lib.load()
    st = lib.owe_hip_trace_batch(None, 0, None, 0, None, None, None)
    assert st == _lib.ERR_INVALID
    assert b"mesh is null" in lib.owe_hip_last_error()
    with pytest.raises(_lib.OweHipError):
        _lib.check(st)
    st = lib.owe_hip_mesh_stats_get(None, None)
    assert st == _lib.ERR_INVALID
