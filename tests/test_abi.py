"""The C-ABI libraries load and export every entry point declared in include/sem2d.h."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from sem2dpack_amd import _lib, capi

HEADER = os.path.join(ROOT, "include", "sem2d.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sem2d_[a-z_0-9]+)\s*\(", txt)))


def test_header_matches_ctypes_description():
    names = [n[len("sem2d_"):] for n in declared()]
    assert sorted(names) == sorted(capi._SIGS.keys())


def test_hip_library_exports_all():
    if not os.path.exists(_lib.HIP_LIB):
        pytest.skip("libsem2d_hip.so not built")
    lib = ctypes.CDLL(_lib.HIP_LIB)
    for n in declared():
        assert hasattr(lib, n), n


def test_oracle_exports_all(oracle_api):
    """The oracle restates the single-scenario path; event timing, the kernel launch
    configuration, the adjoint and the batched ensemble exist only in the HIP library."""
    for n in declared():
        if n[len("sem2d_"):] in capi.GPU_ONLY or n[len("sem2d_"):].startswith(capi.GPU_ONLY_PREFIXES):
            continue
        assert hasattr(oracle_api.lib, "oracle_" + n[len("sem2d_"):]), n


def test_host_library():
    lib = _lib.load_host_lib()
    assert hasattr(lib, "sem2d_genrcm")


def test_missing_hip_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "HIP_LIB", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_hip", None)
    with pytest.raises(_lib.NativeLibraryMissing):
        _lib.load_hip_lib()
