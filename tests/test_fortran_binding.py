"""The Fortran ISO_C_BINDING module compiles with flang and its derived types match the C structs."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "sem2dpack_amd", "fortran", "sem2d_hip.f90")

PROBE = """
program probe
  use iso_c_binding
  use sem2d_hip
  type(sem2d_desc), target :: d
  type(sem2d_fault_desc), target :: f
  print '(I0,1X,I0)', c_sizeof(d), c_sizeof(f)
end program
"""


@pytest.mark.skipif(shutil.which("flang") is None, reason="flang not installed")
def test_module_compiles_and_layout_matches(tmp_path):
    import ctypes
    from sem2dpack_amd import capi
    (tmp_path / "probe.f90").write_text(PROBE)
    subprocess.run(["flang", "-c", SRC], cwd=tmp_path, check=True, capture_output=True)
    from sem2dpack_amd import _lib
    if not os.path.exists(_lib.HIP_LIB):
        pytest.skip("libsem2d_hip.so not built")
    subprocess.run(["flang", "-o", "probe", "probe.f90", "sem2d_hip.o", _lib.HIP_LIB,
                    "-Wl,-rpath," + os.path.dirname(_lib.HIP_LIB), "-L/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath,/opt/rocm/lib"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run([str(tmp_path / "probe")], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == ctypes.sizeof(capi.Desc)
    assert int(out[1]) == ctypes.sizeof(capi.FaultDesc)


def test_module_declares_every_entry_point():
    """Every function of include/sem2d.h has a bind(C) interface in the Fortran module."""
    import re
    hdr = open(os.path.join(ROOT, "include", "sem2d.h")).read()
    names = set(re.findall(r"\b(sem2d_[a-z_0-9]+)\s*\(", hdr))
    mod = open(SRC).read()
    bound = set(re.findall(r"bind\(C, name='(sem2d_[a-z_0-9]+)'\)", mod))
    assert names == bound, names ^ bound
