"""Build tests/native/fortran_driver (Fortran host program over the sem2d_hip module).

flang compiles sem2dpack_amd/fortran/sem2d_hip.f90 and tests/native/fortran_driver.f90
in a temporary directory and links them with sem2dpack_amd/lib/libsem2d_hip.so; the
executable finds the library through an $ORIGIN-relative run path, so the tree can be
moved.  Skipped (returns None) when flang or the HIP library is missing.
"""
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(HERE, "fortran_driver")


def build():
    flang = shutil.which("flang") or shutil.which("amdflang")
    lib = os.path.join(ROOT, "sem2dpack_amd", "lib", "libsem2d_hip.so")
    if flang is None or not os.path.exists(lib):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        mod = os.path.join(ROOT, "sem2dpack_amd", "fortran", "sem2d_hip.f90")
        drv = os.path.join(HERE, "fortran_driver.f90")
        subprocess.run([flang, "-O2", "-c", mod], cwd=tmp, check=True, capture_output=True)
        subprocess.run([flang, "-O2", "-c", drv], cwd=tmp, check=True, capture_output=True)
        out = os.path.join(tmp, "fortran_driver")
        subprocess.run([flang, "-o", out, "fortran_driver.o", "sem2d_hip.o", lib,
                        "-Wl,-rpath,$ORIGIN/../../sem2dpack_amd/lib", "-L/opt/rocm/lib", "-lamdhip64",
                        "-Wl,-rpath,/opt/rocm/lib"], cwd=tmp, check=True, capture_output=True)
        os.replace(out, EXE)
    return EXE


if __name__ == "__main__":
    print(build())
