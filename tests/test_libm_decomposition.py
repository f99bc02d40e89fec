"""The friction solve's asinh/sinh/cosh (sem2dpack_amd/csrc/sem2d_libm.hpp) follow the
C library's decomposition: built for the host, with the C library's log/log1p/exp/expm1
as primitives, they must equal the C library's asinh/sinh/cosh bit for bit, over every
branch and around every branch point (tests/native/libm_check.cpp)."""
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "sem2dpack_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_decomposition_matches_c_library():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "libm_check")
        subprocess.run([HIPCC, "-O2", "-ffp-contract=off", "-I", CSRC, "-o", exe,
                        os.path.join(HERE, "native", "libm_check.cpp")], check=True, cwd=tmp)
        r = subprocess.run([exe], capture_output=True, text=True)
    n, bad = (int(x) for x in r.stdout.split())
    assert n > 3_000_000
    assert bad == 0 and r.returncode == 0
