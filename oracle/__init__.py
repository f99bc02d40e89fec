"""Test infrastructure: the oracle for the HIP time step.

``load()`` returns the C restatement (``sem2d_oracle.c``) bound through the same
ctypes description as the product ABI, under the prefix ``oracle_``.  Only the
tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` use it.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libsem2d_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")


LIB_ULP = os.path.join(HERE, "lib", "libsem2d_oracle_ulp.so")


def load(ulp_perturbed=False):
    """The oracle; ``ulp_perturbed=True``: the build whose libm results are moved by one
    ulp (to measure the reference arithmetic's conditioning)."""
    from sem2dpack_amd import capi
    lib = LIB_ULP if ulp_perturbed else LIB
    if not os.path.exists(lib):
        from . import build_oracle
        build_oracle.build()
    return capi.bind(ctypes.CDLL(lib), prefix="oracle_")
