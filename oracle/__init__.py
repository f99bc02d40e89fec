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


def load(variant="", ulp_perturbed=False):
    """The oracle.  ``variant`` names a build of the perturbation ensemble
    (``build_oracle.ENSEMBLE``); ``ulp_perturbed=True`` is the ``ulp_p`` build."""
    from sem2dpack_amd import capi
    from . import build_oracle
    if ulp_perturbed:
        variant = "ulp_p"
    lib = build_oracle.lib_path(variant)
    if not os.path.exists(lib):
        build_oracle.build()
    return capi.bind(ctypes.CDLL(lib), prefix="oracle_")


def ensemble():
    """[(name, api)] for every perturbed build of the oracle."""
    from . import build_oracle
    return [(v, load(v)) for v in build_oracle.ENSEMBLE]
