"""MI355X-native explicit spectral-element time step for SEM2DPACK inputs.

Host side: ``parfile`` (Par.inp reader), ``builder`` (structured-mesh init),
``solver`` (ctypes driver of the HIP library declared in ``include/sem2d.h``).
"""
from . import parfile  # noqa: F401
