"""Loading of the native libraries built in-tree by ``sem2dpack_amd.build``.

``libsem2d_host.so`` holds host-only init helpers (no HIP); ``libsem2d_hip.so``
holds the HIP kernels behind the C ABI declared in ``include/sem2d.h``.  There
is no fallback: a missing library raises immediately.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libsem2d_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libsem2d_hip.so")

_host = None
_hip = None


class NativeLibraryMissing(RuntimeError):
    pass


def load_host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise NativeLibraryMissing(f"{HOST_LIB} not built: run python -c 'import __graft_entry__ as g; g.build()'")
        lib = ctypes.CDLL(HOST_LIB)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.sem2d_genrcm.argtypes = [ctypes.c_int, ctypes.c_int, ip, ip, ip]
        lib.sem2d_genrcm.restype = ctypes.c_int
        _host = lib
    return _host


def load_hip_lib():
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise NativeLibraryMissing(f"{HIP_LIB} not built: run python -c 'import __graft_entry__ as g; g.build()'")
        from . import capi
        _hip = capi.bind(ctypes.CDLL(HIP_LIB))
    return _hip
