"""ctypes description of the C ABI in ``include/sem2d.h``.

``bind(lib, prefix)`` attaches argument and result types to the ``<prefix>*``
entry points of a loaded library and returns a small namespace object whose
attributes drop the prefix (``api.create``, ``api.step``, ...).
"""
import ctypes as C

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class Desc(C.Structure):
    _fields_ = [
        ("ndof", C.c_int), ("ngll", C.c_int), ("nelem", C.c_int), ("npoin", C.c_int),
        ("nelast", C.c_int), ("coef_mode", C.c_int),
        ("ibool", ip), ("coefs", dp), ("lam_mu", dp), ("geom", dp), ("wgll", dp),
        ("hprime", dp), ("rmass", dp),
        ("dt", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
        ("bulk_fp32", C.c_int),
        ("nx", C.c_int), ("nz", C.c_int), ("ezflt", C.c_int), ("elem_of_cell", ip),
    ]


class FaultDesc(C.Structure):
    _fields_ = [
        ("npoin", C.c_int), ("node1", ip), ("node2", ip), ("n1", dp),
        ("B", dp), ("invM1", dp), ("invM2", dp), ("Z", dp), ("T0", dp), ("cohesion", dp),
        ("coefA2V", C.c_double), ("coefA2D", C.c_double),
        ("allow_opening", C.c_int), ("law", C.c_int), ("kind", C.c_int), ("healing", C.c_int),
        ("p", dp * 6), ("theta0", dp), ("V0", dp),
        ("oit", C.c_int), ("oitd", C.c_int), ("oix1", C.c_int), ("oixn", C.c_int), ("oixd", C.c_int),
    ]


_SIGS = {
    "create": ([C.POINTER(Desc), C.POINTER(C.c_void_p)], C.c_int),
    "set_kelvin_voigt": ([C.c_void_p, dp], C.c_int),
    "set_25d": ([C.c_void_p, dp], C.c_int),
    "set_abso": ([C.c_void_p, C.c_int, ip, dp], C.c_int),
    "set_abso_stacey": ([C.c_void_p, C.c_int, ip, dp], C.c_int),
    "set_abso_normals": ([C.c_void_p, C.c_int, dp], C.c_int),
    "set_dirichlet": ([C.c_void_p, C.c_int, ip, C.c_int], C.c_int),
    "set_force": ([C.c_void_p, C.c_int, dp, ip], C.c_int),
    "set_moment": ([C.c_void_p, C.c_int, ip, ip, dp, dp, ip], C.c_int),
    "set_stf": ([C.c_void_p, C.c_int, C.c_int, dp], C.c_int),
    "set_receivers": ([C.c_void_p, C.c_int, ip, C.c_int, C.c_int, C.c_int], C.c_int),
    "set_fault": ([C.c_void_p, C.POINTER(FaultDesc), C.c_int], C.c_int),
    "set_receivers_interp": ([C.c_void_p, C.c_int, ip, dp, C.c_int, C.c_int, C.c_int], C.c_int),
    "step": ([C.c_void_p, C.c_int], C.c_int),
    "fetch_fields": ([C.c_void_p, dp, dp, dp], C.c_int),
    "set_fields": ([C.c_void_p, dp, dp, dp], C.c_int),
    "fetch_seismo": ([C.c_void_p, dp, ip], C.c_int),
    "fetch_fault": ([C.c_void_p, dp, ip], C.c_int),
    "fetch_fault_state": ([C.c_void_p, dp, dp, dp, dp, dp], C.c_int),
    "fetch_potency": ([C.c_void_p, dp, ip], C.c_int),
    "compute_fint": ([C.c_void_p, dp, dp], C.c_int),
    "compute_fint_noalloc": ([C.c_void_p, dp, dp], C.c_int),
    "set_time_scheme": ([C.c_void_p, C.c_int], C.c_int),
    "set_fault_sides": ([C.c_void_p, C.c_int], C.c_int),
    "set_fault_twf": ([C.c_void_p, C.c_int, dp, dp], C.c_int),
    "set_fault_normal": ([C.c_void_p, C.c_int, C.c_double, C.c_double], C.c_int),
    "set_periodic": ([C.c_void_p, C.c_int, ip, ip], C.c_int),
    "set_hht": ([C.c_void_p, C.c_double], C.c_int),
    "set_abso_wave": ([C.c_void_p, C.c_int, dp, dp, dp, dp, dp], C.c_int),
    "set_abso_periodic": ([C.c_void_p], C.c_int),
    "set_lsf": ([C.c_void_p, C.c_int, ip, ip, dp, dp, dp, dp, C.c_double, C.c_double], C.c_int),
    "fetch_fault_sides": ([C.c_void_p, dp, ip], C.c_int),
    "set_itinfo": ([C.c_void_p, C.c_int], C.c_int),
    "fetch_info": ([C.c_void_p, dp, ip], C.c_int),
    "field_info": ([C.c_void_p, dp, dp, ip], C.c_int),
    "set_tile_config": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "get_step": ([C.c_void_p, ip], C.c_int),
    "last_step_ms": ([C.c_void_p, dp], C.c_int),
    "last_error": ([C.c_void_p], C.c_char_p),
    "destroy": ([C.c_void_p], None),
    # adjoint (HIP library only)
    "adjoint_init": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "adjoint_forward": ([C.c_void_p], C.c_int),
    "adjoint_backward": ([C.c_void_p, dp], C.c_int),
    "adjoint_fetch": ([C.c_void_p, dp, dp, dp, dp], C.c_int),
    "adjoint_timing": ([C.c_void_p, dp, dp], C.c_int),
    "adjoint_plan": ([C.c_void_p, ip, ip], C.c_int),
    "adjoint_release": ([], C.c_int),
    # batched ensemble (HIP library only)
    "batch_create": ([C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)], C.c_int),
    "batch_step": ([C.c_void_p, C.c_int], C.c_int),
    "batch_tune": ([C.c_void_p, C.c_int, ip], C.c_int),
    "batch_set_config": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "batch_config": ([C.c_void_p, ip, ip, ip], C.c_int),
    "batch_fetch_fields": ([C.c_void_p, C.c_int, dp, dp, dp], C.c_int),
    "batch_fetch_seismo": ([C.c_void_p, C.c_int, dp, ip], C.c_int),
    "batch_fetch_fault": ([C.c_void_p, C.c_int, dp, ip], C.c_int),
    "batch_fetch_fault_state": ([C.c_void_p, C.c_int, dp, dp, dp, dp, dp], C.c_int),
    "batch_fetch_potency": ([C.c_void_p, C.c_int, dp, ip], C.c_int),
    "batch_get_step": ([C.c_void_p, ip], C.c_int),
    "batch_last_step_ms": ([C.c_void_p, dp], C.c_int),
    "batch_last_error": ([C.c_void_p], C.c_char_p),
    "batch_destroy": ([C.c_void_p], None),
}

# entry points every implementation of the ABI must export (the CPU oracle has no
# event timing, no kernel launch configuration, no device-scratch variant, no adjoint
# and no batched ensemble)
GPU_ONLY = ("last_step_ms", "set_tile_config", "compute_fint_noalloc")
GPU_ONLY_PREFIXES = ("batch_", "adjoint_")
REQUIRED = [k for k in _SIGS if k not in GPU_ONLY and not k.startswith(GPU_ONLY_PREFIXES)]


class Api:
    def __init__(self, lib, prefix):
        self.lib = lib
        self.prefix = prefix


def bind(lib, prefix="sem2d_"):
    api = Api(lib, prefix)
    for name, (args, res) in _SIGS.items():
        sym = prefix + name
        if not hasattr(lib, sym):
            if name in REQUIRED:
                raise AttributeError(f"{sym} missing from {lib._name}")
            continue
        fn = getattr(lib, sym)
        fn.argtypes = args
        fn.restype = res
        setattr(api, name, fn)
    return api
