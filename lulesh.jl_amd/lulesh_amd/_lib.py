"""ctypes binding of liblulesh_hip.so (the C ABI in include/lulesh_hip.h).

The library is the product: there is no CPU fallback.  If it has not been
built, importing this module raises immediately.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# LULESH_HIP_LIBRARY selects another build of the library (a variant built
# with different kernel options, for instance)
LIB_PATH = os.environ.get("LULESH_HIP_LIBRARY") or os.path.normpath(
    os.path.join(_HERE, "..", "lib", "liblulesh_hip.so"))

c_int, c_int32, c_int64, c_double = ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
c_char_p, c_void_p = ctypes.c_char_p, ctypes.c_void_p
P_double = ctypes.POINTER(c_double)
P_int64 = ctypes.POINTER(c_int64)
P_uint32 = ctypes.POINTER(ctypes.c_uint32)
P_int32 = ctypes.POINTER(ctypes.c_int32)
LULESH_ERR_INVALID_ARG, LULESH_ERR_SCRATCH = 100, 104
NOALLOC_MAX_NX = 68


class DeviceDomain(ctypes.Structure):
    """lulesh_device_domain: nx, quirks and device pointers to the Domain
    fields (lulesh_lagrange_leapfrog_noalloc)."""
    _fields_ = ([("nx", c_int32), ("quirks", c_int32)] +
                [(n, P_double) for n in ("e", "p", "q", "v", "volo", "delv", "vdov", "arealg",
                                         "ss", "elemMass")] +
                [("elemBC", P_uint32)] +
                [(n, P_double) for n in ("x", "y", "z", "xd", "yd", "zd", "nodalMass")])


# name -> (restype, argtypes); mirrors include/lulesh_hip.h
SIGNATURES = {
    "lulesh_lagrange_leapfrog_noalloc": (c_int, [ctypes.POINTER(DeviceDomain), P_double, c_void_p,
                                                 c_int32]),
    "lulesh_create": (c_int, [c_int32, c_int32, c_int32, c_int32, c_int32, c_int32,
                              ctypes.POINTER(c_void_p)]),
    "lulesh_region_download": (c_int, [c_void_p, c_int32, c_char_p, P_int64, c_int64]),
    "lulesh_set_cost_mode": (c_int, [c_void_p, c_int32]),
    "lulesh_destroy": (c_int, [c_void_p]),
    "lulesh_reset": (c_int, [c_void_p]),
    "lulesh_num_elem": (c_int64, [c_void_p]),
    "lulesh_num_node": (c_int64, [c_void_p]),
    "lulesh_last_error": (c_char_p, [c_void_p]),
    "lulesh_error_string": (c_char_p, [c_int]),
    "lulesh_download": (c_int, [c_void_p, c_char_p, P_double, c_int64]),
    "lulesh_upload": (c_int, [c_void_p, c_char_p, P_double, c_int64]),
    "lulesh_download_elem_bc": (c_int, [c_void_p, P_uint32, c_int64]),
    "lulesh_upload_elem_bc": (c_int, [c_void_p, P_uint32, c_int64]),
    "lulesh_get_scalar": (c_int, [c_void_p, c_char_p, P_double]),
    "lulesh_set_scalar": (c_int, [c_void_p, c_char_p, c_double]),
    "lulesh_set_keep_aux": (c_int, [c_void_p, c_int32]),
    "lulesh_time_increment": (c_int, [c_void_p]),
    "lulesh_lagrange_leapfrog": (c_int, [c_void_p]),
    "lulesh_lagrange_nodal": (c_int, [c_void_p]),
    "lulesh_lagrange_elements": (c_int, [c_void_p]),
    "lulesh_calc_time_constraints": (c_int, [c_void_p]),
    "lulesh_run": (c_int, [c_void_p, c_int64, P_int64]),
    "lulesh_verify": (c_int, [c_void_p, P_double]),
    "lulesh_set_timing": (c_int, [c_void_p, c_int32]),
    "lulesh_kernel_time": (c_int, [c_void_p, c_int32, P_double, P_int64]),
    "lulesh_kernel_name": (c_char_p, [c_int32]),
    "lulesh_math_check": (c_int, [c_int32, P_double, P_double, P_double, c_int64, c_int32]),
    "lulesh_batch_create": (c_int, [c_int32, P_int32, P_double, P_double, P_int32, c_int32,
                                    c_int32, ctypes.POINTER(c_void_p)]),
    "lulesh_batch_size": (c_int32, [c_void_p]),
    "lulesh_batch_num_elem": (c_int64, [c_void_p, c_int32]),
    "lulesh_batch_num_node": (c_int64, [c_void_p, c_int32]),
    "lulesh_batch_download": (c_int, [c_void_p, c_int32, c_char_p, P_double, c_int64]),
    "lulesh_batch_upload": (c_int, [c_void_p, c_int32, c_char_p, P_double, c_int64]),
    "lulesh_batch_download_elem_bc": (c_int, [c_void_p, c_int32, P_uint32, c_int64]),
    "lulesh_batch_upload_elem_bc": (c_int, [c_void_p, c_int32, P_uint32, c_int64]),
    "lulesh_batch_get_scalar": (c_int, [c_void_p, c_int32, c_char_p, P_double]),
    "lulesh_batch_set_scalar": (c_int, [c_void_p, c_int32, c_char_p, c_double]),
    "lulesh_batch_run": (c_int, [c_void_p, c_int64, P_int64]),
    "lulesh_batch_verify": (c_int, [c_void_p, c_int32, P_double]),
    "lulesh_set_nodal_depth": (c_int, [c_void_p, c_int32]),
    "lulesh_nodal_depth": (c_int32, [c_void_p]),
    "lulesh_set_nodal_tiles": (c_int, [c_void_p, c_int32, c_int32]),
    "lulesh_nodal_tile": (c_int, [c_void_p, c_int32, P_int32]),
    "lulesh_autotune": (c_int, [c_void_p, P_int32, P_double]),
    "lulesh_grad_run": (c_int, [c_void_p, c_int64, c_int64, P_double, P_double, P_double,
                                P_double]),
    "lulesh_grad_run_revolve": (c_int, [c_void_p, c_int64, c_int64, P_double, P_double, P_double,
                                        P_double, P_int64]),
    "lulesh_revolve_cycles": (c_int64, [c_int64, c_int64]),
    "lulesh_replay": (c_int, [c_void_p, P_double, c_int64]),
    "lulesh_vjp_cycle": (c_int, [c_void_p, P_double]),
    "lulesh_batch_error": (c_int, [c_void_p, c_int32]),
    "lulesh_shadow_upload": (c_int, [c_void_p, c_char_p, P_double, c_int64]),
    "lulesh_shadow_download": (c_int, [c_void_p, c_char_p, P_double, c_int64]),
}

_lib = None


def load():
    """Load and type the library once; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"liblulesh_hip.so not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C lulesh.jl_amd/csrc`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
