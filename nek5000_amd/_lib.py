"""ctypes binding of libnekamd.so (the C ABI declared in include/nekamd.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object is missing (build it with ``python -m nek5000_amd.build``).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnekamd.so")

c_int = ctypes.c_int
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_char = ctypes.c_char
c_char_p = ctypes.c_char_p
c_void_p = ctypes.c_void_p
P = ctypes.POINTER


class NkhParams(ctypes.Structure):
    _fields_ = [
        ("istep", c_int),
        ("param18", c_int),
        ("param22", c_double),
        ("restol", c_double),
        ("precision", c_int),
        ("param42", c_int),
        ("param21", c_double),
        ("param40", c_int),
        ("param93", c_int),
        ("param94", c_int),
        ("param95", c_int),
        ("ifprojfld", c_int),
        ("ifield", c_int),
        ("ldimt_proj", c_int),
        ("ifmvbd", c_int),
    ]


class NkhBatchSolve(ctypes.Structure):
    _fields_ = [
        ("u", c_void_p), ("rhs", c_void_p), ("h1", c_void_p), ("h2", c_void_p),
        ("mask", c_void_p), ("mult", c_void_p), ("tli", c_double),
    ]


class NkhResult(ctypes.Structure):
    _fields_ = [
        ("niter", c_int),
        ("converged", c_int),
        ("rbn0", c_double),
        ("rbn2", c_double),
        ("tol", c_double),
    ]


# name -> (restype, argtypes); every declaration of include/nekamd.h
SIGNATURES = {
    "nkh_init": (c_int, [c_int]),
    "nkh_finalize": (c_int, []),
    "nkh_last_error": (c_char_p, []),
    "nkh_set_stream": (c_int, [c_void_p]),
    "nkh_use_library_stream": (c_int, []),
    "nkh_synchronize": (c_int, []),
    "nkh_zwgll": (c_int, [c_void_p, c_void_p, c_int]),
    "nkh_dgll": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    "nkh_setvert3d": (c_int, [c_void_p, P(c_int64), c_int, c_int, c_void_p]),
    "nkh_re2_header": (c_int, [c_char_p, P(c_int), P(c_int), P(c_int), P(c_int)]),
    "nkh_map_header": (c_int, [c_char_p, P(c_int), P(c_int)]),
    "nkh_map_read": (c_int, [c_char_p, c_int, c_int, c_void_p, c_void_p]),
    "nkh_re2_read": (c_int, [c_char_p, c_int, c_int, c_int] + [c_void_p] * 8),
    "nkh_mesh_vertex_ids": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_double, c_void_p, P(c_int64)]),
    "nkh_mesh_xyz_from_corners": (c_int, [c_int, c_int] + [c_void_p] * 8),
    "nkh_mesh_create": (c_int, [c_int, c_int, P(c_int)]),
    "nkh_mesh_destroy": (c_int, [c_int]),
    "nkh_mesh_geom_from_xyz": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "nkd_mesh_geom_from_xyz": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "nkh_mesh_set_geom": (c_int, [c_int] + [c_void_p] * 7),
    "nkh_mesh_set_deformed": (c_int, [c_int, c_void_p]),
    "nkh_mesh_gs_setup": (c_int, [c_int, c_void_p]),
    "nkh_mesh_info": (c_int, [c_int, P(c_int64), P(c_double), P(c_int64), P(c_int64), P(c_int)]),
    "nkh_mesh_ptrs": (c_int, [c_int, P(c_void_p), P(c_void_p)]),
    "nkd_axhelm": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "nkh_axhelm_tuning": (c_int, [c_int, P(c_int), P(c_int), P(c_int)]),
    "nkh_axhelm_set_tuning": (c_int, [c_int, c_int, c_int]),
    "nkh_axhelm_tuning_fused": (c_int, [c_int, P(c_int), P(c_int)]),
    "nkh_axhelm32_set_variant": (c_int, [c_int, c_int]),
    "nkh_axhelm32_variant": (c_int, [c_int, P(c_int)]),
    "nkh_fdm_set_variant": (c_int, [c_int]),
    "nkh_fdm_variant": (c_int, [P(c_int)]),
    "nkd_dssum": (c_int, [c_int, c_void_p]),
    "nkd_dsop": (c_int, [c_int, c_void_p, c_char]),
    "nkd_setprec": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "nkd_adjoint_coef_grad": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nkd_hmholtz": (c_int, [c_int, c_char_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_double, c_int, c_int, P(NkhParams), P(NkhResult),
                            c_void_p]),
    "nkh_batch_create": (c_int, [c_int, c_void_p, P(c_int)]),
    "nkh_batch_destroy": (c_int, [c_int]),
    "nkd_hmholtz_batch": (c_int, [c_int, c_void_p, c_int, P(NkhParams), c_void_p, P(c_int)]),
    "nkd_cggo": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                         c_double, c_int, P(NkhParams), P(NkhResult), c_void_p]),
    "nkd_cggo_noalloc": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_double, c_int, P(NkhParams), P(NkhResult),
                                 c_void_p]),
    "nkh_mg_setup": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nkd_mg_solve": (c_int, [c_int, c_void_p, c_void_p]),
    "nkd_hmh_flex_cg": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                c_void_p, c_void_p, c_void_p]),
    "nkd_hmh_gmres": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                              c_void_p, c_void_p, c_void_p]),
    "nkh_mg_info": (c_int, [c_int, P(c_int), c_void_p, P(c_int), P(c_double)]),
    "nkh_proj_create": (c_int, [c_int, c_int, P(c_int)]),
    "nkh_proj_destroy": (c_int, [c_int]),
    "nkh_proj_info": (c_int, [c_int, P(c_int), P(c_int)]),
    "nkh_proj_column": (c_int, [c_int, c_int, c_int, P(c_void_p)]),
    "nkd_project1": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                             c_void_p, P(c_int)]),
    "nkd_project2": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nkd_hsolve": (c_int, [c_int, c_char_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_double, c_int, c_int, c_int, P(NkhParams), P(NkhResult),
                           c_void_p]),
    "nkh_timers_enable": (c_int, [c_int]),
    "nkh_timers_reset": (c_int, []),
    "nkh_timers_read": (c_int, [c_char_p, P(c_double), P(c_int64)]),
    "nkh_axhelm": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "nkh_dssum": (c_int, [c_int, c_void_p]),
    "nkh_hmholtz": (c_int, [c_int, c_char_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_double, c_int, c_int, P(NkhParams), P(NkhResult),
                            c_void_p]),
    "nkh_hsolve": (c_int, [c_int, c_char_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_double, c_int, c_int, c_int, P(NkhParams), P(NkhResult),
                           c_void_p]),
}

_lib = None


class NekAmdError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load libnekamd.so and attach prototypes; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NekAmdError(
            f"{path} is missing: the HIP library is the only implementation; "
            "build it with `python -m nek5000_amd.build`")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 and
    # loads it by file name, so it must be loaded first; libnekamd's
    # libamdhip64.so.7 dependency then resolves to that already-loaded copy
    # instead of bringing up a second runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = _lib.nkh_last_error().decode(errors="replace") if _lib else ""
        raise NekAmdError(f"{what} failed with code {rc}: {msg}")
