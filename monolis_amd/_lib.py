"""ctypes binding of libmonolis_hip.so (the C-ABI declared in include/monolis_hip.h).

The library must have been built (``__graft_entry__.build()`` or
``python -m monolis_amd.build``).  There is no fallback: if the shared object
is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmonolis_hip.so")

_lib = None

c_int, c_double, c_void_p, c_size_t, c_float = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float
c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)

# (name, restype, argtypes); every symbol declared in include/*.h
_FLAT = [c_int] * 6 + [c_void_p] * 3 + [c_void_p] * 3 + [c_void_p] * 2 + [c_int] * 3 + \
        [c_int, c_int, c_void_p, c_void_p, c_void_p] * 2
SIGNATURES = [
    ("monolis_solve_R_c_main", None, _FLAT + [c_void_p, c_void_p]),
    ("monolis_matvec_product_R_c_main", None, _FLAT),
    ("monolis_inner_product_R_c_main", None, [c_int, c_void_p, c_void_p, c_void_p, c_int]),
    ("monolis_solve_C_c_main", None, _FLAT + [c_void_p, c_void_p]),
    ("monolis_matvec_product_C_c_main", None, _FLAT),
    ("monolis_inner_product_C_c_main", None, [c_int, c_void_p, c_void_p, c_void_p, c_int]),
    ("monolis_hip_matrix_create_C", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int,
                                            c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_matrix_create_V_C", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_matrix_destroy_C", c_int, [c_void_p]),
    ("monolis_hip_matrix_set_values_C", c_int, [c_void_p, c_void_p]),
    ("monolis_hip_matrix_info_C", c_int, [c_void_p, c_ll_p]),
    ("monolis_hip_matvec_C", c_int, [c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_inner_product_C", c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_solve_C", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    ("monolis_hip_time_matvec_C", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_double_p]),
    ("monolis_hip_matrix_create_by_simple_mesh", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int,
                                                         c_void_p]),
    ("monolis_hip_matrix_get_pattern", c_int, [c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_matrix_get_values_R", c_int, [c_void_p, c_void_p]),
    ("monolis_hip_matrix_add_elements_R", c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    ("monolis_hip_set_Dirichlet_bc_R", c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_matrix_create_by_simple_mesh_C", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int,
                                                           c_void_p]),
    ("monolis_hip_matrix_get_pattern_C", c_int, [c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_matrix_create_by_simple_mesh_V", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_void_p, c_int,
                                                           c_void_p]),
    ("monolis_hip_matrix_create_by_simple_mesh_V_C", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_void_p,
                                                             c_int, c_void_p]),
    ("monolis_hip_matrix_get_values_C", c_int, [c_void_p, c_void_p]),
    ("monolis_hip_matrix_add_elements_C", c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    ("monolis_hip_set_Dirichlet_bc_C", c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_last_error", ctypes.c_char_p, []),
    ("monolis_hip_init", c_int, []),
    ("monolis_hip_stream", c_void_p, []),
    ("monolis_hip_synchronize", c_int, []),
    ("monolis_hip_malloc", c_int, [ctypes.POINTER(c_void_p), c_size_t]),
    ("monolis_hip_free", c_int, [c_void_p]),
    ("monolis_hip_memcpy", c_int, [c_void_p, c_void_p, c_size_t]),
    ("monolis_hip_memset", c_int, [c_void_p, c_int, c_size_t]),
    ("monolis_hip_matrix_create", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int,
                                          c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_matrix_destroy", c_int, [c_void_p]),
    ("monolis_hip_matrix_set_values", c_int, [c_void_p, c_void_p]),
    ("monolis_hip_matrix_info", c_int, [c_void_p, c_ll_p]),
    ("monolis_hip_matvec_R", c_int, [c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_matvec_S", c_int, [c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_inner_product_R", c_int, [c_int, c_void_p, c_void_p, c_double_p]),
    ("monolis_hip_inner_product_R_noalloc", c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_solve_R", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    ("monolis_hip_solve_refined_R", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                            c_int, c_void_p]),
    ("monolis_hip_time_matvec", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_double_p]),
    ("monolis_hip_time_matvec_each", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    ("monolis_hip_stream_copy", c_int, [c_void_p, c_void_p, ctypes.c_longlong, c_int, c_double_p]),
    ("monolis_hip_stream_read", c_int, [c_void_p, ctypes.c_longlong, c_int, c_double_p]),
    ("monolis_hip_num_variants", c_int, []),
    ("monolis_hip_matrix_autotune", c_int, [c_void_p, c_int, c_int, c_int_p, c_double_p]),
    ("monolis_hip_matrix_autotune_solver", c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int_p,
                                                   c_double_p]),
    ("monolis_hip_matrix_set_variant", c_int, [c_void_p, c_int, c_int]),
    ("monolis_hip_batch_create", c_int, [ctypes.POINTER(c_void_p), c_int, c_void_p, c_int, c_int,
                                         c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_batch_destroy", c_int, [c_void_p]),
    ("monolis_hip_batch_solve", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    ("monolis_hip_batch_num_configs", c_int, []),
    ("monolis_hip_batch_config", c_int, [c_int, c_int_p, c_int_p]),
    ("monolis_hip_batch_set_config", c_int, [c_void_p, c_int]),
    ("monolis_hip_batch_schedule", c_int, [c_void_p, c_int, c_int_p, c_int_p, c_int_p]),
    ("monolis_hip_batch_set_wait_limit", c_int, [c_void_p, c_double]),
    ("monolis_hip_batch_autotune", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int_p,
                                           c_double_p]),
    ("monolis_hip_batch_time", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_double_p]),
    ("monolis_hip_solve_multi_R", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p]),
    ("monolis_hip_matrix_create_V", c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_matrix_set_scalar_R", c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double]),
    ("monolis_hip_matrix_add_scalar_R", c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double]),
    ("monolis_hip_matrix_get_scalar_R", c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double_p, c_int_p]),
    ("monolis_hip_solve_transpose_R", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_outer_grad_R", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("monolis_hip_adjoint_R", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                      c_void_p]),
    ("monolis_hip_matvec_multi_R", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("monolis_hip_matvec_multi_S", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("monolis_hip_multi_num_configs", c_int, []),
    ("monolis_hip_multi_config", c_int, [c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p]),
    ("monolis_hip_multi_autotune", c_int, [c_void_p, c_int, c_int, c_int, c_int_p, c_double_p]),
    ("monolis_hip_multi_set_config", c_int, [c_void_p, c_int, c_int]),
    ("monolis_hip_solve_multi_refined_R", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                                  c_double, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
]


class MonolisHipError(RuntimeError):
    pass


def load():
    """Load libmonolis_hip.so and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    # MONOLIS_HIP_LIBRARY selects another build of the same library (e.g. one
    # compiled with different kernel options); it must export the same symbols.
    path = os.environ.get("MONOLIS_HIP_LIBRARY", LIB_PATH)
    if not os.path.exists(path):
        raise MonolisHipError(
            f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m monolis_amd.build` (the HIP path has no fallback)")
    L = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def extra_symbols():
    """Names of further exported entry points registered by extension modules."""
    return [s[0] for s in SIGNATURES]


def register(signatures):
    """Extension modules add their signatures here (before or after load())."""
    for sig in signatures:
        if sig[0] not in [s[0] for s in SIGNATURES]:
            SIGNATURES.append(sig)
            if _lib is not None:
                fn = getattr(_lib, sig[0])
                fn.restype = sig[1]
                fn.argtypes = sig[2]


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().monolis_hip_last_error()
        raise MonolisHipError(f"{what} failed (rc={rc}): {msg.decode() if msg else ''}")
