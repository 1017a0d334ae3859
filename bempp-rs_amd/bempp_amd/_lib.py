"""ctypes binding of libbempp_amd.so (the C ABI declared in include/bempp_amd.h)."""
import ctypes
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
# BEMPP_AMD_LIBRARY selects another build of the same library (e.g. the sanitizer build
# of the host code, make -C bempp-rs_amd/csrc sanitize)
_LIB_FILE = Path(os.environ["BEMPP_AMD_LIBRARY"]) if os.environ.get("BEMPP_AMD_LIBRARY") else _HERE / "libbempp_amd.so"

c_size = ctypes.c_size_t
c_u8 = ctypes.c_uint8
c_int = ctypes.c_int
c_vp = ctypes.c_void_p
c_dp = ctypes.POINTER(ctypes.c_double)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_up = ctypes.POINTER(ctypes.c_size_t)

# name: (restype, argtypes)
_SIGNATURES = {
    "bempp_last_error": (ctypes.c_char_p, []),
    "bempp_version": (ctypes.c_char_p, []),
    "bempp_device_count": (c_int, []),
    "boundary_assembler_options_new": (c_vp, []),
    "boundary_assembler_options_t_free": (None, [c_vp]),
    "boundary_assembler_options_set_regular_quadrature_degree": (None, [c_vp, c_u8, c_size]),
    "boundary_assembler_options_get_regular_quadrature_degree": (c_size, [c_vp, c_u8]),
    "boundary_assembler_options_has_regular_quadrature_degree": (c_int, [c_vp, c_u8]),
    "boundary_assembler_options_set_singular_quadrature_degree": (None, [c_vp, c_u8, c_u8, c_size]),
    "boundary_assembler_options_get_singular_quadrature_degree": (c_size, [c_vp, c_u8, c_u8]),
    "boundary_assembler_options_has_singular_quadrature_degree": (c_int, [c_vp, c_u8, c_u8]),
    "boundary_assembler_options_set_batch_size": (None, [c_vp, c_size]),
    "boundary_assembler_options_get_batch_size": (c_size, [c_vp]),
    "boundary_assembler_options_set_precision": (c_int, [c_vp, c_u8]),
    "boundary_assembler_options_get_precision": (c_u8, [c_vp]),
    "grid_new": (c_vp, [c_vp, c_size, c_vp, c_size]),
    "grid_new_sphere": (c_vp, [c_int, ctypes.c_uint32]),
    "grid_t_free": (None, [c_vp]),
    "grid_npoints": (c_size, [c_vp]),
    "grid_ncells": (c_size, [c_vp]),
    "grid_points_copy": (None, [c_vp, c_vp]),
    "grid_cells_copy": (None, [c_vp, c_vp]),
    "grid_new_cells": (c_vp, [c_vp, c_size, c_vp, c_vp, c_size]),
    "grid_cell_vertices_size": (c_size, [c_vp]),
    "grid_cell_vertices_copy": (None, [c_vp, c_vp]),
    "grid_cell_types_copy": (None, [c_vp, c_vp]),
    "grid_set_points": (c_int, [c_vp, c_vp]),
    "function_space_new": (c_vp, [c_vp, c_size, c_u8]),
    "space_t_free": (None, [c_vp]),
    "space_global_size": (c_size, [c_vp]),
    "space_local_size": (c_size, [c_vp]),
    "space_is_serial": (c_int, [c_vp]),
    "space_cell_dofs_size": (c_size, [c_vp, c_size]),
    "space_has_cell_dofs": (c_int, [c_vp, c_size]),
    "space_cell_dofs": (None, [c_vp, c_size, c_vp]),
    "space_global_dof_index": (c_size, [c_vp, c_size]),
    "space_cell_colouring": (c_size, [c_vp, c_vp]),
    "laplace_boundary_assembler_new": (c_vp, [c_vp, c_u8, c_u8]),
    "helmholtz_boundary_assembler_new": (c_vp, [c_vp, c_vp, c_u8, c_u8]),
    "boundary_assembler_t_free": (None, [c_vp]),
    "boundary_assembler_dtype": (c_u8, [c_vp]),
    "boundary_assembler_assemble_into_memory": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "boundary_assembler_assemble_into_device_memory": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_vp]),
    "boundary_assembler_assemble": (c_vp, [c_vp, c_vp, c_vp]),
    "boundary_assembler_assemble_singular": (c_vp, [c_vp, c_vp, c_vp]),
    "boundary_assembler_assemble_singular_correction": (c_vp, [c_vp, c_vp, c_vp]),
    "potential_assemble_into_memory": (c_int, [c_vp, c_vp, c_vp, c_size, c_vp]),
    "potential_assemble_into_device_memory": (c_int, [c_vp, c_vp, c_vp, c_size, c_vp, c_int, c_vp]),
    "boundary_assembler_shape_gradient": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "boundary_assembler_shape_gradient_noalloc": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "matrix_shape0": (c_size, [c_vp]),
    "matrix_shape1": (c_size, [c_vp]),
    "matrix_is_csr": (c_int, [c_vp]),
    "matrix_dtype": (c_u8, [c_vp]),
    "matrix_dense_data": (c_vp, [c_vp]),
    "matrix_dense_data_copy": (None, [c_vp, c_vp]),
    "matrix_csr_data_size": (c_size, [c_vp]),
    "matrix_csr_indices_size": (c_size, [c_vp]),
    "matrix_csr_indptr_size": (c_size, [c_vp]),
    "matrix_csr_data_copy": (None, [c_vp, c_vp]),
    "matrix_csr_indices_copy": (None, [c_vp, c_vp]),
    "matrix_csr_indptr_copy": (None, [c_vp, c_vp]),
    "matrix_t_free": (None, [c_vp]),
    "boundary_assembler_apply": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_int]),
    "boundary_assembler_set_profiling": (c_int, [c_vp, c_int]),
    "boundary_assembler_kernel_times": (c_int, [c_vp, c_vp]),
    "boundary_assembler_batch_new": (c_vp, [c_vp, c_vp, c_vp, c_size]),
    "boundary_assembler_batch_t_free": (None, [c_vp]),
    "boundary_assembler_batch_assemble": (c_int, [c_vp, c_vp]),
    "boundary_assembler_batch_set_graph": (c_int, [c_vp, c_int]),
    "boundary_assembler_batch_set_launch": (c_int, [c_vp, c_int, c_int, c_int]),
    "boundary_assembler_batch_get_launch": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "boundary_assembler_batch_set_tile_waves": (c_int, [c_vp, c_int]),
    "boundary_assembler_batch_get_tile_waves": (c_int, [c_vp, c_vp]),
    "boundary_assembler_batch_counts": (c_int, [c_vp, c_vp]),
    "boundary_assembler_batch_autotune": (ctypes.c_double, [c_vp, ctypes.c_char_p, c_int, c_vp]),
    "boundary_assembler_batch_set_profiling": (c_int, [c_vp, c_int]),
    "boundary_assembler_batch_kernel_times": (c_int, [c_vp, c_vp]),
    "bempp_triangle_rule": (c_int, [c_int, c_vp, c_vp]),
    "bempp_canonical_pair_rule": (c_int, [c_int, c_int, c_vp]),
}

_lib = None


def lib_path():
    return str(_LIB_FILE)


def lib():
    """Load the native library; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is None:
        # torch bundles its own HIP runtime (same soname as the system one). Loading
        # torch first makes libbempp_amd bind to that runtime, so device memory and
        # streams are shared with torch instead of two runtimes competing for the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not _LIB_FILE.exists():
            raise ImportError(f"{_LIB_FILE} is missing: build it with __graft_entry__.build() "
                              "(make -C bempp-rs_amd/csrc); bempp_amd has no non-native fallback")
        L = ctypes.CDLL(str(_LIB_FILE))
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def release(name, handle):
    """Free a native handle with the C function `name`. At interpreter shutdown (module globals
    already cleared) it does nothing: the process is ending and the HIP runtime may be gone."""
    lib_ = _lib
    if handle and lib_ is not None:
        getattr(lib_, name)(handle)


def last_error():
    return lib().bempp_last_error().decode()


def check(rc):
    if rc != 0:
        raise ValueError(last_error())


def check_ptr(p):
    if not p:
        raise ValueError(last_error())
    return p


def device_count():
    return int(lib().bempp_device_count())


def version():
    return lib().bempp_version().decode()


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)
