"""The argument lists of the ndgrid-style ingest entry points, as the reference's python layer
passes them (python/bempp/function_space.py:143-191), shared by the CPU and GPU tests."""
import ctypes

import numpy as np

from bempp_amd import _lib as L


def ndgrid_args(points, cells, perm, dtype=np.float64, tdim=2, cell_type=2, topo=True):
    """The argument list of single_element_grid_borrowed_create_bempp for a triangle grid whose
    topology numbers vertex p as perm[p] (python/bempp/function_space.py:155-178 passes ndgrid's
    arrays the same way)."""
    sz = ctypes.c_size_t
    keep = []
    pts = np.ascontiguousarray(points, dtype=dtype)
    cls = np.ascontiguousarray(cells, dtype=np.uintp)
    tcells = np.ascontiguousarray(perm[cells], dtype=np.uintp)
    etypes = np.array([0, 1, cell_type], dtype=np.uint8)
    counts = np.array([len(points), 0, len(cells)], dtype=np.uintp)
    keep += [pts, cls, tcells, etypes, counts]
    row2 = (ctypes.c_void_p * 3)(tcells.ctypes.data if topo else None, None, None)
    down = (ctypes.c_void_p * 3)(None, None, ctypes.addressof(row2))
    shape2 = np.array([3, 0, 0], dtype=np.uintp)
    shp = (ctypes.c_void_p * 3)(None, None, shape2.ctypes.data)
    keep += [row2, down, shape2, shp]
    args = [sz(tdim), None, None, etypes.ctypes.data_as(ctypes.c_void_p), counts.ctypes.data_as(ctypes.c_void_p),
            ctypes.cast(down, ctypes.c_void_p), ctypes.cast(shp, ctypes.c_void_p), None, None,
            pts.ctypes.data_as(ctypes.c_void_p), sz(3), sz(len(points)),
            ctypes.c_uint8(0 if dtype == np.float32 else 1), cls.ctypes.data_as(ctypes.c_void_p), sz(3),
            sz(len(cells)), sz(1)]
    return args, keep


def raw_lib():
    lib = ctypes.CDLL(L.lib_path())
    vp, sz, u8 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint8
    lib.single_element_grid_borrowed_create_bempp.restype = vp
    lib.single_element_grid_borrowed_create_bempp.argtypes = [sz, vp, vp, vp, vp, vp, vp, vp, vp, vp, sz, sz, u8, vp,
                                                              sz, sz, sz]
    lib.create_lagrange_family_bempp.restype = vp
    lib.create_lagrange_family_bempp.argtypes = [sz, u8, u8]
    lib.create_raviart_thomas_family_bempp.restype = vp
    lib.create_raviart_thomas_family_bempp.argtypes = [sz, u8, u8]
    lib.function_space.restype = vp
    lib.function_space.argtypes = [vp, vp]
    lib.space_local_space.restype = vp
    lib.space_local_space.argtypes = [vp]
    for f in ("grid_t_free_bempp", "element_family_t_free_bempp", "space_t_free"):
        getattr(lib, f).argtypes = [vp]
    lib.space_get_local_dof_numbers_size.restype = sz
    lib.space_get_local_dof_numbers_size.argtypes = [vp, sz, sz]
    lib.space_get_local_dof_numbers.argtypes = [vp, sz, sz, vp]
    lib.space_global_size.restype = sz
    lib.space_global_size.argtypes = [vp]
    lib.space_dtype.restype = u8
    lib.space_dtype.argtypes = [vp]
    lib.space_is_owned.restype = ctypes.c_int
    lib.space_is_owned.argtypes = [vp, sz]
    lib.space_cell_dofs.argtypes = [vp, sz, vp]
    lib.grid_npoints.restype = sz
    lib.grid_npoints.argtypes = [vp]
    lib.grid_points_copy.argtypes = [vp, vp]
    lib.grid_cells_copy.argtypes = [vp, vp]
    lib.bempp_last_error.restype = ctypes.c_char_p
    return lib
