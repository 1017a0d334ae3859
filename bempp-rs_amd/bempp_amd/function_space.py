"""Function spaces (python/bempp/function_space.py of the reference)."""
import numpy as np

from . import _lib as L
from .element import Continuity, LagrangeElementFamily


class FunctionSpace:
    """A Lagrange space (degree 0 or 1) on a grid."""

    def __init__(self, grid, family, handle):
        self._rs_space = handle
        self._grid = grid  # keeps the grid alive
        self._family = family

    def __del__(self, _release=L.release):
        _release("space_t_free", getattr(self, "_rs_space", None))
        self._rs_space = None

    @property
    def grid(self):
        return self._grid

    @property
    def family(self):
        return self._family

    @property
    def dtype(self):
        return self._family.dtype

    @property
    def local_size(self):
        return int(L.lib().space_local_size(self._rs_space))

    @property
    def global_size(self):
        return int(L.lib().space_global_size(self._rs_space))

    @property
    def is_serial(self):
        return bool(L.lib().space_is_serial(self._rs_space))

    def cell_dofs(self, cell):
        if not L.lib().space_has_cell_dofs(self._rs_space, cell):
            return None
        n = L.lib().space_cell_dofs_size(self._rs_space, cell)
        d = np.zeros(n, dtype=np.uintp)
        L.lib().space_cell_dofs(self._rs_space, cell, L.ptr(d))
        return [int(i) for i in d]

    def global_dof_index(self, local_dof_index):
        return int(L.lib().space_global_dof_index(self._rs_space, local_dof_index))

    def cell_colouring(self):
        """Greedy cell colouring (src/function.rs:281-351): list of cell lists."""
        col = np.zeros(L.lib().grid_ncells(self._grid._handle), dtype=np.int64)
        n = L.lib().space_cell_colouring(self._rs_space, L.ptr(col))
        return [list(np.nonzero(col == c)[0]) for c in range(n)]


def function_space(grid, family):
    """Create a function space (python/bempp/function_space.py:319-332)."""
    if not isinstance(family, LagrangeElementFamily):
        raise ValueError("family must be a LagrangeElementFamily")
    cont = 0 if family.continuity == Continuity.Standard else 1
    h = L.check_ptr(L.lib().function_space_new(grid._handle, family.degree, cont))
    return FunctionSpace(grid, family, h)
