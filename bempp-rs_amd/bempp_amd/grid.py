"""Grids of flat triangles (the role ndgrid plays for the reference)."""
import numpy as np

from . import _lib as L
from .element import ReferenceCellType


class Grid:
    """A grid of flat triangles given by vertex coordinates [nv,3] and cells [nc,3]."""

    def __init__(self, points=None, cells=None, _handle=None):
        if _handle is None:
            points = np.ascontiguousarray(points, dtype=np.float64)
            cells = np.ascontiguousarray(cells, dtype=np.int64)
            if points.ndim != 2 or points.shape[1] != 3 or cells.ndim != 2 or cells.shape[1] != 3:
                raise ValueError("points must be [n,3] and cells [m,3]")
            _handle = L.check_ptr(L.lib().grid_new(L.ptr(points), len(points), L.ptr(cells), len(cells)))
        self._handle = _handle

    def __del__(self, _release=L.release):
        _release("grid_t_free", getattr(self, "_handle", None))
        self._handle = None

    @property
    def topology_dim(self):
        return 2

    @property
    def geometry_dim(self):
        return 3

    @property
    def points(self):
        p = np.zeros((L.lib().grid_npoints(self._handle), 3))
        L.lib().grid_points_copy(self._handle, L.ptr(p))
        return p

    @property
    def cells(self):
        c = np.zeros((L.lib().grid_ncells(self._handle), 3), dtype=np.int64)
        L.lib().grid_cells_copy(self._handle, L.ptr(c))
        return c

    def set_points(self, points):
        """Move the vertices (same topology); cached device geometry is dropped."""
        p = np.ascontiguousarray(points, dtype=np.float64)
        if p.shape != (L.lib().grid_npoints(self._handle), 3):
            raise ValueError("points must keep the grid's shape")
        L.check(L.lib().grid_set_points(self._handle, L.ptr(p)))

    def entity_count(self, cell_type):
        cell_type = ReferenceCellType(cell_type)
        if cell_type == ReferenceCellType.Point:
            return int(L.lib().grid_npoints(self._handle))
        if cell_type == ReferenceCellType.Triangle:
            return int(L.lib().grid_ncells(self._handle))
        if cell_type == ReferenceCellType.Interval:
            c = self.cells
            e = np.sort(np.concatenate([c[:, [1, 2]], c[:, [0, 2]], c[:, [0, 1]]]), axis=1)
            return len(np.unique(e, axis=0))
        return 0
