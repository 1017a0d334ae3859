"""Grids of flat triangles, or of triangles and quadrilaterals with degree-1 geometry (the
role ndgrid plays for the reference)."""
import numpy as np

from . import _lib as L
from .element import ReferenceCellType


class Grid:
    """A grid given by vertex coordinates [nv,3] and cells: an [nc,3] array of triangles, an
    [nc,4] array of quadrilaterals (vertices in tensor order: reference (0,0), (1,0), (0,1),
    (1,1)), or a list of vertex lists of either length (a mixed grid)."""

    def __init__(self, points=None, cells=None, _handle=None):
        if _handle is None:
            points = np.ascontiguousarray(points, dtype=np.float64)
            if points.ndim != 2 or points.shape[1] != 3:
                raise ValueError("points must be [n,3]")
            lists = [list(c) for c in cells] if not isinstance(cells, np.ndarray) else None
            if lists is None and cells.ndim == 2 and cells.shape[1] == 3:
                c = np.ascontiguousarray(cells, dtype=np.int64)
                _handle = L.check_ptr(L.lib().grid_new(L.ptr(points), len(points), L.ptr(c), len(c)))
            else:
                if lists is None:
                    if cells.ndim != 2 or cells.shape[1] != 4:
                        raise ValueError("cells must be [m,3], [m,4] or a list of vertex lists")
                    lists = [list(c) for c in cells]
                if any(len(c) not in (3, 4) for c in lists):
                    raise ValueError("cells must have 3 (triangle) or 4 (quadrilateral) vertices")
                verts = np.array([v for c in lists for v in c], dtype=np.int64)
                types = np.array([int(ReferenceCellType.Triangle) if len(c) == 3 else int(ReferenceCellType.Quadrilateral)
                                  for c in lists], dtype=np.uint8)
                _handle = L.check_ptr(L.lib().grid_new_cells(L.ptr(points), len(points), L.ptr(verts), L.ptr(types),
                                                             len(types)))
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
    def cell_types(self):
        """Per cell: ReferenceCellType.Triangle (2) or Quadrilateral (3)."""
        t = np.zeros(L.lib().grid_ncells(self._handle), dtype=np.uint8)
        L.lib().grid_cell_types_copy(self._handle, L.ptr(t))
        return t

    @property
    def cells(self):
        """[nc,3] for a triangle grid, [nc,4] for a quadrilateral grid, else a list of vertex lists."""
        t = self.cell_types
        v = np.zeros(L.lib().grid_cell_vertices_size(self._handle), dtype=np.int64)
        L.lib().grid_cell_vertices_copy(self._handle, L.ptr(v))
        if np.all(t == int(ReferenceCellType.Triangle)):
            return v.reshape(-1, 3)
        if np.all(t == int(ReferenceCellType.Quadrilateral)):
            return v.reshape(-1, 4)
        out, i = [], 0
        for ti in t:
            n = 3 if ti == int(ReferenceCellType.Triangle) else 4
            out.append([int(x) for x in v[i:i + n]])
            i += n
        return out

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
        if cell_type in (ReferenceCellType.Triangle, ReferenceCellType.Quadrilateral):
            return int(np.count_nonzero(self.cell_types == int(cell_type)))
        if cell_type == ReferenceCellType.Interval:
            edges = set()
            for c in self.cells:
                c = list(c)
                # triangle edges; quadrilateral edges in tensor order: 0-1, 0-2, 1-3, 2-3
                pairs = [(1, 2), (0, 2), (0, 1)] if len(c) == 3 else [(0, 1), (0, 2), (1, 3), (2, 3)]
                edges.update(tuple(sorted((c[a], c[b]))) for a, b in pairs)
            return len(edges)
        return 0
