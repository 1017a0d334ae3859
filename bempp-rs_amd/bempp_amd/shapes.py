"""Sphere meshes (src/shapes.rs:19-114 of the reference)."""
from . import _lib as L
from .grid import Grid


def regular_sphere(refinement_level):
    """Octahedron refined `refinement_level` times, vertices projected to radius 1
    (8*4^L triangles, 4^(L+1)+2 vertices), as bempp::shapes::regular_sphere."""
    return Grid(_handle=L.check_ptr(L.lib().grid_new_sphere(0, int(refinement_level))))


def icosphere(refinement_level):
    """Icosahedron refined by the same scheme (20*4^L triangles, 10*4^L+2 vertices)."""
    return Grid(_handle=L.check_ptr(L.lib().grid_new_sphere(1, int(refinement_level))))


def ellipsoid(refinement_level, semi_axes):
    """Level-L icosphere scaled to an ellipsoid with semi-axes (a, b, c)."""
    import numpy as np

    g = icosphere(refinement_level)
    return Grid(g.points * np.asarray(semi_axes, dtype=np.float64).reshape(1, 3), g.cells)
