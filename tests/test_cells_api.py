"""Grids with quadrilateral cells through the C ABI (grid_new_cells, the ndgrid ingest with 4
points per cell) and the Python Grid: round trips, cell types and entity counts, DOF
numbering against the oracle's (src/function.rs:365-445 over mixed cell types), and the
entry points specialised to triangles refusing such grids with an error (no GPU needed)."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from bempp_amd import _lib as L
from bempp_amd.assembly.boundary import BoundaryAssemblerOptions, OperatorType, create_laplace_assembler
from bempp_amd.assembly.potential import create_laplace_potential_assembler
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.grid import Grid
from ndgrid_helpers import raw_lib
from quad_meshes import cube, cubed_sphere, mixed_screen, split


def test_quadrilateral_and_mixed_grids_round_trip():
    pts, quads = cubed_sphere(3)
    g = Grid(pts, np.array(quads))
    assert g.cells.shape == (54, 4) and np.array_equal(g.cells, np.array(quads))
    assert np.all(g.cell_types == int(ReferenceCellType.Quadrilateral))
    assert g.entity_count(ReferenceCellType.Quadrilateral) == 54 and g.entity_count(ReferenceCellType.Triangle) == 0
    assert g.entity_count(ReferenceCellType.Interval) == 108  # closed surface: V - E + F = 2
    pts, mixed, _, _ = mixed_screen()
    gm = Grid(pts, mixed)
    assert gm.cells == mixed
    assert gm.entity_count(ReferenceCellType.Quadrilateral) == 2 and gm.entity_count(ReferenceCellType.Triangle) == 4
    # a list of triangles only is a triangle grid
    gt = Grid(pts, [c for c in mixed if len(c) == 3])
    assert gt.cells.shape == (4, 3)
    with pytest.raises(ValueError, match="3 .* or 4"):
        Grid(pts, [[0, 1, 3, 4, 5]])
    with pytest.raises(ValueError, match="out of range"):
        Grid(pts, [[0, 1, 3, 40]])


@pytest.mark.parametrize("mesh", ["mixed_screen", "cubed_sphere", "mixed_cube"])
def test_spaces_on_quadrilateral_grids_match_oracle(mesh):
    if mesh == "mixed_screen":
        pts, cells = mixed_screen()[:2]
    elif mesh == "cubed_sphere":
        pts, cells = cubed_sphere(3)
    else:
        pts, c = cube(2)
        cells = split(c, lambda i: i % 3 == 0)
    g = Grid(pts, cells)
    for degree, cont in [(0, Continuity.Discontinuous), (1, Continuity.Standard), (1, Continuity.Discontinuous)]:
        s = function_space(g, create_family(Family.Lagrange, degree, cont))
        size, dofs = O.cells_space(pts, cells, degree, cont == Continuity.Standard)
        assert s.global_size == size
        assert [s.cell_dofs(c) for c in range(len(cells))] == dofs


def test_ndgrid_ingest_of_quadrilateral_grid():
    lib = raw_lib()
    pts, quads = cube(2)
    quads = np.array(quads)
    sz = ctypes.c_size_t
    p = np.ascontiguousarray(pts)
    c = np.ascontiguousarray(quads, dtype=np.uintp)
    etypes = np.array([0, 1, 3], dtype=np.uint8)
    counts = np.array([len(pts), 0, len(quads)], dtype=np.uintp)
    args = [sz(2), None, None, etypes.ctypes.data_as(ctypes.c_void_p), counts.ctypes.data_as(ctypes.c_void_p), None,
            None, None, None, p.ctypes.data_as(ctypes.c_void_p), sz(3), sz(len(pts)), ctypes.c_uint8(1),
            c.ctypes.data_as(ctypes.c_void_p), sz(4), sz(len(quads)), sz(1)]
    gh = lib.single_element_grid_borrowed_create_bempp(*args)
    assert gh, lib.bempp_last_error()
    g = Grid(_handle=gh)
    assert np.array_equal(g.cells, quads)
    etypes[2] = 2  # 4 points per cell but a triangle type
    assert not lib.single_element_grid_borrowed_create_bempp(*args)
    assert "does not match" in lib.bempp_last_error().decode()


def test_triangle_only_entry_points_refuse_quadrilateral_grids():
    """Matrix-free apply, the shape gradient and potentials are specialised to triangles:
    they refuse such grids before any device work."""
    pts, quads = cube(1)
    g = Grid(pts, np.array(quads))
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Quadrilateral, 16)
    a = create_laplace_assembler(OperatorType.SingleLayer, o)
    with pytest.raises(ValueError, match="quadrilateral"):
        a.apply(s, s, np.ones(s.global_size))
    with pytest.raises(ValueError, match="quadrilateral"):
        a.shape_gradient(s, s, np.ones(s.global_size), np.ones(s.global_size))
    pa = create_laplace_potential_assembler(OperatorType.SingleLayer, o)
    with pytest.raises(ValueError, match="quadrilateral"):
        pa.assemble(s, np.array([[3.0, 0.0, 0.0]]))
