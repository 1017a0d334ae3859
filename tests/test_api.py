"""Host-side behaviour of the C ABI and the python mirror (no GPU needed)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import bempp_amd
from bempp_amd import _lib as L
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, Precision,
                                         create_helmholtz_assembler, create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.shapes import icosphere, regular_sphere

import oracle_lib as O
from ndgrid_helpers import ndgrid_args as _ndgrid_args, raw_lib as _raw_lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "bempp_amd.h"


def test_library_exports_every_declared_symbol():
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**([a-z_][a-z0-9_]*)\(", HEADER.read_text(),
                       flags=re.M)
    assert len(names) > 50
    lib = ctypes.CDLL(L.lib_path())
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("otype", [OperatorType.SingleLayer, OperatorType.DoubleLayer,
                                   OperatorType.AdjointDoubleLayer, OperatorType.Hypersingular])
def test_create_assembler(otype):
    """python/test/test_boundary_assembler.py:10-42."""
    o = BoundaryAssemblerOptions()
    assert o.regular_quadrature_degree(ReferenceCellType.Triangle) != 3
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 3)
    assert o.regular_quadrature_degree(ReferenceCellType.Triangle) == 3
    with pytest.raises(ValueError):
        o.set_regular_quadrature_degree(ReferenceCellType.Interval, 3)
    with pytest.raises(ValueError):
        o.regular_quadrature_degree(ReferenceCellType.Interval)
    assert o.singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle) != 3
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 3)
    assert o.singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle) == 3
    with pytest.raises(ValueError):
        o.set_singular_quadrature_degree(ReferenceCellType.Interval, ReferenceCellType.Interval, 3)
    assert o.batch_size() != 4
    o.set_batch_size(4)
    assert o.batch_size() == 4
    a = create_laplace_assembler(otype, o)
    assert a.dtype == np.float64
    with pytest.raises(ValueError):
        o.set_batch_size(8)  # read-only once used (boundary.py:117-118)


def test_defaults_and_precision():
    o = BoundaryAssemblerOptions()
    assert o.regular_quadrature_degree(ReferenceCellType.Triangle) == 37  # src/boundary_assemblers.rs:47
    assert o.regular_quadrature_degree(ReferenceCellType.Quadrilateral) == 37
    assert o.singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle) == 4
    assert o.singular_quadrature_degree(ReferenceCellType.Quadrilateral, ReferenceCellType.Triangle) == 4
    assert o.batch_size() == 128
    assert o.precision() == Precision.FP64
    o.set_precision(Precision.Mixed)
    assert o.precision() == Precision.Mixed


def test_invalid_assemblers():
    """The reference panics on these (src/bindings/boundary_assemblers.rs:153-164,197-208)."""
    with pytest.raises(ValueError):
        create_laplace_assembler(OperatorType.ElectricField)
    with pytest.raises(ValueError):
        create_laplace_assembler(OperatorType.SingleLayer, dtype=np.complex128)
    with pytest.raises(ValueError):
        create_helmholtz_assembler(1.0, OperatorType.SingleLayer, dtype=np.float64)
    assert create_helmholtz_assembler(1.0, OperatorType.DoubleLayer, dtype=np.complex64).dtype == np.complex64


@pytest.mark.parametrize("level", range(4))
def test_spaces(level):
    """python/test/test_function_space.py:8-41 and tests/test_space.rs:137-164."""
    g = regular_sphere(level)
    dp0 = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    p1 = function_space(g, create_family(Family.Lagrange, 1))
    assert dp0.global_size == dp0.local_size == g.entity_count(ReferenceCellType.Triangle)
    assert p1.global_size == g.entity_count(ReferenceCellType.Point)
    assert dp0.is_serial and p1.is_serial
    assert g.topology_dim == 2 and g.geometry_dim == 3
    assert len(dp0.cell_colouring()) == 1
    for cells in p1.cell_colouring():
        verts = g.cells[cells].ravel()
        assert len(verts) == len(np.unique(verts))
    assert dp0.cell_dofs(0) == [0] and dp0.cell_dofs(10 ** 9) is None


def test_spaces_match_oracle():
    """DOF numbering and colouring agree with the oracle's restatement of src/function.rs."""
    g = icosphere(2)
    p, c = g.points, g.cells
    po, co = O.sphere(2, "icosahedron")
    assert np.abs(p - po).max() < 1e-15 and np.array_equal(c, co)  # FMA contraction may differ by 1 ulp
    for deg, cont in [(0, False), (1, True), (1, False)]:
        s = function_space(g, create_family(Family.Lagrange, deg,
                                            Continuity.Standard if cont else Continuity.Discontinuous))
        size, dofs, colour, ncol = O.space(p, c, deg, cont)
        assert s.global_size == size
        assert np.array_equal(np.array([s.cell_dofs(i) for i in range(len(c))]), dofs)
        cols = s.cell_colouring()
        assert len(cols) == ncol
        for k, cells in enumerate(cols):
            assert np.all(colour[cells] == k)


def test_octahedron_matches_reference_shape():
    p, c = O.sphere(3)
    g = regular_sphere(3)
    assert np.abs(g.points - p).max() < 1e-15 and np.array_equal(g.cells, c)


def test_rules_in_orbit_form():
    """The P0 tiles take every rule point as c + t (V_k - c) (c the centroid, V_k a vertex):
    the centroid (t = 0), or an S21 orbit with repeated barycentric coordinate a, t = 1 - 3a,
    whose three points put the distinct coordinate on vertex 0, 1, 2 in that order, all with
    one weight. Checked on the engine's rules against t from the closed-form abscissae (6
    points: a = (8 - sqrt10 +- sqrt(38 - 44 sqrt(2/5)))/18; 7 points: (6 -+ sqrt15)/21)."""
    disc = np.sqrt(38.0 - 44.0 * np.sqrt(0.4))
    a6 = [(8.0 - np.sqrt(10.0) + disc) / 18.0, (8.0 - np.sqrt(10.0) - disc) / 18.0]
    a7 = [(6.0 - np.sqrt(15.0)) / 21.0, (6.0 + np.sqrt(15.0)) / 21.0]
    orbits = {1: [None], 3: [1.0 / 6.0], 6: a6, 7: [None] + a7}
    lib = L.lib()
    for q, orb in orbits.items():
        xi = np.zeros(2 * q)
        w = np.zeros(q)
        assert lib.bempp_triangle_rule(q, L.ptr(xi), L.ptr(w)) == q
        lam = np.column_stack([1.0 - xi[0::2] - xi[1::2], xi[0::2], xi[1::2]])
        p = 0
        for a in orb:
            if a is None:  # centroid
                assert np.abs(lam[p] - 1.0 / 3.0).max() < 1e-15
                p += 1
                continue
            t = 1.0 - 3.0 * a
            for k in range(3):
                want = np.full(3, (1.0 - t) / 3.0)
                want[k] = (1.0 + 2.0 * t) / 3.0
                assert np.abs(lam[p] - want).max() < 1e-15, (q, p)
                assert w[p] == w[p - k]
                p += 1
        assert p == q


def test_rules_match_oracle():
    """The engine's rule tables equal the oracle's (triangle rules and canonical Duffy rules)."""
    lib = L.lib()
    for q in (1, 3, 6, 7, 16, 37):
        xi = np.zeros(2 * q)
        w = np.zeros(q)
        assert lib.bempp_triangle_rule(q, L.ptr(xi), L.ptr(w)) == q
        po, wo = O.regular_rule(q)
        assert np.abs(xi.reshape(-1, 2) - po).max() < 5e-16 and np.abs(w - wo).max() < 5e-17
    assert lib.bempp_triangle_rule(5, None, None) == -1
    canonical = {1: [(0, 0)], 2: [(0, 0), (1, 1)], 3: [(0, 0), (1, 1), (2, 2)]}
    for m, pairs in canonical.items():
        n = lib.bempp_canonical_pair_rule(m, 4, None)
        r = np.zeros((n, 5))
        lib.bempp_canonical_pair_rule(m, 4, L.ptr(r))
        tp, sp, w = O.duffy_rule(pairs, 4)
        assert np.abs(r[:, 0:2] - tp).max() < 5e-16 and np.abs(r[:, 2:4] - sp).max() < 5e-16
        assert np.abs(r[:, 4] / w - 1).max() < 1e-14  # same formula, different product order


def test_bad_grids():
    with pytest.raises(ValueError):
        bempp_amd.grid.Grid(np.zeros((3, 3)), np.array([[0, 1, 5]]))
    with pytest.raises(ValueError):
        bempp_amd.grid.Grid(np.zeros((3, 3)), np.array([[0, 1, 1]]))


def test_no_gpu_fails_loudly():
    if bempp_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    g = regular_sphere(0)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    with pytest.raises(ValueError, match="no HIP device"):
        create_laplace_assembler(OperatorType.SingleLayer, o).assemble(s, s)


def test_noalloc_shape_gradient_rejects_oversize_grids():
    """The no-allocation shape gradient holds 176 B per cell in a 16 MiB static device buffer
    (at most 95325 cells): a larger grid is refused with an error before any device work
    (the device pointers are never touched), with or without a GPU."""
    g = icosphere(7)  # 327680 cells
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    a = create_laplace_assembler(OperatorType.SingleLayer, o)
    dummy = ctypes.c_void_p(0x1000)
    rc = L.lib().boundary_assembler_shape_gradient_noalloc(a._rs_assembler, s._rs_space, s._rs_space, dummy, dummy,
                                                            dummy, None)
    assert rc != 0
    assert "exceed its static scratch buffer" in L.last_error() and "95325" in L.last_error()


def test_batch_without_gpu_fails_cleanly():
    """The batched entry point reports an error (no crash, no CPU fallback) without a GPU."""
    if L.device_count() > 0:
        pytest.skip("a GPU is visible")
    g = icosphere(1)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    a = create_helmholtz_assembler(2.0, OperatorType.SingleLayer)
    ah = (ctypes.c_void_p * 1)(a._rs_assembler)
    sh = (ctypes.c_void_p * 1)(s._rs_space)
    oh = (ctypes.c_void_p * 1)(0x1000)  # never dereferenced: device selection fails first
    assert not L.lib().boundary_assembler_batch_new(ah, sh, oh, 1)
    assert L.last_error()
    assert not L.lib().boundary_assembler_batch_new(ah, sh, oh, 0)
    assert "at least one" in L.last_error()


def test_handles_held_by_module_globals_exit_cleanly():
    """Objects still alive at interpreter shutdown are not freed through a half-torn-down module."""
    import subprocess
    import sys

    pkg = Path(bempp_amd.__file__).resolve().parent.parent
    code = ("import torch\n"
            "from bempp_amd.assembly.boundary import BoundaryAssemblerOptions, OperatorType, "
            "create_helmholtz_assembler\n"
            "from bempp_amd.element import Continuity, Family, create_family\n"
            "from bempp_amd.function_space import function_space\n"
            "from bempp_amd.shapes import icosphere\n"
            "import bempp_amd._lib as L\n"
            "g = icosphere(1)\n"
            "s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))\n"
            "for k in [1.0, 2.0]:\n"
            "    o = BoundaryAssemblerOptions()\n"
            "    a = create_helmholtz_assembler(k, OperatorType.SingleLayer, o)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       env={**__import__("os").environ, "PYTHONPATH": str(pkg)})
    assert r.returncode == 0, r.stderr
    assert "Exception ignored" not in r.stderr, r.stderr


def _entity_dofs(lib, sp, dim, index):
    n = lib.space_get_local_dof_numbers_size(sp, dim, index)
    out = np.zeros(n, dtype=np.uintp)
    lib.space_get_local_dof_numbers(sp, dim, index, out.ctypes.data)
    return [int(x) for x in out]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ndgrid_grid_ingest(dtype):
    """single_element_grid_borrowed_create_bempp / create_lagrange_family_bempp /
    function_space (the entry points python/bempp/function_space.py binds): the grid equals the
    one built from plain arrays, the spaces number DOFs identically, and entity DOFs follow the
    topology's vertex numbering (here a random relabelling of the points)."""
    lib = _raw_lib()
    g0 = icosphere(2)
    pts, cells = g0.points, g0.cells
    perm = np.random.default_rng(5).permutation(len(pts))
    args, keep = _ndgrid_args(pts, cells, perm, dtype)
    gh = lib.single_element_grid_borrowed_create_bempp(*args)
    assert gh, lib.bempp_last_error()
    got = np.zeros_like(pts)
    lib.grid_points_copy(gh, got.ctypes.data)
    np.testing.assert_array_equal(got, np.asarray(pts, dtype=dtype).astype(np.float64))
    gc = np.zeros_like(cells)
    lib.grid_cells_copy(gh, gc.ctypes.data)
    np.testing.assert_array_equal(gc, cells)
    for degree, cont, nb in [(0, 1, 1), (1, 0, 3), (1, 1, 3)]:
        fam = lib.create_lagrange_family_bempp(degree, cont, 3)
        sp = lib.function_space(gh, fam)
        assert sp, lib.bempp_last_error()
        assert lib.space_dtype(sp) == 3
        ref = function_space(g0, create_family(Family.Lagrange, degree, Continuity(cont)))
        assert lib.space_global_size(sp) == ref.global_size
        for c in (0, 17, len(cells) - 1):
            d = np.zeros(nb, dtype=np.uintp)
            lib.space_cell_dofs(sp, c, d.ctypes.data)
            assert [int(x) for x in d] == ref.cell_dofs(c)
        if degree == 1 and cont == 0:  # vertex DOFs, addressed by topology vertex
            for p in (0, 5, len(pts) - 1):
                c, m = np.argwhere(cells == p)[0]
                assert _entity_dofs(lib, sp, 0, int(perm[p])) == [ref.cell_dofs(int(c))[m]]
            assert _entity_dofs(lib, sp, 2, 3) == []
        else:  # cell-interior DOFs
            assert _entity_dofs(lib, sp, 2, 7) == ref.cell_dofs(7)
            assert _entity_dofs(lib, sp, 0, 7) == []
        assert lib.space_is_owned(sp, 0) == 1
        loc = lib.space_local_space(sp)
        assert lib.space_global_size(loc) == ref.global_size
        lib.space_t_free(loc)
        lib.space_t_free(sp)
        lib.element_family_t_free_bempp(fam)
    lib.grid_t_free_bempp(gh)


def test_ndgrid_grid_ingest_errors():
    lib = _raw_lib()
    g0 = icosphere(1)
    pts, cells = g0.points, g0.cells
    ident = np.arange(len(pts))
    for kwargs, msg in [({"tdim": 3}, "topological dimension"), ({"cell_type": 3}, "triangle"),
                        ({}, None)]:
        args, keep = _ndgrid_args(pts, cells, ident, **kwargs)
        gh = lib.single_element_grid_borrowed_create_bempp(*args)
        if msg is None:
            assert gh
            lib.grid_t_free_bempp(gh)
        else:
            assert not gh and msg in lib.bempp_last_error().decode()
    # topology that is not a relabelling of the geometry (a cell's vertex order swapped)
    args, keep = _ndgrid_args(pts, cells, ident)
    tc = keep[2]
    tc[0, [0, 1]] = tc[0, [1, 0]]
    assert not lib.single_element_grid_borrowed_create_bempp(*args)
    assert "relabelling" in lib.bempp_last_error().decode()
    assert not lib.create_raviart_thomas_family_bempp(1, 0, 1)
    assert "not supported" in lib.bempp_last_error().decode()
