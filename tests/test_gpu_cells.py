"""GPU dense assembly on grids with quadrilateral cells (cells.hip) against the oracle's
general-cell assembly (tests/test_oracle_cells.py pins it): flat squares (cube), non-planar
bilinear cells (cubed sphere), mixed triangle/quadrilateral grids (a cube with some squares
split, the reference's mixed screen of tests/mixed_assembly.rs.bak:15-31); all four operators,
Laplace and Helmholtz, P0 / P1 (Q1 on quadrilaterals) / DP1; relative Frobenius error
<= 1e-12."""
import numpy as np
import pytest

import oracle_lib as O
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, Precision, create_helmholtz_assembler,
                                         create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.grid import Grid
from quad_meshes import cube, cubed_sphere, mixed_screen, split

pytestmark = pytest.mark.gpu

OPS = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer,
       "adjoint_double_layer": OperatorType.AdjointDoubleLayer, "hypersingular": OperatorType.Hypersingular}
SPACES = {"dp0": (0, False), "p1": (1, True), "dp1": (1, False)}


def _mesh(name):
    if name == "cube":
        pts, c = cube(2)
        return pts, c
    if name == "cubed_sphere":
        return cubed_sphere(3)
    if name == "mixed_cube":
        pts, c = cube(2)
        return pts, split(c, lambda i: i % 3 == 0)
    return mixed_screen()[:2]


def _options(nq_tri=6, nq_quad=16, nsing=4, mixed=False):
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, nq_tri)
    o.set_regular_quadrature_degree(ReferenceCellType.Quadrilateral, nq_quad)
    for t in [(ReferenceCellType.Triangle, ReferenceCellType.Triangle),
              (ReferenceCellType.Quadrilateral, ReferenceCellType.Quadrilateral),
              (ReferenceCellType.Triangle, ReferenceCellType.Quadrilateral),
              (ReferenceCellType.Quadrilateral, ReferenceCellType.Triangle)]:
        o.set_singular_quadrature_degree(*t, nsing)
    if mixed:
        o.set_precision(Precision.Mixed)
    return o


def _space(g, name):
    deg, cont = SPACES[name]
    return function_space(g, create_family(Family.Lagrange, deg, Continuity.Standard if cont else Continuity.Discontinuous))


def _assembler(kind, op, k, o, dtype=None):
    if kind == "laplace":
        return create_laplace_assembler(OPS[op], o, dtype=dtype or np.float64)
    return create_helmholtz_assembler(k, OPS[op], o, dtype=dtype or np.complex128)


@pytest.mark.parametrize("mesh", ["cube", "cubed_sphere", "mixed_cube", "mixed_screen"])
@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 2.0)])
@pytest.mark.parametrize("op", list(OPS))
def test_quadrilateral_grids_match_oracle(mesh, kind, k, op):
    pts, cells = _mesh(mesh)
    g = Grid(pts, cells)
    o = _options()
    for sp in ["dp0", "p1", "dp1"]:
        s = _space(g, sp)
        print(f"[cells] {mesh} {kind} {op} {sp} ({len(cells)} cells)", flush=True)
        A = _assembler(kind, op, k, o).assemble(s, s)
        R = O.assemble_cells(kind, op, pts, cells, test=SPACES[sp], trial=SPACES[sp], k=k, nq_tri=6, nq_quad=16,
                             nsing=4)
        nr = np.linalg.norm(R)
        if nr == 0.0:  # Laplace hypersingular on P0: identically zero
            assert np.abs(A).max() == 0.0
            continue
        err = np.linalg.norm(A - R) / nr
        print(f"[cells]   rel err {err:.2e}", flush=True)
        assert err <= 1e-12


@pytest.mark.parametrize("test,trial", [("dp0", "p1"), ("p1", "dp1")])
def test_quadrilateral_mixed_spaces_and_other_rules(test, trial):
    pts, cells = _mesh("mixed_cube")
    g = Grid(pts, cells)
    o = _options(nq_tri=7, nq_quad=9, nsing=5)
    for kind, op, k in [("helmholtz", "double_layer", 3.0), ("laplace", "hypersingular", 0.0)]:
        A = _assembler(kind, op, k, o).assemble(_space(g, trial), _space(g, test))
        R = O.assemble_cells(kind, op, pts, cells, test=SPACES[test], trial=SPACES[trial], k=k, nq_tri=7, nq_quad=9,
                             nsing=5)
        if np.linalg.norm(R) == 0.0:
            assert np.abs(A).max() == 0.0
            continue
        assert np.linalg.norm(A - R) / np.linalg.norm(R) <= 1e-12


def test_quadrilateral_device_output_dtypes_and_modes():
    """assemble_device overwrites, assemble_into_device_memory accumulates; float32 /
    complex64 outputs; the mixed-precision option runs FP64 on these grids."""
    import torch

    pts, cells = _mesh("cubed_sphere")
    g = Grid(pts, cells)
    s = _space(g, "p1")
    R = O.assemble_cells("helmholtz", "single_layer", pts, cells, test=(1, True), trial=(1, True), k=2.0)
    a = _assembler("helmholtz", "single_layer", 2.0, _options())
    D = a.assemble_device(s, s)
    torch.cuda.synchronize()
    assert np.linalg.norm(D.cpu().numpy() - R) / np.linalg.norm(R) <= 1e-12
    a.assemble_into_device(s, s, D, accumulate=True)
    torch.cuda.synchronize()
    assert np.linalg.norm(D.cpu().numpy() - 2 * R) / np.linalg.norm(R) <= 1e-12
    a32 = _assembler("helmholtz", "single_layer", 2.0, _options(), dtype=np.complex64)
    A32 = a32.assemble(s, s)
    assert A32.dtype == np.complex64 and np.linalg.norm(A32 - R) / np.linalg.norm(R) <= 1e-6
    am = _assembler("helmholtz", "single_layer", 2.0, _options(mixed=True))
    assert np.linalg.norm(am.assemble(s, s) - R) / np.linalg.norm(R) <= 1e-12
    L = O.assemble_cells("laplace", "double_layer", pts, cells)
    l32 = create_laplace_assembler(OperatorType.DoubleLayer, _options(), dtype=np.float32).assemble(
        _space(g, "dp0"), _space(g, "dp0"))
    assert l32.dtype == np.float32 and np.linalg.norm(l32 - L) / np.linalg.norm(L) <= 1e-6


@pytest.mark.parametrize("k,tier", [(1.2, 1), (0.6, 2), (-1.2, 1), (-0.6, 2)])
@pytest.mark.parametrize("op", list(OPS))
def test_quadrilateral_centred_phase_tiers(k, tier, op):
    """Small |k| x cell size on a 384-cell cubed sphere: the regular kernel takes the phase
    relative to each cell pair's centroid distance (tier 1: |k| (rad_test + rad_trial) <= 0.5,
    tier 2: <= 0.25, rad the largest centroid-to-vertex distance)."""
    pts, cells = cubed_sphere(8)
    rad = max(np.linalg.norm(pts[c] - pts[c].mean(axis=0), axis=1).max() for c in cells)
    bound = abs(k) * 2 * rad
    assert (bound <= 0.25) == (tier == 2) and bound <= 0.5
    g = Grid(pts, cells)
    for sp in ["dp0", "p1"]:
        s = _space(g, sp)
        A = _assembler("helmholtz", op, k, _options(nq_quad=9)).assemble(s, s)
        R = O.assemble_cells("helmholtz", op, pts, cells, test=SPACES[sp], trial=SPACES[sp], k=k, nq_quad=9)
        err = np.linalg.norm(A - R) / np.linalg.norm(R)
        print(f"[cells] tier {tier} k={k} {op} {sp}: rel err {err:.2e}", flush=True)
        assert err <= 1e-12


def test_quadrilateral_rule_must_be_square():
    pts, cells = _mesh("cube")
    g = Grid(pts, cells)
    s = _space(g, "dp0")
    a = _assembler("laplace", "single_layer", 0.0, BoundaryAssemblerOptions())  # default: 37 points
    with pytest.raises(ValueError, match="square number"):
        a.assemble(s, s)


def test_larger_cubed_sphere():
    """864 non-planar cells (12 x 12 per face): several trial chunks and thousands of
    adjacent pairs, Helmholtz double layer on P1."""
    pts, cells = cubed_sphere(12)
    g = Grid(pts, cells)
    s = _space(g, "p1")
    A = _assembler("helmholtz", "double_layer", 4.0, _options(nq_quad=9)).assemble(s, s)
    R = O.assemble_cells("helmholtz", "double_layer", pts, cells, test=(1, True), trial=(1, True), k=4.0, nq_quad=9)
    err = np.linalg.norm(A - R) / np.linalg.norm(R)
    print(f"[cells] cubed sphere 12: rel err {err:.2e}", flush=True)
    assert err <= 1e-12


def _adjacent_pattern(pts, cells, test, trial):
    """Row-major keys row * ncols + col of the DOF pairs of adjacent cells (sorted, unique)."""
    nt, tdofs = O.cells_space(pts, cells, *SPACES[test])
    ns, sdofs = O.cells_space(pts, cells, *SPACES[trial])
    keys = set()
    for a in range(len(cells)):
        for b in range(len(cells)):
            if set(cells[a]) & set(cells[b]):
                keys.update(r * ns + c for r in tdofs[a] for c in sdofs[b])
    return np.array(sorted(keys), dtype=np.int64), nt, ns


@pytest.mark.parametrize("mesh", ["cubed_sphere", "mixed_cube", "mixed_screen"])
@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 2.0)])
@pytest.mark.parametrize("op", list(OPS))
def test_quadrilateral_singular_csr(mesh, kind, k, op):
    """assemble_singular and assemble_singular_correction (src/boundary_assemblers.rs:122-169)
    on grids with quadrilaterals: CSR pattern = the DOF pairs of adjacent cells (a triangle's
    missing fourth function leaves no entry), columns strictly ascending per row, the same
    bytes on a second call, values against the oracle's adjacent-pair parts (pair rules /
    regular rule without coincident points)."""
    pts, cells = _mesh(mesh)
    g = Grid(pts, cells)
    o = _options()
    a = _assembler(kind, op, k, o)
    for test, trial in [("dp0", "dp0"), ("p1", "p1"), ("dp1", "p1"), ("dp0", "dp1")]:
        st, ss = _space(g, test), _space(g, trial)
        keys, nt, ns = _adjacent_pattern(pts, cells, test, trial)
        for part, fn in [("singular", a.assemble_singular), ("correction", a.assemble_singular_correction)]:
            print(f"[cells-csr] {mesh} {kind} {op} {test}x{trial} {part}", flush=True)
            csr = fn(ss, st)
            again = fn(ss, st)
            assert np.array_equal(csr.indptr, again.indptr) and np.array_equal(csr.indices, again.indices)
            assert np.array_equal(csr.data.view(np.uint8), again.data.view(np.uint8))
            R = O.assemble_cells(kind, op, pts, cells, test=SPACES[test], trial=SPACES[trial], k=k, nq_tri=6,
                                 nq_quad=16, nsing=4, part=part)
            assert csr.shape == (nt, ns) and csr.indptr[0] == 0 and csr.indptr[-1] == csr.nnz
            if np.abs(R).max() <= 1e-10:  # every entry at most 1e-10: the empty matrix
                assert csr.nnz == 0
                continue
            row_of = np.repeat(np.arange(nt), np.diff(csr.indptr))
            assert np.all(np.diff(csr.indices)[np.diff(row_of) == 0] > 0)
            assert np.array_equal(row_of * ns + csr.indices, keys)
            err = np.linalg.norm(csr.toarray() - R) / np.linalg.norm(R)
            print(f"[cells-csr]   nnz {csr.nnz}, rel err {err:.2e}", flush=True)
            assert err <= 1e-12


def test_cached_device_data_follows_rules_and_points():
    """The general-cell device data (point records, tables, adjacent pairs, DOFs) is cached per
    grid and space: on one grid object, two regular rules each match the oracle, and after
    set_points the assembly follows the moved points."""
    pts, cells = cubed_sphere(4)
    g = Grid(pts, cells)
    s = _space(g, "p1")
    for nq in (16, 9, 16):
        A = _assembler("helmholtz", "double_layer", 2.0, _options(nq_quad=nq)).assemble(s, s)
        R = O.assemble_cells("helmholtz", "double_layer", pts, cells, test=(1, True), trial=(1, True), k=2.0,
                             nq_quad=nq)
        assert np.linalg.norm(A - R) / np.linalg.norm(R) <= 1e-12, nq
    moved = pts * np.array([1.0, 1.2, 0.9])
    g.set_points(moved)
    a = _assembler("helmholtz", "double_layer", 2.0, _options())
    A = a.assemble(s, s)
    R = O.assemble_cells("helmholtz", "double_layer", moved, cells, test=(1, True), trial=(1, True), k=2.0)
    err = np.linalg.norm(A - R) / np.linalg.norm(R)
    print(f"[cells] after set_points: rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    S = a.assemble_singular(s, s).toarray()
    RS = O.assemble_cells("helmholtz", "double_layer", moved, cells, test=(1, True), trial=(1, True), k=2.0,
                          part="singular")
    assert np.linalg.norm(S - RS) / np.linalg.norm(RS) <= 1e-12


@pytest.mark.parametrize("sp", ["dp0", "dp1"])
@pytest.mark.parametrize("op", ["single_layer", "double_layer"])
def test_quadrilateral_device_overwrite_and_accumulate(sp, op):
    """Spaces without DOFs shared between cells: every regular entry has one writer, stored
    without atomics — overwriting a device buffer that held NaN, then accumulating onto it;
    symmetric (single layer) and non-symmetric (double layer) assembly."""
    import torch

    pts, cells = _mesh("mixed_cube")
    g = Grid(pts, cells)
    s = _space(g, sp)
    a = _assembler("helmholtz", op, 2.0, _options())
    R = O.assemble_cells("helmholtz", op, pts, cells, test=SPACES[sp], trial=SPACES[sp], k=2.0)
    n = s.global_size
    D = torch.full((n, n), float("nan"), dtype=torch.complex128, device="cuda").t()
    a.assemble_into_device(s, s, D)
    torch.cuda.synchronize()
    assert np.linalg.norm(D.cpu().numpy() - R) / np.linalg.norm(R) <= 1e-12
    a.assemble_into_device(s, s, D, accumulate=True)
    torch.cuda.synchronize()
    assert np.linalg.norm(D.cpu().numpy() - 2 * R) / np.linalg.norm(R) <= 1e-12


def test_quadrilateral_bench_size_rows():
    """The measured quadrilateral workload (9 600-cell cubed sphere, 40 x 40 per face; Helmholtz
    SL P0, k = 5, 4 x 4 Gauss points, 4-point Duffy rules) on the device: the first, a middle and
    the last 64 rows against the oracle (the last rows hold mostly transposed stores of the
    symmetric evaluation)."""
    import torch

    pts, cells = cubed_sphere(40)
    g = Grid(pts, np.array(cells))
    s = _space(g, "dp0")
    a = _assembler("helmholtz", "single_layer", 5.0, _options(nq_quad=16, nsing=4))
    D = a.assemble_device(s, s)
    torch.cuda.synchronize()
    n = len(cells)
    for r0 in (0, n // 2, n - 64):
        got = D[r0:r0 + 64].cpu().numpy()
        ref = O.assemble_cells("helmholtz", "single_layer", pts, cells, k=5.0, nq_quad=16, nsing=4,
                               rows=(r0, r0 + 64))
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"[cells] 9600 cells, rows {r0}..{r0 + 63}: rel err {err:.2e}", flush=True)
        assert err <= 1e-12
    del D
    torch.cuda.empty_cache()


def test_batch_refuses_quadrilateral_grids():
    """Batches are specialised to triangles: an item on a grid with quadrilateral cells is
    refused when the batch is built, before any launch; a triangle item beside it is not
    what fails."""
    import torch

    from bempp_amd.assembly.batch import BatchAssembler
    from bempp_amd.shapes import icosphere

    pts, quads = cube(1)
    sq = _space(Grid(pts, np.array(quads)), "dp0")
    st = _space(icosphere(1), "dp0")
    a = _assembler("laplace", "single_layer", 0.0, _options())
    with pytest.raises(ValueError, match="quadrilateral"):
        BatchAssembler([a, a], [st, sq])
    b = BatchAssembler([a], [st])  # the triangle item alone builds and assembles
    b.assemble()
    torch.cuda.synchronize()
    assert np.isfinite(b.outputs[0].cpu().numpy()).all()
