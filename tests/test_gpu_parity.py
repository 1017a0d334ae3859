"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the same
inputs, relative Frobenius error <= 1e-12 in FP64."""
import numpy as np
import pytest

import oracle_lib as O
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler,
                                         create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.grid import Grid
from bempp_amd.shapes import icosphere, regular_sphere

pytestmark = pytest.mark.gpu

TOL = 1e-12
OPS = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer,
       "adjoint_double_layer": OperatorType.AdjointDoubleLayer, "hypersingular": OperatorType.Hypersingular}
SPACES = {"dp0": (0, False), "p1": (1, True), "dp1": (1, False)}


def _space(g, name):
    deg, cont = SPACES[name]
    return function_space(g, create_family(Family.Lagrange, deg, Continuity.Standard if cont else Continuity.Discontinuous))


def _options(nq=6, ns=4):
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, nq)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, ns)
    return o


def _assembler(kind, op, k, o):
    if kind == "laplace":
        return create_laplace_assembler(OPS[op], o)
    return create_helmholtz_assembler(k, OPS[op], o)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _check(kind, op, grid, test, trial, k=0.0, nq=6, ns=4, tol=TOL):
    p, c = grid.points, grid.cells
    a = _assembler(kind, op, k, _options(nq, ns))
    ts, ss = _space(grid, test), _space(grid, trial)
    print(f"[parity] {kind} {op} test={test} trial={trial} ncells={len(c)} k={k} nq={nq}", flush=True)
    A = a.assemble(ss, ts)
    R = O.assemble(kind, op, p, c, test=SPACES[test], trial=SPACES[trial], k=k, nq=nq, nsing=ns)
    if np.linalg.norm(R) == 0.0:
        # Laplace hypersingular with a P0 factor: the curl of a constant vanishes, A == 0
        print("[parity]   operator is identically zero", flush=True)
        assert np.abs(A).max() == 0.0
        return A, R
    err = _rel(A, R)
    print(f"[parity]   rel err {err:.3e}", flush=True)
    assert err <= tol, err
    return A, R


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 5.0)])
@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("space", ["dp0", "p1"])
def test_operators_icosphere_l2(kind, k, op, space):
    _check(kind, op, icosphere(2), space, space, k=k)


def test_laplace_sl_p0_icosphere_l3():
    """Laplace single-layer P0 on the level-3 icosphere (1280 triangles), 6-point / 4-point rules."""
    _check("laplace", "single_layer", icosphere(3), "dp0", "dp0")


def test_helmholtz_sl_p0_icosphere_l3():
    _check("helmholtz", "single_layer", icosphere(3), "dp0", "dp0", k=5.0)


@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("space", ["dp0", "p1"])
def test_helmholtz_centred_phase(op, space):
    """k = 2.5 on the level-3 icosphere: |k| (rad_i + rad_j) <= 0.43, so the regular kernels
    take the phase relative to each cell pair's centroid distance (small-argument sin/cos)."""
    _check("helmholtz", op, icosphere(3), space, space, k=2.5)


@pytest.mark.parametrize("op,space", [("single_layer", "dp0"), ("double_layer", "p1"), ("hypersingular", "p1")])
def test_helmholtz_small_phase_adjacent(op, space):
    """k = 2.5 on the level-4 icosphere (5120 cells): |k| |x - y| <= 0.43 on every adjacent
    pair, so the Duffy kernel also uses the small-argument sin/cos."""
    _check("helmholtz", op, icosphere(4), space, space, k=2.5)


@pytest.mark.parametrize("test,trial", [("dp0", "p1"), ("p1", "dp0"), ("dp1", "dp1"), ("dp1", "p1")])
@pytest.mark.parametrize("op", ["single_layer", "double_layer", "hypersingular"])
def test_mixed_spaces(test, trial, op):
    _check("helmholtz", op, regular_sphere(1), test, trial, k=3.0)


@pytest.mark.parametrize("nq,ns", [(1, 2), (3, 3), (7, 5)])
def test_other_rules(nq, ns):
    _check("helmholtz", "double_layer", regular_sphere(1), "p1", "p1", k=2.0, nq=nq, ns=ns)


def test_goldens_on_gpu():
    """The GPU reproduces the bempp-cl golden matrices (tests/compare_to_bempp_cl.rs)."""
    import json
    from pathlib import Path

    cases = json.loads((Path(__file__).parent / "golden" / "bempp_cl_octahedron_l0.json").read_text())
    g = regular_sphere(0)
    for m in cases:
        ref = np.array(m["from_cl"])
        if ref.ndim == 3:
            ref = ref[..., 0] + 1j * ref[..., 1]
        perm = m["perm"]
        sp = "p1" if m["degree"] == 1 else "dp0"
        a = _assembler(m["kind"], m["operator"], m["k"], _options())
        s = _space(g, sp)
        A = a.assemble(s, s)
        assert np.abs(A - ref[np.ix_(perm, perm)]).max() <= m["epsilon"], m["name"]


def test_accumulates_into_memory():
    """assemble_into_memory adds into the caller's buffer (src/boundary_assemblers.rs:176-214)."""
    g = icosphere(1)
    s = _space(g, "p1")
    a = _assembler("helmholtz", "single_layer", 2.0, _options())
    A = a.assemble(s, s)
    B = np.full_like(A, 1.0 + 2.0j, order="F")
    a.assemble_into_memory(s, s, B)
    assert np.abs(B - (A + 1.0 + 2.0j)).max() < 1e-13  # atomics: summation order varies


def test_singular_csr_matches_oracle():
    """assemble_singular: CSR of the adjacent-pair part equals the oracle's COO summed."""
    g = icosphere(1)
    p, c = g.points, g.cells
    for kind, op, sp, k in [("laplace", "single_layer", "dp0", 0.0), ("helmholtz", "double_layer", "p1", 2.0),
                            ("helmholtz", "hypersingular", "p1", 2.0)]:
        a = _assembler(kind, op, k, _options())
        s = _space(g, sp)
        csr = a.assemble_singular(s, s)
        rows, cols, vals = O.assemble_singular_coo(kind, op, p, c, test=SPACES[sp], trial=SPACES[sp], k=k)
        ref = np.zeros(csr.shape, dtype=vals.dtype)
        np.add.at(ref, (rows, cols), vals)
        assert _rel(csr.toarray(), ref) < TOL
        assert np.all(np.diff(csr.indptr) >= 0) and csr.has_sorted_indices


@pytest.mark.parametrize("level,kind,op,sp,k,dt,tol", [
    (1, "laplace", "single_layer", "dp0", 0.0, np.float32, 2e-6),
    (1, "helmholtz", "hypersingular", "p1", 2.0, np.complex64, 2e-6),
    (2, "laplace", "double_layer", "p1", 0.0, np.float32, 2e-6),
    (3, "helmholtz", "adjoint_double_layer", "p1", 1.5, np.complex128, TOL),
    (3, "laplace", "hypersingular", "dp1", 0.0, np.float64, TOL),
])
def test_singular_csr_structure(level, kind, op, sp, k, dt, tol):
    """The device-built CSR (src/boundary_assemblers.rs:122-155): its pattern is exactly the
    distinct (row, column) pairs of the oracle's COO (duplicates summed, none dropped), columns
    strictly ascending within each row, row pointers from 0 to nnz, the same bytes on a second
    call, values in the assembler dtype equal to the oracle's COO summed."""
    g = icosphere(level)
    p, c = g.points, g.cells
    o = _options()
    a = (create_laplace_assembler(OPS[op], o, dtype=dt) if kind == "laplace"
         else create_helmholtz_assembler(k, OPS[op], o, dtype=dt))
    s = _space(g, sp)
    print(f"[csr] level {level} {kind} {op} {sp} {np.dtype(dt).name}", flush=True)
    csr = a.assemble_singular(s, s)
    again = a.assemble_singular(s, s)
    assert csr.dtype == np.dtype(dt)
    assert np.array_equal(csr.indptr, again.indptr) and np.array_equal(csr.indices, again.indices)
    assert np.array_equal(csr.data.view(np.uint8), again.data.view(np.uint8))
    n = s.global_size
    assert csr.shape == (n, n) and csr.indptr[0] == 0 and csr.indptr[-1] == csr.nnz
    assert np.all(np.diff(csr.indptr) >= 0)
    row_of = np.repeat(np.arange(n), np.diff(csr.indptr))
    same_row = np.diff(row_of) == 0
    assert np.all(np.diff(csr.indices)[same_row] > 0)
    rows, cols, vals = O.assemble_singular_coo(kind, op, p, c, test=SPACES[sp], trial=SPACES[sp], k=k)
    assert len(rows) > csr.nnz or sp != "p1"  # continuous spaces: duplicates were summed
    assert np.array_equal(np.unique(rows * n + cols), row_of * n + csr.indices)
    ref = np.zeros((n, n), dtype=vals.dtype)
    np.add.at(ref, (rows, cols), vals)
    err = _rel(csr.toarray().astype(ref.dtype), ref)
    print(f"[csr]   nnz {csr.nnz} of {len(rows)} COO entries, rel err {err:.2e}", flush=True)
    assert err < tol


def test_separate_grids_have_no_singular_part():
    """Different grids: no neighbour skipping and no Duffy pairs (src/boundary_assemblers.rs:240-243,794-821)."""
    g1, g2 = icosphere(1), icosphere(1)
    pts = g2.points * 0.5 + np.array([3.0, 0.0, 0.0])
    g2.set_points(pts)
    a = _assembler("laplace", "single_layer", 0.0, _options())
    A = a.assemble(_space(g2, "dp0"), _space(g1, "dp0"))
    xi = O.regular_rule(6)
    # brute force all pairs with the regular rule
    def mapped(p, c):
        v = p[c]
        jd = np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
        x = v[:, None, 0] + xi[0][None, :, 0, None] * (v[:, None, 1] - v[:, None, 0]) + \
            xi[0][None, :, 1, None] * (v[:, None, 2] - v[:, None, 0])
        return x, xi[1][None] * jd[:, None]
    x, wx = mapped(g1.points, g1.cells)
    y, wy = mapped(pts, g2.cells)
    r = np.linalg.norm(x[:, :, None, None] - y[None, None], axis=-1)
    ref = np.einsum("ipjq,ip,jq->ij", 1 / (4 * np.pi * r), wx, wy)
    assert _rel(A, ref) < TOL
    assert a.assemble_singular(_space(g2, "dp0"), _space(g1, "dp0")).nnz == 0


def test_device_output_torch():
    """Device-resident assembly into a torch tensor equals the host path; accumulate adds."""
    import torch

    g = icosphere(2)
    s = _space(g, "dp0")
    a = _assembler("helmholtz", "single_layer", 5.0, _options())
    A = a.assemble(s, s)
    D = a.assemble_device(s, s)
    torch.cuda.synchronize()
    assert np.abs(D.cpu().numpy() - A).max() < 1e-16
    a.assemble_into_device(s, s, D, accumulate=True)
    torch.cuda.synchronize()
    assert np.abs(D.cpu().numpy() - 2 * A).max() < 1e-15


def test_deterministic_p0():
    """P0 x P0 writes each entry from exactly one pair with plain stores: bitwise repeatable."""
    g = icosphere(2)
    s = _space(g, "dp0")
    a = _assembler("helmholtz", "double_layer", 5.0, _options())
    assert np.array_equal(a.assemble(s, s), a.assemble(s, s))


def _mixed_options(nq=6, ns=4):
    from bempp_amd.assembly.boundary import Precision

    o = _options(nq, ns)
    o.set_precision(Precision.Mixed)
    return o


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 2.5), ("helmholtz", 5.0), ("helmholtz", -2.5),
                                    ("helmholtz", 0.0)])
@pytest.mark.parametrize("op", list(OPS))
def test_mixed_precision_p0(kind, k, op):
    """Mixed precision (FP32 far pairs, FP64 near and adjacent pairs) on the level-4 icosphere
    (5120 cells; |k| (rad_i + rad_j) <= 0.25 for k = 2.5, <= 0.5 for k = 5: both short phase
    series; negative k; Helmholtz with k = 0, which the tiles evaluate with the Laplace
    kernels): relative Frobenius error <= 1e-6
    against the FP64 oracle, and no entry dropped or counted twice (largest entrywise error
    relative to the largest entry <= 1e-5)."""
    g = icosphere(4)
    s = _space(g, "dp0")
    a = _assembler(kind, op, k, _mixed_options())
    print(f"[mixed] {kind} {op} k={k}", flush=True)
    A = a.assemble(s, s)
    R = O.assemble(kind, op, g.points, g.cells, k=k)
    if np.linalg.norm(R) == 0.0:
        assert np.abs(A).max() == 0.0
        return
    err = _rel(A, R)
    emax = np.abs(A - R).max() / np.abs(R).max()
    print(f"[mixed]   rel Frobenius err {err:.3e}, max entry err {emax:.3e}", flush=True)
    assert err <= 1e-6 and emax <= 1e-5
    assert err > 1e-12  # the FP32 path ran


@pytest.mark.parametrize("op", list(OPS))
def test_mixed_precision_wide_phase(op):
    """Coarse stretched mesh with a large wavenumber (level-3 icosphere scaled to semi-axes
    0.7, 1.3, 1.8; k = 8, so |k| (rad_i + rad_j) > 0.5 and the FP32 phase k (r - R) needs
    its own range reduction): relative Frobenius error <= 1e-6 against the FP64 oracle."""
    g0 = icosphere(3)
    pts = g0.points * np.array([0.7, 1.3, 1.8])
    g = Grid(pts, g0.cells)
    s = _space(g, "dp0")
    a = _assembler("helmholtz", op, 8.0, _mixed_options())
    print(f"[mixed-wide] {op}", flush=True)
    A = a.assemble(s, s)
    R = O.assemble("helmholtz", op, pts, g0.cells, k=8.0)
    err = _rel(A, R)
    emax = np.abs(A - R).max() / np.abs(R).max()
    print(f"[mixed-wide]   rel Frobenius err {err:.3e}, max entry err {emax:.3e}", flush=True)
    assert err <= 1e-6 and emax <= 1e-5
    assert err > 1e-12


@pytest.mark.parametrize("nq", [1, 3, 7])
@pytest.mark.parametrize("op", list(OPS))
def test_mixed_precision_rules(nq, op):
    """Mixed-precision P0 tiles with the 1-, 3- and 7-point regular rules, whose trial points
    the tiles take in orbit form c + t (V_k - c) (centroid only; one orbit; centroid and two
    orbits): Helmholtz k = 2.5 on the level-3 icosphere, <= 1e-6 against the FP64 oracle
    with the same rule."""
    g = icosphere(3)
    s = _space(g, "dp0")
    a = _assembler("helmholtz", op, 2.5, _mixed_options(nq))
    print(f"[mixed-rule] {op} nq={nq}", flush=True)
    A = a.assemble(s, s)
    R = O.assemble("helmholtz", op, g.points, g.cells, k=2.5, nq=nq)
    err = _rel(A, R)
    emax = np.abs(A - R).max() / np.abs(R).max()
    print(f"[mixed-rule]   rel Frobenius err {err:.3e}, max entry err {emax:.3e}", flush=True)
    assert 1e-13 < err <= 1e-6 and emax <= 1e-5


@pytest.mark.parametrize("op", ["single_layer", "double_layer"])
def test_mixed_precision_complex64_output(op):
    """The complex64-mixed form: FP32 far field, FP64 near/adjacent pairs and accumulation,
    complex64 storage; k = 2.5 is exact in FP32. Level-4 icosphere, <= 1e-6 against the
    FP64 oracle."""
    g = icosphere(4)
    s = _space(g, "dp0")
    a = create_helmholtz_assembler(2.5, OPS[op], _mixed_options(), dtype=np.complex64)
    A = a.assemble(s, s)
    assert A.dtype == np.complex64
    R = O.assemble("helmholtz", op, g.points, g.cells, k=2.5)
    err = _rel(A.astype(np.complex128), R)
    print(f"[mixed-c64] {op}: rel Frobenius err {err:.3e}", flush=True)
    assert 1e-12 < err <= 1e-6


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 2.5), ("helmholtz", 8.0)])
@pytest.mark.parametrize("op", list(OPS))
def test_mixed_precision_p1(kind, k, op):
    """Mixed precision with P1 spaces (FP32 far field in the LDS-accumulating tiles, FP64 near
    pairs with atomics, FP64 Duffy pairs), level-3 icosphere: <= 1e-6 against the FP64 oracle.
    k = 8 exercises the range-reduced FP32 phase."""
    g = icosphere(3)
    s = _space(g, "p1")
    a = _assembler(kind, op, k, _mixed_options())
    A = a.assemble(s, s)
    R = O.assemble(kind, op, g.points, g.cells, test=(1, True), trial=(1, True), k=k)
    err = _rel(A, R)
    print(f"[mixed-p1] {kind} {op} k={k}: rel Frobenius err {err:.3e}", flush=True)
    if kind == "laplace" and op == "hypersingular":
        assert err <= 1e-12  # assembled by the FP64 orbit-form tiles, faster than FP32 here
        return
    assert err <= 1e-6
    assert err > 1e-13  # the FP32 path ran


@pytest.mark.parametrize("test,trial", [("dp0", "p1"), ("p1", "dp0")])
def test_mixed_precision_mixed_spaces(test, trial):
    g = icosphere(3)
    a = _assembler("helmholtz", "double_layer", 2.5, _mixed_options())
    A = a.assemble(_space(g, trial), _space(g, test))
    R = O.assemble("helmholtz", "double_layer", g.points, g.cells, test=SPACES[test], trial=SPACES[trial], k=2.5)
    err = _rel(A, R)
    print(f"[mixed-p0p1] {test} x {trial}: rel Frobenius err {err:.3e}", flush=True)
    assert 1e-13 < err <= 1e-6


@pytest.mark.parametrize("op", ["single_layer", "double_layer"])
def test_mixed_precision_shuffled_cells(op):
    """Mixed precision on a level-4 icosphere whose cells are randomly renumbered: 64-cell
    blocks are then scattered over the sphere, so the per-tile all-far bound flags few tiles
    and most pairs take the per-pair adjacency and near tests; the sorted order flags most
    tiles. Both must give the same matrix (up to the renumbering) within the mixed tolerance,
    and match the FP64 oracle."""
    g0 = icosphere(4)
    perm = np.random.default_rng(11).permutation(len(g0.cells))
    cells = np.asarray(g0.cells)[perm]
    g = Grid(g0.points, cells)
    s, s0 = _space(g, "dp0"), _space(g0, "dp0")
    a = _assembler("helmholtz", op, 2.5, _mixed_options())
    A = a.assemble(s, s)
    A0 = a.assemble(s0, s0)[np.ix_(perm, perm)]
    R = O.assemble("helmholtz", op, g.points, cells, k=2.5)
    err, err0 = _rel(A, R), _rel(A0, R)
    print(f"[mixed-shuffled] {op}: shuffled {err:.3e}, sorted {err0:.3e}", flush=True)
    assert 1e-12 < err <= 1e-6 and err0 <= 1e-6
    assert np.abs(A - R).max() / np.abs(R).max() <= 1e-5


@pytest.mark.parametrize("nq", [1, 3, 7])
@pytest.mark.parametrize("kind,k,level", [("laplace", 0.0, 3), ("helmholtz", 1.2, 3), ("helmholtz", 2.5, 3),
                                          ("helmholtz", 5.0, 2)],
                         ids=["laplace", "helmholtz_tier2", "helmholtz_tier1", "helmholtz_full_reduction"])
@pytest.mark.parametrize("op", list(OPS))
def test_p0_tiles_other_rules(nq, kind, k, level, op):
    """The FP64 P0 x P0 tiles take both rules' points in orbit form c + t (V_k - c): with one
    point (the centroid alone, t = 0), three points (one orbit) and seven points (the centroid
    and two orbits, three weight classes), for all four operators, Laplace and Helmholtz with
    each phase mode (k = 1.2 and 2.5 on level 3: centred phase tiers 2 and 1; k = 5 on level 2:
    full reduction per point pair), against the oracle with the same rule."""
    _check(kind, op, icosphere(level), "dp0", "dp0", k=k, nq=nq)


@pytest.mark.parametrize("space", ["dp0", "p1"])
@pytest.mark.parametrize("kind,op,k", [("laplace", "single_layer", 0.0), ("laplace", "double_layer", 0.0),
                                       ("helmholtz", "single_layer", 2.0), ("helmholtz", "adjoint_double_layer", 2.0),
                                       ("helmholtz", "hypersingular", 2.0)])
def test_fp64_path_single_precision_outputs(kind, op, k, space):
    """The reference's factories are generic over f32 (Laplace) and c32 (Helmholtz): the FP64
    path computes in FP64 and stores float32 / complex64. P0 x P0 entries are stored once
    (the tile's value, or the Duffy value added to the tile's zero), so they equal the FP64
    matrix rounded to single precision exactly; P1 entries are summed with single-precision
    atomics (<= 1e-6). assemble_into_memory adds into the caller's single-precision buffer."""
    g = icosphere(2)
    s = _space(g, space)
    o = _options()
    if kind == "laplace":
        a32, a64 = create_laplace_assembler(OPS[op], o, dtype=np.float32), create_laplace_assembler(OPS[op], o)
    else:
        a32 = create_helmholtz_assembler(k, OPS[op], o, dtype=np.complex64)
        a64 = create_helmholtz_assembler(k, OPS[op], o)
    A32 = a32.assemble(s, s)
    A64 = a64.assemble(s, s)
    assert A32.dtype == (np.float32 if kind == "laplace" else np.complex64)
    R = O.assemble(kind, op, g.points, g.cells, test=SPACES[space], trial=SPACES[space], k=k)
    assert _rel(A64, R) <= TOL
    err = _rel(A32.astype(A64.dtype), R)
    print(f"[fp64-f32out] {kind} {op} {space}: rel err vs oracle {err:.3e}", flush=True)
    assert err <= 1e-6
    if space == "dp0":
        assert np.array_equal(A32, A64.astype(A32.dtype))
    B = np.full_like(A32, 0.5 if kind == "laplace" else 0.5 - 0.25j, order="F")
    B0 = B.copy()
    a32.assemble_into_memory(s, s, B)
    if space == "dp0":
        assert np.array_equal(B, B0 + A32)  # one single-precision add per entry
    else:
        assert np.abs(B - (B0 + A32)).max() <= 1e-6 * np.abs(A32).max() + 1e-6


@pytest.mark.parametrize("op", ["single_layer", "double_layer", "hypersingular"])
def test_p1_basis_placement_vertex_relabelling(op):
    """The GPU maps each adjacent pair onto canonical Sauter-Schwab rules by permuting the
    cells' vertices and carries the P1 basis along: rotating every cell's vertex list (orientation
    kept) exercises other permutation codes but leaves the operator unchanged up to the Duffy
    quadrature error (10 points per dimension: below 1e-10); both match the oracle."""
    g0 = icosphere(1)
    shift = np.random.default_rng(3).integers(0, 3, len(g0.cells))
    c2 = np.array([np.roll(row, -s) for row, s in zip(g0.cells, shift)])
    out = []
    for g in (g0, Grid(g0.points, c2)):
        s = _space(g, "p1")
        A = _assembler("helmholtz", op, 3.0, _options(6, 10)).assemble(s, s)
        R = O.assemble("helmholtz", op, g.points, g.cells, test=(1, True), trial=(1, True), k=3.0, nsing=10)
        assert _rel(A, R) <= TOL
        vdof = np.zeros(len(g.points), dtype=np.int64)
        for c in range(len(g.cells)):
            vdof[g.cells[c]] = s.cell_dofs(c)
        out.append(A[np.ix_(vdof, vdof)])
    diff = np.abs(out[0] - out[1]).max() / np.abs(out[0]).max()
    print(f"[placement] {op}: relabelled vs original {diff:.2e}", flush=True)
    assert diff < 1e-10


def test_oversized_host_assembly_fails_cleanly():
    """A dense matrix larger than the free device memory (level-7 icosphere, 327680 cells:
    859 GB in float64) is refused with an error before anything is staged or copied; the
    library keeps working afterwards."""
    import ctypes

    from bempp_amd import _lib as L

    g = icosphere(7)
    s = _space(g, "dp0")
    a = _assembler("laplace", "single_layer", 0.0, _options())
    dummy = np.zeros(1)
    rc = L.lib().boundary_assembler_assemble_into_memory(a._rs_assembler, s._rs_space, s._rs_space,
                                                         ctypes.c_void_p(dummy.ctypes.data))
    assert rc != 0
    assert "bytes of device memory" in L.last_error()
    small = icosphere(1)
    sp = _space(small, "dp0")
    A = a.assemble(sp, sp)
    R = O.assemble("laplace", "single_layer", small.points, small.cells)
    assert _rel(A, R) <= TOL


@pytest.mark.parametrize("k", [2.5, 5.0, -2.5, -5.0, 1e-120, 0.0])
@pytest.mark.parametrize("op", ["single_layer", "hypersingular"])
def test_fp64_p0_scaled_orbit_sums(op, k):
    """The FP64 P0 tiles' Helmholtz single-layer / P0 hypersingular sums in lengths scaled by
    |k|/sqrt(2) (level-4 icosphere: k = +-2.5 phase tier 2, +-5 tier 1; the sine changes sign
    with k) and the full-reduction path that |k| < 1e-100 takes: relative Frobenius error
    <= 1e-12 against the oracle."""
    g = icosphere(4)
    s = _space(g, "dp0")
    a = _assembler("helmholtz", op, k, _options())
    print(f"[scaled] {op} k={k}", flush=True)
    A = a.assemble(s, s)
    R = O.assemble("helmholtz", op, g.points, g.cells, k=k)
    scale = np.abs(R).max()
    if scale == 0.0:  # P0 hypersingular at k = 0: -k^2 n.n vanishes
        assert np.abs(A).max() == 0.0
        return
    err = _rel(A / scale, R / scale)  # entries of order k^2 = 1e-240 at k = 1e-120
    print(f"[scaled]   rel err {err:.2e}", flush=True)
    assert err <= 1e-12
