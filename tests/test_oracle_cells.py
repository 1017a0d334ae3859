"""Pins of the oracle's general-cell assembly (grids of triangles and quadrilaterals,
oracle/bem_oracle.cpp `oracle_assemble_cells`): agreement with the triangle oracle on
triangle grids, closed forms for unit-square pairs, exact polynomial integration by the
quadrilateral pair rules, the Gauss identity and hypersingular row sums on closed quadrilateral
and mixed surfaces, and the reference's own mixed-grid tests
(tests/mixed_assembly.rs.bak:77-240)."""
import numpy as np
import pytest
from scipy import integrate

import oracle_lib as O
from quad_meshes import cube, cubed_sphere, mixed_screen, orient_outward, split

OPS = ["single_layer", "double_layer", "adjoint_double_layer", "hypersingular"]


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 3.0)])
@pytest.mark.parametrize("op", OPS)
def test_general_oracle_matches_triangle_oracle(kind, k, op):
    pts, cells = O.sphere(2)
    for sp in [(0, False), (1, True), (1, False)]:
        A = O.assemble(kind, op, pts, cells, test=sp, trial=sp, k=k)
        B = O.assemble_cells(kind, op, pts, cells, test=sp, trial=sp, k=k)
        assert np.linalg.norm(A - B) <= 1e-14 * np.linalg.norm(A), sp


def _unit_square_values():
    """int int 1/|x - y| over pairs of coplanar unit squares: coincident (closed form
    4 ln(1 + sqrt 2) - 4 (sqrt 2 - 1) / 3), sharing an edge, sharing a vertex, the last two
    as 2-D integrals of the difference variable u = x - y weighted by the overlap of the
    squares' indicator functions (bounded integrands)."""
    C = 4.0 * np.log(1.0 + np.sqrt(2.0)) - 4.0 / 3.0 * (np.sqrt(2.0) - 1.0)
    E = integrate.dblquad(lambda v, u: (1 - abs(u + 1)) * (1 - abs(v)) / np.hypot(u, v), -2, 0, -1, 1,
                          epsabs=1e-14, epsrel=1e-14)[0]
    V = integrate.dblquad(lambda v, u: (1 - abs(u + 1)) * (1 - abs(v + 1)) / np.hypot(u, v), -2, 0, -2, 0,
                          epsabs=1e-14, epsrel=1e-14)[0]
    return C, E, V


def test_unit_square_pairs_against_closed_forms():
    """Laplace SL P0 on three unit squares (coincident, edge and vertex pairs): the
    quadrilateral pair rules converge to the exact integrals."""
    C, E, V = _unit_square_values()
    pts = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0], [1, 1, 0], [2, 1, 0], [0, 2, 0], [1, 2, 0], [2, 2, 0]],
                   float)
    cells = [[0, 1, 3, 4], [1, 2, 4, 5], [4, 5, 7, 8]]
    errs = []
    for n in (4, 6, 8, 10):
        A = O.assemble_cells("laplace", "single_layer", pts, cells, nq_quad=64, nsing=n) * 4 * np.pi
        errs.append([abs(A[0, 0] - C) / C, abs(A[0, 1] - E) / E, abs(A[0, 2] - V) / V])
    errs = np.array(errs)
    assert np.all(np.diff(np.log(errs[:, :2]), axis=0) < -2.0)  # exponential convergence
    assert errs[-1, 0] < 5e-9 and errs[-1, 1] < 3e-10 and errs[-1, 2] < 1e-13


@pytest.mark.parametrize("cells,pair,total", [
    ([[0, 1, 3, 4]], (0, 0), 1.0),                 # coincident quadrilateral
    ([[0, 1, 3, 4], [1, 2, 4, 5]], (0, 1), 1.0),   # edge, quadrilateral x quadrilateral
    ([[0, 1, 3, 4], [4, 5, 7, 8]], (1, 0), 1.0),   # vertex
    ([[0, 1, 3, 4], [1, 2, 4]], (1, 0), 0.5),      # edge, triangle x quadrilateral
    ([[0, 1, 3, 4], [4, 5, 8]], (0, 1), 0.5),      # vertex, quadrilateral x triangle
])
def test_quadrilateral_pair_rules_integrate_polynomials(cells, pair, total):
    """Each pair rule is a rule on (test reference cell) x (trial reference cell): weights sum
    to the product of the reference areas and products of monomials integrate exactly."""
    pts = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0], [1, 1, 0], [2, 1, 0], [0, 2, 0], [1, 2, 0], [2, 2, 0]],
                   float)
    tp, sp, w = O.cell_pair_rule(pts, cells, pair[0], pair[1], 4)
    assert abs(w.sum() - total) < 1e-14 and np.all(w > 0)

    def exact(cell, a, b):  # int over the reference cell of x^a y^b
        if len(cell) == 4:
            return 1.0 / ((a + 1) * (b + 1))
        from math import factorial
        return factorial(a) * factorial(b) / factorial(a + b + 2)

    for a, b, c, d in [(1, 0, 0, 0), (0, 2, 1, 0), (1, 1, 0, 2), (2, 0, 1, 1), (0, 1, 2, 1)]:
        got = np.sum(w * tp[:, 0] ** a * tp[:, 1] ** b * sp[:, 0] ** c * sp[:, 1] ** d)
        ref = exact(cells[pair[0]], a, b) * exact(cells[pair[1]], c, d)
        assert abs(got - ref) < 1e-14, (a, b, c, d)


def _fd_areas(pts, cells, m=12):
    """Cell areas with a central-difference Jacobian of the degree-1 map."""
    g, gw = np.polynomial.legendre.leggauss(m)
    g, gw = 0.5 * (g + 1), 0.5 * gw
    out = []
    for c in cells:
        V = pts[c]

        def x(s, t):
            if len(c) == 4:
                return V[0] + s * (V[1] - V[0]) + t * (V[2] - V[0]) + s * t * (V[3] - V[1] - V[2] + V[0])
            return V[0] + s * (V[1] - V[0]) + t * (V[2] - V[0])

        h, a = 1e-6, 0.0
        for i in range(m):
            for j in range(m):
                s, t, jac = g[i], g[j], 1.0
                if len(c) == 3:
                    s, jac = s * (1 - t), 1 - t  # the square onto the triangle
                xs = (x(s + h, t) - x(s - h, t)) / (2 * h)
                xt = (x(s, t + h) - x(s, t - h)) / (2 * h)
                a += gw[i] * gw[j] * np.linalg.norm(np.cross(xs, xt)) * jac
        out.append(a)
    return np.array(out)


@pytest.mark.parametrize("name", ["cube", "cubed_sphere", "mixed_cube"])
def test_gauss_identity_on_closed_quadrilateral_surfaces(name):
    """Laplace DL P0 row sums -> -|cell| / 2 on a closed surface (flat squares, non-planar
    bilinear cells, triangles and squares mixed), converging with the Duffy order; the areas
    come from a finite-difference Jacobian. Laplace HS P1 row sums vanish."""
    pts, cells = {"cube": cube(2), "cubed_sphere": cubed_sphere(3),
                  "mixed_cube": (cube(2)[0], split(cube(2)[1], lambda i: i % 2 == 0))}[name]
    orient_outward(pts, cells)
    area = _fd_areas(pts, cells)
    errs = []
    for n in (4, 6):
        A = O.assemble_cells("laplace", "double_layer", pts, cells, nq_tri=37, nq_quad=64, nsing=n)
        errs.append(np.abs(A.sum(axis=1) + area / 2).max() / area.max())
    assert errs[1] < 1e-5 and errs[1] < errs[0] / 10
    H = O.assemble_cells("laplace", "hypersingular", pts, cells, test=(1, True), trial=(1, True))
    assert np.abs(H.sum(axis=1)).max() < 1e-13 * np.abs(H).max()


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 3.0)])
@pytest.mark.parametrize("op", OPS)
def test_mixed_grid_blocks_equal_single_element_grids(kind, k, op):
    """tests/mixed_assembly.rs.bak:77-176: on the mixed screen, the quadrilateral block of the
    DP0 / DP1 matrix equals the matrix of the grid of the two quadrilaterals, the triangle
    block that of the grid of the four triangles."""
    pts, mixed, quads, tris = mixed_screen()
    for sp, nq, nt in [((0, False), 2, 4), ((1, False), 8, 12)]:
        M = O.assemble_cells(kind, op, pts, mixed, test=sp, trial=sp, k=k)
        Q = O.assemble_cells(kind, op, pts, quads, test=sp, trial=sp, k=k)
        T = O.assemble_cells(kind, op, pts, tris, test=sp, trial=sp, k=k)
        np.testing.assert_allclose(M[:nq, :nq], Q, rtol=0, atol=1e-14 * np.abs(Q).max())
        np.testing.assert_allclose(M[nq:, nq:], T, rtol=0, atol=1e-14 * np.abs(T).max())


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 3.0)])
@pytest.mark.parametrize("op", OPS)
def test_mixed_grid_p1_is_sum_of_dp1(kind, k, op):
    """tests/mixed_assembly.rs.bak:178-225: P1 entries on the mixed screen are sums of DP1
    entries over the DP1 functions of each vertex (the combinations listed there)."""
    pts, mixed, _, _ = mixed_screen()
    P = O.assemble_cells(kind, op, pts, mixed, test=(1, True), trial=(1, True), k=k)
    D = O.assemble_cells(kind, op, pts, mixed, test=(1, False), trial=(1, False), k=k)
    combos = [[0], [1, 8, 11], [2, 4], [3, 5, 13, 14, 17], [6], [7, 19], [9], [10, 12, 15], [16, 18]]
    for i, ci in enumerate(combos):
        for j, cj in enumerate(combos):
            assert abs(D[np.ix_(ci, cj)].sum() - P[i, j]) <= 1e-13 * np.abs(P).max()


def test_cells_space_numbering():
    """P1 DOFs on the mixed screen follow first appearance over the cells (the numbering the
    reference's combinations above assume); DP1 numbers cell by cell."""
    pts, mixed, _, _ = mixed_screen()
    n, dofs = O.cells_space(pts, mixed, 1, True)
    assert n == 9 and dofs[0] == [0, 1, 2, 3] and dofs[1] == [2, 3, 4, 5] and dofs[2] == [1, 6, 7]
    n, dofs = O.cells_space(pts, mixed, 1, False)
    assert n == 20 and dofs[1] == [4, 5, 6, 7] and dofs[5] == [17, 18, 19]
    with pytest.raises(ValueError, match="square number"):
        O.assemble_cells("laplace", "single_layer", pts, mixed, nq_quad=37)


def _cube_mixed():
    pts, c = cube(2)
    return pts, split(c, lambda i: i % 3 == 0)


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 2.5)])
@pytest.mark.parametrize("op", OPS)
def test_parts_split_the_operator(kind, k, op):
    """The oracle's parts: whole operator = regular (non-adjacent pairs) + singular (adjacent
    pairs, pair rules); the singular part lives on the DOF pairs of adjacent cells."""
    pts, cells = _cube_mixed()
    for sp in [(0, False), (1, True)]:
        A = O.assemble_cells(kind, op, pts, cells, test=sp, trial=sp, k=k)
        R = O.assemble_cells(kind, op, pts, cells, test=sp, trial=sp, k=k, part="regular")
        S = O.assemble_cells(kind, op, pts, cells, test=sp, trial=sp, k=k, part="singular")
        assert np.abs(A - R - S).max() <= 1e-15 * max(np.abs(A).max(), 1e-300)
        _, dofs = O.cells_space(pts, cells, *sp)
        allowed = np.zeros(A.shape, dtype=bool)
        for a in range(len(cells)):
            for b in range(len(cells)):
                if set(cells[a]) & set(cells[b]):
                    allowed[np.ix_(dofs[a], dofs[b])] = True
        assert not np.any(S[~allowed])


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 2.5)])
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("mesh", ["cubed_sphere", "mixed_cube"])
def test_regular_and_correction_parts_against_numpy(mesh, kind, k, op):
    """The regular rule written out in numpy (bilinear map, Q1 / P1 bases, surface curls,
    kernel zero at zero distance): the correction part (adjacent pairs under the regular rule)
    equals the numpy sums over adjacent cell pairs, regular + correction the sums over all."""
    import fmm_helpers as F

    pts, cells = cubed_sphere(2) if mesh == "cubed_sphere" else _cube_mixed()
    for deg, cont in [(0, False), (1, True)]:
        full, adj = F.cells_regular_parts(kind, op, k, pts, cells, deg, cont, nq_tri=6, m=4)
        kw = dict(test=(deg, cont), trial=(deg, cont), k=k, nq_tri=6, nq_quad=16)
        C = O.assemble_cells(kind, op, pts, cells, part="correction", **kw)
        R = O.assemble_cells(kind, op, pts, cells, part="regular", **kw)
        if np.linalg.norm(full) == 0.0:  # Laplace hypersingular on P0
            assert np.abs(C).max() == 0.0 and np.abs(R).max() == 0.0
            continue
        assert np.linalg.norm(C - adj) <= 1e-14 * np.linalg.norm(adj)
        assert np.linalg.norm(R + C - full) <= 1e-14 * np.linalg.norm(full)


def test_row_range_equals_full_rows():
    """The oracle's row range (P0 test spaces): rows r0..r1-1 of the full matrix, bit for bit;
    a P1 test space is refused."""
    pts, cells = _cube_mixed()
    for kind, op, k in [("helmholtz", "single_layer", 2.0), ("laplace", "double_layer", 0.0)]:
        A = O.assemble_cells(kind, op, pts, cells, k=k, trial=(1, True))
        for r in [(0, 5), (10, 30), (len(cells) - 7, len(cells))]:
            B = O.assemble_cells(kind, op, pts, cells, k=k, trial=(1, True), rows=r)
            assert np.array_equal(B, A[r[0]:r[1]])
    with pytest.raises(ValueError, match="P0 test space"):
        O.assemble_cells("laplace", "single_layer", pts, cells, test=(1, True), rows=(0, 3))
