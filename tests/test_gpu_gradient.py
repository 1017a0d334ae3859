"""Matrix-free apply and the shape gradient d(v^T A(X) u)/dX on the GPU, checked against
the CPU oracle: apply against the oracle's dense matrix, the gradient against central
finite differences (h = 1e-6) of the oracle's bilinear form along random vertex directions."""
import numpy as np
import pytest

import oracle_lib as O
import shapes_helpers as S
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler,
                                         create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.grid import Grid
from bempp_amd.shapes import icosphere

pytestmark = pytest.mark.gpu


def _opts():
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    return o


def _asm(kind, k, op=OperatorType.SingleLayer):
    return create_laplace_assembler(op, _opts()) if kind == "laplace" else create_helmholtz_assembler(k, op, _opts())


def _oracle_form(kind, k, pts, cells, u, v, block=4096):
    """v^T A u with the oracle's dense assembly, in row blocks (bounded memory)."""
    total = 0.0
    for r0 in range(0, len(cells), block):
        r1 = min(len(cells), r0 + block)
        Ab = O.assemble(kind, "single_layer", pts, cells, k=k, rows=(r0, r1))
        total += v[r0:r1] @ (Ab @ u)
    return total


@pytest.mark.parametrize("kind,k,op", [("helmholtz", 3.0, "single_layer"), ("helmholtz", 5.0, "double_layer"),
                                       ("laplace", 0.0, "adjoint_double_layer"), ("laplace", 0.0, "hypersingular")])
def test_apply_matches_dense(kind, k, op):
    opmap = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer,
             "adjoint_double_layer": OperatorType.AdjointDoubleLayer, "hypersingular": OperatorType.Hypersingular}
    g = icosphere(3)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    a = _asm(kind, k, opmap[op])
    rng = np.random.default_rng(1)
    u = S.complex_normal(s.global_size, rng) if kind == "helmholtz" else rng.standard_normal(s.global_size)
    y = a.apply(s, s, u)
    R = O.assemble(kind, op, g.points, g.cells, k=k)
    ref = R @ u
    if np.linalg.norm(ref) == 0.0:
        assert np.abs(y).max() == 0.0
        return
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"[apply] {kind} {op}: rel err {err:.2e}", flush=True)
    assert err < 1e-12


def _check_gradient(kind, k, level, seed, ndir, tol=1e-6, h=1e-6):
    g0 = icosphere(level)
    pts = S.perturb_radially(g0.points, seed)
    cells = g0.cells
    g = Grid(pts, cells)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    rng = np.random.default_rng(seed)
    if kind == "helmholtz":
        u, v = S.complex_normal(len(cells), rng), S.complex_normal(len(cells), rng)
    else:
        u, v = rng.standard_normal(len(cells)), rng.standard_normal(len(cells))
    a = _asm(kind, k)
    G = a.shape_gradient(s, s, u, v)
    worst = 0.0
    for d in range(ndir):
        D = rng.standard_normal(pts.shape)
        D /= np.linalg.norm(D)
        fd = (_oracle_form(kind, k, pts + h * D, cells, u, v) - _oracle_form(kind, k, pts - h * D, cells, u, v)) / (2 * h)
        an = np.sum(G * D)
        rel = abs(an - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"[grad] {kind} L{level} dir {d}: analytic {an:.10e} fd {fd:.10e} rel {rel:.2e}", flush=True)
    assert worst <= tol


@pytest.mark.parametrize("kind,k", [("helmholtz", 3.0), ("laplace", 0.0)])
def test_shape_gradient_fd_l3(kind, k):
    """Level-3 perturbed icosphere (1280 cells), 8 random unit directions in R^{3 Nv}."""
    _check_gradient(kind, k, 3, seed=11, ndir=8)


@pytest.mark.parametrize("kind,k", [("helmholtz", 4.0), ("laplace", 0.0)])
def test_shape_gradient_every_component_l1(kind, k):
    """Every component of the gradient (42 vertices x 3 coordinates of a perturbed level-1
    icosphere) against central differences of the oracle's bilinear form, so that a wrong
    vertex, coordinate or cell-side term shows in its own component."""
    g0 = icosphere(1)
    pts = S.perturb_radially(g0.points, 7)
    cells = g0.cells
    g = Grid(pts, cells)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    rng = np.random.default_rng(7)
    if kind == "helmholtz":
        u, v = S.complex_normal(len(cells), rng), S.complex_normal(len(cells), rng)
    else:
        u, v = rng.standard_normal(len(cells)), rng.standard_normal(len(cells))
    G = _asm(kind, k).shape_gradient(s, s, u, v)
    h = 1e-6
    fd = np.zeros_like(G)
    for vtx in range(len(pts)):
        for c in range(3):
            D = np.zeros_like(pts)
            D[vtx, c] = h
            fd[vtx, c] = (_oracle_form(kind, k, pts + D, cells, u, v) - _oracle_form(kind, k, pts - D, cells, u, v)) / (2 * h)
    err = np.linalg.norm(G - fd) / np.linalg.norm(fd)
    emax = np.abs(G - fd).max() / np.abs(fd).max()
    print(f"[grad] {kind} L1 all {G.size} components: rel err {err:.2e}, max {emax:.2e}", flush=True)
    assert err <= 1e-8 and emax <= 1e-7  # the differences themselves agree to ~6e-10 (h vs 2h)


def test_shape_gradient_fd_l5_config():
    """The shape-gradient configuration: Helmholtz SL P0, k = 3, level-5 icosphere radially
    perturbed by 0.05 sum_{l<=4} a_lm Y_lm (seeded), u, v complex normal; 3 directions."""
    _check_gradient("helmholtz", 3.0, 5, seed=2024, ndir=3)


def test_autograd_function():
    """torch.autograd: loss = Re L + 0.3 Im L, gradients w.r.t. points (finite differences
    of the oracle), u and v (analytic: g conj(A^T v), g conj(A u))."""
    import torch

    from bempp_amd.autograd import bilinear_form

    g0 = icosphere(2)
    pts = S.perturb_radially(g0.points, 5)
    g = Grid(pts, g0.cells)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    a = _asm("helmholtz", 2.0)
    rng = np.random.default_rng(3)
    un, vn = S.complex_normal(len(g0.cells), rng), S.complex_normal(len(g0.cells), rng)
    P = torch.tensor(pts, requires_grad=True)
    u = torch.tensor(un, requires_grad=True)
    v = torch.tensor(vn, requires_grad=True)
    L = bilinear_form(P, u, v, a, s)
    loss = L.real + 0.3 * L.imag
    loss.backward()
    A = O.assemble("helmholtz", "single_layer", pts, g0.cells, k=2.0)
    assert abs(L.item() - vn @ A @ un) / abs(vn @ A @ un) < 1e-12
    # for a real loss of complex L torch passes g = dloss/dRe L + i dloss/dIm L = 1 + 0.3i,
    # and a complex leaf z receives g conj(dL/dz)
    gL = 1.0 + 0.3j
    assert np.allclose(v.grad.numpy(), gL * np.conj(A @ un), rtol=1e-10, atol=1e-13)
    assert np.allclose(u.grad.numpy(), gL * np.conj(A.T @ vn), rtol=1e-10, atol=1e-13)
    h = 1e-6
    for d in range(2):
        D = rng.standard_normal(pts.shape)
        D /= np.linalg.norm(D)
        Lp = vn @ O.assemble("helmholtz", "single_layer", pts + h * D, g0.cells, k=2.0) @ un
        Lm = vn @ O.assemble("helmholtz", "single_layer", pts - h * D, g0.cells, k=2.0) @ un
        fd = ((Lp - Lm) / (2 * h))
        fd_loss = fd.real + 0.3 * fd.imag
        an = float(np.sum(P.grad.numpy() * D))
        assert abs(an - fd_loss) <= 1e-6 * abs(fd_loss), (an, fd_loss)


def test_apply_transpose():
    """A^T u: exact transpose including the (non-symmetric) Duffy part."""
    g = icosphere(2)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    a = _asm("helmholtz", 4.0)
    u = S.complex_normal(s.global_size, np.random.default_rng(8))
    R = O.assemble("helmholtz", "single_layer", g.points, g.cells, k=4.0)
    y = a.apply(s, s, u, transpose=True)
    assert np.linalg.norm(y - R.T @ u) / np.linalg.norm(R.T @ u) < 1e-12
