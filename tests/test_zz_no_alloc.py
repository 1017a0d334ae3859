"""The shape gradient entry point that allocates no device memory
(boundary_assembler_shape_gradient_noalloc), called through the C ABI, against the
existing entry point (boundary_assembler_shape_gradient) on the same inputs: two calls in
a row on one GPU, the second on a smaller grid (so the shared static scratch holds the
first call's larger contents), Laplace (real gradient) and Helmholtz (complex gradient);
relative difference <= 1e-12 (both run the same kernels; the vertex scatter sums with
atomics in a varying order). A grid whose device data was never built is refused instead
of being uploaded."""
import ctypes

import numpy as np
import pytest

from bempp_amd import _lib as L
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler,
                                         create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.grid import Grid
from bempp_amd.shapes import icosphere

pytestmark = pytest.mark.gpu


def _dp0(g):
    return function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))


def _options():
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    return o


def _perturbed(level, seed):
    g0 = icosphere(level)
    rng = np.random.default_rng(seed)
    pts = g0.points * (1.0 + 0.05 * rng.standard_normal((len(g0.points), 1)))
    return Grid(pts, g0.cells)


def _vectors(n, dtype, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n)
    if np.dtype(dtype).kind == "c":
        x = x + 1j * rng.standard_normal(n)
    return x.astype(dtype)


def _noalloc(a, s, u, v, stream=None):
    import torch

    gdt = torch.float64 if np.dtype(a.dtype).kind == "f" else torch.complex128
    du = torch.from_numpy(u).cuda()
    dv = torch.from_numpy(v).cuda()
    npts = L.lib().grid_npoints(s.grid._handle)
    grad = torch.full((npts, 3), 7.0, dtype=gdt, device="cuda")  # stale values: the call overwrites
    torch.cuda.synchronize()
    rc = L.lib().boundary_assembler_shape_gradient_noalloc(
        a._rs_assembler, s._rs_space, s._rs_space, ctypes.c_void_p(du.data_ptr()), ctypes.c_void_p(dv.data_ptr()),
        ctypes.c_void_p(grad.data_ptr()), ctypes.c_void_p(stream))
    return rc, grad, (du, dv)


@pytest.mark.parametrize("kind", ["laplace", "helmholtz"])
def test_noalloc_shape_gradient_matches(kind):
    import torch

    if kind == "laplace":
        a = create_laplace_assembler(OperatorType.SingleLayer, _options())
        dtype = np.float64
    else:
        a = create_helmholtz_assembler(3.0, OperatorType.SingleLayer, _options())
        dtype = np.complex128
    big, small = _perturbed(3, 1), _perturbed(2, 2)
    cases = []
    for g, seed in ((big, 3), (small, 4)):
        s = _dp0(g)
        u, v = _vectors(len(g.cells), dtype, seed), _vectors(len(g.cells), dtype, seed + 10)
        ref = a.shape_gradient(s, s, u, v)  # the existing entry point (builds the device data)
        cases.append((g, s, u, v, ref))
    print(f"[noalloc] {kind}: cells {[len(c[0].cells) for c in cases]}", flush=True)
    outs = []
    for g, s, u, v, ref in cases:  # two calls in a row, the second on the smaller grid
        rc, grad, keep = _noalloc(a, s, u, v)
        assert rc == 0, L.last_error()
        outs.append((grad, keep))
    torch.cuda.synchronize()
    for (g, s, u, v, ref), (grad, _) in zip(cases, outs):
        got = grad.cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"[noalloc]   {len(g.cells)} cells: rel diff {err:.3e}", flush=True)
        assert err <= 1e-12


def test_noalloc_needs_built_device_data():
    """A grid whose device data does not exist yet is refused (the call would have to allocate)."""
    a = create_laplace_assembler(OperatorType.SingleLayer, _options())
    g = _perturbed(1, 5)
    s = _dp0(g)
    u = _vectors(len(g.cells), np.float64, 6)
    rc, grad, _ = _noalloc(a, s, u, u)
    assert rc != 0
    assert "allocates no device memory" in L.last_error()
    # after an ordinary call it works
    ref = a.shape_gradient(s, s, u, u)
    rc, grad, _ = _noalloc(a, s, u, u)
    assert rc == 0, L.last_error()
    import torch

    torch.cuda.synchronize()
    got = grad.cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
