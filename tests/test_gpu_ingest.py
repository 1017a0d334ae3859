"""GPU assembly on grids and spaces built through the ndgrid-style ingest entry points
(single_element_grid_borrowed_create_bempp, create_lagrange_family_bempp, function_space:
src/bindings/ndgrid.rs:16-56, src/bindings/ndelement.rs:18-25, src/bindings/space.rs:20-42),
as the reference's python layer calls them (python/bempp/function_space.py:143-191): f32 or
f64 points and a topology whose vertex numbering is a random relabelling of the points.
Against the oracle on the same (rounded) coordinates, relative Frobenius error <= 1e-12."""
import numpy as np
import pytest

import oracle_lib as O
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler,
                                         create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import FunctionSpace
from bempp_amd.grid import Grid
from bempp_amd.shapes import icosphere
from ndgrid_helpers import ndgrid_args, raw_lib

pytestmark = pytest.mark.gpu

OPS = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer,
       "adjoint_double_layer": OperatorType.AdjointDoubleLayer, "hypersingular": OperatorType.Hypersingular}


def _options():
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    return o


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_assembly_on_ingested_grid(dtype):
    import torch

    lib = raw_lib()
    g0 = icosphere(3)
    pts, cells = g0.points, g0.cells
    perm = np.random.default_rng(11).permutation(len(pts))
    args, keep = ndgrid_args(pts, cells, perm, dtype)
    gh = lib.single_element_grid_borrowed_create_bempp(*args)
    assert gh, lib.bempp_last_error()
    grid = Grid(_handle=gh)  # frees the handle with the grid
    p64 = np.asarray(pts, dtype=dtype).astype(np.float64)
    cases = [(0, Continuity.Discontinuous, "helmholtz", "single_layer", 5.0),
             (1, Continuity.Standard, "helmholtz", "double_layer", 2.0),
             (1, Continuity.Standard, "laplace", "hypersingular", 0.0),
             (1, Continuity.Discontinuous, "laplace", "adjoint_double_layer", 0.0)]
    for degree, cont, kind, op, k in cases:
        fam = lib.create_lagrange_family_bempp(degree, int(cont), 1)
        assert fam, lib.bempp_last_error()
        sh = lib.function_space(gh, fam)
        assert sh, lib.bempp_last_error()
        lib.element_family_t_free_bempp(fam)
        space = FunctionSpace(grid, create_family(Family.Lagrange, degree, cont), sh)
        a = (create_laplace_assembler(OPS[op], _options()) if kind == "laplace"
             else create_helmholtz_assembler(k, OPS[op], _options()))
        print(f"[ingest] {np.dtype(dtype).name} points: {kind} {op} degree {degree} {cont.name}", flush=True)
        A = a.assemble(space, space)
        D = a.assemble_device(space, space)
        torch.cuda.synchronize()
        sp = (degree, cont == Continuity.Standard)
        R = O.assemble(kind, op, p64, cells, test=sp, trial=sp, k=k)
        err = np.linalg.norm(A - R) / np.linalg.norm(R)
        errd = np.linalg.norm(D.cpu().numpy() - R) / np.linalg.norm(R)
        print(f"[ingest]   rel err host {err:.2e}, device {errd:.2e}", flush=True)
        assert err <= 1e-12 and errd <= 1e-12
        del D
        csr = a.assemble_singular(space, space)
        rows, cols, vals = O.assemble_singular_coo(kind, op, p64, cells, test=sp, trial=sp, k=k)
        S = np.zeros(csr.shape, dtype=vals.dtype)
        np.add.at(S, (rows, cols), vals)
        assert np.linalg.norm(csr.toarray() - S) / np.linalg.norm(S) <= 1e-12
