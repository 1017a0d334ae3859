"""Singular correction (python/bempp/assembly/boundary.py:157-169) on the GPU: the
regular-rule sum over each adjacent cell pair. Checked against the same sum written out in
numpy, and through the identity of the reference's FMM prototype (tests/fmm.rs.bak:22-95):
dense = P^T K P - correction + singular, against the oracle's dense matrix."""
import numpy as np
import pytest

import fmm_helpers as F
import oracle_lib as O
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler,
                                         create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.shapes import icosphere

pytestmark = pytest.mark.gpu

OPS = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer,
       "adjoint_double_layer": OperatorType.AdjointDoubleLayer}


@pytest.mark.parametrize("kind,op,k", [("laplace", "single_layer", 0.0), ("helmholtz", "single_layer", 3.0),
                                       ("helmholtz", "double_layer", 3.0), ("laplace", "adjoint_double_layer", 0.0)])
@pytest.mark.parametrize("space_name", ["dp0", "p1"])
def test_correction_brute_force_and_fmm_identity(kind, op, k, space_name):
    g = icosphere(1)
    pts, cells = g.points, g.cells
    deg = (0, False) if space_name == "dp0" else (1, True)
    s = function_space(g, create_family(Family.Lagrange, deg[0],
                                        Continuity.Standard if deg[1] else Continuity.Discontinuous))
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    a = create_laplace_assembler(OPS[op], o) if kind == "laplace" else create_helmholtz_assembler(k, OPS[op], o)
    C = a.assemble_singular_correction(s, s).toarray()
    S = a.assemble_singular(s, s).toarray()
    dofs = np.array([s.cell_dofs(c) for c in range(len(cells))], dtype=np.int64)
    PKP, C_ref = F.regular_parts(kind, op, k, pts, cells, dofs, s.global_size)
    errC = np.linalg.norm(C - C_ref) / np.linalg.norm(C_ref)
    print(f"[correction] {kind} {op} {space_name}: vs numpy sum {errC:.2e}", flush=True)
    assert errC < 1e-12
    dense = O.assemble(kind, op, pts, cells, test=deg, trial=deg, k=k)
    err = np.linalg.norm(PKP - C + S - dense) / np.linalg.norm(dense)
    print(f"[correction]   P^T K P - C + S vs dense: {err:.2e}", flush=True)
    assert err < 1e-12
