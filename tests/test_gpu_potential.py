"""Potential operators on the GPU against the oracle (which reproduces the four bempp-cl
golden potentials of tests/compare_to_bempp_cl.rs:208-320 to 1e-16 with the 6-point rule)."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from bempp_amd.assembly.boundary import BoundaryAssemblerOptions, OperatorType
from bempp_amd.assembly.potential import create_helmholtz_potential_assembler, create_laplace_potential_assembler
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.shapes import icosphere, regular_sphere

pytestmark = pytest.mark.gpu

OPS = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer}


def _opts():
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    return o


def _pot(kind, op, k):
    if kind == "laplace":
        return create_laplace_potential_assembler(OPS[op], _opts())
    return create_helmholtz_potential_assembler(k, OPS[op], _opts())


def test_potential_goldens():
    cases = json.loads((Path(__file__).parent / "golden" / "bempp_cl_potentials_octahedron_l0.json").read_text())
    g = regular_sphere(0)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    for c in cases:
        ref = np.array(c["from_cl"])
        if ref.ndim == 3:
            ref = ref[..., 0] + 1j * ref[..., 1]
        P = _pot(c["kind"], c["operator"], c["k"]).assemble(s, np.array(c["points"]))
        err = np.abs(P.T - ref).max()
        print(f"[potential] {c['name']}: max |P - bempp-cl| {err:.2e}", flush=True)
        assert err <= 1e-12


@pytest.mark.parametrize("kind,k", [("laplace", 0.0), ("helmholtz", 4.0)])
@pytest.mark.parametrize("op", ["single_layer", "double_layer"])
@pytest.mark.parametrize("space", [(0, False), (1, True)])
def test_potential_vs_oracle(kind, k, op, space):
    g = icosphere(3)
    fam = create_family(Family.Lagrange, space[0], Continuity.Standard if space[1] else Continuity.Discontinuous)
    s = function_space(g, fam)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((500, 3))
    x *= (rng.uniform(0.2, 3.0, 500) / np.linalg.norm(x, axis=1))[:, None]  # inside and outside the sphere
    P = _pot(kind, op, k).assemble(s, x)
    R = O.potential(kind, op, g.points, g.cells, x, space=space, k=k)
    err = np.linalg.norm(P - R) / np.linalg.norm(R)
    print(f"[potential] {kind} {op} {space}: rel err {err:.2e}", flush=True)
    assert err <= 1e-12


def test_potential_gauss_law_and_device_output():
    """Double-layer potential of a unit density: -1 inside the closed surface, 0 outside
    (exact for flat triangles up to quadrature error); the device entry point writes the
    same matrix into a torch tensor and accumulates when asked."""
    import torch

    g = icosphere(4)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    x = np.array([[0.0, 0.0, 0.0], [0.1, -0.2, 0.3], [2.0, 0.0, 0.0], [0.0, -3.0, 1.0]])
    a = _pot("laplace", "double_layer", 0.0)
    P = a.assemble(s, x)
    assert np.allclose(P.sum(axis=1), [-1.0, -1.0, 0.0, 0.0], atol=1e-7)
    xt = torch.tensor(x, device="cuda")
    out = torch.empty((s.global_size, len(x)), dtype=torch.float64, device="cuda").t()
    a.assemble_into_device(s, xt, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), P)
    a.assemble_into_device(s, xt, out, accumulate=True)
    torch.cuda.synchronize()
    assert np.allclose(out.cpu().numpy(), 2 * P, rtol=1e-15, atol=0)
