"""GPU parity at the benchmark configurations' own sizes, against the CPU oracle.

The full matrices are too large for the oracle at these sizes (the level-5 P0 matrix has
4.2e8 entries, the level-6 one 6.7e9), so each assembly is compared on row blocks: the
first 64 rows, 64 rows straddling the middle tile boundary, and the last 64 rows. The
first tile row is computed directly by the symmetric P0/P1 tiles; the last tile rows are
almost entirely written as mirrored (transposed) stores of other tiles. Tolerances:
relative Frobenius error <= 1e-12 per block in FP64, <= 1e-6 in mixed precision (with
> 1e-12 showing that the FP32 path ran).

Configurations (same meshes and rules as the benchmark: icospheres, 6-point regular rule,
4-point Duffy rule):
  - Helmholtz single layer P0, level 5 (20480 cells), k = 5: complex128 FP64 and
    complex64-mixed (the headline metric's workload);
  - Laplace double layer and hypersingular P1, level 5 (10242 DOFs): FP64 and mixed;
  - Helmholtz double, adjoint double and hypersingular P1, level 5, k = 5: complex128
    FP64 and complex64-mixed;
  - Laplace single layer P0, level 6 (81920 cells, a 53.7 GB matrix): FP64, the last rows'
    entries lying beyond element offset 2^32;
  - the 64-item batch (seed 0): complex128 and complex64-mixed, rows of the smallest, a
    median and the largest item.
"""
import numpy as np
import pytest

import oracle_lib as O
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, Precision, create_helmholtz_assembler,
                                         create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.shapes import icosphere

pytestmark = pytest.mark.gpu

OPS = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer,
       "adjoint_double_layer": OperatorType.AdjointDoubleLayer, "hypersingular": OperatorType.Hypersingular}
ROWS = 64

_GRIDS = {}


def _grid(level):
    if level not in _GRIDS:
        _GRIDS[level] = icosphere(level)
    return _GRIDS[level]


def _space(g, p1):
    fam = create_family(Family.Lagrange, 1, Continuity.Standard) if p1 else \
        create_family(Family.Lagrange, 0, Continuity.Discontinuous)
    return function_space(g, fam)


def _options(mixed):
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    if mixed:
        o.set_precision(Precision.Mixed)
    return o


def _blocks(n):
    mid = n // 2 - ROWS // 2 + 7  # not tile aligned: straddles two tile rows
    return [(0, ROWS), (mid, mid + ROWS), (n - ROWS, n)]


def _compare_blocks(D, kind, op, g, space, k, mixed, tag):
    """D: the device matrix (column-major torch tensor). Returns the worst block error."""
    n = D.shape[0]
    worst = 0.0
    # mixed precision assembles the Laplace hypersingular operator on P1 x P1 in FP64 (its
    # orbit-form FP64 tiles run faster than the FP32 ones)
    if mixed and kind == "laplace" and op == "hypersingular" and space[0] == 1:
        mixed = False
    for rb, re in _blocks(n):
        got = D[rb:re].cpu().numpy()
        got = got.astype(np.complex128 if np.iscomplexobj(got) else np.float64)
        ref = O.assemble(kind, op, g.points, g.cells, test=space, trial=space, k=k, rows=(rb, re))
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        emax = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"[config] {tag} rows {rb}..{re - 1}: rel Frobenius err {err:.3e}, max entry err {emax:.3e}",
              flush=True)
        worst = max(worst, err)
        if mixed:
            assert err <= 1e-6 and emax <= 1e-5, (rb, err, emax)
        else:
            assert err <= 1e-12, (rb, err)
    if mixed:
        assert worst > 1e-12  # the FP32 far field ran
    return worst


def _assemble(kind, op, k, mixed, p1, level):
    import torch

    g = _grid(level)
    s = _space(g, p1)
    if kind == "laplace":
        a = create_laplace_assembler(OPS[op], _options(mixed))
    else:
        a = create_helmholtz_assembler(k, OPS[op], _options(mixed), dtype=np.complex64 if mixed else np.complex128)
    print(f"[config] assembling {kind} {op} {'P1' if p1 else 'P0'} level {level} k={k} "
          f"{'mixed' if mixed else 'fp64'} ({s.global_size} DOFs)", flush=True)
    D = a.assemble_device(s, s)
    torch.cuda.synchronize()
    return g, D


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64_c128", "mixed_c64"])
def test_metric_helmholtz_sl_p0_l5(mixed):
    """The headline workload: Helmholtz single layer P0 on the level-5 icosphere, k = 5."""
    import torch

    g, D = _assemble("helmholtz", "single_layer", 5.0, mixed, False, 5)
    _compare_blocks(D, "helmholtz", "single_layer", g, (0, False), 5.0, mixed, f"metric {'mixed' if mixed else 'fp64'}")
    del D
    torch.cuda.empty_cache()


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
@pytest.mark.parametrize("kind,op,k", [("laplace", "double_layer", 0.0), ("laplace", "hypersingular", 0.0),
                                       ("helmholtz", "double_layer", 5.0), ("helmholtz", "adjoint_double_layer", 5.0),
                                       ("helmholtz", "hypersingular", 5.0)])
def test_p1_operators_l5(kind, op, k, mixed):
    """P1 x P1 on the level-5 icosphere (10242 DOFs): the LDS-accumulating tiles, their
    mirrored flush (hypersingular) and the P1 Duffy pairs, against the row-restricted oracle."""
    import torch

    g, D = _assemble(kind, op, k, mixed, True, 5)
    _compare_blocks(D, kind, op, g, (1, True), k, mixed, f"{kind} {op} P1 L5 {'mixed' if mixed else 'fp64'}")
    del D
    torch.cuda.empty_cache()


def test_laplace_sl_p0_l6():
    """Laplace single layer P0 on the level-6 icosphere (81920 cells, 53.7 GB): rows at the end
    of the matrix address entries far beyond 2^31 (and 2^32) elements."""
    import torch

    g, D = _assemble("laplace", "single_layer", 0.0, False, False, 6)
    n = D.shape[0]
    assert (n - 1) + n * (n - 1) > 2 ** 32
    _compare_blocks(D, "laplace", "single_layer", g, (0, False), 0.0, False, "laplace SL P0 L6")
    del D
    torch.cuda.empty_cache()


@pytest.mark.parametrize("mixed", [False, True], ids=["c128", "mixed_c64"])
def test_batch_64_items_seed0(mixed):
    """The 64-item batched workload (seed 0, levels 2..5, ellipsoids, k in [1, 10]) in one
    persistent launch: rows of the smallest, a median and the largest item (first, middle
    and last row blocks, the last ones mirrored) against the oracle."""
    import torch

    from bempp_amd.assembly.batch import BatchAssembler
    from bempp_amd.workloads import batch_items

    items = batch_items(0, 64, np.complex64 if mixed else np.complex128, mixed)
    sizes = [len(g.cells) for g, *_ in items]
    print(f"[config] batch of {len(items)} ({'mixed c64' if mixed else 'c128'}): {sum(n * n for n in sizes)} entries",
          flush=True)
    b = BatchAssembler([it[2] for it in items], [it[1] for it in items])
    outs = b.assemble()
    torch.cuda.synchronize()
    order = np.argsort(sizes, kind="stable")
    for idx in (order[0], order[len(order) // 2], order[-1]):
        g, s, a, k, lev = items[idx]
        _compare_blocks(outs[idx], "helmholtz", "single_layer", g, (0, False), k, mixed,
                        f"batch item {idx} (level {lev}, k={k:.4f})")
    del b, outs, items
    torch.cuda.empty_cache()
