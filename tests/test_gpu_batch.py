"""Batched assembly on the GPU: every item of a persistent batched launch equals the
single-operator assembly of the same item and the CPU oracle; graph replays equal the
eager launch; every launch shape gives the same result; the autotuner's cache is
reused; mixed-precision batches stay within 1e-6 of the FP64 oracle."""
import numpy as np
import pytest

import oracle_lib as O
from bempp_amd.assembly.batch import BatchAssembler
from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, Precision,
                                         create_helmholtz_assembler, create_laplace_assembler)
from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
from bempp_amd.function_space import function_space
from bempp_amd.shapes import ellipsoid, icosphere

pytestmark = pytest.mark.gpu

OPS = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer,
       "adjoint_double_layer": OperatorType.AdjointDoubleLayer, "hypersingular": OperatorType.Hypersingular}


def _options(mixed=False):
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    if mixed:
        o.set_precision(Precision.Mixed)
    return o


def _dp0(g):
    return function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))


def _items(seed, n, levels, kind="helmholtz", op="single_layer", mixed=False, dtype=None):
    """n ellipsoids (levels drawn from `levels`, semi-axes in [0.5, 2], k in [1, 10])."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        lev = int(rng.choice(levels))
        axes = rng.uniform(0.5, 2.0, 3)
        k = float(rng.uniform(1.0, 10.0))
        if dtype == np.complex64:
            k = float(np.float32(k))  # a complex64 assembler holds k in FP32; the oracle uses the same k
        g = ellipsoid(lev, axes)
        if kind == "laplace":
            a = create_laplace_assembler(OPS[op], _options(mixed))
            k = 0.0
        else:
            a = create_helmholtz_assembler(k, OPS[op], _options(mixed), dtype=dtype or np.complex128)
        out.append((g, _dp0(g), a, k))
    return out


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("kind,op", [("helmholtz", "single_layer"), ("helmholtz", "double_layer"),
                                     ("helmholtz", "hypersingular"), ("laplace", "single_layer"),
                                     ("laplace", "adjoint_double_layer")])
def test_batch_matches_single_and_oracle(kind, op):
    import torch

    items = _items(7, 6, [1, 2, 3], kind, op)
    print(f"[batch] {kind} {op}: sizes {[len(g.cells) for g, *_ in items]}", flush=True)
    b = BatchAssembler([a for *_, a, _ in items], [s for _, s, _, _ in items])
    print(f"[batch]   counts {b.counts()}", flush=True)
    outs = b.assemble()
    torch.cuda.synchronize()
    for (g, s, a, k), o in zip(items, outs):
        single = a.assemble_device(s, s)
        torch.cuda.synchronize()
        got = o.cpu().numpy()
        assert np.array_equal(got, single.cpu().numpy())  # same tile and Duffy code, same order
        R = O.assemble(kind, op, g.points, g.cells, k=k)
        if np.linalg.norm(R) == 0.0:
            assert np.abs(got).max() == 0.0
            continue
        err = _rel(got, R)
        print(f"[batch]   n={len(g.cells)} k={k:.3f} rel err {err:.2e}", flush=True)
        assert err <= 1e-12


def test_batch_launch_shapes_agree():
    import torch

    items = _items(11, 5, [1, 2, 3])
    b = BatchAssembler([a for *_, a, _ in items], [s for _, s, _, _ in items])
    b.set_launch(2, 8, 0)
    ref = [o.cpu().clone() for o in b.assemble()]
    shapes = [(bpc, chunk, order, 4) for bpc in (0, 1, 4) for chunk in (4, 8) for order in (0, 1)]
    shapes += [(bpc, 8, order, waves) for bpc in (0, 2) for order in (0, 1) for waves in (2, 8)]
    for shape in shapes:
        for o in b.outputs:
            o.fill_(float("nan"))
        b.set_launch(*shape)
        assert b.launch() == shape
        outs = b.assemble()
        torch.cuda.synchronize()
        for r, o in zip(ref, outs):
            assert torch.equal(r, o.cpu()), shape


def test_batch_graph_replays():
    """Replays of the captured graph, with eager torch work and synchronizations in between,
    each equal the eager launch (outputs are cleared before every replay)."""
    import torch

    items = _items(3, 4, [1, 2])
    b = BatchAssembler([a for *_, a, _ in items], [s for _, s, _, _ in items])
    ref = [o.cpu().clone() for o in b.assemble()]
    b.set_graph(True)
    for rep in range(4):
        for o in b.outputs:
            o.zero_()
        x = torch.randn(512, 512, device="cuda", dtype=torch.float64)
        (x @ x).sum().item()  # unrelated eager work
        torch.cuda.synchronize()
        outs = b.assemble()
        torch.cuda.synchronize()
        for r, o in zip(ref, outs):
            assert torch.equal(r, o.cpu()), rep
    b.set_launch(1, 4, 1)  # a new shape recaptures
    for o in b.outputs:
        o.zero_()
    outs = b.assemble()
    torch.cuda.synchronize()
    for r, o in zip(ref, outs):
        assert torch.equal(r, o.cpu())


def test_batch_autotune_cache(tmp_path):
    import torch

    items = _items(5, 4, [1, 2, 3])
    b = BatchAssembler([a for *_, a, _ in items], [s for _, s, _, _ in items])
    ref = [o.cpu().clone() for o in b.assemble()]
    cache = tmp_path / "tune.jsonl"
    ms, cached = b.autotune(cache, reps=2)
    assert not cached and ms > 0
    shape = b.launch()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and '"key"' in lines[0]
    b2 = BatchAssembler([a for *_, a, _ in items], [s for _, s, _, _ in items])
    ms2, cached2 = b2.autotune(cache, reps=2)
    assert cached2 and ms2 == pytest.approx(ms, rel=1e-5) and b2.launch() == shape
    outs = b2.assemble()
    torch.cuda.synchronize()
    for r, o in zip(ref, outs):
        assert torch.equal(r, o.cpu())


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batch_mixed_precision(dtype):
    """Mixed precision batches (FP32 far field, FP64 near/adjacent pairs; items with
    |k| (rad_i + rad_j) > 0.5 use the range-reduced FP32 phase): <= 1e-6 per item
    against the FP64 oracle."""
    import torch

    items = _items(13, 5, [2, 3], mixed=True, dtype=dtype)
    b = BatchAssembler([a for *_, a, _ in items], [s for _, s, _, _ in items])
    c = b.counts()
    print(f"[batch-mixed] counts {c}", flush=True)
    assert c["near_pairs"] > 0
    outs = b.assemble()
    torch.cuda.synchronize()
    for (g, s, a, k), o in zip(items, outs):
        R = O.assemble("helmholtz", "single_layer", g.points, g.cells, k=k)
        err = _rel(o.cpu().numpy().astype(np.complex128), R)
        print(f"[batch-mixed]   n={len(g.cells)} k={k:.3f} rel err {err:.2e}", flush=True)
        assert err <= 1e-6


def test_batch_errors():
    import torch

    g1, g2 = icosphere(1), icosphere(2)
    s1, s2 = _dp0(g1), _dp0(g2)
    sl = create_helmholtz_assembler(2.0, OperatorType.SingleLayer, _options())
    dl = create_helmholtz_assembler(2.0, OperatorType.DoubleLayer, _options())
    with pytest.raises(ValueError, match="share"):
        BatchAssembler([sl, dl], [s1, s2])
    p1 = function_space(g1, create_family(Family.Lagrange, 1, Continuity.Standard))
    with pytest.raises(ValueError, match="P0"):
        BatchAssembler([sl], [p1], outputs=[torch.empty((p1.global_size,) * 2, dtype=torch.complex128,
                                                        device="cuda").t()])
    b = BatchAssembler([sl, sl], [s1, s2])
    b.assemble()
    g2.set_points(g2.points * 1.5)
    with pytest.raises(ValueError, match="changed"):
        b.assemble()
    with pytest.raises(ValueError, match="chunk"):
        b.set_launch(2, 5, 0)
    with pytest.raises(ValueError, match="wavefronts"):
        b.set_launch(2, 8, 0, waves=3)
    with pytest.raises(ValueError, match="chunk 8"):
        b.set_launch(2, 4, 0, waves=8)
