"""Benchmark: assembled entries/s of the dense Helmholtz single-layer P0 operator on the
level-5 icosphere (20480 flat triangles, k = 5, 6-point regular rule, 4-point Duffy rule),
output device-resident (complex128, column-major 20480 x 20480 = 6.7 GB).

A step is one full dense assembly (regular-pair kernel + adjacent-pair kernel) into the
preallocated device matrix, with the mesh already resident in HBM. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "bempp-rs_amd"), str(ROOT / "tests")]

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
# nominal flops per regular quadrature-point pair of Helmholtz SL (x-y: 3, r^2: 5, sqrt: 1,
# 1/r: 1, kr: 1, sin, cos: 2, e^{ikr}/r: 2, weight: 1, complex accumulate: 4) and per pair epilogue
FLOP_PER_QPP = 20
FLOP_PER_PAIR_EPILOGUE = 4
# rocprofv3 summary of the FP64 tile kernel at the default workload (level 5, k = 5): its
# "traffic_bytes_per_launch:" line is FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 of one dispatch,
# the factors being what the counters read for a 1 GiB device copy (--copy-calibration),
# measured with the same build; it is rewritten whenever the kernel changes.
TILE_PROFILE = ROOT / "profiles" / "regular_p0_tiled_helmholtz_sl_p0_l5_k5.md"


def profiled_figure(key):
    """The number on the committed counter summary's "key: value" line, or None."""
    import re

    try:
        m = re.search(rf"^{key}:\s*([0-9.eE+]+)", TILE_PROFILE.read_text(), flags=re.M)
    except OSError:
        m = None
    return float(m.group(1)) if m else None


def profiled_traffic():
    """(bytes per launch, source) from the committed counter summary, or (None, reason)."""
    v = profiled_figure("traffic_bytes_per_launch")
    if v is None:
        return None, "no counter summary"
    return v, (f"{TILE_PROFILE.relative_to(ROOT)}: rocprofv3 --pmc FETCH_SIZE x 2.00 + WRITE_SIZE "
               "x 1.00 per dispatch (1 GiB copy calibration)")


def adjacent_ordered_pairs(cells, nv):
    import numpy as np

    vc = [[] for _ in range(nv)]
    for c, verts in enumerate(cells):
        for v in verts:
            vc[v].append(c)
    pairs = set()
    for lst in vc:
        for a in lst:
            for b in lst:
                pairs.add((a, b))
    return len(pairs)


def cpu_baseline(points, cells, k, target_s, threads):
    import numpy as np

    import oracle_lib as O

    O.use_native_build()  # the AVX-512 build on hosts that have it (as target-cpu=native would)

    n = len(cells)
    rows = 64
    t = 0.0
    while True:
        t0 = time.perf_counter()
        O.assemble("helmholtz", "single_layer", points, cells, k=k, fast=True, nthreads=threads, rows=(0, rows))
        t = time.perf_counter() - t0
        if t >= target_s or rows >= n:
            break
        rows = min(n, max(rows * 2, int(rows * target_s / max(t, 1e-3) * 1.05)))
    return rows * n / t, rows, t


def batch_items(seed, n, dtype, mixed, levels=(2, 5)):
    """The batched workload (bempp_amd.workloads.batch_spec): n ellipsoids, level uniform in
    levels (default {2..5}), semi-axes uniform in [0.5, 2], k uniform in [1, 10]."""
    from bempp_amd.workloads import batch_items as items

    return items(seed, n, dtype, mixed, levels=levels)


def batch_leg(seed, n, steps, mixed, levels=(2, 5)):
    """One persistent (graph-replayed, autotuned) launch for all items, timed, against one
    assembly launch per item; sample rows of three items checked against the oracle."""
    import numpy as np
    import torch

    import oracle_lib as O
    from bempp_amd.assembly.batch import BatchAssembler

    dtype = np.complex64 if mixed else np.complex128
    items = batch_items(seed, n, dtype, mixed, levels)
    entries = sum(len(g.cells) ** 2 for g, *_ in items)
    t0 = time.perf_counter()
    b = BatchAssembler([it[2] for it in items], [it[1] for it in items])
    setup_s = time.perf_counter() - t0
    tune_ms, _ = b.autotune(None, reps=2)
    b.set_graph(True)
    b.assemble()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.assemble()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    tuned = b.launch()
    # the untuned default shape, for the tuning gain
    b.set_launch()
    b.assemble()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.assemble()
    torch.cuda.synchronize()
    default_ms = (time.perf_counter() - t0) * 1e3 / steps
    b.set_launch(*tuned)
    b.assemble()
    torch.cuda.synchronize()
    # sample rows of the smallest, a median and the largest item against the FP64 oracle
    order = np.argsort([len(g.cells) for g, *_ in items], kind="stable")
    errs = []
    for idx in (order[0], order[len(order) // 2], order[-1]):
        g, s, a, k, _ = items[idx]
        rows = min(32, len(g.cells))
        ref = O.assemble("helmholtz", "single_layer", g.points, g.cells, k=k, rows=(0, rows))
        got = b.outputs[idx][:rows].cpu().numpy().astype(np.complex128)
        errs.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
    # one launch sequence per item into the same outputs, for comparison
    for (g, s, a, k, _), o in zip(items, b.outputs):
        a.assemble_into_device(s, s, o)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for (g, s, a, k, _), o in zip(items, b.outputs):
            a.assemble_into_device(s, s, o)
    torch.cuda.synchronize()
    per_item_ms = (time.perf_counter() - t0) * 1e3 / steps
    res = {"ms_per_step": ms, "value": entries / (ms * 1e-3), "unit": "entries/s",
           "per_item_launches_ms": per_item_ms, "speedup_vs_per_item_launches": per_item_ms / ms,
           "launch_shape": dict(zip(("blocks_per_cu", "chunk", "order", "waves"), tuned)), "autotune_best_ms": tune_ms,
           "default_shape_ms": default_ms,
           "counts": b.counts(), "setup_s": setup_s, "sample_rel_err_vs_oracle": max(errs),
           "output": "complex64" if mixed else "complex128"}
    del b, items
    torch.cuda.empty_cache()
    return entries, res


def shape_gradient_leg(level, warmup, steps):
    """Times the matrix-free shape gradient (regular-pair and adjacent-pair adjoint kernels and
    the vertex scatter) with device inputs; one step is one full gradient."""
    import ctypes

    import numpy as np
    import torch

    from bempp_amd import _lib as L
    from bempp_amd.assembly.boundary import BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler
    from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
    from bempp_amd.function_space import function_space
    from bempp_amd.grid import Grid
    from bempp_amd.shapes import icosphere
    from bempp_amd.workloads import complex_normal, perturb_radially

    g0 = icosphere(level)
    g = Grid(perturb_radially(g0.points, 0), g0.cells)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    a = create_helmholtz_assembler(3.0, OperatorType.SingleLayer, o)
    rng = np.random.default_rng(0)
    u, v = complex_normal(len(g.cells), rng), complex_normal(len(g.cells), rng)
    ref = a.shape_gradient(s, s, u, v)  # the allocating entry point (also builds the device data)
    du, dv = torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda()
    grad = torch.empty((len(g.points), 3), dtype=torch.complex128, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    args = (a._rs_assembler, s._rs_space, s._rs_space, ctypes.c_void_p(du.data_ptr()), ctypes.c_void_p(dv.data_ptr()),
            ctypes.c_void_p(grad.data_ptr()), ctypes.c_void_p(stream))
    for _ in range(max(1, warmup // 4)):
        L.check(L.lib().boundary_assembler_shape_gradient_noalloc(*args))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        L.check(L.lib().boundary_assembler_shape_gradient_noalloc(*args))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    diff = float(np.linalg.norm(grad.cpu().numpy() - ref) / np.linalg.norm(ref))
    n = len(g.cells)
    print(f"[bench] shape gradient L{level}: {ms:.3f} ms", file=sys.stderr, flush=True)
    return {"workload": f"d(v^T A(X) u)/dX, Helmholtz SL P0, k = 3, level-{level} icosphere perturbed radially by "
                        "0.05 sum_{l<=4} a_lm Y_lm (seed 0)", "cells": n, "ms_per_gradient": ms,
            "cell_pairs_per_s": n * n / (ms * 1e-3), "rel_diff_vs_allocating_entry_point": diff}


def singular_csr_leg(level, k, steps):
    """assemble_singular (the adjacent-pair part as a CSR matrix in host memory: Duffy kernel in
    COO mode, then keys, radix sort, duplicate sums and row pointers on the device) on the
    level-`level` icosphere: Helmholtz SL P0 (no duplicates) and Laplace HS P1 (duplicates)."""
    import numpy as np

    from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler,
                                             create_laplace_assembler)
    from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
    from bempp_amd.function_space import function_space
    from bempp_amd.shapes import icosphere

    g = icosphere(level)
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    out = {}
    for name, a, fam in [("helmholtz_sl_dp0", create_helmholtz_assembler(k, OperatorType.SingleLayer, o),
                          create_family(Family.Lagrange, 0, Continuity.Discontinuous)),
                         ("laplace_hs_p1", create_laplace_assembler(OperatorType.Hypersingular, o),
                          create_family(Family.Lagrange, 1, Continuity.Standard))]:
        s = function_space(g, fam)
        a.assemble_singular(s, s)  # builds the grid's and space's device data
        times = []
        for _ in range(steps):
            t0 = time.perf_counter()
            m = a.assemble_singular(s, s)
            times.append((time.perf_counter() - t0) * 1e3)
        out[name] = {"ms_per_call": float(np.median(times)), "nnz": int(m.nnz), "shape": list(m.shape)}
        print(f"[bench] singular CSR {name} L{level}: {np.median(times):.2f} ms, nnz {m.nnz}", file=sys.stderr,
              flush=True)
    return out


def quadrilateral_leg(m, k, steps):
    """Dense assembly on a grid of quadrilaterals: the cube's faces cut into m x m squares and
    projected onto the unit sphere (6 m^2 non-planar bilinear cells), Helmholtz SL P0, 4 x 4
    Gauss points per cell, 4-point Duffy rules (the general-cell kernels, FP64, device output)."""
    import numpy as np
    import torch

    from bempp_amd.assembly.boundary import BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler
    from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
    from bempp_amd.function_space import function_space
    from bempp_amd.grid import Grid

    g1 = np.linspace(-1.0, 1.0, m + 1)
    pts, index, cells = [], {}, []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            a, b = (axis + 1) % 3, (axis + 2) % 3
            if sign < 0:
                a, b = b, a
            for j in range(m):
                for i in range(m):
                    quad = []
                    for dj, di in ((0, 0), (0, 1), (1, 0), (1, 1)):
                        p = [0.0, 0.0, 0.0]
                        p[axis], p[a], p[b] = sign, g1[i + di], g1[j + dj]
                        key = tuple(round(x, 12) for x in p)
                        if key not in index:
                            index[key] = len(pts)
                            pts.append(p)
                        quad.append(index[key])
                    cells.append(quad)
    pts = np.array(pts)
    pts /= np.linalg.norm(pts, axis=1)[:, None]
    grid = Grid(pts, np.array(cells))
    space = function_space(grid, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Quadrilateral, 16)
    o.set_singular_quadrature_degree(ReferenceCellType.Quadrilateral, ReferenceCellType.Quadrilateral, 4)
    a = create_helmholtz_assembler(k, OperatorType.SingleLayer, o)
    n = space.global_size
    out = torch.empty((n, n), dtype=torch.complex128, device="cuda").t()
    a.assemble_into_device(space, space, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a.assemble_into_device(space, space, out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    print(f"[bench] quadrilateral grid ({n} cells): {ms:.2f} ms", file=sys.stderr, flush=True)
    return {"workload": f"Helmholtz SL P0, k = {k:g}, cubed sphere of {n} non-planar quadrilaterals "
                        f"({m} x {m} per face), 4 x 4 Gauss points, 4-point Duffy rules, complex128 device output",
            "cells": n, "ms_per_assembly": ms, "entries_per_s": n * n / (ms * 1e-3),
            "point_pairs_per_s": n * n * 256 / (ms * 1e-3)}


def copy_calibration():
    """A plain device copy of 1 GiB (read 1 GiB + write 1 GiB per dispatch), three times:
    the reference point that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_dispatch": n * 8, "dispatches": 3}))


# nominal flop per regular cell pair with the 6-point rule (36 point pairs), by (kernel,
# operator, space): the counting rule of FLOP_PER_QPP (+ - x = 1, FMA = 2, sqrt, div, sin,
# cos = 1) applied to each integrand and its basis products
NOMINAL_FLOP_PER_PAIR = {("laplace", "single_layer", "dp0"): 470, ("helmholtz", "single_layer", "dp0"): 724,
                         ("laplace", "double_layer", "p1"): 1485, ("laplace", "adjoint_double_layer", "p1"): 1485,
                         ("laplace", "hypersingular", "p1"): 524, ("helmholtz", "double_layer", "p1"): 2646,
                         ("helmholtz", "adjoint_double_layer", "p1"): 2646, ("helmholtz", "hypersingular", "p1"): 2408}


def config_suite(steps, warmup, only=None):
    """Every dense configuration of the workload set, timed on the GPU (HIP events around the
    kernels, one launch sequence per step, output in HBM) with 32 sample rows checked against
    the FP64 oracle: one JSON line per configuration."""
    import numpy as np
    import torch

    import oracle_lib as O
    from bempp_amd.assembly.boundary import (BoundaryAssemblerOptions, OperatorType, Precision,
                                             create_helmholtz_assembler, create_laplace_assembler)
    from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
    from bempp_amd.function_space import function_space
    from bempp_amd.shapes import icosphere

    ops = {"single_layer": OperatorType.SingleLayer, "double_layer": OperatorType.DoubleLayer,
           "adjoint_double_layer": OperatorType.AdjointDoubleLayer, "hypersingular": OperatorType.Hypersingular}
    cases = [("laplace", "single_layer", "dp0", 3, 0.0, "fp64", np.float64),
             ("laplace", "double_layer", "p1", 5, 0.0, "fp64", np.float64),
             ("laplace", "hypersingular", "p1", 5, 0.0, "fp64", np.float64),
             ("laplace", "double_layer", "p1", 5, 0.0, "mixed", np.float64),
             ("laplace", "hypersingular", "p1", 5, 0.0, "mixed", np.float64),
             ("helmholtz", "single_layer", "dp0", 5, 5.0, "fp64", np.complex128),
             ("helmholtz", "single_layer", "dp0", 5, 5.0, "mixed", np.complex64),
             ("helmholtz", "double_layer", "p1", 5, 5.0, "fp64", np.complex128),
             ("helmholtz", "double_layer", "p1", 5, 5.0, "mixed", np.complex64),
             ("helmholtz", "adjoint_double_layer", "p1", 5, 5.0, "fp64", np.complex128),
             ("helmholtz", "adjoint_double_layer", "p1", 5, 5.0, "mixed", np.complex64),
             ("helmholtz", "hypersingular", "p1", 5, 5.0, "fp64", np.complex128),
             ("helmholtz", "hypersingular", "p1", 5, 5.0, "mixed", np.complex64),
             ("laplace", "single_layer", "dp0", 6, 0.0, "fp64", np.float64)]
    grids = {}
    for kind, op, sp, level, k, prec, dt in cases:
        name = f"{kind}_{op}_{sp}_icosphere_l{level}" + (f"_k{k:g}" if kind == "helmholtz" else "")
        if only and not any(o in f"{name}_{prec}" for o in only.split(",")):
            continue
        g = grids.setdefault(level, icosphere(level))
        space = function_space(g, create_family(Family.Lagrange, 1 if sp == "p1" else 0,
                                                Continuity.Standard if sp == "p1" else Continuity.Discontinuous))
        o = BoundaryAssemblerOptions()
        o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
        o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
        if prec == "mixed":
            o.set_precision(Precision.Mixed)
        a = (create_laplace_assembler(ops[op], o, dtype=dt) if kind == "laplace"
             else create_helmholtz_assembler(k, ops[op], o, dtype=dt))
        a.set_profiling(True)
        n = space.global_size
        tdt = {np.float64: torch.float64, np.complex128: torch.complex128, np.complex64: torch.complex64}[dt]
        out = torch.empty((n, n), dtype=tdt, device="cuda").t()
        for _ in range(warmup):
            a.assemble_into_device(space, space, out)
        torch.cuda.synchronize()
        times = []
        t0 = time.perf_counter()
        for _ in range(steps):
            a.assemble_into_device(space, space, out)
            times.append(a.kernel_times())
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        fp64_diff = None
        if prec == "mixed":  # whole-matrix difference to the FP64 GPU assembly (parity-tested against the oracle)
            o64 = BoundaryAssemblerOptions()
            o64.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
            o64.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
            a64 = (create_laplace_assembler(ops[op], o64) if kind == "laplace"
                   else create_helmholtz_assembler(k, ops[op], o64))
            ref64 = a64.assemble_device(space, space)
            diff = (out.to(ref64.dtype) - ref64).norm() / ref64.norm()
            fp64_diff = float(diff)
            del ref64
        rows = min(32, n)
        deg = (1, True) if sp == "p1" else (0, False)
        ref = O.assemble(kind, op, g.points, g.cells, test=deg, trial=deg, k=k, rows=(0, rows))
        got = out[:rows].cpu().numpy().astype(ref.dtype)
        err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref)) if np.linalg.norm(ref) > 0 else 0.0
        reg_ms = float(np.median([t[0] for t in times]))
        nreg = len(g.cells) ** 2 - adjacent_ordered_pairs(g.cells, len(g.points))
        nominal = NOMINAL_FLOP_PER_PAIR.get((kind, op, sp))
        peak = FP64_PEAK_TFLOPS if prec == "fp64" else FP32_PEAK_TFLOPS
        rate = nreg * nominal / (reg_ms * 1e-3) / 1e12 if nominal else None
        line = {"config": name,
                "precision": prec, "output": np.dtype(dt).name, "dofs": n, "entries": n * n, "ms_per_assembly": ms,
                "entries_per_s": n * n / (ms * 1e-3),
                "nominal_tflops": rate, "nominal_frac": rate / peak if rate else None,
                "nominal_basis": f"{nreg} regular cell pairs x {nominal} flop / regular-kernel time, peak {peak}",
                "regular_kernels_ms": reg_ms,
                "adjacent_kernel_ms": float(np.median([t[1] for t in times])),
                "sample_rows": rows, "sample_rel_err_vs_oracle": err, "rel_diff_vs_fp64_gpu": fp64_diff}
        print(json.dumps(line), flush=True)
        del out
        torch.cuda.empty_cache()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # the first launches after idle run at lower clocks (kernel traces: ~6 dispatches to steady state)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--k", type=float, default=5.0)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-items", type=int, default=64)
    ap.add_argument("--batch-seed", type=int, default=0)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--configs", action="store_true",
                    help="time every dense configuration (one JSON line each) instead of the metric")
    ap.add_argument("--only", default=None,
                    help="with --configs: comma-separated substrings of '<config>_<precision>' to run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB copy used to calibrate the HBM byte counters")
    ap.add_argument("--quadrilateral", action="store_true",
                    help="only the quadrilateral-grid leg (9600-cell cubed sphere), one JSON line")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.quadrilateral:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps(quadrilateral_leg(40, args.k, max(1, args.steps))))
        return
    if args.configs:
        config_suite(max(1, min(args.steps, 5)), max(1, min(args.warmup, 3)), args.only)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine assembles on one GPU; multi-GPU runs are not part of this build "
                                   "(sharding by trial-column blocks is a design note only)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import oracle_lib as O
    from bempp_amd.assembly.boundary import BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler
    from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
    from bempp_amd.function_space import function_space
    from bempp_amd.shapes import icosphere

    torch.cuda.set_device(0)
    grid = icosphere(args.level)
    space = function_space(grid, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    n = space.global_size
    opts = BoundaryAssemblerOptions()
    opts.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    opts.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    asm = create_helmholtz_assembler(args.k, OperatorType.SingleLayer, opts)
    asm.set_profiling(True)
    out = torch.empty((n, n), dtype=torch.complex128, device="cuda").t()  # column-major n x n
    stream = torch.cuda.current_stream()
    print(f"[bench] icosphere L{args.level}: {n} cells, {n * n} entries; warmup {args.warmup}", file=sys.stderr,
          flush=True)
    # the first call on a new grid also builds its device data on the host and uploads it
    # (geometry, rule points, adjacent pairs, tile list): the cold first-call time
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    asm.assemble_into_device(space, space, out, stream=stream)
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(args.warmup):
        asm.assemble_into_device(space, space, out, stream=stream)
    torch.cuda.synchronize()

    reg_ms, sing_ms = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        asm.assemble_into_device(space, space, out, stream=stream)
        r, s = asm.kernel_times()  # waits for this step's kernels
        reg_ms.append(r)
        sing_ms.append(s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    entries = n * n
    value = entries / (ms_per_step * 1e-3)
    print(f"[bench] {ms_per_step:.3f} ms/step, regular kernel {np.mean(reg_ms):.3f} ms, adjacent kernel "
          f"{np.mean(sing_ms):.3f} ms", file=sys.stderr, flush=True)

    # rows kept for the oracle spot check, which runs on the host after all GPU legs (an idle GPU
    # would drop its clocks before the next leg)
    rows = 64
    got = out[:rows].cpu().numpy()

    # the first call above also loaded the library's code objects into the HIP runtime (once per
    # process); the first call on a second grid object of the same mesh is the per-grid set-up
    # (geometry, rule points, adjacent pairs, tile list, uploads) plus one assembly
    grid2 = icosphere(args.level)
    space2 = function_space(grid2, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    asm.assemble_into_device(space2, space2, out, stream=stream)
    torch.cuda.synchronize()
    new_grid_ms = (time.perf_counter() - t0) * 1e3
    asm.kernel_times()
    print(f"[bench] first call {cold_ms:.1f} ms (process), {new_grid_ms:.1f} ms (new grid)", file=sys.stderr,
          flush=True)

    # mixed precision: FP32 far pairs, FP64 near/adjacent pairs and accumulation; the
    # complex64-mixed form (k = 5 is exact in FP32) and, for comparison, complex128 output
    from bempp_amd.assembly.boundary import Precision

    mixed = {}
    for mdt, tdt in ((np.complex64, torch.complex64), (np.complex128, torch.complex128)):
        mopts = BoundaryAssemblerOptions()
        mopts.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
        mopts.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
        mopts.set_precision(Precision.Mixed)
        masm = create_helmholtz_assembler(args.k, OperatorType.SingleLayer, mopts, dtype=mdt)
        masm.set_profiling(True)
        mout = out if mdt == np.complex128 else torch.empty((n, n), dtype=tdt, device="cuda").t()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        masm.assemble_into_device(space, space, mout, stream=stream)
        torch.cuda.synchronize()
        mcold = (time.perf_counter() - t0) * 1e3  # first mixed call: builds the FP32 records and near pairs
        for _ in range(args.warmup):
            masm.assemble_into_device(space, space, mout, stream=stream)
        torch.cuda.synchronize()
        mreg, msing = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            masm.assemble_into_device(space, space, mout, stream=stream)
            r, s_ = masm.kernel_times()
            mreg.append(r)
            msing.append(s_)
        torch.cuda.synchronize()
        mms = (time.perf_counter() - t0) * 1e3 / args.steps
        mgot = mout[:rows].cpu().numpy().astype(np.complex128)
        key = "complex64" if mdt == np.complex64 else "complex128"
        mixed[key] = {"value": entries / (mms * 1e-3), "ms_per_step": mms, "speedup_vs_fp64": ms_per_step / mms,
                      "sample_rows": mgot, "regular_kernels_ms": float(np.mean(mreg)),
                      "adjacent_kernel_ms": float(np.mean(msing)), "first_call_ms": mcold,
                      "written_bytes_per_entry": 8 if mdt == np.complex64 else 16}
        print(f"[bench] mixed {key}: {mms:.3f} ms/step, far+near kernels {np.mean(mreg):.3f} ms, adjacent "
              f"{np.mean(msing):.3f} ms", file=sys.stderr, flush=True)
        if mout is not out:
            del mout
            torch.cuda.empty_cache()

    # shape gradient d(v^T A(X) u)/dX of the single layer (k = 3) on the level-5 icosphere moved
    # radially by 0.05 sum a_lm Y_lm (l <= 4, seed 0), u, v complex normal: matrix-free, device
    # inputs through the entry point that allocates nothing
    gradient = shape_gradient_leg(args.level, args.warmup, args.steps)
    singular_csr = singular_csr_leg(args.level, args.k, max(3, args.steps))
    quadrilaterals = quadrilateral_leg(40, args.k, 3)

    # the oracle spot check of the sample rows (first 64 rows: test cells 0..63 against all trial cells)
    ref = O.assemble("helmholtz", "single_layer", grid.points, grid.cells, k=args.k, fast=False, rows=(0, rows))
    sample_err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    for key, m in mixed.items():
        m["sample_rel_err_vs_oracle"] = float(np.linalg.norm(m.pop("sample_rows") - ref) / np.linalg.norm(ref))
        print(f"[bench] mixed {key}: sample rel err {m['sample_rel_err_vs_oracle']:.2e}", file=sys.stderr, flush=True)

    n_adj = adjacent_ordered_pairs(grid.cells, len(grid.points))
    n_reg = entries - n_adj
    flops_reg = n_reg * (36 * FLOP_PER_QPP + FLOP_PER_PAIR_EPILOGUE)
    achieved = flops_reg / (np.mean(reg_ms) * 1e-3) / 1e12
    # the tile kernel writes every entry once (complex128): algorithmic write rate
    write_gbps = entries * 16 / (np.mean(reg_ms) * 1e-3) / 1e9
    for m in mixed.values():  # the same nominal flops and entries over the far+near kernels' time
        m["roofline_fp32"] = {
            "achieved": flops_reg / (m["regular_kernels_ms"] * 1e-3) / 1e12, "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": flops_reg / (m["regular_kernels_ms"] * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
            "hbm_write_GBps": entries * m["written_bytes_per_entry"] / (m["regular_kernels_ms"] * 1e-3) / 1e9,
            "hbm_write_frac": entries * m["written_bytes_per_entry"] / (m["regular_kernels_ms"] * 1e-3) / 1e9
            / HBM_PEAK_GBPS}

    batched = None
    if not args.no_batch:
        del out
        torch.cuda.empty_cache()
        batched = {"workload": f"{args.batch_items} Helmholtz SL P0 assemblies: icospheres of level uniform in 2..5 "
                               f"scaled to ellipsoids (semi-axes uniform in [0.5, 2]), k uniform in [1, 10], "
                               f"numpy seed {args.batch_seed}"}
        for bmixed in (False, True):
            e, res = batch_leg(args.batch_seed, args.batch_items, args.batch_steps, bmixed)
            batched["entries"] = e
            batched["mixed_complex64" if bmixed else "fp64_complex128"] = res
            print(f"[bench] batch {'mixed c64' if bmixed else 'fp64 c128'}: {res['ms_per_step']:.2f} ms, per-item "
                  f"launches {res['per_item_launches_ms']:.2f} ms, sample err {res['sample_rel_err_vs_oracle']:.2e}",
                  file=sys.stderr, flush=True)
        # many small meshes, where one launch sequence per item is dominated by launch overhead
        small = {"workload": f"{args.batch_items} Helmholtz SL P0 assemblies on level-2 icospheres (320 cells) "
                             f"scaled to ellipsoids (semi-axes uniform in [0.5, 2]), k uniform in [1, 10], "
                             f"numpy seed {args.batch_seed}"}
        for bmixed in (False, True):
            e, res = batch_leg(args.batch_seed, args.batch_items, max(args.batch_steps, 20), bmixed, levels=(2, 2))
            small["entries"] = e
            small["mixed_complex64" if bmixed else "fp64_complex128"] = res
            print(f"[bench] small-mesh batch {'mixed c64' if bmixed else 'fp64 c128'}: {res['ms_per_step']:.3f} ms, "
                  f"per-item launches {res['per_item_launches_ms']:.3f} ms", file=sys.stderr, flush=True)
        batched["small_meshes"] = small

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        v, crow, ct = cpu_baseline(grid.points, grid.cells, args.k, args.cpu_seconds, threads)
        cpu = {"value": v, "unit": "entries/s", "cores": threads, "kind": "port", "isa": O.isa(),
               "sample": f"rows 0..{crow} of the same operator ({crow}x{n} entries incl. their Duffy pairs) "
                         f"in {ct:.1f} s, oracle built {O.isa()}, OpenMP {threads} threads"}

    traffic, traffic_src = profiled_traffic()
    on_profile = (args.level, args.k) == (5, 5.0)
    hw_flop = profiled_figure("fp64_flop_per_launch") if on_profile else None
    line = {
        "metric": "assembled entries/s, Helmholtz SL P0 20480 tris: fp64 vs mixed @<=1e-6 rel err",
        "value": value, "unit": "entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"helmholtz_single_layer_p0_icosphere_l{args.level}_k{args.k:g}", "cells": n,
                   "entries": entries, "regular_points": 6, "duffy_gauss_points": 4, "precision": "fp64",
                   "output": "device-resident complex128 column-major"},
        "sample_rel_err_vs_oracle": sample_err,
        "mixed": dict(mixed["complex64"], output="complex64", complex128_output=mixed["complex128"],
                      scheme="far cell pairs (R >= 3 (rad_i + rad_j)) in FP32: point pairs from FP32 centroid "
                             "offsets, summed in FP32 per weight class, FP32 per-cell-pair tail (class weights, "
                             "length scaling, rotation by e^{ikR}); FP64 centroid distance R and reduction of kR; "
                             "near non-adjacent and adjacent (Duffy) pairs in FP64"),
        "cold_first_call_ms": cold_ms,
        "setup_ms": cold_ms - ms_per_step,
        "new_grid_first_call_ms": new_grid_ms,
        "new_grid_setup_ms": new_grid_ms - ms_per_step,
        "kernels_ms": {"regular": float(np.mean(reg_ms)), "adjacent": float(np.mean(sing_ms))},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic if on_profile else None,
                     "traffic_source": traffic_src if on_profile else None,
                     # the FP64 flops the hardware executed (SQ_INSTS_VALU_FLOPS_FP64 x 64 of one
                     # dispatch, from the same summary) over this run's mean kernel time
                     "counter_tflops": hw_flop / (np.mean(reg_ms) * 1e-3) / 1e12 if hw_flop else None,
                     "counter_frac": hw_flop / (np.mean(reg_ms) * 1e-3) / 1e12 / FP64_PEAK_TFLOPS
                     if hw_flop else None,
                     "hbm_write_GBps": write_gbps, "hbm_write_frac": write_gbps / HBM_PEAK_GBPS,
                     "basis": f"{n_reg} regular pairs x (36 x {FLOP_PER_QPP} + {FLOP_PER_PAIR_EPILOGUE}) nominal "
                              "flop per launch / mean regular-kernel HIP-event time"},
        "cpu_baseline": cpu,
        "batched": batched,
        "shape_gradient": gradient,
        "singular_csr": singular_csr,
        "quadrilateral_grid": quadrilaterals,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
