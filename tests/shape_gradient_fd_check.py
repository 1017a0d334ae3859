"""Shape-gradient check of the full configuration: d(v^T A(X) u)/dX for Helmholtz SL P0,
k = 3, level-5 icosphere radially perturbed by 0.05 sum_{l<=4} a_lm Y_lm (seed 2024), u and v
complex normal (same seed), against central finite differences (h = 1e-6) of the oracle's
FP64 assembly along 32 random unit directions in R^{3 N_v}.

Two steps, because the finite differences cost two full CPU assemblies per direction:
  python tests/shape_gradient_fd_check.py gpu OUT.npz     # GPU: gradient (and the inputs)
  python tests/shape_gradient_fd_check.py fd OUT.npz      # CPU: finite differences, report
The pytest suite runs the same comparison on 3 directions (test_shape_gradient_fd_l5_config).
"""
import argparse
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path[:0] = [str(HERE.parent / "bempp-rs_amd"), str(HERE)]

import shapes_helpers as S  # noqa: E402

SEED, LEVEL, K, H, NDIR = 2024, 5, 3.0, 1e-6, 32


def inputs():
    import oracle_lib as O

    pts0, cells = O.sphere(LEVEL, "icosahedron")
    pts = S.perturb_radially(pts0, SEED)
    rng = np.random.default_rng(SEED)
    u, v = S.complex_normal(len(cells), rng), S.complex_normal(len(cells), rng)
    return pts, cells, u, v


def gpu(out):
    from bempp_amd.assembly.boundary import BoundaryAssemblerOptions, OperatorType, create_helmholtz_assembler
    from bempp_amd.element import Continuity, Family, ReferenceCellType, create_family
    from bempp_amd.function_space import function_space
    from bempp_amd.grid import Grid

    pts, cells, u, v = inputs()
    g = Grid(pts, cells)
    s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
    o = BoundaryAssemblerOptions()
    o.set_regular_quadrature_degree(ReferenceCellType.Triangle, 6)
    o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, 4)
    a = create_helmholtz_assembler(K, OperatorType.SingleLayer, o)
    t0 = time.perf_counter()
    G = a.shape_gradient(s, s, u, v)
    dt = time.perf_counter() - t0
    L = v @ a.apply(s, s, u)
    np.savez(out, pts=pts, cells=cells, u=u, v=v, grad=G, form=L, seconds=dt)
    print(f"gradient of v^T A u (= {L:.6e}) for {len(pts)} vertices in {dt:.3f} s -> {out}", flush=True)


def fd(inp, ndir, nthreads):
    import oracle_lib as O

    d = np.load(inp)
    pts, cells, u, v, G = d["pts"], d["cells"], d["u"], d["v"], d["grad"]
    ref_pts, ref_cells, ref_u, ref_v = inputs()
    assert np.array_equal(pts, ref_pts) and np.array_equal(cells, ref_cells) and np.array_equal(u, ref_u)

    def form(p):
        total = 0.0
        for r0 in range(0, len(cells), 4096):
            r1 = min(len(cells), r0 + 4096)
            A = O.assemble("helmholtz", "single_layer", p, cells, k=K, rows=(r0, r1), nthreads=nthreads)
            total += v[r0:r1] @ (A @ u)
        return total

    rng = np.random.default_rng(SEED + 1)
    worst = 0.0
    for i in range(ndir):
        D = rng.standard_normal(pts.shape)
        D /= np.linalg.norm(D)
        t0 = time.perf_counter()
        fdv = (form(pts + H * D) - form(pts - H * D)) / (2 * H)
        an = np.sum(G * D)
        rel = abs(an - fdv) / abs(fdv)
        worst = max(worst, rel)
        print(f"direction {i:2d}: analytic {an:.10e} fd {fdv:.10e} rel {rel:.2e} ({time.perf_counter() - t0:.1f} s)",
              flush=True)
    print(f"worst relative difference over {ndir} directions: {worst:.3e}", flush=True)
    return worst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("step", choices=["gpu", "fd"])
    ap.add_argument("file")
    ap.add_argument("--directions", type=int, default=NDIR)
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.step == "gpu":
        gpu(args.file)
    else:
        worst = fd(args.file, args.directions, args.threads)
        sys.exit(0 if worst <= 1e-6 else 1)


if __name__ == "__main__":
    main()
