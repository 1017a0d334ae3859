"""Synthetic workloads of the benchmark configurations, shared by the benchmark and the tests
so that both draw identical inputs.

- `batch_spec(seed, n)`: the batched workload, n ellipsoids with icosphere level uniform in
  {2, 3, 4, 5}, semi-axes uniform in [0.5, 2] and wavenumber uniform in [1, 10]
  (numpy default_rng(seed), drawn in that order per item).
- `batch_items(...)`: the same as (grid, DP0 space, assembler, k, level) tuples.
- `perturb_radially(points, seed)`: the shape-gradient workload's surface, r = 1 + 0.05 sum of
  the real orthonormal spherical harmonics Y_lm (l <= 4) with coefficients N(0, 1) from a seed.
"""
import numpy as np


def batch_spec(seed, n, levels=(2, 5)):
    """[(level, semi_axes[3], k)] for n items; level uniform in levels[0]..levels[1]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        lev = int(rng.integers(levels[0], levels[1] + 1))
        axes = rng.uniform(0.5, 2.0, 3)
        k = float(rng.uniform(1.0, 10.0))
        out.append((lev, axes, k))
    return out


def batch_items(seed, n, dtype=np.complex128, mixed=False, nq=6, nsing=4, levels=(2, 5)):
    """Helmholtz single-layer P0 assemblers on the batch_spec meshes: [(grid, space, assembler,
    k, level)]. A complex64 assembler holds k in FP32; the k returned is that FP32 value, so a
    reference assembly with it uses the same wavenumber."""
    from .assembly.boundary import BoundaryAssemblerOptions, OperatorType, Precision, create_helmholtz_assembler
    from .element import Continuity, Family, ReferenceCellType, create_family
    from .function_space import function_space
    from .shapes import ellipsoid

    items = []
    for lev, axes, k in batch_spec(seed, n, levels):
        if np.dtype(dtype) == np.complex64:
            k = float(np.float32(k))
        o = BoundaryAssemblerOptions()
        o.set_regular_quadrature_degree(ReferenceCellType.Triangle, nq)
        o.set_singular_quadrature_degree(ReferenceCellType.Triangle, ReferenceCellType.Triangle, nsing)
        if mixed:
            o.set_precision(Precision.Mixed)
        g = ellipsoid(lev, axes)
        s = function_space(g, create_family(Family.Lagrange, 0, Continuity.Discontinuous))
        items.append((g, s, create_helmholtz_assembler(k, OperatorType.SingleLayer, o, dtype=dtype), k, lev))
    return items


def real_sph_harm(l, m, theta, phi):
    """Real orthonormal Y_lm; theta polar angle, phi azimuth."""
    from scipy.special import sph_harm_y

    if m == 0:
        return np.real(sph_harm_y(l, 0, theta, phi))
    y = sph_harm_y(l, abs(m), theta, phi)
    return np.sqrt(2.0) * (-1) ** m * (np.imag(y) if m < 0 else np.real(y))


def perturb_radially(points, seed, eps=0.05, lmax=4):
    """Vertices moved radially to r = 1 + eps sum_{l <= lmax, m} a_lm Y_lm, a_lm ~ N(0, 1)."""
    rng = np.random.default_rng(seed)
    r = np.linalg.norm(points, axis=1)
    theta = np.arccos(np.clip(points[:, 2] / r, -1, 1))
    phi = np.arctan2(points[:, 1], points[:, 0])
    f = np.zeros(len(points))
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            f += rng.standard_normal() * real_sph_harm(l, m, theta, phi)
    return points * (1.0 + eps * f)[:, None] / r[:, None]


def complex_normal(n, rng):
    """u ~ CN(0, 1): real and imaginary parts N(0, 1/2) each."""
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2.0)
