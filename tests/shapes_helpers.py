"""Perturbed-sphere meshes for the shape-gradient checks: r = 1 + eps * sum_{l<=4,m} a_lm Y_lm
(real orthonormal spherical harmonics, a_lm ~ N(0,1) from a seed), applied radially to the
vertices of a level-L icosphere."""
import numpy as np
from scipy.special import sph_harm_y


def real_sph_harm(l, m, theta, phi):
    """Real orthonormal Y_lm; theta polar angle, phi azimuth."""
    if m == 0:
        return np.real(sph_harm_y(l, 0, theta, phi))
    y = sph_harm_y(l, abs(m), theta, phi)
    return np.sqrt(2.0) * (-1) ** m * (np.imag(y) if m < 0 else np.real(y))


def perturb_radially(points, seed, eps=0.05, lmax=4):
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
    """u ~ CN(0,1) with real and imaginary parts N(0, 1/2) each."""
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2.0)
