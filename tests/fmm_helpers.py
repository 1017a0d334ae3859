"""Regular-rule quantities written out in numpy for the singular-correction checks: the
mapped quadrature points of every cell, the kernel between all of them (zero at zero
distance, as a point-to-point fast multipole evaluation has it) and the basis-to-quadrature
map P (tests/fmm.rs.bak:22-95 builds the same three objects)."""
import numpy as np

import oracle_lib as O


def quadrature(points, cells, npts=6):
    """Mapped regular points x[c, p], weights w_p jdet_c, unit normals, reference points."""
    xi, w = O.regular_rule(npts)
    v = points[cells]
    j0, j1 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    cr = np.cross(j0, j1)
    jdet = np.linalg.norm(cr, axis=1)
    x = v[:, None, 0] + xi[None, :, 0, None] * j0[:, None] + xi[None, :, 1, None] * j1[:, None]
    return x, w[None, :] * jdet[:, None], cr / jdet[:, None], xi


def kernel_matrix(kind, op, k, x, nrm):
    """K[i, j] between all points (rows: test points, columns: trial points), 0 where r = 0."""
    d = x[:, None, :] - x[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = (np.exp(1j * k * r) if kind == "helmholtz" else 1.0) / (4 * np.pi * r)
        if op == "single_layer":
            K = g
        else:
            f = g * ((1 - 1j * k * r) if kind == "helmholtz" else 1.0) / r ** 2
            if op == "double_layer":
                K = f * np.einsum("ijc,jc->ij", d, nrm)
            else:
                K = -f * np.einsum("ijc,ic->ij", d, nrm)
    K[r == 0.0] = 0.0
    return K


def basis_to_quadrature(dofs, ndofs, xi, wj):
    """P[(c, p), dof] = phi_b(xi_p) w_p jdet_c summed over the cell's basis functions b."""
    nc, nb = dofs.shape
    Q = len(xi)
    phi = np.ones((Q, 1)) if nb == 1 else np.stack([1 - xi[:, 0] - xi[:, 1], xi[:, 0], xi[:, 1]], axis=1)
    P = np.zeros((nc * Q, ndofs))
    for c in range(nc):
        for b in range(nb):
            P[c * Q:(c + 1) * Q, dofs[c, b]] += phi[:, b] * wj[c]
    return P


def adjacency_mask(cells, Q):
    nc = len(cells)
    adjacent = np.zeros((nc, nc), dtype=bool)
    for c in range(nc):
        adjacent[c] = np.isin(cells, cells[c]).any(axis=1)
    return np.kron(adjacent, np.ones((Q, Q), dtype=bool))


def regular_parts(kind, op, k, points, cells, dofs, ndofs, npts=6):
    """(P^T K P, regular-rule sums of the adjacent pairs only)."""
    x, wj, n, xi = quadrature(points, cells, npts)
    Q = x.shape[1]
    P = basis_to_quadrature(dofs, ndofs, xi, wj)
    K = kernel_matrix(kind, op, k, x.reshape(-1, 3), np.repeat(n, Q, axis=0))
    return P.T @ K @ P, P.T @ np.where(adjacency_mask(cells, Q), K, 0.0) @ P
