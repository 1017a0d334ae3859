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


def _cell_points(points, cell, xi):
    """Mapped points, Jacobians |x_s x x_t|, unit normals, basis values and J-scaled surface
    curls (t_t d_s phi - t_s d_t phi) of a triangle or bilinear quadrilateral at reference points xi."""
    V = points[cell]
    s, t = xi[:, 0:1], xi[:, 1:2]
    e1, e2 = V[1] - V[0], V[2] - V[0]
    tw = V[3] - V[1] - V[2] + V[0] if len(cell) == 4 else np.zeros(3)
    X = V[0] + s * e1 + t * e2 + s * t * tw
    ts, tt = e1 + t * tw, e2 + s * tw
    cr = np.cross(ts, tt)
    J = np.linalg.norm(cr, axis=1)
    s, t = s[:, 0], t[:, 0]
    if len(cell) == 4:
        phi = np.stack([(1 - s) * (1 - t), s * (1 - t), (1 - s) * t, s * t], axis=1)
        ds = np.stack([-(1 - t), 1 - t, -t, t], axis=1)
        dt = np.stack([-(1 - s), -s, 1 - s, s], axis=1)
    else:
        phi = np.stack([1 - s - t, s, t], axis=1)
        ds = np.tile([-1.0, 1.0, 0.0], (len(s), 1))
        dt = np.tile([-1.0, 0.0, 1.0], (len(s), 1))
    curl = tt[:, None, :] * ds[:, :, None] - ts[:, None, :] * dt[:, :, None]
    return X, J, cr / J[:, None], phi, curl


def cells_regular_parts(kind, op, k, points, cells, degree, continuous, nq_tri=6, m=4):
    """On a grid of triangles and quadrilaterals (m x m Gauss-Legendre points per
    quadrilateral, the oracle's nq_tri-point rule per triangle): (P^T K P over all cell pairs,
    the same over adjacent pairs only), K zero at zero distance; degree 0 or 1 (P1 / Q1)."""
    g, gw = np.polynomial.legendre.leggauss(m)
    g, gw = 0.5 * (g + 1.0), 0.5 * gw
    qxi = np.array([[a, b] for a in g for b in g])
    qw = np.array([wa * wb for wa in gw for wb in gw])
    txi, tw = O.regular_rule(nq_tri)
    ndofs, dofs = O.cells_space(points, cells, degree, continuous)
    X, WJ, N, W, owner, PHI, CURL = [], [], [], [], [], [], []
    for c, cell in enumerate(cells):
        xi, w = (qxi, qw) if len(cell) == 4 else (txi, tw)
        x, J, n, phi, curl = _cell_points(points, np.asarray(cell), xi)
        if degree == 0:
            phi, curl = np.ones((len(w), 1)), np.zeros((len(w), 1, 3))
        for lst, v in ((X, x), (WJ, w * J), (N, n), (W, w), (PHI, phi), (CURL, curl)):
            lst.append(v)
        owner += [c] * len(w)
    X, N = np.concatenate(X), np.concatenate(N)
    owner = np.array(owner)
    K = kernel_matrix(kind, "single_layer" if op == "hypersingular" else op, k, X, N)
    npt = len(X)
    P = np.zeros((npt, ndofs))
    Pc = np.zeros((3, npt, ndofs))
    row = 0
    for c in range(len(cells)):
        nq = len(WJ[c])
        for b, dof in enumerate(dofs[c]):
            P[row:row + nq, dof] += PHI[c][:, b] * WJ[c]
            Pc[:, row:row + nq, dof] += (CURL[c][:, b, :] * W[c][:, None]).T
        row += nq
    adjacent = np.array([[bool(set(cells[a]) & set(cells[b])) for b in range(len(cells))] for a in range(len(cells))])
    mask = adjacent[owner][:, owner]

    def form(Km):
        if op != "hypersingular":
            return P.T @ Km @ P
        out = sum(Pc[d].T @ Km @ Pc[d] for d in range(3))
        if kind == "helmholtz":
            out = out - k * k * sum((P * N[:, d:d + 1]).T @ Km @ (P * N[:, d:d + 1]) for d in range(3))
        return out

    return form(K), form(np.where(mask, K, 0.0))
