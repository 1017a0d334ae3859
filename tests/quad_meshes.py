"""Test meshes with quadrilateral cells (tensor vertex order V0..V3 at reference (0,0), (1,0),
(0,1), (1,1), as ndgrid orders them; src/shapes.rs:178-223 builds its screens the same way).

- `cube(m)`: the surface of [-1, 1]^3, m x m flat square quadrilaterals per face, normals
  (t_xi x t_eta) pointing outwards.
- `cubed_sphere(m)`: the same vertices projected onto the unit sphere: non-planar bilinear
  quadrilaterals.
- `split(cells, which)`: the quadrilaterals selected by `which` replaced by the triangles
  (V0, V1, V3), (V0, V3, V2) (same orientation): mixed grids.
- `mixed_screen()`: the mixed grid of the reference's tests/mixed_assembly.rs.bak:15-31
  (two quadrilaterals and four triangles on [0, 1]^2) and its single-type parts.
"""
import numpy as np


def cube(m):
    pts, index, cells = [], {}, []

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in index:
            index[key] = len(pts)
            pts.append(p)
        return index[key]

    g = np.linspace(-1.0, 1.0, m + 1)
    for axis in range(3):
        for sign in (-1.0, 1.0):
            # face param (u, v) with e_u x e_v = sign e_axis
            a, b = (axis + 1) % 3, (axis + 2) % 3
            if sign < 0:
                a, b = b, a
            for j in range(m):
                for i in range(m):
                    quad = []
                    for dj, di in ((0, 0), (0, 1), (1, 0), (1, 1)):
                        p = np.zeros(3)
                        p[axis] = sign
                        p[a] = g[i + di]
                        p[b] = g[j + dj]
                        quad.append(vid(p))
                    cells.append(quad)
    return np.array(pts), cells


def cubed_sphere(m):
    pts, cells = cube(m)
    return pts / np.linalg.norm(pts, axis=1)[:, None], cells


def split(cells, which):
    out = []
    for i, c in enumerate(cells):
        if len(c) == 4 and which(i):
            out += [[c[0], c[1], c[3]], [c[0], c[3], c[2]]]
        else:
            out.append(list(c))
    return out


def mixed_screen():
    pts = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.5, 0.0], [0.5, 0.5, 0.0],
                    [1.0, 0.5, 0.0], [0.0, 1.0, 0.0], [0.5, 1.0, 0.0], [1.0, 1.0, 0.0]])
    quads = [[0, 1, 3, 4], [3, 4, 6, 7]]
    tris = [[1, 2, 5], [1, 5, 4], [4, 5, 8], [4, 8, 7]]
    return pts, quads + tris, quads, tris


def orient_outward(pts, cells):
    """Cells of a closed surface around the origin, checked: the centroid-to-normal direction
    (t_xi x t_eta at the reference centre) points away from the origin."""
    for c in cells:
        v = pts[c]
        if len(c) == 3:
            n = np.cross(v[1] - v[0], v[2] - v[0])
        else:
            n = np.cross(0.5 * (v[1] - v[0] + v[3] - v[2]), 0.5 * (v[2] - v[0] + v[3] - v[1]))
        assert np.dot(n, v.mean(axis=0)) > 0
