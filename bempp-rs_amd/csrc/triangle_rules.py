"""Derives the fully symmetric triangle rules of 16 points (degree 8) and 37 points
(degree 13) used for regular quadrature, and prints them as orbit tables for host.cpp.

The reference takes its regular rules from ndelement (src/boundary_assemblers.rs:419-464);
its defaults are 37 points (:43-56) and its example and benchmark use 16. ndelement's
tables are not available here, so the rules are derived from their moment equations:

- orbits on the reference triangle in barycentric coordinates: the centroid (S3), S21 orbits
  (a, a, 1 - 2a) of 3 points and S111 orbits (a, b, 1 - a - b) of 6 points;
- a fully symmetric rule integrates every polynomial of degree <= d exactly iff it integrates
  the S3-invariant ones, spanned by e2^i e3^j (e2 = l1 l2 + l2 l3 + l3 l1, e3 = l1 l2 l3,
  2i + 3j <= d); their exact integrals follow from int l1^a l2^b l3^c = a! b! c! / (a+b+c+2)!;
- 16 points: centroid + 3 S21 + 1 S111 (10 unknowns, 10 equations at degree 8);
  37 points: centroid + 4 S21 + 4 S111 (21 unknowns, 21 equations at degree 13);
- the equations are taken in an orthonormal basis of the invariants (Gram matrix of exact
  integrals, Cholesky at 60 digits), solved by Levenberg-Marquardt from seeded random starts,
  filtered for positive weights and interior points, polished by Newton's method at 60 digits;
- degree 8 has one such solution (Dunavant's degree-8 rule); degree 13 has several, and the
  one with the smallest truncation error -- the norm of the orthonormal degree-14 residuals,
  the criterion Witherden and Vincent rank their rules by -- is taken.

Whether these equal ndelement's tables digit for digit is not known (parity unpinned); the
rules are pinned by exact integration of all monomials of degree <= d and failure at d + 1.

Usage: python triangle_rules.py   (about 5 minutes; prints the C++ table)
"""
import sys
from math import factorial

import mpmath as mp
import numpy as np
import sympy as sp

mp.mp.dps = 60
RULES = [(16, 8, (1, 3, 1)), (37, 13, (1, 4, 4))]


def invariants(d):
    return [(i, j) for j in range(d // 3 + 1) for i in range((d - 3 * j) // 2 + 1)]


_cache = {}


def inv_integral(i, j):
    """Exact integral of e2^i e3^j over the reference triangle (area 1/2)."""
    if (i, j) not in _cache:
        l1, l2, l3 = sp.symbols("l1 l2 l3")
        poly = sp.Poly(sp.expand((l1 * l2 + l2 * l3 + l3 * l1) ** i * (l1 * l2 * l3) ** j), l1, l2, l3)
        tot = sp.Rational(0)
        for (a, b, c), coef in poly.terms():
            tot += coef * sp.Rational(factorial(a) * factorial(b) * factorial(c), factorial(a + b + c + 2))
        _cache[(i, j)] = tot
    v = _cache[(i, j)]
    return mp.mpf(v.p) / v.q


def orthonormal(d):
    """Invariant exponents, L^-1 (rows: orthonormal basis in terms of e2^i e3^j), moments."""
    ij = invariants(d)
    n = len(ij)
    G = mp.matrix(n, n)
    for a in range(n):
        for b in range(n):
            G[a, b] = inv_integral(ij[a][0] + ij[b][0], ij[a][1] + ij[b][1])
    Linv = mp.cholesky(G) ** -1
    return ij, Linv, [inv_integral(i, j) for i, j in ij]


def orbits_of(x, n0, n21, n111):
    """[(barycentric point, weight, multiplicity)] from the unknown vector."""
    out, k = [], 0
    if n0:
        out.append(((mp.mpf(1) / 3,) * 3, x[0], 1))
        k = 1
    for _ in range(n21):
        a, w = x[k], x[k + 1]
        k += 2
        out.append(((a, a, 1 - 2 * a), w, 3))
    for _ in range(n111):
        a, b, w = x[k], x[k + 1], x[k + 2]
        k += 3
        out.append(((a, b, 1 - a - b), w, 6))
    return out


def residual_mp(x, struct, ij, Linv, mom):
    r = [-m for m in mom]
    for (l1, l2, l3), w, m in orbits_of(x, *struct):
        e2, e3 = l1 * l2 + l2 * l3 + l3 * l1, l1 * l2 * l3
        for t, (i, j) in enumerate(ij):
            r[t] += m * w * e2 ** i * e3 ** j
    return list(Linv * mp.matrix(r))


def search(d, struct, starts, seed=0, iters=400):
    """Levenberg-Marquardt on many starts at once (float64, orthonormal residuals)."""
    n0, n21, n111 = struct
    ij, Linv, mom = orthonormal(d)
    T = np.array([[float(Linv[a, b]) for b in range(len(ij))] for a in range(len(ij))])
    I = np.array([float(m) for m in mom])
    P = np.array([i for i, j in ij], float)
    Qe = np.array([j for i, j in ij], float)
    nv = n0 + 2 * n21 + 3 * n111
    rng = np.random.default_rng(seed)
    wavg = 0.5 / (n0 + 3 * n21 + 6 * n111)
    X = np.zeros((starts, nv))
    k = 0
    if n0:
        X[:, 0] = wavg * rng.uniform(0.3, 2, starts)
        k = 1
    for i in range(n21):
        X[:, k] = (i + rng.uniform(0.05, 0.95, starts)) * 0.5 / n21
        X[:, k + 1] = wavg * rng.uniform(0.3, 2, starts)
        k += 2
    for _ in range(n111):
        L = np.sort(rng.dirichlet([1, 1, 1], starts), axis=1)
        X[:, k], X[:, k + 1], X[:, k + 2] = L[:, 0], L[:, 1], wavg * rng.uniform(0.3, 2, starts)
        k += 3

    def fg(l1, l2, l3):
        E2 = (l1 * l2 + l2 * l3 + l3 * l1)[:, None]
        E3 = (l1 * l2 * l3)[:, None]
        f = E2 ** P * E3 ** Qe
        d2 = P * np.where(P > 0, E2 ** np.maximum(P - 1, 0), 0) * E3 ** Qe
        d3 = Qe * E2 ** P * np.where(Qe > 0, E3 ** np.maximum(Qe - 1, 0), 0)
        g = [d2 * (l2 + l3)[:, None] + d3 * (l2 * l3)[:, None], d2 * (l1 + l3)[:, None] + d3 * (l1 * l3)[:, None],
             d2 * (l1 + l2)[:, None] + d3 * (l1 * l2)[:, None]]
        return f, g

    def rj(X):
        r = np.tile(-I, (len(X), 1))
        J = np.zeros((len(X), len(ij), nv))
        k = 0
        if n0:
            c = np.full(len(X), 1 / 3)
            f, _ = fg(c, c, c)
            r += X[:, :1] * f
            J[:, :, 0] = f
            k = 1
        for _ in range(n21):
            a, w = X[:, k], X[:, k + 1]
            f, g = fg(a, a, 1 - 2 * a)
            r += 3 * w[:, None] * f
            J[:, :, k] = 3 * w[:, None] * (g[0] + g[1] - 2 * g[2])
            J[:, :, k + 1] = 3 * f
            k += 2
        for _ in range(n111):
            a, b, w = X[:, k], X[:, k + 1], X[:, k + 2]
            f, g = fg(a, b, 1 - a - b)
            r += 6 * w[:, None] * f
            J[:, :, k] = 6 * w[:, None] * (g[0] - g[2])
            J[:, :, k + 1] = 6 * w[:, None] * (g[1] - g[2])
            J[:, :, k + 2] = 6 * f
            k += 3
        return r @ T.T, np.einsum("ab,nbv->nav", T, J)

    lam = np.full(starts, 1e-3)
    r, J = rj(X)
    cost = (r ** 2).sum(1)
    with np.errstate(all="ignore"):
        for _ in range(iters):
            JT = np.transpose(J, (0, 2, 1))
            A = JT @ J
            g = (JT @ r[:, :, None])[:, :, 0]
            D = np.einsum("bii->bi", A)
            A2 = A + lam[:, None, None] * np.eye(nv)[None] * (D[:, :, None] + 1e-12)
            try:
                dx = -np.linalg.solve(A2, g[:, :, None])[:, :, 0]
            except np.linalg.LinAlgError:
                break
            rn, Jn = rj(X + dx)
            cn = (rn ** 2).sum(1)
            ok = np.isfinite(cn) & (cn < cost)
            X[ok] += dx[ok]
            r[ok], J[ok], cost[ok] = rn[ok], Jn[ok], cn[ok]
            lam = np.clip(np.where(ok, lam * 0.3, lam * 4), 1e-15, 1e10)
    return [X[b] for b in np.where(cost < 1e-20)[0]]


def canonical(x, struct):
    """Orbit parameters sorted: S21 by a; S111 as (l_min, l_mid) sorted; weights kept."""
    n0, n21, n111 = struct
    k = n0
    c0 = [x[0]] if n0 else []
    s21, s111 = [], []
    for _ in range(n21):
        s21.append((x[k], x[k + 1]))
        k += 2
    for _ in range(n111):
        lam = sorted([x[k], x[k + 1], 1 - x[k] - x[k + 1]])
        s111.append((lam[0], lam[1], x[k + 2]))
        k += 3
    s21.sort()
    s111.sort()
    out = list(c0)
    for a, w in s21:
        out += [a, w]
    for a, b, w in s111:
        out += [a, b, w]
    return out


def positive_interior(x, struct):
    orbs = orbits_of([mp.mpf(v) for v in x], *struct)
    for (l1, l2, l3), w, m in orbs:
        if w <= 0 or min(l1, l2, l3) <= 0:
            return False
        if m == 3 and abs(l1 - mp.mpf(1) / 3) < 1e-6:
            return False
        if m == 6 and min(abs(l1 - l2), abs(l2 - l3), abs(l1 - l3)) < 1e-6:
            return False
    return True


def polish(x, struct, ij, Linv, mom):
    f = lambda *v: residual_mp(list(v), struct, ij, Linv, mom)  # noqa: E731
    return list(mp.findroot(f, [mp.mpf(v) for v in x], tol=mp.mpf(10) ** -50, maxsteps=100))


def truncation(x, struct, d):
    """Norm of the orthonormal residuals new at degree d + 1."""
    ij, Linv, mom = orthonormal(d + 1)
    r = residual_mp(x, struct, ij, Linv, mom)
    new = [t for t, (i, j) in enumerate(ij) if 2 * i + 3 * j == d + 1]
    return mp.sqrt(sum(r[t] ** 2 for t in new))


def derive(npts, d, struct, starts=6000):
    ij, Linv, mom = orthonormal(d)
    found = []
    for x in search(d, struct, starts):
        c = canonical(list(x), struct)
        if not positive_interior(c, struct):
            continue
        if any(np.allclose(c, f, atol=1e-8) for f in found):
            continue
        found.append(c)
    polished = []
    for c in found:
        p = canonical(polish(c, struct, ij, Linv, mom), struct)
        if positive_interior(p, struct) and max(abs(v) for v in residual_mp(p, struct, ij, Linv, mom)) < 1e-45:
            if not any(max(abs(a - b) for a, b in zip(p, q)) < 1e-30 for q in polished):
                polished.append(p)
    ranked = sorted(polished, key=lambda p: truncation(p, struct, d))
    print(f"// {npts} points, degree {d}: {len(ranked)} positive interior solution(s); truncation norms "
          + ", ".join(mp.nstr(truncation(p, struct, d), 5) for p in ranked), file=sys.stderr)
    return ranked[0]


def emit(npts, d, struct, x):
    n0, n21, n111 = struct
    lines = [f"// {npts} points, degree {d}: centroid x {n0}, S21 x {n21}, S111 x {n111}"]
    k = 0
    if n0:
        lines.append(f"{{0, {mp.nstr(mp.mpf(1) / 3, 21)}, 0.0, {mp.nstr(x[0], 21)}}},")
        k = 1
    for _ in range(n21):
        lines.append(f"{{1, {mp.nstr(x[k], 21)}, 0.0, {mp.nstr(x[k + 1], 21)}}},")
        k += 2
    for _ in range(n111):
        lines.append(f"{{2, {mp.nstr(x[k], 21)}, {mp.nstr(x[k + 1], 21)}, {mp.nstr(x[k + 2], 21)}}},")
        k += 3
    return "\n".join(lines)


if __name__ == "__main__":
    for npts, d, struct in RULES:
        print(emit(npts, d, struct, derive(npts, d, struct)))
