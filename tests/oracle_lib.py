"""ctypes access to the CPU oracle (oracle/_build/libbem_oracle.so) for tests and the benchmark's
cpu_baseline leg. The oracle is the checker: nothing in the product imports this module."""
import ctypes
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
# BEMPP_ORACLE_LIBRARY selects another build (e.g. the sanitizer build, make -C oracle sanitize)
LIB_PATH = (Path(os.environ["BEMPP_ORACLE_LIBRARY"]) if os.environ.get("BEMPP_ORACLE_LIBRARY")
            else ROOT / "oracle" / "_build" / "libbem_oracle.so")
# the same oracle built for AVX-512 hosts (the timed CPU baseline)
LIB_PATH_AVX512 = ROOT / "oracle" / "_build" / "libbem_oracle_avx512.so"

OPS = {"single_layer": 0, "double_layer": 1, "adjoint_double_layer": 2, "hypersingular": 3}
KINDS = {"laplace": 0, "helmholtz": 1}

_lib = None
_path = LIB_PATH


def host_has_avx512():
    try:
        return " avx512f " in " " + Path("/proc/cpuinfo").read_text().replace("\n", " ") + " "
    except OSError:
        return False


def use_native_build():
    """Switch to the AVX-512 build when the host supports it and it is built; returns its ISA."""
    global _lib, _path
    if host_has_avx512() and LIB_PATH_AVX512.exists():
        _path, _lib = LIB_PATH_AVX512, None
    return isa()


def isa():
    return lib().oracle_isa().decode()


def lib():
    global _lib
    if _lib is None:
        if not _path.exists():
            raise RuntimeError(f"oracle library missing: {_path} (run __graft_entry__.build())")
        L = ctypes.CDLL(str(_path))
        i64, i32, dbl, vp = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_isa.restype = ctypes.c_char_p
        L.oracle_assemble_dense.argtypes = [i32, i32, dbl, vp, i64, vp, i64, i32, i32, i32, i32, i32, i32, i32,
                                            i32, i32, i64, i64, vp]
        L.oracle_assemble_singular_coo.argtypes = [i32, i32, dbl, vp, i64, vp, i64, i32, i32, i32, i32, i32, i32,
                                                   i32, vp, vp, vp, vp]
        L.oracle_sphere.argtypes = [i32, i32, vp, vp, vp, vp]
        L.oracle_space.argtypes = [vp, i64, vp, i64, i32, i32, vp, vp, vp, vp]
        L.oracle_regular_rule.argtypes = [i32, vp, vp]
        L.oracle_gauss_legendre01.argtypes = [i32, vp, vp]
        L.oracle_duffy_rule.argtypes = [i32, vp, vp, i32, vp, vp, vp]
        L.oracle_potential.argtypes = [i32, i32, dbl, vp, i64, vp, i64, i32, i32, i32, vp, i64, vp]
        L.oracle_assemble_cells.argtypes = [i32, i32, dbl, vp, i64, vp, vp, i64, i32, i32, i32, i32, i32, i32, i32,
                                            i32, i32, i64, i64, vp]
        L.oracle_cells_space.argtypes = [vp, i64, vp, vp, i64, i32, i32, vp, vp]
        L.oracle_cell_pair_rule.argtypes = [vp, i64, vp, vp, i64, i64, i64, i32, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise ValueError(lib().oracle_last_error().decode())


def sphere(level, seed="octahedron"):
    """Unit sphere mesh: seed 'octahedron' (src/shapes.rs:19-114) or 'icosahedron'."""
    s = 0 if seed == "octahedron" else 1
    nv, nc = ctypes.c_int64(), ctypes.c_int64()
    _check(lib().oracle_sphere(s, level, ctypes.byref(nv), ctypes.byref(nc), None, None))
    pts = np.zeros((nv.value, 3))
    cells = np.zeros((nc.value, 3), dtype=np.int64)
    _check(lib().oracle_sphere(s, level, None, None, _p(pts), _p(cells)))
    return pts, cells


def space(pts, cells, degree, continuous):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    nb = 1 if degree == 0 else 3
    size, ncol = ctypes.c_int64(), ctypes.c_int64()
    dofs = np.zeros((len(cells), nb), dtype=np.int64)
    colour = np.zeros(len(cells), dtype=np.int64)
    _check(lib().oracle_space(_p(pts), len(pts), _p(cells), len(cells), degree, int(continuous),
                              ctypes.byref(size), _p(dofs), _p(colour), ctypes.byref(ncol)))
    return size.value, dofs, colour, ncol.value


def _space_size(pts, cells, degree, continuous):
    return space(pts, cells, degree, continuous)[0]


def assemble(kind, op, pts, cells, test=(0, False), trial=(0, False), k=0.0, nq=6, nsing=4, batch=128,
             nthreads=0, fast=False, rows=None, out=None):
    """Dense assembly (accumulated into `out` if given). test/trial = (degree, continuous).
    rows = (first, end): only the test-DOF rows first..end-1 (any test space); `out` then has
    end - first rows."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    nt = _space_size(pts, cells, *test)
    ns = _space_size(pts, cells, *trial)
    rb, re = (0, nt) if rows is None else (int(rows[0]), min(int(rows[1]), nt))
    nt = re - rb
    dtype = np.float64 if kind == "laplace" else np.complex128
    if out is None:
        out = np.zeros((nt, ns), dtype=dtype, order="F")
    assert out.flags.f_contiguous and out.dtype == dtype and out.shape == (nt, ns)
    _check(lib().oracle_assemble_dense(KINDS[kind], OPS[op], float(k), _p(pts), len(pts), _p(cells), len(cells),
                                       test[0], int(test[1]), trial[0], int(trial[1]), nq, nsing, batch, nthreads,
                                       int(fast), rb, re, _p(out)))
    return out


def assemble_singular_coo(kind, op, pts, cells, test=(0, False), trial=(0, False), k=0.0, nsing=4, batch=128,
                          nthreads=0):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    cnt = ctypes.c_int64()
    args = [KINDS[kind], OPS[op], float(k), _p(pts), len(pts), _p(cells), len(cells), test[0], int(test[1]),
            trial[0], int(trial[1]), nsing, batch, nthreads]
    _check(lib().oracle_assemble_singular_coo(*args, ctypes.byref(cnt), None, None, None))
    n = cnt.value
    rows = np.zeros(n, dtype=np.int64)
    cols = np.zeros(n, dtype=np.int64)
    vals = np.zeros(n, dtype=np.float64 if kind == "laplace" else np.complex128)
    _check(lib().oracle_assemble_singular_coo(*args, ctypes.byref(cnt), _p(rows), _p(cols), _p(vals)))
    return rows, cols, vals


def potential(kind, op, pts, cells, xs, space=(0, False), k=0.0, nq=6):
    """Potential operator matrix [npoints, ndofs] (see oracle_potential)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, 3)
    ndofs = _space_size(pts, cells, *space)
    out = np.zeros((len(xs), ndofs), dtype=np.float64 if kind == "laplace" else np.complex128, order="F")
    _check(lib().oracle_potential(KINDS[kind], OPS[op], float(k), _p(pts), len(pts), _p(cells), len(cells),
                                  space[0], int(space[1]), nq, _p(xs), len(xs), _p(out)))
    return out


def regular_rule(npts):
    p = np.zeros(2 * max(npts, 1))
    w = np.zeros(max(npts, 1))
    q = lib().oracle_regular_rule(npts, _p(p), _p(w))
    if q < 0:
        raise ValueError(lib().oracle_last_error().decode())
    return p.reshape(-1, 2)[:q], w[:q]


def gauss_legendre01(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_gauss_legendre01(n, _p(x), _p(w))
    return x, w


def duffy_rule(pairs, n):
    tl = np.array([a for a, _ in pairs], dtype=np.int32)
    sl = np.array([b for _, b in pairs], dtype=np.int32)
    npts = lib().oracle_duffy_rule(len(pairs), _p(tl), _p(sl), n, None, None, None)
    if npts < 0:
        raise ValueError(lib().oracle_last_error().decode())
    tp = np.zeros((npts, 2))
    sp = np.zeros((npts, 2))
    w = np.zeros(npts)
    lib().oracle_duffy_rule(len(pairs), _p(tl), _p(sl), n, _p(tp), _p(sp), _p(w))
    return tp, sp, w


def _cells_arrays(cells):
    """cells: a list of vertex lists (3: triangle, 4: quadrilateral in tensor order) or an
    [n, 3] / [n, 4] array -> (flat vertices, types)."""
    cl = [list(map(int, c)) for c in cells]
    flat = np.array([v for c in cl for v in c], dtype=np.int64)
    types = np.array([2 if len(c) == 3 else 3 for c in cl], dtype=np.uint8)
    if any(len(c) not in (3, 4) for c in cl):
        raise ValueError("cells must have 3 or 4 vertices")
    return flat, types


def cells_space(pts, cells, degree, continuous):
    """(size, per-cell DOF lists) of a space on a grid of triangles and quadrilaterals."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    flat, types = _cells_arrays(cells)
    size = ctypes.c_int64()
    n = len(types) if degree == 0 else len(flat)
    dofs = np.zeros(n, dtype=np.int64)
    _check(lib().oracle_cells_space(_p(pts), len(pts), _p(flat), _p(types), len(types), degree, int(continuous),
                                    ctypes.byref(size), _p(dofs)))
    out, i = [], 0
    for t in types:
        nb = 1 if degree == 0 else (3 if t == 2 else 4)
        out.append([int(x) for x in dofs[i:i + nb]])
        i += nb
    return size.value, out


def assemble_cells(kind, op, pts, cells, test=(0, False), trial=(0, False), k=0.0, nq_tri=6, nq_quad=16, nsing=4,
                   nthreads=0, part="all", rows=None):
    """Dense assembly on a grid of triangles and quadrilaterals (the general-cell oracle);
    part: "all", "regular" (non-adjacent pairs), "singular" (adjacent pairs) or "correction"
    (adjacent pairs under the regular rule); rows=(r0, r1): those rows only (P0 test space)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    flat, types = _cells_arrays(cells)
    nt = cells_space(pts, cells, *test)[0]
    ns = cells_space(pts, cells, *trial)[0]
    r0, r1 = rows if rows is not None else (0, 0)
    dtype = np.float64 if kind == "laplace" else np.complex128
    out = np.zeros((r1 - r0 if rows is not None else nt, ns), dtype=dtype, order="F")
    _check(lib().oracle_assemble_cells(KINDS[kind], OPS[op], float(k), _p(pts), len(pts), _p(flat), _p(types),
                                       len(types), test[0], int(test[1]), trial[0], int(trial[1]), nq_tri, nq_quad,
                                       nsing, nthreads or os.cpu_count() or 1,
                                       {"all": 0, "regular": 1, "singular": 2, "correction": 3}[part], r0, r1,
                                       _p(out)))
    return out


def cell_pair_rule(pts, cells, test_cell, trial_cell, nsing):
    """Duffy rule of an adjacent cell pair in the cells' reference coordinates: (test points,
    trial points, weights)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    flat, types = _cells_arrays(cells)
    n = ctypes.c_int64()
    args = (_p(pts), len(pts), _p(flat), _p(types), len(types), int(test_cell), int(trial_cell), nsing)
    _check(lib().oracle_cell_pair_rule(*args, ctypes.byref(n), None, None, None))
    tp, sp, w = np.zeros((n.value, 2)), np.zeros((n.value, 2)), np.zeros(n.value)
    _check(lib().oracle_cell_pair_rule(*args, ctypes.byref(n), _p(tp), _p(sp), _p(w)))
    return tp, sp, w
