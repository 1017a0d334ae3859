"""Extract the bempp-cl golden matrices held in the reference's tests as JSON data.

Source (read as text, nothing executed): tests/compare_to_bempp_cl.rs:19-206 (boundary
operators) and :208-320 (potential operators, commented out there) of the reference. Each `from_cl` literal is a dense matrix that bempp-cl assembled on the
octahedron sphere of level 0 (8 triangles, 6 vertices). P1 matrices are compared
through the vertex permutation perm = [0, 5, 2, 4, 3, 1] (compare_to_bempp_cl.rs:102,195):
matrix[i][j] == from_cl[perm[i]][perm[j]]. Tolerances are absolute (approx epsilon).

Run: python tests/golden/make_bempp_cl_goldens.py /root/reference
"""
import json
import re
import sys
from pathlib import Path

CASES = [
    # (test fn name, kind, operator, space degree, k, epsilon)
    ("test_laplace_single_layer_dp0_dp0", "laplace", "single_layer", 0, 0.0, 1e-3),
    ("test_laplace_double_layer_dp0_dp0", "laplace", "double_layer", 0, 0.0, 1e-4),
    ("test_laplace_adjoint_double_layer_dp0_dp0", "laplace", "adjoint_double_layer", 0, 0.0, 1e-4),
    ("test_laplace_hypersingular_p1_p1", "laplace", "hypersingular", 1, 0.0, 1e-4),
    ("test_helmholtz_single_layer_dp0_dp0", "helmholtz", "single_layer", 0, 3.0, 1e-4),
    ("test_helmholtz_double_layer_dp0_dp0", "helmholtz", "double_layer", 0, 3.0, 1e-4),
    ("test_helmholtz_adjoint_double_layer_dp0_dp0", "helmholtz", "adjoint_double_layer", 0, 3.0, 1e-4),
    ("test_helmholtz_hypersingular_p1_p1", "helmholtz", "hypersingular", 1, 3.0, 1e-3),
]

# Potential operators (the commented tests at compare_to_bempp_cl.rs:208-320): DP0 on the
# octahedron sphere of level 0, evaluated at the points (2,0,0), (0,2,0), (0,0,2);
# from_cl[i][j] is (dof i, point j) and the reference compares it with matrix[j][i].
POTENTIAL_CASES = [
    ("test_laplace_single_layer_potential_dp0", "laplace", "single_layer", 0.0),
    ("test_helmholtz_single_layer_potential_dp0", "helmholtz", "single_layer", 3.0),
    ("test_laplace_double_layer_potential_dp0", "laplace", "double_layer", 0.0),
    ("test_helmholtz_double_layer_potential_dp0", "helmholtz", "double_layer", 3.0),
]
POTENTIAL_POINTS = [[2.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 2.0]]

NUM = r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?"


def parse_matrix(literal):
    rows = []
    for row in re.findall(r"\[([^\[\]]*)\]", literal):
        cplx = re.findall(r"c64::new\((" + NUM + r"),\s*(" + NUM + r")\)", row)
        if cplx:
            rows.append([[float(a), float(b)] for a, b in cplx])
        else:
            rows.append([float(x) for x in re.findall(NUM, row)])
    return rows


def main(ref_root):
    text = (Path(ref_root) / "tests" / "compare_to_bempp_cl.rs").read_text()
    out = []
    for name, kind, op, degree, k, eps in CASES:
        start = text.index("fn " + name + "(")
        body = text[start:]
        lit = body[body.index("let from_cl = [") + len("let from_cl = "):]
        lit = lit[: lit.index("];") + 1]
        m = parse_matrix(lit)
        perm = [0, 5, 2, 4, 3, 1] if degree == 1 else list(range(len(m)))
        out.append({"name": name, "kind": kind, "operator": op, "degree": degree,
                    "continuity": "standard" if degree == 1 else "discontinuous",
                    "k": k, "epsilon": eps, "perm": perm, "from_cl": m,
                    "source": "tests/compare_to_bempp_cl.rs:" + name})
    dest = Path(__file__).with_name("bempp_cl_octahedron_l0.json")
    dest.write_text(json.dumps(out, indent=1))
    print("wrote", dest, len(out), "matrices")
    pot = []
    for name, kind, op, k in POTENTIAL_CASES:
        start = text.index("fn " + name + "(")
        body = text[start:]
        lit = body[body.index("let from_cl = [") + len("let from_cl = "):]
        lit = lit[: lit.index("];") + 1]
        pot.append({"name": name, "kind": kind, "operator": op, "degree": 0, "continuity": "discontinuous",
                    "k": k, "epsilon": 1e-3, "points": POTENTIAL_POINTS, "from_cl": parse_matrix(lit),
                    "source": "tests/compare_to_bempp_cl.rs:" + name + " (commented out in the reference)"})
    dest = Path(__file__).with_name("bempp_cl_potentials_octahedron_l0.json")
    dest.write_text(json.dumps(pot, indent=1))
    print("wrote", dest, len(pot), "potential matrices")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
