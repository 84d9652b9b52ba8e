"""Writes extrap_corner.json: vectors of the three-axis corner system of the
free-surface boundary condition (Algorithm_Extrapolate_Boundary_Condition.h,
CreateMatrix :548-576 and SolveMatrix :577-778), for every sign combination
(s1, s2, s3), on seeded random neighbourhoods and anisotropic spacings.

The expected values come from a mechanical transliteration of the
reference's text, not from a hand restatement: the statements of
SolveMatrix's right-hand side (b[0..2]) and CreateMatrix's m_value[0..8] are
read line by line and each C++ expression is rewritten by fixed text rules
into a Python expression (this->m_1_h.Get(k) -> ih[k-1];
cell.GetNeighbour(a,b)...Speed_GetRef().Get(k) -> the face k-1 of the cell at
the summed offset; if / else blocks by their braces).  The matrix is then
solved as UMFPACK reads it, compressed columns (value[3c + r] is row r of
column c, offsets 0,3,6,9 and row indices k % 3 as CreateMatrix gives them),
with numpy.linalg.solve.  The reference's source is read as text; nothing of
it is compiled or run.

Run from the repository root where the reference's sources are present:
    python tests/golden/make_corner.py [/path/to/reference]
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _method(src, name):
    i = src.index(f"void {name}(")
    j = src.index("{", i)
    depth, k = 0, j
    while True:
        if src[k] == "{":
            depth += 1
        elif src[k] == "}":
            depth -= 1
            if depth == 0:
                return src[j + 1:k]
        k += 1


_NEIGH = re.compile(r"cell((?:\.GetNeighbour\(\d,-?s?\d\))*)\.Speed_GetRef\(\)\.Get\((\d)\)")


def _expr(e):
    e = e.replace("this->m_1_h.Get(", "ih(")

    def face(m):
        steps = re.findall(r"GetNeighbour\((\d),(-?s?\d)\)", m.group(1))
        off = ["0", "0", "0"]
        for a, b in steps:
            off[int(a) - 1] += f"+({b})"
        return f"U({int(m.group(2)) - 1},{off[0]},{off[1]},{off[2]})"

    return _NEIGH.sub(face, e)


def transliterate(body, stop):
    """C++ statements of `body` (up to the line starting with `stop`) as a
    Python function body, by brace depth.  The solver cache test and the
    offset / index loops (blocks that set no b or m_value) are dropped."""
    out, ind, skip = [], 1, 0
    lines = [ln.strip() for ln in body.split("\n")]
    for line in lines:
        if stop and line.startswith(stop):
            break
        if not line or line.startswith("//"):
            continue
        if skip:
            skip += line == "{"
            skip -= line == "}"
            if skip == 1 and line == "}":
                skip = 0
            continue
        if line.startswith("if(m_solver.count") or line.startswith("for("):
            skip = 1
            continue
        if line == "{":
            ind += 1
            continue
        if line == "}":
            ind -= 1
            continue
        if line.startswith("if("):
            out.append("    " * ind + "if " + line[2:] + ":")
            continue
        if line == "else":
            out.append("    " * ind + "else:")
            continue
        if line.startswith("type_speed_data_value") or line.startswith("int id=") \
                or re.match(r"m_solver\[id\]\.(m_offset|m_indice|m_n)=", line) \
                or line.startswith("m_solver[id].m_value=new"):
            continue
        assert line.endswith(";"), line
        line = line[:-1].replace("m_solver[id].", "")
        out.append("    " * ind + _expr(line))
    return "\n".join(out)


def build(ref):
    with open(os.path.join(ref, "src", "Algorithm_Extrapolate_Boundary_Condition.h")) as f:
        src = f.read()
    rhs = transliterate(_method(src, "SolveMatrix"), "int id=SignToInt")
    mat = transliterate(_method(src, "CreateMatrix"), "m_solver[SignToInt")
    code = ("def rhs(s1, s2, s3, ih, U):\n    b = [0.0, 0.0, 0.0]\n" + rhs + "\n    return b\n\n"
            "def matrix(s1, s2, s3, ih):\n    m_value = [0.0] * 9\n" + mat + "\n    return m_value\n")
    ns = {}
    exec(compile(code, "transliteration", "exec"), ns)
    return ns["rhs"], ns["matrix"], code


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    rhs, matrix, code = build(ref)
    rng = np.random.default_rng(2024)
    cases = []
    for s3 in (-1, 1):
        for s2 in (-1, 1):
            for s1 in (-1, 1):
                for trial in range(4):
                    h = [0.01, 0.01, 0.01] if trial == 0 else list(rng.uniform(0.005, 0.02, 3))
                    nb = rng.standard_normal((3, 3, 3, 3))  # [c, dz+1, dy+1, dx+1]

                    def U(c, dx, dy, dz):
                        return nb[c, dz + 1, dy + 1, dx + 1]

                    def ih(k):
                        return 1.0 / h[k - 1]

                    b = np.array(rhs(s1, s2, s3, ih, U))
                    v = matrix(s1, s2, s3, ih)
                    A = np.array(v).reshape(3, 3).T  # compressed columns: A[r, c] = v[3c + r]
                    x = np.linalg.solve(A, b)
                    cases.append({"signs": [s1, s2, s3], "h": h, "nb": nb.reshape(-1).tolist(), "res": x.tolist(),
                                  "b": b.tolist(), "value": v})
    with open(os.path.join(HERE, "extrap_corner.json"), "w") as f:
        json.dump({"source": "src/Algorithm_Extrapolate_Boundary_Condition.h", "layout": "nb[c][dz+1][dy+1][dx+1]",
                   "cases": cases}, f, separators=(",", ":"))
    print(code)
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
