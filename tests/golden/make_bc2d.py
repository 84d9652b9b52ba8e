"""Writes extrap_bc_2d.json: the known answers of the reference's 2-D tests of
the free-surface velocity boundary condition
(test/test_Extrapolate_Boundary.h, every test_* method), as data.

Each test builds cells on a 2-D grid (h = 0.1 on both axes), sets each one's
two face velocities (a forced Set of the Data_Speed_Data, i.e. the cell's low
x face and low y face) and its type (SetFluid / SetAir), runs the
algorithm's Do(), and asserts face values: TS_ASSERT_DELTA(value, eps = 1e-10),
std::isnan or !std::isnan.  Later blocks of the same test change some cells
and run Do() again on the state the last Do() left.  The fixture keeps, per
test, the list of these phases: {"set": {"(x, y)": {"speed": [u, v],
"type": "Fluid" | "Air"}}, "asserts": [[[x, y], component (1 = x, 2 = y),
"eq" | "nan" | "notnan", value or null]]}.

Run from the repository root where the reference's sources are present:
    python tests/golden/make_bc2d.py [/path/to/reference]
The reference's header is read as text; nothing of it is compiled or run.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse(src):
    tests = re.split(r"\n    void (test_\w+)\(\)", src)
    out = {}
    for name, body in zip(tests[1::2], tests[2::2]):
        v = [0, 0]
        sp = [0.0, 0.0]
        phases = []
        cur = {"set": {}, "asserts": []}
        ran = False
        for line in body.split("\n"):
            line = line.strip()
            m = re.match(r"v\.Set\((\d),(-?\d+)\);", line)
            if m:
                v[int(m.group(1)) - 1] = int(m.group(2))
                continue
            m = re.match(r"speed\.Set\((\d),(-?[\d.]+)\);", line)
            if m:
                sp[int(m.group(1)) - 1] = float(m.group(2))
                continue
            if "Speed_GetRef().Set(Data_Speed_Data" in line:
                if ran:  # a Set after a Do(): the next phase starts
                    phases.append(cur)
                    cur = {"set": {}, "asserts": []}
                    ran = False
                cur["set"].setdefault(str(tuple(v)), {})["speed"] = list(sp)
                continue
            m = re.search(r"CellType_GetRef\(\)\.Set(Fluid|Air)\(\)", line)
            if m:
                cur["set"].setdefault(str(tuple(v)), {})["type"] = m.group(1)
                continue
            if line.startswith("alg.Do()"):
                ran = True
                continue
            m = re.match(r"TS_ASSERT_DELTA\(.*\[v\]\.Speed_GetRef\(\)\.Get\((\d)\),(-?[\d.]+),eps\);", line)
            if m:
                cur["asserts"].append([list(v), int(m.group(1)), "eq", float(m.group(2))])
                continue
            m = re.match(r"TS_ASSERT\((!?)std::isnan\(.*\[v\]\.Speed_GetRef\(\)\.Get\((\d)\)\)\);", line)
            if m:
                cur["asserts"].append([list(v), int(m.group(2)), "notnan" if m.group(1) else "nan", None])
                continue
            if "TS_ASSERT" in line:
                raise ValueError(f"{name}: unparsed assertion {line!r}")
        phases.append(cur)
        out[name] = phases
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "test", "test_Extrapolate_Boundary.h")) as f:
        out = parse(f.read())
    with open(os.path.join(HERE, "extrap_bc_2d.json"), "w") as f:
        json.dump({"source": "test/test_Extrapolate_Boundary.h", "h": [0.1, 0.1], "eps": 1e-10, "tests": out},
                  f, separators=(",", ":"))
    for n, ph in out.items():
        print(n, [(len(p["set"]), len(p["asserts"])) for p in ph])


if __name__ == "__main__":
    main()
