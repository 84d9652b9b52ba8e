"""The free-surface velocity boundary condition
(Algorithm_Extrapolate_Boundary_Condition.h) on the CPU oracle.

Pins:
  * the reference's own 2-D known answers, test/test_Extrapolate_Boundary.h
    (7 tests, 565 assertions over 19 Do() calls; fixture extrap_bc_2d.json),
    replayed phase by phase on the x-y plane of a 3-D grid;
  * the mathematics in 3-D: a face cell's air face gives it zero divergence,
    a three-axis corner's 3x3 solve gives it zero divergence, the tangential
    faces of a face-cell pair give zero shear, a 45-degree corner copies its
    opposite faces; a uniform field and a rigid rotation (zero strain, zero
    divergence) pass through every rule unchanged where the reference's
    discretisation is consistent for them;
  * the defined order: the result does not depend on where the shape sits
    relative to the bricks (a shift changes every enumeration order);
  * what the rules may touch: faces between two non-air cells never change,
    faces between two air cells end NaN, const faces keep their values, and
    in 2-D the z faces are untouched.
"""
import numpy as np
import pytest

from tests.bc_cases import AIRS, Grid, State, air_neighbours, blob, divergence, from_mask, load_kats, random_fields
from mac_amd import workloads as W
from oracle import oracle as O

KATS = load_kats()


def run_oracle(g, h, dim=3, slot=0):
    s = O.Sim(h=h, gravity=(0, 0, 0), viscosity=0.0, particles_per_inflow_cell=0)
    s.upload(g.keys, g.flags, g.layer, u=g.u, v=g.v, w=g.w)
    s.surface_bc(slot, dim)
    return State(s.download())


def replay_kat(name, run):
    """Replays one reference test: each phase applies its Sets (forced: the
    cell's low x and low y faces, and SetFluid / SetAir) to the state the last
    Do() left, runs the condition in 2-D, and checks the phase's assertions
    (TS_ASSERT_DELTA to eps = 1e-10, isnan, !isnan).  `run(grid, h)` returns
    the state after the condition."""
    h = tuple(KATS["h"]) + (KATS["h"][0],)
    state = {}
    failures = []
    for pi, ph in enumerate(KATS["tests"][name]):
        for k, d in ph["set"].items():
            x, y = (int(t) for t in k.strip("()").split(","))
            c = state.setdefault((x, y), [W.AIR, 0.0, 0.0])
            if "speed" in d:
                c[1], c[2] = d["speed"]
            if "type" in d:
                c[0] = W.FLUID if d["type"] == "Fluid" else W.AIR
        g = Grid({(x, y, 0): (m, a, b, 0.0) for (x, y), (m, a, b) in state.items()})
        st = run(g, h)
        for (x, y) in state:
            state[(x, y)][1] = st.face(0, x, y, 0)
            state[(x, y)][2] = st.face(1, x, y, 0)
        for (xy, comp, kind, val) in ph["asserts"]:
            got = state[tuple(xy)][comp]
            ok = {"eq": lambda: abs(got - val) <= KATS["eps"], "nan": lambda: np.isnan(got),
                  "notnan": lambda: not np.isnan(got)}[kind]()
            if not ok:
                failures.append((pi, xy, comp, kind, val, got))
    return failures


@pytest.mark.parametrize("name", sorted(KATS["tests"]))
def test_reference_2d_kats(name):
    assert replay_kat(name, lambda g, h: run_oracle(g, h, dim=2)) == []


def test_kat_fixture_is_complete():
    n = sum(len(p["asserts"]) for ph in KATS["tests"].values() for p in ph)
    assert len(KATS["tests"]) == 7 and n == 565


# ---------------------------------------------------------------- 3-D box --
H = (0.1, 0.1, 0.1)


def _box(fields, shape=(6, 5, 4), origin=(0, 0, 0), h=H):
    cells = from_mask(np.ones(shape, bool), origin, pad=2, fields=fields)
    g = Grid(cells)
    return cells, g, run_oracle(g, h)


def air_neighbours_of_fluid(cells, c):
    """True if a fluid cell is among the six neighbours of c."""
    for i in range(3):
        for sg in (-1, 1):
            n = (c[0] + sg * (i == 0), c[1] + sg * (i == 1), c[2] + sg * (i == 2))
            if n in cells and cells[n][0] == W.FLUID:
                return True
    return False


def test_box_rules_random_field():
    """Random N(0,1) faces on a 6 x 5 x 4 fluid box in air."""
    cells, g, st = _box(random_fields(3, (6, 5, 4), 2))
    before = State({"keys": g.keys, "u": g.u, "v": g.v, "w": g.w})
    scale = 10.0 / H[0]
    fluid = [c for c, v in cells.items() if v[0] == W.FLUID]
    for c in cells:
        for i in range(3):
            lo = (c[0] - (i == 0), c[1] - (i == 1), c[2] - (i == 2))
            mats = [cells[c][0], cells[lo][0] if lo in cells else None]
            got = st.face(i, *c)
            if mats[1] is not None and mats[0] not in AIRS and mats[1] not in AIRS:
                assert got == before.face(i, *c), ("non-air face changed", c, i)
            # between two air cells away from the fluid (the air layer next to
            # it takes the tangential faces of the face-cell pairs): NaN
            far = not air_neighbours_of_fluid(cells, c) and (lo not in cells or not air_neighbours_of_fluid(cells, lo))
            if mats[0] in AIRS and (mats[1] is None or mats[1] in AIRS) and far:
                assert np.isnan(got), ("air face not NaN", c, i)
    for c in fluid:
        an = air_neighbours(cells, *c)
        if len(an) in (1, 3):  # face cells and three-axis corners: zero divergence
            assert abs(divergence(st, H, *c)) <= 1e-12 * scale, (c, an)
        if len(an) == 2:  # 45-degree corners: the air faces copy the opposite faces
            for (i, sg) in an:
                e = [int(i == k) for k in range(3)]
                lo, hi = st.face(i, *c), st.face(i, c[0] + e[0], c[1] + e[1], c[2] + e[2])
                assert lo == hi, (c, i)
        for i in range(3):  # every face of a fluid cell is finite
            assert np.isfinite(st.face(i, *c)) and np.isfinite(st.face(i, c[0] + (i == 0), c[1] + (i == 1), c[2] + (i == 2)))
    # zero shear on the tangential faces of face-cell pairs (B, A = B - e_t)
    # on the same face of the box: s (val - v3) / h_d + (v2 - v1) / h_t = 0
    n = 0
    for B in fluid:
        an = air_neighbours(cells, *B)
        if len(an) != 1:
            continue
        d, s = an[0]
        for t in range(3):
            if t == d:
                continue
            A = (B[0] - (t == 0), B[1] - (t == 1), B[2] - (t == 2))
            if A not in cells or cells[A][0] != W.FLUID or air_neighbours(cells, *A) != [(d, s)]:
                continue
            e = [int(d == k) for k in range(3)]
            p = e if s == 1 else [0, 0, 0]
            X = (B[0] + s * e[0], B[1] + s * e[1], B[2] + s * e[2])
            v1 = st.face(d, A[0] + p[0], A[1] + p[1], A[2] + p[2])
            v2 = st.face(d, B[0] + p[0], B[1] + p[1], B[2] + p[2])
            v3 = st.face(t, *B)
            val = st.face(t, *X)
            assert abs(s * (val - v3) / H[d] + (v2 - v1) / H[t]) <= 1e-11 * scale, (B, t)
            n += 1
    assert n > 40


def _uniform(x, y, z):
    return (0.3, -1.7, 2.2)


def _rotation(x, y, z):
    """Translation plus rigid rotation at the face centres (zero strain)."""
    om, t = np.array([0.7, -1.1, 0.4]), np.array([0.3, -1.7, 2.2])
    out = []
    for comp in range(3):
        X = np.array([x, y, z], float) * np.array(H)
        X[comp] -= 0.5 * H[comp]
        out.append((t + np.cross(om, X))[comp])
    return tuple(out)


def _rotation_at(comp, c):
    return _rotation(*c)[comp]


@pytest.mark.parametrize("field", [_uniform, _rotation], ids=["uniform", "rigid_rotation"])
def test_box_preserves_zero_strain_fields(field):
    """Every finite face equals the field, except the faces the three-axis
    corners with s1 = -1 or s2 = -1 write and the faces of their cells: there
    the reference's right-hand side is not the one its matrix implies for
    these fields (e.g. term 9 of b[1] carries -0.5/h_y for both signs of s1,
    Algorithm_Extrapolate_Boundary_Condition.h:693-703), so no restatement of
    it keeps them.  At the corners with s1 = s2 = +1 (either s3) the 3x3
    solve returns the field."""
    shape = (6, 5, 4)
    cells, g, st = _box(field, shape)
    skip = set()
    for c, v in cells.items():
        if v[0] != W.FLUID:
            continue
        an = dict((i, sg) for i, sg in air_neighbours(cells, *c))
        if len(an) == 3 and (an[0] == -1 or an[1] == -1):
            for i in range(3):
                skip.add((i, c))
                skip.add((i, (c[0] + (i == 0), c[1] + (i == 1), c[2] + (i == 2))))
    checked = 0
    for c in cells:
        for i in range(3):
            got = st.face(i, *c)
            if (i, c) in skip or np.isnan(got):
                continue
            assert got == pytest.approx(field(*c)[i], abs=1e-13), (c, i)
            checked += 1
    assert checked > 400
    # the corners with s1 = s2 = +1 were solved and checked, not skipped
    for c, sg in (((5, 4, 0), (1, 1, -1)), ((5, 4, 3), (1, 1, 1))):
        for i in range(3):
            f = tuple(c[k] + (k == i and sg[i] == 1) for k in range(3))
            assert (i, f) not in skip
            assert st.face(i, *f) == pytest.approx(field(*f)[i], abs=1e-13)


def test_thin_lines_fill_their_nan_faces():
    """1-cell-thick lines along each axis and an L: the line ends' faces
    toward air with fluid behind them are not in the domain
    (Get_Is_Speed_In_Domain), get NaN and are filled by apply_nan_extrap,
    which leaves the end cell divergence free."""
    m = np.zeros((12, 12, 12), bool)
    m[1:6, 2, 2] = True       # along x
    m[8, 1:6, 8] = True       # along y
    m[2, 8, 4:9] = True       # along z
    m[6:10, 9, 1] = True      # an L in the x-y plane
    m[9, 9:11, 1] = True
    cells = from_mask(m, pad=2, fields=random_fields(5, m.shape, 2))
    g = Grid(cells)
    st = run_oracle(g, H)
    ends = 0
    for c, v in cells.items():
        if v[0] != W.FLUID:
            continue
        for i in range(3):
            assert np.isfinite(st.face(i, *c)) and np.isfinite(st.face(i, c[0] + (i == 0), c[1] + (i == 1), c[2] + (i == 2))), c
        if len(air_neighbours(cells, *c)) == 5:
            assert abs(divergence(st, H, *c)) <= 1e-11 * 10 / H[0], c
            ends += 1
    assert ends >= 7
    # a uniform field passes unchanged (every mean apply_nan_extrap takes is
    # of equal values, and the field is divergence free)
    st = run_oracle(Grid(from_mask(m, pad=2, fields=_uniform)), H)
    for c, v in cells.items():
        for i in range(3):
            got = st.face(i, *c)
            if not np.isnan(got):
                assert got == pytest.approx(_uniform(*c)[i], abs=1e-13), (c, i)


@pytest.mark.parametrize("shift", [(3, 5, 2), (-8, 13, 7)])
def test_result_independent_of_brick_placement(shift):
    """A random blob with random faces, shifted against the brick grid: the
    same faces, bit for bit (the phases read and write in a defined order)."""
    m = blob(7, (14, 11, 9), 0.45)
    f = random_fields(8, m.shape, 2)
    a = from_mask(m, (0, 0, 0), 2, f)
    b = from_mask(m, shift, 2, f)
    sa, sb = run_oracle(Grid(a), H), run_oracle(Grid(b), H)
    for c in a:
        cb = (c[0] + shift[0], c[1] + shift[1], c[2] + shift[2])
        for i in range(3):
            x, y = sa.face(i, *c), sb.face(i, *cb)
            assert (np.isnan(x) and np.isnan(y)) or x == y, (c, i, x, y)


def test_blob_touches_only_faces_next_to_air():
    m = blob(9, (14, 11, 9), 0.5)
    cells = from_mask(m, (0, 0, 0), 2, random_fields(10, m.shape, 2))
    g = Grid(cells)
    st = run_oracle(g, H)
    before = State({"keys": g.keys, "u": g.u, "v": g.v, "w": g.w})
    nonair = 0
    for c, v in cells.items():
        for i in range(3):
            lo = (c[0] - (i == 0), c[1] - (i == 1), c[2] - (i == 2))
            if lo in cells and v[0] not in AIRS and cells[lo][0] not in AIRS:
                assert st.face(i, *c) == before.face(i, *c)
                nonair += 1
        if v[0] == W.FLUID:  # every fluid cell's faces end finite
            for i in range(3):
                hi = (c[0] + (i == 0), c[1] + (i == 1), c[2] + (i == 2))
                assert np.isfinite(st.face(i, *c)) and np.isfinite(st.face(i, *hi)), c
    assert nonair > 500


def test_const_faces_kept_and_2d_leaves_z():
    m = np.ones((5, 4, 3), bool)
    cells = from_mask(m, (0, 0, 0), 2, random_fields(11, m.shape, 2))
    const = {(5, 1, 1): (0,), (2, 4, 1): (1,), (-1, -1, -1): (0, 1, 2)}
    g = Grid(cells, const)
    st = run_oracle(g, H)
    for c, comps in const.items():
        for i in comps:
            assert st.face(i, *c) == (g.u, g.v, g.w)[i][g.slot(*c)]
    st2 = run_oracle(Grid(cells), H, dim=2)
    g2 = Grid(cells)
    for c in cells:
        assert st2.face(2, *c) == g2.w[g2.slot(*c)]


def test_corner_system_against_transliteration():
    """The three-axis corner system for all eight sign combinations, isotropic
    and anisotropic spacings, against extrap_corner.json: the reference's
    SolveMatrix / CreateMatrix text transliterated mechanically and solved in
    compressed columns (make_corner.py)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "extrap_corner.json")) as f:
        fx = json.load(f)
    assert len(fx["cases"]) == 32
    for c in fx["cases"]:
        got = O.corner_solve(c["h"], c["signs"], np.array(c["nb"]))
        assert np.allclose(got, c["res"], rtol=1e-12, atol=1e-12 * np.abs(c["res"]).max()), c["signs"]
