"""CPU pins of the oracle's viscous free-surface projection
(Algorithms_Solve_Pressure_Dirichlet with Policy_Dirichlet_Boundary)."""
import numpy as np

from mac_amd import workloads as W
from oracle import oracle as O


def one_cell(seed=0, h=(1.0, 1.0, 1.0), solid_side=False):
    """One fluid cell at (2,2,2) of a single brick with its six neighbours
    existing as Air (or the -z one Solid with its shared face const)."""
    flags = np.zeros(512, np.uint8)
    t = lambda x, y, z: x + 8 * y + 64 * z
    flags[t(2, 2, 2)] = W.EXISTS | W.FLUID
    for d in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
        flags[t(2 + d[0], 2 + d[1], 2 + d[2])] = W.EXISTS | W.AIR
    if solid_side:
        flags[t(2, 2, 1)] = W.EXISTS | W.SOLID
    rng = np.random.default_rng(seed)
    u, v, w = (rng.normal(size=512) for _ in range(3))
    if solid_side:
        flags[t(2, 2, 2)] |= W.CONST[2]  # the face between the solid and the fluid cell
        w[t(2, 2, 2)] = 0.0
    return {"keys": np.zeros((1, 3), np.int32), "flags": flags, "u": u, "v": v, "w": w, "h": h}, t


def test_zero_viscosity_is_the_plain_projection():
    for seed in (1, 2):
        g = W.random_grid(seed, 2, 2, 2)
        a = O.project(g, rtol=1e-13)
        b = O.project_dirichlet(g, 0.0, rtol=1e-13)
        for k in ("p", "u", "v", "w"):
            assert np.array_equal(a[k], b[k]), k


def test_one_cell_closed_form():
    """One fluid cell between six Dirichlet air cells: each air cell's
    pressure is -2 nu (du_i/dx_i across the air cell) times the direction
    (+-1) from the air cell to the fluid, and the cell's pressure solves
    sum_n h^-2 (p_n - p_c) = div u, so p_c = (sum_n p_n h_n^-2 - div) / sum h^-2."""
    nu = 0.37
    h = (1.0, 0.5, 0.25)
    g, t = one_cell(3, h)
    U = (g["u"], g["v"], g["w"])
    ih = [1 / hh for hh in h]
    c = (2, 2, 2)
    pn, acc = {}, 0.0
    for i in range(3):
        for s in (-1, 1):
            n = list(c)
            n[i] += s
            n = tuple(n)
            n_plus = list(n)
            n_plus[i] += 1
            grad = (U[i][t(*n_plus)] - U[i][t(*n)]) * ih[i]  # d u_i / d x_i across the air cell
            # the fluid lies at -s from the air cell: p = -2 nu grad (-s)
            pn[n] = 2 * nu * grad * s
            acc += pn[n] * ih[i] ** 2
    div = sum((U[i][t(*[c[k] + (k == i) for k in range(3)])] - U[i][t(*c)]) * ih[i] for i in range(3))
    pc = (acc - div) / sum(2 * x ** 2 for x in ih)
    r = O.project_dirichlet(g, nu, rtol=1e-15)
    assert abs(r["p"][t(*c)] - pc) <= 1e-12 * max(1.0, abs(pc))
    for n, val in pn.items():
        assert abs(r["p"][t(*n)] - val) <= 1e-14 * max(1.0, abs(val))
    assert r["max_div"] < 1e-12
    # the velocity update uses the air pressures: low x face of the cell
    assert abs(r["u"][t(*c)] - (g["u"][t(*c)] - (pc - pn[(1, 2, 2)]) * ih[0])) < 1e-12


def test_solid_neighbour_counts_as_air():
    """GetIsAir is Air or Solid: a solid cell touching the fluid gets an air
    pressure too, and its term enters the right-hand side even though the
    face is Neumann in the matrix."""
    nu = 0.5
    g, t = one_cell(4, solid_side=True)
    r = O.project_dirichlet(g, nu, rtol=1e-15)
    assert r["p"][t(2, 2, 1)] != 0.0
    g0 = dict(g)
    r0 = O.project_dirichlet(g0, 0.0, rtol=1e-15)
    assert r0["p"][t(2, 2, 1)] == 0.0


def test_divergence_free_and_linear_in_viscosity():
    """div u = 0 in the domain after the update whatever the air pressures,
    and p is affine in nu (p_air is linear in nu, the solve is linear)."""
    g = W.sphere_box(24, 5, nspheres=4, rmin=3.0, rmax=7.0, slab=5)
    res = [O.project_dirichlet(g, nu, rtol=1e-14) for nu in (0.0, 0.3, 0.6)]
    for r in res:
        assert r["max_div"] < 1e-9
    d1 = res[1]["p"] - res[0]["p"]
    d2 = res[2]["p"] - res[1]["p"]
    assert np.abs(d1).max() > 0
    assert np.allclose(d1, d2, rtol=1e-8, atol=1e-10 * np.abs(d1).max())
