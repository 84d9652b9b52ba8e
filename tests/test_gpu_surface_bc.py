"""The free-surface velocity boundary condition on the device
(mac_sim_surface_bc, surface_bc.h) against the CPU oracle, bit for bit, and
against the reference's own 2-D known answers.

Cases: the 565 assertions of test/test_Extrapolate_Boundary.h replayed
through the device in 2-D; a fluid box, thin lines and seeded random blobs in
3-D (isotropic and anisotropic spacing, shifted against the bricks, const
faces); the acceleration slot after the forces; whole time steps with the
condition chained before every extrapolation (the RK4 "all" chain of
test/Jet3.cpp); and, at 256^3 (the union-of-spheres projection workload), the
properties that hold at any size: every fluid cell's faces finite, faces
between non-air cells unchanged, air faces away from the fluid NaN.
Bar: identical bytes (NaN where the oracle has NaN)."""
import numpy as np
import pytest

from mac_amd import workloads as W
from mac_amd.lib import Sim
from oracle import oracle as O
from tests.bc_cases import AIRS, Grid, State, blob, from_mask, load_kats, random_fields
from tests.test_surface_bc import replay_kat, run_oracle

pytestmark = pytest.mark.gpu

KATS = load_kats()


def run_device(g, h, dim=3, slot=0):
    s = Sim(h=h, gravity=(0, 0, 0), viscosity=0.0, particles_per_inflow_cell=0)
    s.upload(g.keys, g.flags, g.layer, u=g.u, v=g.v, w=g.w)
    s.surface_bc(slot, dim)
    d = s.download()
    s.close()
    return State(d)


def same_faces(a, b, where):
    assert np.array_equal(a.d["keys"], b.d["keys"]), where
    for k in ("u", "v", "w"):
        x, y = a.d[k], b.d[k]
        bad = ~((x == y) | (np.isnan(x) & np.isnan(y)))
        assert not bad.any(), (where, k, int(bad.sum()), np.flatnonzero(bad)[:5])


@pytest.mark.parametrize("name", sorted(KATS["tests"]))
def test_reference_2d_kats_on_device(name):
    assert replay_kat(name, lambda g, h: run_device(g, h, dim=2)) == []


H = (0.1, 0.1, 0.1)
HA = (0.01, 0.013, 0.008)


@pytest.mark.parametrize("h", [H, HA], ids=["isotropic", "anisotropic"])
def test_box_and_lines_vs_oracle(h):
    cells = from_mask(np.ones((6, 5, 4), bool), (0, 0, 0), 2, random_fields(3, (6, 5, 4), 2))
    g = Grid(cells)
    same_faces(run_device(g, h), run_oracle(g, h), "box")
    m = np.zeros((12, 12, 12), bool)
    m[1:6, 2, 2] = True
    m[8, 1:6, 8] = True
    m[2, 8, 4:9] = True
    m[6:10, 9, 1] = True
    m[9, 9:11, 1] = True
    g = Grid(from_mask(m, pad=2, fields=random_fields(5, m.shape, 2)))
    same_faces(run_device(g, h), run_oracle(g, h), "lines")


@pytest.mark.parametrize("seed,shape,fill,shift", [
    (7, (14, 11, 9), 0.45, (0, 0, 0)),
    (7, (14, 11, 9), 0.45, (3, 5, 2)),
    (21, (24, 20, 17), 0.3, (-8, 13, 7)),
    (22, (24, 20, 17), 0.6, (5, -3, 1)),
    (23, (40, 33, 29), 0.5, (0, 0, 0)),
])
def test_random_blobs_vs_oracle(seed, shape, fill, shift):
    m = blob(seed, shape, fill)
    g = Grid(from_mask(m, shift, 2, random_fields(seed + 1, shape, 2)))
    for h in (H, HA):
        same_faces(run_device(g, h), run_oracle(g, h), (seed, h))


def test_2d_mode_and_const_faces_vs_oracle():
    m = blob(31, (20, 18, 6), 0.5)
    cells = from_mask(m, (1, 2, 0), 2, random_fields(32, m.shape, 2))
    rng = np.random.default_rng(33)
    keys = list(cells)
    const = {keys[i]: tuple(int(c) for c in np.flatnonzero(rng.random(3) < 0.5)) for i in rng.choice(len(keys), 60, replace=False)}
    g = Grid(cells, const)
    same_faces(run_device(g, H, dim=2), run_oracle(g, H, dim=2), "2-D")
    same_faces(run_device(g, H, dim=3), run_oracle(g, H, dim=3), "3-D const")


def test_acceleration_slot_vs_oracle():
    """Slot 3: the forces of a random blob (gravity and convection) put in
    the acceleration, then the condition on it, on both sides."""
    m = blob(41, (18, 15, 13), 0.5)
    g = Grid(from_mask(m, (0, 0, 0), 2, random_fields(42, m.shape, 2)))
    kw = dict(h=H, gravity=(0.0, 0.0, -9.81), viscosity=0.0, particles_per_inflow_cell=0)
    dev, orc = Sim(**kw), O.Sim(**kw)
    for s in (dev, orc):
        s.upload(g.keys, g.flags, g.layer, u=g.u, v=g.v, w=g.w)
        s.forces(0)
        s.surface_bc(3, 3)
    a, b = dev.download(), orc.download()
    dev.close()
    for k in ("ax", "ay", "az", "u", "v", "w"):
        x, y = a[k], b[k]
        assert np.array_equal(x, y, equal_nan=True), k


def _close_fields(dg, do, where, tol=1e-8):
    ex = (dg["flags"] & W.EXISTS) > 0
    for k in ("u", "v", "w"):
        x, y = dg[k][ex], do[k][ex]
        assert np.array_equal(np.isnan(x), np.isnan(y)), (where, k)
        f = ~np.isnan(y)
        assert np.linalg.norm(x[f] - y[f]) <= tol * max(np.linalg.norm(y[f]), 1.0), (where, k)


def test_jet3_rk4all_depth_with_surface_bc_vs_oracle():
    """The Jet3 chain on a small lateral jet (RK4 "all", depth 2/3, particle
    floor, cube fill, viscosity in the forces and the projections) with the
    condition before every extrapolation of the speeds and of the
    acceleration, as test/Jet3.cpp:541-550,636-639 chains it: 4 steps against
    the oracle."""
    from mac_amd import jets as J
    from tests.test_gpu_sim import same_cells, same_depth, same_particles, upload_both
    jet = J.lateral_jet(32, 24, 24, h=0.01, nozzle=8, inflow=1.5, seed=3)
    kw = dict(jet["params"], layer_max=2, integrator=2, seed=7, particles_per_inflow_cell=0, fill_cube_n=1,
              depth_interior=2, depth_remove=3, floor_axis=2, floor_value=5.5 * 0.01, viscosity=1.307e-3, surface_bc=1)
    cells = np.argwhere(np.ones((10, 10, 10), bool)) + np.array([1, 7, 7])
    rng = np.random.default_rng(2)
    pts = np.repeat(cells, 2, axis=0) * 0.01 + rng.uniform(-0.004, 0.004, size=(2 * len(cells), 3))
    pair = (Sim(**kw, proj={"rtol": 1e-13, "viscosity": 1.307e-3}), O.Sim(**kw, proj={"viscosity": 1.307e-3}))
    off = Sim(**dict(kw, surface_bc=0), proj={"rtol": 1e-13, "viscosity": 1.307e-3})
    upload_both(jet, pair + (off,), particles=pts)
    differs = 0.0
    for step in range(4):
        sg = pair[0].step()
        so = pair[1].step()
        off.step()
        assert sg["dt"] == pytest.approx(so["dt"], rel=1e-9)
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        same_depth(pair, dg)
        _close_fields(dg, do, step)
        same_particles(dg["particles"], do["particles"], 1e-9)
        assert sg["max_div"] < 1e-7
        d0 = off.download()
        if d0["keys"].shape == dg["keys"].shape and np.array_equal(d0["flags"], dg["flags"]):
            ex = (dg["flags"] & W.EXISTS) > 0
            differs = max(differs, float(np.abs(d0["u"][ex] - dg["u"][ex]).max()))
        else:
            differs = np.inf
    print(f"[jet3-bc] largest face change from the condition: {differs:.3e}", flush=True)
    assert differs > 1e-6  # the condition is not a no-op on this run
    pair[0].close()
    off.close()


def test_union_of_spheres_256_properties():
    """At 256^3 (40 seeded spheres and a floor slab, ~45% fluid, a solid
    wall, which the condition counts as air, N(0,1) faces): every fluid
    cell's faces finite, faces between two non-air cells unchanged, faces
    between two air cells with no fluid neighbour NaN."""
    n = 256
    g = W.sphere_box(n, seed=0)
    s = Sim(h=(0.01, 0.01, 0.01), gravity=(0, 0, 0), viscosity=0.0, particles_per_inflow_cell=0)
    s.upload(g["keys"], g["flags"], None, u=g["u"], v=g["v"], w=g["w"])
    s.surface_bc(0, 3)
    d = s.download()
    s.close()
    assert np.array_equal(d["keys"], g["keys"])
    nbk = n // 8

    def dense(a):
        return W.bricks_to_dense(a.reshape(-1, 512), nbk, nbk, nbk)  # [z, y, x]

    f = dense(g["flags"])
    mat = f & 7
    AI = np.isin(mat, AIRS)
    NA = ~AI
    FL = mat == W.FLUID
    near = FL.copy()
    for a2 in range(3):
        for sh in (-1, 1):
            near |= np.roll(FL, sh, a2)
    for comp, k in enumerate("uvw"):
        before, after = dense(g[k]), dense(d[k])
        ax = 2 - comp  # the component's axis in [z, y, x]
        hi = tuple(slice(1, None) if q == ax else slice(None) for q in range(3))
        lo = tuple(slice(None, -1) if q == ax else slice(None) for q in range(3))
        both = NA[hi] & NA[lo]
        assert np.array_equal(after[hi][both], before[hi][both]), k
        assert np.isfinite(after[FL]).all(), k
        assert np.isfinite(after[hi][FL[lo]]).all(), k
        cst = (dense(g["flags"])[hi] & W.CONST[comp]) > 0  # const faces keep their value
        far = AI[hi] & AI[lo] & ~near[hi] & ~near[lo] & ~cst
        assert np.isnan(after[hi][far]).all(), k
        print(f"[bc256] {k}: {int(both.sum())} non-air faces unchanged, {int(far.sum())} far air faces NaN", flush=True)
