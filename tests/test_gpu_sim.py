"""GPU parity of the time-step loop (mac_sim_* through the C-ABI) against the
CPU oracle on identical seeded inputs.

Bars: chunk keys, flag bytes (cell types, existence) and layers bit-exact;
particle -> key binning and the seeded inflow particles bit-exact;
accelerations, the time step and interpolated velocities bit-exact (same
operation order, no contraction on either side); after projections,
velocities within 1e-8 and particle positions within 1e-9 relative."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from mac_amd import jets as J
from mac_amd import workloads as W
from mac_amd.lib import Sim
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def both(**kw):
    return Sim(**kw, proj={"rtol": 1e-13}), O.Sim(**kw)


def upload_both(g, pair, layer=None, particles=None):
    for s in pair:
        s.upload(g["keys"], g["flags"], layer, g["u"], g["v"], g["w"], particles)


def same_cells(dg, do):
    assert np.array_equal(dg["keys"], do["keys"])
    assert np.array_equal(dg["flags"], do["flags"])
    ex = (dg["flags"] & W.EXISTS) > 0
    assert np.array_equal(dg["layer"][ex], do["layer"][ex])


def same_particles(pg, po, tol):
    assert pg.shape == po.shape
    if pg.shape[0] == 0:
        return
    d, _ = cKDTree(po).query(pg)
    assert d.max() <= tol, d.max()


def random_state(seed, nb=(3, 3, 2)):
    g = W.random_grid(seed, *nb, fluid_p=0.5, solid_p=0.05, const_p=0.1)
    rng = np.random.default_rng(seed)
    layer = rng.integers(-1, 3, size=g["flags"].size).astype(np.int8)
    return g, layer


@pytest.mark.parametrize("seed", [1, 2])
def test_forces_bitexact(seed):
    g, layer = random_state(seed)
    for nu in (0.0, 1.3e-3):
        pair = both(h=(0.1, 0.2, 0.15), gravity=(0.0, -2.0, -9.81), viscosity=nu)
        upload_both(g, pair, layer)
        for s in pair:
            s.forces(0)
        dg, do = pair[0].download(), pair[1].download()
        for k in ("ax", "ay", "az"):
            assert np.array_equal(dg[k], do[k]), (k, nu, np.abs(dg[k] - do[k]).max())


@pytest.mark.parametrize("seed", [1, 2])
def test_dt_and_interpolation_bitexact(seed):
    g, layer = random_state(seed)
    pair = both(h=(0.1, 0.2, 0.15), cfl=0.5)
    upload_both(g, pair, layer)
    # same field on both sides: the max is order-free and sqrt and the division
    # are correctly rounded, so dt is bit-exact
    assert pair[0].compute_dt() == pair[1].compute_dt()
    rng = np.random.default_rng(seed)
    lo = np.array([-1.0, -1.0, -1.0]) * np.array([0.1, 0.2, 0.15])
    hi = np.array([25.0, 25.0, 17.0]) * np.array([0.1, 0.2, 0.15])
    pos = rng.uniform(lo, hi, size=(4000, 3))
    assert np.array_equal(pair[0].interpolate(pos), pair[1].interpolate(pos))


def test_topology_single_particle_kat():
    """TestAlgorithms_Update_CellType_Layer.h:43-150 on the device."""
    pair = both(h=(1.0, 1.0, 1.0), layer_max=1, gravity=(0, 0, 0), particles_per_inflow_cell=0)
    for s in pair:
        s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), particles=np.array([[0.1, -0.2, 0.3]]))
        s.topology()
    same_cells(pair[0].download(), pair[1].download())


@pytest.mark.parametrize("seed", [3, 4])
def test_topology_random_particles(seed):
    """Particle -> key binning, layers, brick creation and deletion on a cloud of
    particles spread over many (partly missing) bricks."""
    g, _ = random_state(seed, (2, 2, 2))
    rng = np.random.default_rng(seed)
    parts = rng.uniform(-6.0, 22.0, size=(600, 3))
    for lmax in (1, 3):
        pair = both(h=(1.0, 1.0, 1.0), layer_max=lmax, particles_per_inflow_cell=8, seed=seed)
        upload_both(g, pair, particles=parts)
        for s in pair:
            s.topology()
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        same_particles(dg["particles"], do["particles"], 0.0)


@pytest.mark.parametrize("integrator,prk", [(0, 2), (1, 4)])
def test_jet_steps_vs_oracle(integrator, prk):
    """A small vertical jet (config-1 layout) stepped on both sides."""
    jet = J.vertical_jet(16, 16, 24, h=0.01, nozzle=4, inflow=1.0)
    kw = dict(jet["params"], layer_max=2, integrator=integrator, particle_rk=prk, seed=7)
    pair = both(**kw)
    upload_both(jet, pair)
    for step in range(6):
        sg = pair[0].step()
        so = pair[1].step()
        # dt of a step reads the last step's projected speeds, which agree to 1e-8
        assert sg["dt"] == pytest.approx(so["dt"], rel=1e-9)
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        ex = (dg["flags"] & W.EXISTS) > 0
        for k in ("u", "v", "w"):
            ref = do[k][ex]
            assert np.linalg.norm(dg[k][ex] - ref) <= 1e-8 * max(np.linalg.norm(ref), 1.0), (step, k)
        same_particles(dg["particles"], do["particles"], 1e-9)
        assert sg["max_div"] < 1e-7


def _same_speeds(dg, do):
    for k in ("u", "v", "w"):
        assert np.array_equal(dg[k], do[k], equal_nan=True), (k, np.nanmax(np.abs(dg[k] - do[k])))


@pytest.mark.parametrize("seed", [5, 6])
def test_extrapolate_bitexact_random_layers(seed):
    """Algorithms_Extrapolate on an arbitrary state (layers not from a BFS):
    the device's frontier steps reproduce the oracle's sets bit for bit."""
    g, layer = random_state(seed, (3, 3, 3))
    pair = both()
    upload_both(g, pair, layer=layer)
    for s in pair:
        s.extrapolate()
    dg, do = pair[0].download(), pair[1].download()
    same_cells(dg, do)
    _same_speeds(dg, do)


def test_extrapolate_bitexact_after_jet_steps():
    jet = J.vertical_jet(16, 16, 24, h=0.01, nozzle=4, inflow=1.0)
    kw = dict(jet["params"], layer_max=3, seed=9)
    pair = both(**kw)
    upload_both(jet, pair)
    for _ in range(5):
        for s in pair:
            s.step()
    for s in pair:
        s.topology()
        s.extrapolate()
    dg, do = pair[0].download(), pair[1].download()
    same_cells(dg, do)
    ex = (dg["flags"] & W.EXISTS) > 0
    for k in ("u", "v", "w"):
        # the velocities before extrapolation agree to the solver tolerance, and
        # extrapolation only averages them
        a, b = dg[k][ex], do[k][ex]
        fin = np.isfinite(b)
        assert np.array_equal(np.isfinite(a), fin), k
        assert np.linalg.norm(a[fin] - b[fin]) <= 1e-8 * max(1.0, np.linalg.norm(b[fin])), k


# ------------------------------------------------------------ depth chain ----
def same_depth(pair, dg):
    ex = (dg["flags"] & W.EXISTS) > 0
    assert np.array_equal(pair[0].get_depth()[ex], pair[1].get_depth()[ex])


@pytest.mark.parametrize("shape", ["ball", "random"])
def test_depth_topology_bitexact(shape):
    """Algorithms_Initialize_Layer_And_Depth -> ..._Layer_Initial_With_Particle_Depth
    -> Algorithms_Create_Fluid_Particle_Depth -> Algorithms_Update_CellType_Layer_Depth
    on a blob of particles: first call makes Interior cells, then holes are
    punched in the particle list (deep particles go, shallow Interior cells
    are refilled) and two more calls follow.  Flags, layers, depth and the
    particle multiset bit-exact."""
    n = 20
    g = np.indices((n, n, n))
    if shape == "ball":
        mask = ((g - (n - 1) / 2.0) ** 2).sum(0) <= (n / 2.0) ** 2
    else:
        mask = np.random.default_rng(9).random((n, n, n)) < 0.98
    pts = np.argwhere(mask).astype(np.float64) * 0.5 + 0.01
    pair = both(h=(0.5, 0.5, 0.5), layer_max=3, gravity=(0, 0, 0), particles_per_inflow_cell=3, depth_interior=3,
                depth_remove=4, seed=5)
    for s in pair:
        s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), particles=pts)
        s.topology()
    dg, do = pair[0].download(), pair[1].download()
    same_cells(dg, do)
    same_depth(pair, dg)
    assert ((dg["flags"][(dg["flags"] & W.EXISTS) > 0] & 7) == W.INTERIOR).sum() > 50
    keep = np.random.default_rng(1).random(len(pts)) > 0.3
    dep = pair[0].get_depth()
    for s in pair:
        s.upload(dg["keys"], dg["flags"], dg["layer"], particles=pts[keep])
        s.set_depth(dep)
    for it in range(2):
        for s in pair:
            s.topology()
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        same_depth(pair, dg)
        same_particles(dg["particles"], do["particles"], 0.0)


def test_depth_floor_bitexact():
    """Algorithm_DelParticle: particles below the floor never mark a cell."""
    rng = np.random.default_rng(4)
    pts = rng.uniform(-4.0, 12.0, size=(3000, 3))
    pair = both(h=(1.0, 1.0, 1.0), layer_max=2, gravity=(0, 0, 0), particles_per_inflow_cell=0, depth_interior=2,
                depth_remove=2, floor_axis=1, floor_value=0.25)
    for s in pair:
        s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), particles=pts)
        s.topology()
    dg, do = pair[0].download(), pair[1].download()
    same_cells(dg, do)
    same_depth(pair, dg)
    same_particles(dg["particles"], do["particles"], 0.0)
    assert len(dg["particles"]) == int((pts[:, 1] >= 0.25).sum())


def test_depth_jet_steps_vs_oracle():
    """A thick lateral jet stepped with the depth chain (Interior cells,
    deep-particle removal, Interior refill, floor) on both sides."""
    jet = J.lateral_jet(32, 24, 24, h=0.01, nozzle=8, inflow=1.5, seed=3)
    kw = dict(jet["params"], layer_max=2, integrator=1, particle_rk=2, seed=7, particles_per_inflow_cell=4,
              depth_interior=2, depth_remove=3, floor_axis=2, floor_value=5.5 * 0.01)
    # a slab of fluid in front of the nozzle, two jittered particles per cell
    cells = np.argwhere(np.ones((14, 12, 12), bool)) + np.array([1, 6, 6])
    rng = np.random.default_rng(2)
    pts = np.repeat(cells, 2, axis=0) * 0.01 + rng.uniform(-0.004, 0.004, size=(2 * len(cells), 3))
    pair = both(**kw)
    upload_both(jet, pair, particles=pts)
    n_interior = 0
    for step in range(8):
        sg = pair[0].step()
        so = pair[1].step()
        # dt of a step reads the last step's projected speeds, which agree to 1e-8
        assert sg["dt"] == pytest.approx(so["dt"], rel=1e-9)
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        same_depth(pair, dg)
        ex = (dg["flags"] & W.EXISTS) > 0
        n_interior = max(n_interior, int(((dg["flags"][ex] & 7) == W.INTERIOR).sum()))
        for k in ("u", "v", "w"):
            ref = do[k][ex]
            assert np.linalg.norm(dg[k][ex] - ref) <= 1e-8 * max(np.linalg.norm(ref), 1.0), (step, k)
        same_particles(dg["particles"], do["particles"], 1e-9)
        assert sg["max_div"] < 1e-7
    assert n_interior > 0


# ---------------------------------------------------- RK4 "all" integrator ----
def _blob(h):
    cells = np.argwhere(np.ones((8, 8, 8), bool)) + 4
    pts = cells * h + np.array([1, 3, -5]) * (h / 64)
    keys = np.array(sorted({tuple(k) for k in (cells // 8).tolist()}, key=lambda k: (k[2], k[1], k[0])), np.int32)
    return keys, pts


@pytest.mark.parametrize("case", ["uniform", "free_fall"])
def test_rk4all_closed_forms_vs_oracle(case):
    """The two closed forms the oracle is pinned by, stepped on the device:
    cells bit-exact, grid speeds and particles equal to the oracle's."""
    h, dt = 0.125, 0.0625
    keys, pts = _blob(h)
    n = len(keys) * 512
    g = (0, 0, -8.0) if case == "free_fall" else (0, 0, 0)
    vel = (0.0, 0.0, 0.0) if case == "free_fall" else (0.25, -0.5, 0.125)
    pair = both(h=(h, h, h), gravity=g, integrator=2, particles_per_inflow_cell=0, layer_max=2, cfl=1e9, dt_min=dt,
                dt_max=dt)
    for s in pair:
        s.upload(keys, np.zeros(n, np.uint8), u=np.full(n, vel[0]), v=np.full(n, vel[1]), w=np.full(n, vel[2]),
                 particles=pts)
    for step in range(2):
        for s in pair:
            s.step()
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        ex = (dg["flags"] & W.EXISTS) > 0
        for k in ("u", "v", "w"):
            assert np.allclose(dg[k][ex], do[k][ex], rtol=0, atol=1e-13), (step, k)
        same_particles(dg["particles"], do["particles"], 1e-13)


def test_rk4all_depth_jet_vs_oracle():
    """The Jet3 configuration on a small lateral jet: RK4 "all" with the
    depth chain (DepthInterior 2, DepthRemove 3), the particle floor, the
    cube-lattice fill (n = 1), viscosity in the forces and the viscous air
    pressures in every projection, against the oracle step by step."""
    jet = J.lateral_jet(32, 24, 24, h=0.01, nozzle=8, inflow=1.5, seed=3)
    kw = dict(jet["params"], layer_max=2, integrator=2, seed=7, particles_per_inflow_cell=0, fill_cube_n=1,
              depth_interior=2, depth_remove=3, floor_axis=2, floor_value=5.5 * 0.01, viscosity=1.307e-3)
    cells = np.argwhere(np.ones((10, 10, 10), bool)) + np.array([1, 7, 7])
    rng = np.random.default_rng(2)
    pts = np.repeat(cells, 2, axis=0) * 0.01 + rng.uniform(-0.004, 0.004, size=(2 * len(cells), 3))
    pair = (Sim(**kw, proj={"rtol": 1e-13, "viscosity": 1.307e-3}), O.Sim(**kw, proj={"viscosity": 1.307e-3}))
    upload_both(jet, pair, particles=pts)
    n_interior = 0
    for step in range(4):
        sg = pair[0].step()
        so = pair[1].step()
        # dt of a step reads the last step's projected speeds, which agree to 1e-8
        assert sg["dt"] == pytest.approx(so["dt"], rel=1e-9)
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        same_depth(pair, dg)
        ex = (dg["flags"] & W.EXISTS) > 0
        n_interior = max(n_interior, int(((dg["flags"][ex] & 7) == W.INTERIOR).sum()))
        for k in ("u", "v", "w"):
            ref = do[k][ex]
            assert np.linalg.norm(dg[k][ex] - ref) <= 1e-8 * max(np.linalg.norm(ref), 1.0), (step, k)
        same_particles(dg["particles"], do["particles"], 1e-9)
        assert sg["max_div"] < 1e-7
    assert n_interior > 0


@pytest.mark.parametrize("integrator,prk", [(0, 2), (2, 4)])
def test_box_deletion_vs_oracle(integrator, prk):
    """Particles leaving the box are deleted: a 16x8x8-cell block moving along
    +x through the box's upper x bound, 8 shuffled particles per cell (so the
    deleted ones are spread over the whole particle array), two steps on both
    sides: the kept particles equal the oracle's, none is outside the box and
    none is NaN."""
    h, dt = 0.125, 0.0625
    cells = np.argwhere(np.ones((16, 8, 8), bool)) + 4
    rng = np.random.default_rng(11)
    pts = np.repeat(cells, 8, axis=0) * h + rng.uniform(-0.49 * h, 0.49 * h, size=(8 * len(cells), 3))
    pts = pts[rng.permutation(len(pts))]
    keys = np.array(sorted({tuple(k) for k in (cells // 8).tolist()}, key=lambda k: (k[2], k[1], k[0])), np.int32)
    n = len(keys) * 512
    hi = (12.3 * h, 100.0, 100.0)
    kw = dict(h=(h, h, h), gravity=(0, 0, 0), integrator=integrator, particle_rk=prk, particles_per_inflow_cell=0,
              layer_max=2, cfl=1e9, dt_min=dt, dt_max=dt, box_lo=(-100.0, -100.0, -100.0), box_hi=hi)
    pair = both(**kw)
    for s in pair:
        s.upload(keys, np.zeros(n, np.uint8), u=np.full(n, 1.0), v=np.zeros(n), w=np.zeros(n), particles=pts)
    counts = []
    for step in range(2):
        for s in pair:
            s.step()
        pg, po = pair[0].download()["particles"], pair[1].download()["particles"]
        assert not np.isnan(pg).any()
        assert (pg[:, 0] < hi[0]).all()
        same_particles(pg, po, 1e-12)
        counts.append(len(pg))
    assert len(pts) > counts[0] > counts[1] > 0


@pytest.mark.parametrize("integrator", [0, 2])
def test_empty_state_steps_like_the_oracle(integrator):
    """No bricks and no particles: a step changes nothing, and the time step
    is the oracle's (the upper bound, nothing moves)."""
    pair = both(h=(0.1, 0.1, 0.1), gravity=(0, 0, -9.81), particles_per_inflow_cell=0, integrator=integrator,
                particle_rk=4 if integrator == 2 else 2)
    for s in pair:
        s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), u=np.zeros(0), v=np.zeros(0), w=np.zeros(0))
    for _ in range(2):
        sg, so = pair[0].step(), pair[1].step()
        assert sg["dt"] == so["dt"]
        dg, do = pair[0].download(), pair[1].download()
        assert dg["keys"].shape == do["keys"].shape == (0, 3)
        assert dg["particles"].shape == do["particles"].shape == (0, 3)


@pytest.mark.parametrize("integrator,prk", [(0, 2), (2, 4)])
def test_pulsed_inflow_jet_vs_oracle(integrator, prk):
    """The vertical jet with the seeded inflow pulsation (5% at a period of 4
    steps): the nozzle faces bit-exact, cells bit-exact, velocities and
    particles within the loop's bars."""
    jet = J.vertical_jet(16, 16, 24, h=0.01, nozzle=4, inflow=1.0)
    kw = dict(jet["params"], layer_max=2, integrator=integrator, particle_rk=prk, seed=7, inflow_speed=(0.0, 0.0, -1.0),
              inflow_amp=0.05, inflow_period=4)
    pair = both(**kw)
    upload_both(jet, pair)
    for step in range(5):
        sg, so = pair[0].step(), pair[1].step()
        # dt of a step reads the last step's projected speeds, which agree to 1e-8
        assert sg["dt"] == pytest.approx(so["dt"], rel=1e-9)
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        f = dg["flags"]
        noz = ((f & 0x80) > 0) & ((f & 7) == W.INFLOW)
        assert noz.sum() == 16 and np.array_equal(dg["w"][noz], do["w"][noz])
        ex = (f & W.EXISTS) > 0
        for k in ("u", "v", "w"):
            ref = do[k][ex]
            assert np.linalg.norm(dg[k][ex] - ref) <= 1e-8 * max(np.linalg.norm(ref), 1.0), (step, k)
        same_particles(dg["particles"], do["particles"], 1e-9)


def test_small_breakup_jet_vs_oracle():
    """The large-jet layout (pulsed nozzle, falling column, pool on a solid
    floor, open sides) at 32^3: three steps against the oracle."""
    b = J.breakup_jet(nx=32, ny=32, nz=32, nozzle=4, pool_depth=6, margin=4, period=4)
    kw = dict(b["params"], layer_max=2, integrator=0, particle_rk=2, seed=3, particles_per_inflow_cell=8)
    pair = both(**kw)
    upload_both(b, pair, particles=b["particles"])
    for step in range(3):
        sg, so = pair[0].step(), pair[1].step()
        # dt of a step reads the last step's projected speeds, which agree to 1e-8
        assert sg["dt"] == pytest.approx(so["dt"], rel=1e-9)
        dg, do = pair[0].download(), pair[1].download()
        same_cells(dg, do)
        ex = (dg["flags"] & W.EXISTS) > 0
        for k in ("u", "v", "w"):
            ref = do[k][ex]
            assert np.linalg.norm(dg[k][ex] - ref) <= 1e-8 * max(np.linalg.norm(ref), 1.0), (step, k)
        same_particles(dg["particles"], do["particles"], 1e-9)


def test_resume_from_checkpoint_on_device(tmp_path):
    """The device time step resumed from the output set (grid file with layers,
    depth and the step counter, particle file) after 4 steps of the
    depth-chain jet with a pulsed inflow equals the uninterrupted device run
    4 steps later, bit for bit."""
    from tests.test_vtk_io import resume_matches, same_state
    same_state(*resume_matches(lambda kw: Sim(**kw, proj={"rtol": 1e-12}), tmp_path))
