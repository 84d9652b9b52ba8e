"""CPU checks pinning the time-step oracle (sim_oracle.cpp) with the
reference's known answers and with properties the mathematics fixes."""
import numpy as np
import pytest

from mac_amd import jets as J
from mac_amd import workloads as W
from oracle import oracle as O


def single_cell(layer_max=1):
    """One fluid particle in cell (0,0,0); nothing else exists."""
    s = O.Sim(h=(1.0, 1.0, 1.0), layer_max=layer_max, gravity=(0, 0, 0), particles_per_inflow_cell=0)
    s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), particles=np.array([[0.1, -0.2, 0.3]]))
    s.topology()
    return s.download()


def cell_layers(d):
    out = {}
    for b, k in enumerate(d["keys"]):
        for t in range(512):
            if d["flags"][b * 512 + t] & W.EXISTS:
                out[(8 * k[0] + (t & 7), 8 * k[1] + ((t >> 3) & 7), 8 * k[2] + (t >> 6))] = (
                    int(d["layer"][b * 512 + t]), int(d["flags"][b * 512 + t] & 7))
    return out


def test_layer_kat_2d_slice():
    """TestAlgorithms_Update_CellType_Layer.h:43-150 (LayerMax = 1, one fluid
    cell): the 4 in-plane neighbours get layer 1 and (2,0), (1,1), (1,-1),
    (0,2), (-1,1) get layer 2 (the +1-only extra layer).  In 3-D the z = 0
    slice reproduces the 2-D answer."""
    cl = cell_layers(single_cell(1))
    sl = {(x, y): v for (x, y, z), v in cl.items() if z == 0}
    assert sl[(0, 0)] == (0, W.FLUID)
    for c in [(1, 0), (-1, 0), (0, 1), (0, -1)]:
        assert sl[c] == (1, W.AIR)
    for c in [(2, 0), (1, 1), (1, -1), (0, 2), (-1, 1)]:
        assert sl[c] == (2, W.AIR)
    assert len(sl) == 10


def test_layer_bfs_matches_distance():
    """Layers are the L1 distance to the fluid up to layer_max, and the extra
    layer is exactly the empty +e_i neighbours of the last layer."""
    for lmax in (1, 2, 3):
        cl = cell_layers(single_cell(lmax))
        for (x, y, z), (lay, mat) in cl.items():
            dist = abs(x) + abs(y) + abs(z)
            if dist <= lmax:
                assert lay == dist
            else:
                assert lay == lmax + 1
                assert any(abs(x - (i == 0)) + abs(y - (i == 1)) + abs(z - (i == 2)) == lmax for i in range(3))
        brute = {(x, y, z) for x in range(-5, 6) for y in range(-5, 6) for z in range(-5, 6)
                 if abs(x) + abs(y) + abs(z) <= lmax}
        extra = {(x + (i == 0), y + (i == 1), z + (i == 2)) for (x, y, z) in brute
                 if abs(x) + abs(y) + abs(z) == lmax for i in range(3)} - brute
        assert set(cl) == brute | extra


def test_delete_frees_cells_and_chunks():
    """Algorithms_Delete_MacCell: when the particle moves away, the old cells
    are erased and empty chunks freed."""
    s = O.Sim(h=(1.0, 1.0, 1.0), layer_max=1, gravity=(0, 0, 0), particles_per_inflow_cell=0)
    s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), particles=np.array([[0.0, 0.0, 0.0]]))
    s.topology()
    d0 = s.download()
    # the origin chunk and the three chunks reached by the -x, -y, -z neighbours
    assert {tuple(k) for k in d0["keys"]} == {(0, 0, 0), (-1, 0, 0), (0, -1, 0), (0, 0, -1)}
    s.upload(d0["keys"], d0["flags"], d0["layer"], particles=np.array([[40.0, 40.0, 40.0]]))
    s.topology()
    d1 = s.download()
    assert {tuple(int(c) for c in k) for k in d1["keys"]} == {(5, 5, 5), (4, 5, 5), (5, 4, 5), (5, 5, 4)}


def test_inflow_fill_jitter_inside_cell():
    jet = J.vertical_jet(8, 8, 16, h=0.5, nozzle=2)
    s = O.Sim(**jet["params"], layer_max=2)
    s.upload(jet["keys"], jet["flags"], u=jet["u"], v=jet["v"], w=jet["w"])
    s.topology()
    d = s.download()
    parts = d["particles"]
    assert parts.shape == (4 * 8, 3)
    keys = np.rint(parts / 0.5).astype(int)
    assert set(map(tuple, keys)) == {(x, y, 15) for x in (3, 4) for y in (3, 4)}


def test_interpolation_reproduces_linear_fields():
    """Staggered trilinear interpolation is exact for any linear field
    (Policy_Speed_Interpolation_Linear_Symmetric)."""
    g = W.random_grid(3, 3, 3, 3, fluid_p=1.0, solid_p=0.0, const_p=0.0)
    g["flags"][:] = W.FLUID | W.EXISTS
    keys = g["keys"]
    n = keys.shape[0]
    t = np.arange(512)
    gx = (keys[:, 0:1] * 8 + (t & 7)[None, :]).reshape(-1).astype(float)
    gy = (keys[:, 1:2] * 8 + ((t >> 3) & 7)[None, :]).reshape(-1).astype(float)
    gz = (keys[:, 2:3] * 8 + (t >> 6)[None, :]).reshape(-1).astype(float)
    h = 0.5
    A = np.array([[0.3, -1.2, 0.7, 0.1], [2.0, 0.5, -0.4, -0.3], [-0.6, 0.2, 1.1, 0.9]])
    # component k lives on the low face of cell (x,y,z): position (x,y,z) - 0.5 e_k, times h
    comps = []
    for k in range(3):
        X = np.stack([gx - 0.5 * (k == 0), gy - 0.5 * (k == 1), gz - 0.5 * (k == 2)]) * h
        comps.append(A[k, 0] * X[0] + A[k, 1] * X[1] + A[k, 2] * X[2] + A[k, 3])
    s = O.Sim(h=(h, h, h))
    s.upload(keys, g["flags"], u=comps[0], v=comps[1], w=comps[2])
    rng = np.random.default_rng(0)
    pos = rng.uniform(2.0, 20.0, size=(200, 3)) * h
    vel = s.interpolate(pos)
    exact = pos @ A[:, :3].T + A[:, 3]
    assert np.allclose(vel, exact, rtol=0, atol=1e-12)


def test_upwind_of_uniform_field_is_zero_and_gravity_on_fluid():
    """First-order upwind of a uniform velocity is 0; gravity lands on the
    faces of fluid cells and of non-empty cells above fluid."""
    jet = J.vertical_jet(16, 16, 16, h=0.1, nozzle=2, floor=False)
    g = W.random_grid(5, 2, 2, 2, fluid_p=1.0, solid_p=0.0, const_p=0.0)
    flags = np.full(g["flags"].size, W.FLUID | W.EXISTS, np.uint8)
    n = flags.size
    s = O.Sim(h=(0.1, 0.1, 0.1), gravity=(0, 0, -9.81))
    s.upload(g["keys"], flags, np.zeros(n, np.int8), u=np.full(n, 0.7), v=np.full(n, -0.2), w=np.full(n, 0.4))
    s.forces(0)
    d = s.download()
    # interior cells (away from the box boundary, where missing neighbours read 0)
    dense = {k: W.bricks_to_dense(d[k].reshape(-1, 512), 2, 2, 2) for k in ("ax", "ay", "az")}
    inner = (slice(2, -2),) * 3
    assert np.all(dense["ax"][inner] == 0.0) and np.all(dense["ay"][inner] == 0.0)
    assert np.allclose(dense["az"][inner], -9.81, rtol=0, atol=1e-12)
    del jet


def _all_fluid(seed=7, nb=2, h=0.1):
    g = W.random_grid(seed, nb, nb, nb, fluid_p=1.0, solid_p=0.0, const_p=0.0)
    flags = np.full(g["flags"].size, W.FLUID | W.EXISTS, np.uint8)
    keys = g["keys"]
    t = np.arange(512)
    X = (keys[:, 0:1] * 8 + (t & 7)[None, :]).reshape(-1) * h
    Y = (keys[:, 1:2] * 8 + ((t >> 3) & 7)[None, :]).reshape(-1) * h
    Z = (keys[:, 2:3] * 8 + (t >> 6)[None, :]).reshape(-1) * h
    return keys, flags, X, Y, Z


def test_dt_closed_form_uniform_flow():
    """Algorithms_Calculate_Time_Step.h:35-66: dt = cfl / sqrt(max_c sum_i
    (u_i ih_i)^2); for a uniform field that is cfl / |u / h|, clamped."""
    h = (0.1, 0.2, 0.05)
    keys, flags, X, _, _ = _all_fluid(h=1.0)
    n = flags.size
    a, b, c = 0.7, -1.3, 0.4
    for cfl, lo, hi in ((0.5, 0.0, 1e30), (0.9, 0.0, 1e-3), (0.5, 0.5, 1e30)):
        s = O.Sim(h=h, cfl=cfl, dt_min=lo, dt_max=hi, gravity=(0, 0, 0))
        s.upload(keys, flags, np.zeros(n, np.int8), u=np.full(n, a), v=np.full(n, b), w=np.full(n, c))
        dt = s.compute_dt()
        exact = cfl / np.sqrt((a / h[0]) ** 2 + (b / h[1]) ** 2 + (c / h[2]) ** 2)
        assert abs(dt - min(max(exact, lo), hi)) <= 1e-15 * exact, (dt, exact)


def test_viscosity_of_quadratic_field_is_exact():
    """Algorithms_Viscosity / Policy_Laplacian_Speed (7-point): nu * lap of
    u = Y^2 + 2 Z^2 is 6 nu in the interior; with v = w = 0 and u independent
    of x the upwind convection term vanishes."""
    h, nu = 0.1, 0.37
    keys, flags, X, Y, Z = _all_fluid(h=h)
    n = flags.size
    s = O.Sim(h=(h, h, h), viscosity=nu, gravity=(0, 0, 0))
    s.upload(keys, flags, np.zeros(n, np.int8), u=Y ** 2 + 2 * Z ** 2, v=np.zeros(n), w=np.zeros(n))
    s.forces(0)
    d = s.download()
    ax = W.bricks_to_dense(d["ax"].reshape(-1, 512), 2, 2, 2)
    inner = (slice(2, -2),) * 3
    assert np.allclose(ax[inner], 6 * nu, rtol=0, atol=1e-10)
    assert np.all(W.bricks_to_dense(d["ay"].reshape(-1, 512), 2, 2, 2)[inner] == 0.0)


@pytest.mark.parametrize("prk", [2, 4])
def test_particles_in_uniform_flow_move_exactly(prk):
    """Algorithms_Move_Particles + Policy_Advance_ODE_RungeKutta: RK2 and RK4
    of a uniform velocity field move every particle by dt * u."""
    h = 0.1
    keys, flags, *_ = _all_fluid(h=h)
    n = flags.size
    vel = (0.3, -0.25, 0.15)
    s = O.Sim(h=(h, h, h), gravity=(0, 0, 0), particle_rk=prk, particles_per_inflow_cell=0)
    rng = np.random.default_rng(1)
    pos = rng.uniform(4 * h, 12 * h, size=(50, 3))
    s.upload(keys, flags, np.zeros(n, np.int8), u=np.full(n, vel[0]), v=np.full(n, vel[1]), w=np.full(n, vel[2]),
             particles=pos)
    dt = 0.01
    s.move_particles(dt)
    out = s.download()["particles"]
    assert np.allclose(out, pos + dt * np.array(vel), rtol=0, atol=1e-14)


def _jet_state(steps=4):
    jet = J.vertical_jet(16, 16, 24, h=0.01, nozzle=4, inflow=1.0)
    s = O.Sim(**jet["params"], layer_max=3, seed=5)
    s.upload(jet["keys"], jet["flags"], u=jet["u"], v=jet["v"], w=jet["w"])
    for _ in range(steps):
        s.step()
    s.topology()
    return s


def test_extrapolation_of_uniform_field_is_uniform():
    """Algorithms_Extrapolate: every extrapolated face is a mean of faces that
    already hold the value, so a uniform field stays uniform (non-const faces
    of every cell with a layer), and fluid-fluid faces are never written."""
    s = _jet_state()
    d = s.download()
    ex = (d["flags"] & W.EXISTS) > 0
    vals = (0.25, -1.5, 0.75)
    n = d["flags"].size
    s.upload(d["keys"], d["flags"], d["layer"], *[np.where(ex, vals[i], 0.0) for i in range(3)])
    s.extrapolate()
    e = s.download()
    for i, k in enumerate(("u", "v", "w")):
        touched = ex & (d["layer"] >= 0)
        got = e[k][touched]
        assert np.all((got == vals[i]) | np.isnan(got)), k
        assert np.isfinite(got).mean() > 0.99
    del n


def test_extrapolation_leaves_lower_layers_and_is_idempotent():
    """Faces between two cells of layer 0 are read, never written; running the
    extrapolation twice gives the same field (it only reads faces that touch a
    lower layer, which it does not write)."""
    s = _jet_state()
    d = s.download()
    rng = np.random.default_rng(3)
    ex = (d["flags"] & W.EXISTS) > 0
    comps = [np.where(ex, rng.normal(size=d["flags"].size), 0.0) for _ in range(3)]
    s.upload(d["keys"], d["flags"], d["layer"], *comps)
    s.extrapolate()
    e1 = s.download()
    s.extrapolate()
    e2 = s.download()
    for k in ("u", "v", "w"):
        assert np.array_equal(e1[k], e2[k], equal_nan=True), k
    # the -x face of a layer-0 cell whose -x neighbour is layer 0 is untouched
    keys = d["keys"]
    lay = d["layer"].reshape(-1, 512)
    t = np.arange(512)
    for b in range(keys.shape[0]):
        for tt in t[(lay[b] == 0) & ((t & 7) > 0)]:
            if lay[b, tt - 1] == 0:
                assert e1["u"][b * 512 + tt] == comps[0][b * 512 + tt]


# ------------------------------------------------------------ depth chain ----
def _blob_state(mask, h=1.0):
    """One particle at the centre of every cell of `mask` (x, y, z indexing),
    nothing uploaded but the particles."""
    pts = np.argwhere(mask).astype(np.float64) * h
    return pts


def _depth_by_key(s):
    d = s.download()
    dep = s.get_depth()
    out = {}
    for b, k in enumerate(d["keys"]):
        for t in range(512):
            f = d["flags"][b * 512 + t]
            if f & W.EXISTS:
                key = (8 * k[0] + (t & 7), 8 * k[1] + ((t >> 3) & 7), 8 * k[2] + (t >> 6))
                out[key] = (int(d["layer"][b * 512 + t]), int(dep[b * 512 + t]), int(f & 7))
    return out, d


def _taxicab_depth(mask):
    """Depth closed form: a cell's 6-connected distance, inside the mask, to
    the cells that touch a non-mask cell is its taxicab distance to the
    nearest non-mask cell minus one (the cells of a monotone path to the
    nearest outside cell are all inside)."""
    from scipy import ndimage
    pad = np.pad(mask, 1)
    return ndimage.distance_transform_cdt(pad, metric="taxicab")[1:-1, 1:-1, 1:-1] - 1


@pytest.mark.parametrize("shape", ["box", "ball", "random"])
def test_depth_is_taxicab_distance_to_the_surface(shape):
    """Algorithms_Update_CellType_Layer_Depth.h:31-133: depth 0 on the
    surface, then the BFS level; depth < DepthInterior gives Fluid, the rest
    Interior; air cells keep depth -1."""
    n = 14
    g = np.indices((n, n, n))
    if shape == "box":
        mask = np.ones((n, n, n), bool)
    elif shape == "ball":
        mask = ((g - (n - 1) / 2.0) ** 2).sum(0) <= (n / 2.0) ** 2
    else:
        rng = np.random.default_rng(3)
        mask = rng.random((n, n, n)) < 0.8
    want = _taxicab_depth(mask)
    s = O.Sim(h=(1.0, 1.0, 1.0), layer_max=2, gravity=(0, 0, 0), particles_per_inflow_cell=0, depth_interior=3,
              depth_remove=5)
    s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), particles=_blob_state(mask))
    s.topology()
    cells, _ = _depth_by_key(s)
    for (x, y, z), (lay, dep, mat) in cells.items():
        inside = 0 <= x < n and 0 <= y < n and 0 <= z < n and mask[x, y, z]
        if inside:
            assert lay == 0
            assert dep == want[x, y, z]
            assert mat == (W.FLUID if dep < 3 else W.INTERIOR)
        else:
            assert lay > 0 and dep == -1 and mat == W.AIR
    if shape == "box":  # closed form: min distance to a face
        x, y, z = g
        box = np.minimum.reduce([x, n - 1 - x, y, n - 1 - y, z, n - 1 - z])
        assert np.array_equal(box, want)


def test_depth_interior_particles_removed_and_refilled():
    """Algorithms_Layer_Initial_With_Particle_Depth.h:29-46: a particle in an
    Interior cell deeper than DepthRemove is erased; the cell keeps layer 0.
    Algorithms_Create_Fluid_Particle_Depth.h:29-44: an Interior cell left
    without particles and no deeper than DepthRemove is refilled."""
    n, ppc = 16, 2
    mask = np.ones((n, n, n), bool)
    depth = _taxicab_depth(mask)
    s = O.Sim(h=(1.0, 1.0, 1.0), layer_max=2, gravity=(0, 0, 0), particles_per_inflow_cell=ppc, depth_interior=3,
              depth_remove=5, seed=11)
    pts = _blob_state(mask)
    s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), particles=pts)
    s.topology()
    d = s.download()
    # empty one Interior cell of depth 4 (refilled) and one of depth 7 (not)
    hole4 = (4, 8, 8)
    hole7 = (7, 7, 8)
    assert depth[hole4] == 4 and depth[hole7] == 7
    keep = ~np.all(pts == np.array(hole4, float), axis=1) & ~np.all(pts == np.array(hole7, float), axis=1)
    s.upload(d["keys"], d["flags"], d["layer"], particles=pts[keep])
    s.set_depth(_depth_by_key_array(d, depth))
    s.topology()
    cells, d2 = _depth_by_key(s)
    left = d2["particles"]
    keys = np.rint(left).astype(int)
    deep = depth[keys[:, 0], keys[:, 1], keys[:, 2]]
    # no particle is left in a cell deeper than 5, except the refill of hole4
    assert (deep <= 5).all()
    n_expected = int(((depth <= 5) & mask).sum()) - 1 + ppc
    assert len(left) == n_expected
    at4 = np.all(keys == np.array(hole4), axis=1)
    assert at4.sum() == ppc
    for c in (hole4, hole7):
        assert cells[c][0] == 0 and cells[c][2] == W.INTERIOR
    # the geometry did not change, so neither did the depth
    for (x, y, z), (lay, dep, mat) in cells.items():
        if 0 <= min(x, y, z) and max(x, y, z) < n:
            assert dep == depth[x, y, z]


def _depth_by_key_array(d, depth):
    """The depth field in download order for cells of `depth` (others -1)."""
    out = np.full(len(d["keys"]) * 512, -1, np.int16)
    n = depth.shape[0]
    for b, k in enumerate(d["keys"]):
        for t in range(512):
            x, y, z = 8 * k[0] + (t & 7), 8 * k[1] + ((t >> 3) & 7), 8 * k[2] + (t >> 6)
            if 0 <= min(x, y, z) and max(x, y, z) < n:
                out[b * 512 + t] = depth[x, y, z]
    return out


def test_floor_erases_particles_before_marking():
    """Algorithm_DelParticle (test/Jet3.cpp:259-289): particles below the
    floor are gone before any cell is marked, so they create no cell."""
    pts = np.array([[0.0, 0.0, 0.0], [0.0, -3.0, 0.0], [5.0, -0.4, 0.0], [9.0, -0.6, 0.0]])
    s = O.Sim(h=(1.0, 1.0, 1.0), layer_max=1, gravity=(0, 0, 0), particles_per_inflow_cell=0, floor_axis=1,
              floor_value=-0.5)
    s.upload(np.zeros((0, 3), np.int32), np.zeros(0, np.uint8), particles=pts)
    s.topology()
    cells, d = _depth_by_key(s)
    assert np.array_equal(d["particles"], pts[[0, 2]])
    fluid = {k for k, v in cells.items() if v[0] == 0}
    assert fluid == {(0, 0, 0), (5, 0, 0)}


# ---------------------------------------------------- RK4 "all" integrator ----
def test_rk4all_uniform_flow_is_kept_and_particles_move_by_dt_u():
    """Algorithms_RungeKutta_all on a uniform flow with no gravity and no
    viscosity: every stage's acceleration is 0 (upwind convection of a
    constant field, extrapolated into the air layers with dyadic values so
    that the means are exact), its projection is 0, so the grid speed is
    unchanged and the particle stages add dt (1/6 + 1/3 + 1/3 + 1/6) u."""
    h = 0.125
    vel = np.array([0.25, -0.5, 0.125])
    cells = np.argwhere(np.ones((8, 8, 8), bool)) + 4
    pts = cells * h + np.array([1, 3, -5]) * (h / 64)
    s = O.Sim(h=(h, h, h), gravity=(0, 0, 0), integrator=2, particles_per_inflow_cell=0, layer_max=2, cfl=1e9,
              dt_min=0.0625, dt_max=0.0625)
    nb_keys = np.array(sorted({tuple(k) for k in (cells // 8).tolist()}, key=lambda k: (k[2], k[1], k[0])), np.int32)
    n = len(nb_keys) * 512
    s.upload(nb_keys, np.zeros(n, np.uint8), u=np.full(n, vel[0]), v=np.full(n, vel[1]), w=np.full(n, vel[2]), particles=pts)
    st = s.step()
    assert st["dt"] == 0.0625
    d = s.download()
    ex = (d["flags"] & W.EXISTS) > 0
    fluid = ex & np.isin(d["flags"] & 7, [W.FLUID])
    for k, comp in zip(("u", "v", "w"), vel):
        assert np.all(d[k][fluid] == comp)
    assert np.allclose(d["particles"], pts + 0.0625 * vel, rtol=0, atol=1e-15)


def test_rk4all_free_fall_follows_the_reference_staging():
    """A blob at rest under gravity g: every stage's acceleration is g on
    the z faces (its divergence vanishes, so the projection keeps it), the
    grid speed ends at g dt, and the particles, which read the stage speed
    after the grid loop has written the next stage's value (stage 2 reads
    g dt/2, stages 3 and 4 read g dt), move by dt^2 g (1/2 (1/3) + 1 (1/3) +
    1 (1/6)) = (2/3) g dt^2 -- not the g dt^2 / 2 of free fall, a property
    of Algorithms_RungeKutta_all.h:84-137."""
    h, g, dt = 0.125, -8.0, 0.0625
    cells = np.argwhere(np.ones((8, 8, 8), bool)) + 4
    pts = cells * h + np.array([1, 3, -5]) * (h / 64)
    s = O.Sim(h=(h, h, h), gravity=(0, 0, g), integrator=2, particles_per_inflow_cell=0, layer_max=2, cfl=1e9,
              dt_min=dt, dt_max=dt)
    keys = np.array(sorted({tuple(k) for k in (cells // 8).tolist()}, key=lambda k: (k[2], k[1], k[0])), np.int32)
    s.upload(keys, np.zeros(len(keys) * 512, np.uint8), particles=pts)
    s.step()
    d = s.download()
    fluid = ((d["flags"] & W.EXISTS) > 0) & ((d["flags"] & 7) == W.FLUID)
    assert fluid.sum() == len(cells)
    assert np.allclose(d["w"][fluid], g * dt, rtol=0, atol=1e-15)  # the 1/6 and 1/3 terms round
    assert np.all(d["u"][fluid] == 0.0) and np.all(d["v"][fluid] == 0.0)
    want = pts + np.array([0.0, 0.0, 2.0 / 3.0 * g * dt * dt])
    assert np.allclose(d["particles"], want, rtol=0, atol=1e-15)


def test_inflow_pulsation_is_a_sinusoid_of_the_step():
    """The pulsed inflow face speed v_n = v0 (1 + a sin(2 pi n / T + phi)):
    with T = 8, v_n + v_{n+4} = 2 v0 (half a period apart), (v_n - v0)^2 +
    (v_{n+2} - v0)^2 = (a v0)^2 (a quarter apart) and v_{n+8} = v_n, the same
    on every Inflow cell's const face."""
    jet = J.vertical_jet(16, 16, 24, h=0.01, nozzle=4, inflow=1.0)
    v0, a, T = -1.5, 0.05, 8
    kw = dict(jet["params"], layer_max=2, integrator=0, particle_rk=2, seed=11, particles_per_inflow_cell=0,
              inflow_speed=(0.0, 0.0, v0), inflow_amp=a, inflow_period=T)
    o = O.Sim(**kw)
    o.upload(jet["keys"], jet["flags"], u=jet["u"], v=jet["v"], w=jet["w"])
    vals = []
    for _ in range(12):
        o.step()
        d = o.download()
        f = d["flags"]
        sel = ((f & 0x80) > 0) & ((f & 7) == W.INFLOW)
        assert sel.sum() == 16
        w = np.unique(d["w"][sel])
        assert w.size == 1
        vals.append(float(w[0]))
    v = np.array(vals)
    assert np.allclose(v[:8] + v[4:12], 2 * v0, rtol=0, atol=1e-14)
    assert np.allclose((v[:10] - v0) ** 2 + (v[2:12] - v0) ** 2, (a * v0) ** 2, rtol=1e-10, atol=0)
    assert np.allclose(v[8:12], v[0:4], rtol=0, atol=1e-14)
    assert np.ptp(v) > 1.5 * a * abs(v0)
