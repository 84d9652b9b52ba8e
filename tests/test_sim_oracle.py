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
