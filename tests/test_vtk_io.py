"""VTK output and checkpoint round trips (mac_amd/vtk_io.py), on states of
the CPU oracle's time step."""
import xml.etree.ElementTree as ET

import numpy as np

from mac_amd import jets as J
from mac_amd import vtk_io as V
from mac_amd import workloads as W
from oracle import oracle as O


def _state(steps=3):
    jet = J.vertical_jet(16, 16, 24, h=0.01, nozzle=4, inflow=1.0)
    s = O.Sim(**jet["params"], layer_max=2, seed=3, depth_interior=2, depth_remove=3)
    s.upload(jet["keys"], jet["flags"], None, jet["u"], jet["v"], jet["w"], None)
    for _ in range(steps):
        s.step()
    return s, s.download(), jet["params"]["h"]


def test_grid_checkpoint_round_trip(tmp_path):
    """InputGridPressure(OutputGridPressure(state)) gives back every existing
    cell bit for bit (flags, layers, speeds, pressure, depth) and the step
    counter."""
    s, d, h = _state()
    dep = s.get_depth()
    path = str(tmp_path / "press_3.vtu")
    n = V.write_grid(path, d, h, dep, step=s.get_step(), time=0.125)
    r = V.read_grid(path)
    gk, idx = V._cells(d)
    assert n == len(idx) > 0
    assert r["step"] == s.get_step() == 3 and r["time"] == 0.125
    lut = {tuple(k): c for k, c in zip(gk.tolist(), idx.tolist())}
    gr, ir = V._cells(r)
    assert len(ir) == n
    for k, c in zip(gr.tolist(), ir.tolist()):
        c0 = lut[tuple(k)]
        assert r["flags"][c] == d["flags"][c0]
        assert r["layer"][c] == d["layer"][c0]
        for f in ("u", "v", "w", "p"):
            assert r[f][c] == d[f][c0]
        assert r["depth"][c] == dep[c0]
    v = V.read_vtu(path)
    assert (v["types"] == V.VOXEL).all() and len(v["offsets"]) == n
    assert np.allclose(v["points"][v["connectivity"][:8]][1] - v["points"][v["connectivity"][:8]][0], [h[0], 0, 0])


def test_speed_and_particle_files(tmp_path):
    s, d, h = _state()
    names = V.write_speed(str(tmp_path / "speed"), 3, d, h)
    assert [n.rsplit("/", 1)[1] for n in names] == ["speed_1_3.vtu", "speed_2_3.vtu", "speed_3_3.vtu"]
    lay = (d["flags"] & W.EXISTS > 0) & (d["layer"] != -1)
    for comp, n in enumerate(names):
        v = V.read_vtu(n)
        arr = v["point_data"]["speed" + "xyz"[comp]]
        assert len(arr) == int(lay.sum())
        assert np.array_equal(np.sort(arr), np.sort(d["uvw"[comp]][lay]))
        # points sit on the component's faces: (key - 1/2 e_comp) h
        frac = v["points"][:, comp] / h[comp] + 0.5
        assert np.allclose(frac, np.round(frac))
    p = V.write_particles(str(tmp_path / "topo_"), 3, d["particles"])
    assert np.array_equal(V.read_particles(p), d["particles"])


def test_animation_collection(tmp_path):
    s, d, h = _state(1)
    anim = V.Animation(str(tmp_path / "Out_Animation.pvd"))
    for step in (5, 10):
        files = V.write_step(str(tmp_path) + "/out_", step, 0.01 * step, d, h, anim)
        assert len(files) == 5
    root = ET.parse(str(tmp_path / "Out_Animation.pvd")).getroot()
    ds = root.findall("Collection/DataSet")
    assert len(ds) == 10
    assert [d_.get("part") for d_ in ds[:5]] == ["1", "2", "3", "4", "5"]
    assert float(ds[5].get("timestep")) == 0.1


def _pulsed_jet():
    jet = J.vertical_jet(16, 16, 24, h=0.01, nozzle=4, inflow=1.0)
    kw = dict(jet["params"], layer_max=2, seed=3, depth_interior=2, depth_remove=3, inflow_speed=(0.0, 0.0, -1.0),
              inflow_amp=0.05, inflow_period=7)
    return jet, kw


def resume_matches(make_sim, tmp_path, first=4, then=4):
    """Steps `first`, writes the output set, reads it into a new simulation
    (upload, depth, step counter) and steps both `then` more times; returns
    (uninterrupted, resumed) downloads."""
    jet, kw = _pulsed_jet()
    a = make_sim(kw)
    a.upload(jet["keys"], jet["flags"], None, jet["u"], jet["v"], jet["w"], None)
    t = 0.0
    for _ in range(first):
        t += a.step()["dt"]
    d = a.download()
    files = V.write_step(str(tmp_path / "out_"), a.get_step(), t, d, kw["h"], depth=a.get_depth())
    r = V.read_grid(files[0])
    parts = V.read_particles(files[-1])
    assert r["step"] == first and r["time"] == t
    b = make_sim(kw)
    b.upload(r["keys"], r["flags"], r["layer"], r["u"], r["v"], r["w"], parts)
    b.set_depth(r["depth"])
    b.set_step(r["step"])
    for _ in range(then):
        a.step()
        b.step()
    return a.download(), b.download(), a.get_depth(), b.get_depth()


def same_state(da, db, dep_a, dep_b):
    assert np.array_equal(da["keys"], db["keys"]) and np.array_equal(da["flags"], db["flags"])
    ex = (da["flags"] & W.EXISTS) > 0
    assert np.array_equal(da["layer"][ex], db["layer"][ex]) and np.array_equal(dep_a[ex], dep_b[ex])
    for k in ("u", "v", "w", "p"):
        assert np.array_equal(da[k][ex], db[k][ex]), k
    pa = da["particles"][np.lexsort(da["particles"].T)]
    pb = db["particles"][np.lexsort(db["particles"].T)]
    assert np.array_equal(pa, pb)


def test_resume_from_checkpoint_matches_uninterrupted_run(tmp_path):
    """Algorithms_Output's files as a checkpoint (src/Algorithms_Output.h:27-37
    reloads speed, pressure and topology): a run resumed from them after 4
    steps of the depth-chain jet with a pulsed inflow equals the run that
    never stopped, bit for bit, 4 steps later (cells, layers, depth,
    velocities, pressure, particles)."""
    same_state(*resume_matches(lambda kw: O.Sim(**kw), tmp_path))
