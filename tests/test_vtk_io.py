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
    """InputGridPressure(OutputGridPressure(state)) gives back every cell that
    has its 8 corners in the file, bit for bit (flags, speeds, pressure,
    depth).  Cells whose far corners are not points are not in the file, as
    in the reference's format."""
    s, d, h = _state()
    dep = s.get_depth()
    path = str(tmp_path / "press_3.vtu")
    n = V.write_grid(path, d, h, dep)
    r = V.read_grid(path)
    assert n > 0
    gk, idx = V._cells(d)
    lut = {tuple(k): c for k, c in zip(gk.tolist(), idx.tolist())}
    gr, ir = V._cells(r)
    assert len(ir) == n
    for k, c in zip(gr.tolist(), ir.tolist()):
        c0 = lut[tuple(k)]
        assert r["flags"][c] == d["flags"][c0]
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
