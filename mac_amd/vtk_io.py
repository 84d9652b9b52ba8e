"""VTK XML output and checkpoint input of the time-step loop's state.

Files follow the reference's output policies:

* ``write_grid`` -- Policy_Output_Grid_Pressure_ID (OutputGridPressure,
  InputGridPressure): points at the low corner (key - 1/2) h of every
  existing cell plus the +e_i corners of cells without a +e_i neighbour (the
  reference's), then the remaining corners of existing cells, so that every
  existing cell is one VTK_VOXEL; cell arrays Pressure, Fluid, Type
  (MaterialToInt), Solid, Interior, Key, Speed, SpeedConst, SpeedAverage.  It
  is also the checkpoint: ``read_grid`` returns the bricks (keys, flag bytes,
  layers, speeds, pressure, depth) and the step counter to resume from;
  ``Flags``, ``Layer``, ``Depth`` and the field data ``Step`` and ``Time`` are
  added for that.
* ``write_speed`` -- Policy_Output_Grid_Speed_ID: one file per component,
  points at the face centres of the cells with a layer, VTK_VOXEL over faces
  whose 7 +e neighbours have a layer, point array speedx/speedy/speedz.
* ``write_particles`` / ``read_particles`` -- Policy_Output_Particle.
* ``Animation`` -- Policy_Output_Animation's .pvd collection.

Cells are written in ascending (z, y, x) chunk order and offset order inside a
chunk; the reference writes hash order, and both carry the keys.
Arrays are stored as base64 binary blocks with a UInt64 byte-count header.
"""
import base64
import os
import xml.etree.ElementTree as ET

import numpy as np

EXISTS = 0x80
VOXEL = 11
VERTEX = 1
_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64",
              np.dtype(np.uint8): "UInt8", np.dtype(np.int16): "Int16"}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


def _encode(a):
    a = np.ascontiguousarray(a)
    raw = a.tobytes()
    return base64.b64encode(np.uint64(len(raw)).tobytes() + raw).decode("ascii")


def _decode(text, dtype, ncomp):
    raw = base64.b64decode(text.strip())
    n = int(np.frombuffer(raw[:8], np.uint64)[0])
    a = np.frombuffer(raw[8:8 + n], dtype).copy()
    return a.reshape(-1, ncomp) if ncomp > 1 else a


def _array(parent, name, a, ncomp=1):
    a = np.asarray(a)
    e = ET.SubElement(parent, "DataArray", type=_VTK_TYPES[a.dtype], Name=name, NumberOfComponents=str(ncomp),
                      format="binary")
    e.text = _encode(a)
    return e


def _write_vtu(path, points, conn, offsets, types, cell_data=(), point_data=(), field_data=()):
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="1.0", byte_order="LittleEndian", header_type="UInt64")
    grid = ET.SubElement(root, "UnstructuredGrid")
    if field_data:
        fd = ET.SubElement(grid, "FieldData")
        for name, a in field_data:
            e = _array(fd, name, np.atleast_1d(a))
            e.set("NumberOfTuples", str(np.atleast_1d(a).size))
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(len(points)), NumberOfCells=str(len(types)))
    pd = ET.SubElement(piece, "PointData")
    for name, a, nc in point_data:
        _array(pd, name, a, nc)
    cd = ET.SubElement(piece, "CellData")
    for name, a, nc in cell_data:
        _array(cd, name, a, nc)
    pts = ET.SubElement(piece, "Points")
    _array(pts, "Points", np.asarray(points, np.float64).reshape(-1), 3)
    cells = ET.SubElement(piece, "Cells")
    _array(cells, "connectivity", np.asarray(conn, np.int64))
    _array(cells, "offsets", np.asarray(offsets, np.int64))
    _array(cells, "types", np.asarray(types, np.uint8))
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def read_vtu(path):
    """dict with points (n,3), connectivity, offsets, types, and the point and
    cell arrays by name (binary blocks of this module, or ascii)."""
    piece = ET.parse(path).getroot().find("UnstructuredGrid/Piece")

    def arrays(node):
        out = {}
        if node is None:
            return out
        for e in node.findall("DataArray"):
            dt = _NP_TYPES[e.get("type")]
            nc = int(e.get("NumberOfComponents", "1"))
            if e.get("format") == "ascii":
                a = np.array(e.text.split(), dtype=dt)
                out[e.get("Name")] = a.reshape(-1, nc) if nc > 1 else a
            else:
                out[e.get("Name")] = _decode(e.text, dt, nc)
        return out

    cells = arrays(piece.find("Cells"))
    return {"points": arrays(piece.find("Points"))["Points"], "connectivity": cells["connectivity"],
            "offsets": cells["offsets"], "types": cells["types"], "point_data": arrays(piece.find("PointData")),
            "cell_data": arrays(piece.find("CellData")),
            "field_data": arrays(ET.parse(path).getroot().find("UnstructuredGrid/FieldData"))}


def _cells(state):
    """Global keys (n,3) and flat indices of every existing cell, in (z, y, x)
    chunk order then offset order."""
    keys, flags = np.asarray(state["keys"]), np.asarray(state["flags"])
    nb = len(keys)
    t = np.arange(512)
    off = np.stack([t & 7, (t >> 3) & 7, t >> 6], 1)
    gk = (keys[:, None, :] * 8 + off[None, :, :]).reshape(-1, 3)
    idx = np.nonzero(flags[: nb * 512] & EXISTS)[0]
    return gk[idx], idx


_CORNERS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]


def write_grid(path, state, h, depth=None, step=None, time=None):
    """Policy_Output_Grid_Pressure_ID::OutputGridPressure of a downloaded
    state (mac_amd.lib.Sim.download / oracle.Sim.download).  The points are
    the reference's (the low corner of every existing cell, then the +e_i
    corners of cells without a +e_i neighbour) followed by every other corner
    of an existing cell, so that every existing cell has its voxel and the
    file is a complete checkpoint; Layer, Depth, Flags and, as field data, the
    step counter and the time are added for the resume."""
    h = np.asarray(h, np.float64)
    gk, idx = _cells(state)
    pmap = {}
    pts = []

    def add(k):
        if k not in pmap:
            pmap[k] = len(pts)
            pts.append((np.asarray(k, np.float64) - 0.5) * h)

    exist = {tuple(k) for k in gk.tolist()}
    for k in gk.tolist():
        k = tuple(k)
        add(k)
        for i in range(3):
            n = list(k)
            n[i] += 1
            n = tuple(n)
            if n not in exist:
                add(n)
    for k in gk.tolist():
        for d in _CORNERS:
            add(tuple(k[j] + d[j] for j in range(3)))
    conn, cell_idx, cell_keys = [], [], []
    for k, c in zip(gk.tolist(), idx.tolist()):
        corners = [tuple(k[j] + d[j] for j in range(3)) for d in _CORNERS]
        if all(q in pmap for q in corners):
            conn.extend(pmap[q] for q in corners)
            cell_idx.append(c)
            cell_keys.append(k)
    ci = np.asarray(cell_idx, np.int64)
    flags = np.asarray(state["flags"])[ci]
    mat = (flags & 7).astype(np.int32)
    speed = np.stack([np.asarray(state[k])[ci] for k in ("u", "v", "w")], 1)
    const = np.stack([((flags >> (3 + i)) & 1).astype(np.int32) for i in range(3)], 1)
    # SpeedAverage: mean of the two faces where the +e_i neighbour exists
    lut = {tuple(k): c for k, c in zip(gk.tolist(), idx.tolist())}
    avg = speed.copy()
    for r, k in enumerate(cell_keys):
        for i, comp in enumerate(("u", "v", "w")):
            n = list(k)
            n[i] += 1
            c2 = lut.get(tuple(n))
            if c2 is not None:
                avg[r, i] = (speed[r, i] + np.asarray(state[comp])[c2]) / 2
    cell_data = [("Pressure", np.asarray(state["p"])[ci], 1), ("Fluid", np.isin(mat, (1, 3, 4, 5)).astype(np.int32), 1),
                 ("Type", mat, 1), ("Solid", (mat == 2).astype(np.int32), 1), ("Interior", (mat == 4).astype(np.int32), 1),
                 ("Key", np.asarray(cell_keys, np.int32).reshape(-1), 3), ("Speed", speed.reshape(-1), 3),
                 ("SpeedConst", const.reshape(-1), 3), ("SpeedAverage", avg.reshape(-1), 3),
                 ("Flags", flags.astype(np.uint8), 1)]
    if depth is not None:
        cell_data.append(("Depth", np.asarray(depth, np.int16)[ci], 1))
    if "layer" in state:
        cell_data.append(("Layer", np.asarray(state["layer"]).astype(np.int16)[ci], 1))
    field = []
    if step is not None:
        field.append(("Step", np.int64(step)))
    if time is not None:
        field.append(("Time", np.float64(time)))
    n = len(cell_idx)
    _write_vtu(path, np.asarray(pts).reshape(-1, 3), conn, 8 * np.arange(1, n + 1), np.full(n, VOXEL), cell_data,
               field_data=field)
    return n


def read_grid(path):
    """InputGridPressure: the bricks of the cells a grid file holds, as
    dict(keys, flags, layer, u, v, w, p, depth, step, time) for Sim.upload /
    set_depth / set_step (step and time None when the file has none)."""
    vt = read_vtu(path)
    d = vt["cell_data"]
    gk = d["Key"].astype(np.int64)
    bk = np.floor_divide(gk, 8)
    uniq = sorted({tuple(k) for k in bk.tolist()}, key=lambda k: (k[2], k[1], k[0]))
    bid = {k: i for i, k in enumerate(uniq)}
    nb = len(uniq)
    out = {"keys": np.asarray(uniq, np.int32).reshape(-1, 3), "flags": np.zeros(nb * 512, np.uint8),
           "p": np.zeros(nb * 512), "depth": np.full(nb * 512, -1, np.int16), "layer": np.full(nb * 512, -1, np.int8)}
    fdat = vt["field_data"]
    out["step"] = int(fdat["Step"][0]) if "Step" in fdat else None
    out["time"] = float(fdat["Time"][0]) if "Time" in fdat else None
    for c in ("u", "v", "w"):
        out[c] = np.zeros(nb * 512)
    o = gk - bk * 8
    flat = np.asarray([bid[tuple(k)] for k in bk.tolist()], np.int64) * 512 + o[:, 0] + 8 * o[:, 1] + 64 * o[:, 2]
    out["flags"][flat] = d["Flags"]
    out["p"][flat] = d["Pressure"]
    for i, c in enumerate(("u", "v", "w")):
        out[c][flat] = d["Speed"][:, i]
    if "Depth" in d:
        out["depth"][flat] = d["Depth"]
    if "Layer" in d:
        out["layer"][flat] = d["Layer"].astype(np.int8)
    return out


def write_speed(prefix, step, state, h):
    """Policy_Output_Grid_Speed_ID::OutputGridSpeed: files prefix_<c>_<step>.vtu
    for c = 1, 2, 3.  Returns the file names."""
    h = np.asarray(h, np.float64)
    keys, flags = np.asarray(state["keys"]), np.asarray(state["flags"])
    layer = np.asarray(state["layer"])
    gk, idx = _cells(state)
    sel = layer[idx] != -1
    gk, idx = gk[sel], idx[sel]
    pmap = {tuple(k): n for n, k in enumerate(gk.tolist())}
    names = []
    for comp, name in enumerate(("speedx", "speedy", "speedz")):
        pts = gk.astype(np.float64) * h
        pts[:, comp] -= 0.5 * h[comp]
        conn = []
        for k in gk.tolist():
            corners = [(k[0] + d[0], k[1] + d[1], k[2] + d[2]) for d in _CORNERS]
            if all(q in pmap for q in corners):
                conn.extend(pmap[q] for q in corners)
        n = len(conn) // 8
        path = f"{prefix}_{comp + 1}_{step}.vtu"
        _write_vtu(path, pts, conn, 8 * np.arange(1, n + 1), np.full(n, VOXEL),
                   point_data=[(name, np.asarray(state["uvw"[comp]])[idx], 1)])
        names.append(path)
    return names


def write_particles(prefix, step, particles):
    """Policy_Output_Particle::OutputTopology: prefix<step>.vtu, one vertex per
    particle."""
    pts = np.asarray(particles, np.float64).reshape(-1, 3)
    n = len(pts)
    path = f"{prefix}{step}.vtu"
    _write_vtu(path, pts, np.arange(n), np.arange(1, n + 1), np.full(n, VERTEX))
    return path


def read_particles(path):
    """Policy_Output_Particle::InputTopology."""
    return read_vtu(path)["points"].reshape(-1, 3)


class Animation:
    """Policy_Output_Animation: a .pvd collection, one DataSet per file and
    output time (the file is kept closed and valid after every output)."""

    def __init__(self, path):
        self.path = path
        self.entries = []

    def output(self, files, t):
        for part, f in enumerate(files, start=1):
            self.entries.append((t, part, os.path.relpath(f, os.path.dirname(os.path.abspath(self.path)))))
        with open(self.path, "w") as fh:
            fh.write('<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n<Collection>\n')
            for t, part, f in self.entries:
                fh.write(f' <DataSet timestep="{t!r}" group="" part="{part}" file="{f}" />\n')
            fh.write("</Collection>\n</VTKFile>\n")


def write_step(prefix, step, t, state, h, animation=None, depth=None):
    """Algorithms_Output's set for one output step: the grid file (checkpoint),
    the three speed files and the particle file, added to the animation."""
    files = [f"{prefix}press_{step}.vtu"]
    write_grid(files[0], state, h, depth, step=step, time=t)
    files += write_speed(f"{prefix}speed", step, state, h)
    files.append(write_particles(f"{prefix}topo_", step, state["particles"]))
    if animation is not None:
        animation.output(files, t)
    return files
