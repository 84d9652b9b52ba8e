"""Shared cases of the free-surface velocity boundary condition: cell sets on
a brick grid, their face fields, and lookups into a downloaded state.  Used by
the oracle tests (test_surface_bc.py) and the device tests
(test_gpu_surface_bc.py)."""
import json
import os

import numpy as np

from mac_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
AIRS = (W.AIR, W.SOLID)  # Material_To_Type_Fluid maps Solid to Air


def load_kats():
    with open(os.path.join(HERE, "golden", "extrap_bc_2d.json")) as f:
        return json.load(f)


class Grid:
    """cells: dict (x, y, z) -> (material, u, v, w) on a brick grid; cells not
    listed do not exist (their bricks may)."""

    def __init__(self, cells, const=None):
        bkeys = sorted({(x >> 3, y >> 3, z >> 3) for (x, y, z) in cells}, key=lambda k: (k[2], k[1], k[0]))
        self.index = {k: i for i, k in enumerate(bkeys)}
        nb = len(bkeys)
        self.keys = np.array(bkeys, np.int32).reshape(-1, 3)
        self.flags = np.zeros(nb * 512, np.uint8)
        self.layer = np.full(nb * 512, -1, np.int8)
        self.u, self.v, self.w = (np.zeros(nb * 512) for _ in range(3))
        for (x, y, z), (mat, a, b, c) in cells.items():
            s = self.slot(x, y, z)
            self.flags[s] = mat | W.EXISTS
            self.layer[s] = 0 if mat == W.FLUID else 1
            self.u[s], self.v[s], self.w[s] = a, b, c
        for (x, y, z), comps in (const or {}).items():
            for i in comps:
                self.flags[self.slot(x, y, z)] |= W.CONST[i]
        self.cells = cells

    def slot(self, x, y, z):
        return self.index[(x >> 3, y >> 3, z >> 3)] * 512 + (x & 7) + 8 * (y & 7) + 64 * (z & 7)

    def fields(self):
        return {"keys": self.keys, "flags": self.flags, "layer": self.layer, "u": self.u, "v": self.v, "w": self.w}


class State:
    """A downloaded state, looked up by cell coordinates."""

    def __init__(self, d):
        self.d = d
        self.index = {tuple(k): i for i, k in enumerate(d["keys"].tolist())}

    def slot(self, x, y, z):
        b = self.index.get((x >> 3, y >> 3, z >> 3))
        return None if b is None else b * 512 + (x & 7) + 8 * (y & 7) + 64 * (z & 7)

    def face(self, comp, x, y, z):
        s = self.slot(x, y, z)
        return None if s is None else self.d["uvw"[comp]][s]


def from_mask(fluid, origin=(0, 0, 0), pad=2, fields=None):
    """Cells of a boolean fluid mask (indexed [x, y, z]) placed at origin,
    inside an air shell `pad` cells thick; fields(x, y, z) -> (u, v, w) in the
    mask's own coordinates (default 0)."""
    nx, ny, nz = fluid.shape
    ox, oy, oz = origin
    cells = {}
    for z in range(-pad, nz + pad):
        for y in range(-pad, ny + pad):
            for x in range(-pad, nx + pad):
                fl = 0 <= x < nx and 0 <= y < ny and 0 <= z < nz and fluid[x, y, z]
                uvw = fields(x, y, z) if fields else (0.0, 0.0, 0.0)
                cells[(x + ox, y + oy, z + oz)] = (W.FLUID if fl else W.AIR,) + tuple(uvw)
    return cells


def random_fields(seed, shape, pad):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    f = rng.standard_normal((3, nx + 2 * pad, ny + 2 * pad, nz + 2 * pad))
    return lambda x, y, z: tuple(f[:, x + pad, y + pad, z + pad])


def blob(seed, shape, fill=0.5):
    """A seeded random fluid mask: box-filtered noise thresholded at `fill`."""
    rng = np.random.default_rng(seed)
    n = rng.standard_normal(shape)
    s = n.copy()
    for ax in range(3):
        s = s + np.roll(n, 1, ax) + np.roll(n, -1, ax)
    return s > np.quantile(s, 1.0 - fill)


def air_neighbours(cells, x, y, z, dim=3):
    """IsBoundary's (axis, side) list of a non-air cell, axis-major."""
    out = []
    for i in range(dim):
        for sg in (-1, 1):
            n = (x + sg * (i == 0), y + sg * (i == 1), z + sg * (i == 2))
            if n in cells and cells[n][0] in AIRS:
                out.append((i, sg))
    return out


def divergence(st, h, x, y, z):
    return sum((st.face(i, x + (i == 0), y + (i == 1), z + (i == 2)) - st.face(i, x, y, z)) / h[i] for i in range(3))
