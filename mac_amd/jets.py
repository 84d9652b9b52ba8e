"""Seeded jet set-ups for the time-step loop (brick format of include/mac_amd.h).

vertical_jet(): a box of nx x ny x nz cells (h per axis), open sides, solid
floor and a square inflow nozzle centred in the top layer.  Only the floor
and the nozzle exist at the start; fluid enters through the nozzle, whose
empty cells are refilled with jittered particles every step.
lateral_jet(): a horizontal nozzle on the x = 0 wall at mid-height, no floor.
"""
import numpy as np

from .workloads import AIR, CONST, EXISTS, INFLOW, SOLID, uniform

NODELETE = 0x40


def _bricks_of(cells):
    """cells: dict (x,y,z) -> (flag, speed).  Returns keys, flags and the speed
    component array."""
    bkeys = sorted({(x >> 3, y >> 3, z >> 3) for (x, y, z) in cells}, key=lambda k: (k[2], k[1], k[0]))
    index = {k: i for i, k in enumerate(bkeys)}
    nb = len(bkeys)
    flags = np.zeros((nb, 512), np.uint8)
    w = np.zeros((nb, 512))
    for (x, y, z), (f, wv) in cells.items():
        b = index[(x >> 3, y >> 3, z >> 3)]
        t = (x & 7) + 8 * (y & 7) + 64 * (z & 7)
        flags[b, t] = f
        w[b, t] = wv
    return np.array(bkeys, np.int32), flags.reshape(-1), w.reshape(-1)


def vertical_jet(nx=32, ny=32, nz=64, h=0.01, nozzle=4, inflow=1.0, floor=True):
    """Config-1 style jet: returns dict(keys, flags, u, v, w, params) where
    params are the keyword arguments of mac_amd.lib.Sim / oracle.Sim for the
    box (particle deletion outside it)."""
    cells = {}
    if floor:
        for x in range(nx):
            for y in range(ny):
                # the floor: solid, never deleted; the layer above keeps the
                # floor face const (no flow through it, Neumann for p)
                cells[(x, y, 0)] = (SOLID | EXISTS | NODELETE, 0.0)
                cells[(x, y, 1)] = (AIR | EXISTS | NODELETE | CONST[2], 0.0)
    x0, y0 = (nx - nozzle) // 2, (ny - nozzle) // 2
    for x in range(x0, x0 + nozzle):
        for y in range(y0, y0 + nozzle):
            # inflow cell: its low z face carries the prescribed downward speed
            cells[(x, y, nz - 1)] = (INFLOW | EXISTS | NODELETE | CONST[2], -float(inflow))
    keys, flags, w = _bricks_of(cells)
    n = flags.size
    params = dict(h=(h, h, h), gravity=(0.0, 0.0, -9.81), box_lo=(-0.5 * h, -0.5 * h, -0.5 * h),
                  box_hi=((nx - 0.5) * h, (ny - 0.5) * h, (nz - 0.5) * h))
    return {"keys": keys, "flags": flags, "u": np.zeros(n), "v": np.zeros(n), "w": w, "params": params}


def lateral_jet(nx=192, ny=96, nz=96, h=0.01, nozzle=6, inflow=None, seed=0):
    """Lateral jet: a horizontal square nozzle of `nozzle` x `nozzle` cells on
    the x = 0 wall at mid-height, flowing along +x; no floor, open sides, so
    particles leaving the box are deleted.  The inflow speed is uniform in
    [0.5, 2.0] m/s from the seed unless given.  Returns the dict of
    vertical_jet (params for a box that deletes particles leaving it)."""
    if inflow is None:
        inflow = 0.5 + 1.5 * float(uniform(seed, 7, np.zeros(1, dtype=np.uint64))[0])
    cells = {}
    y0, z0 = (ny - nozzle) // 2, (nz - nozzle) // 2
    for y in range(y0, y0 + nozzle):
        for z in range(z0, z0 + nozzle):
            # inflow cell: its low x face carries the prescribed speed into the box
            cells[(0, y, z)] = (INFLOW | EXISTS | NODELETE | CONST[0], float(inflow))
    keys, flags, u = _bricks_of(cells)
    n = flags.size
    params = dict(h=(h, h, h), gravity=(0.0, 0.0, -9.81), box_lo=(-0.5 * h, -0.5 * h, -0.5 * h),
                  box_hi=((nx - 0.5) * h, (ny - 0.5) * h, (nz - 0.5) * h))
    return {"keys": keys, "flags": flags, "u": u, "v": np.zeros(n), "w": np.zeros(n), "params": params,
            "inflow": float(inflow)}


def fluid_block(dims=(96, 96, 192), offset=(80, 80, 32), box=256, h=0.01, ppc=8, vel=(1.0, 0.0, 0.0), seed=0):
    """A large moving block of fluid for measuring the time step at scale: the
    cells of a dims-sized block at `offset` inside a box^3 domain are Fluid,
    each with `ppc` particles at seeded uniform positions inside the cell,
    speed `vel` on their faces; particles leaving the box are deleted.  Only
    the fluid's bricks exist at the start; the first cell-type update builds
    the air layers."""
    dims, offset = np.asarray(dims), np.asarray(offset)
    lo_b, hi_b = offset // 8, (offset + dims - 1) // 8
    bk = np.stack(np.meshgrid(*[np.arange(lo_b[i], hi_b[i] + 1) for i in range(3)], indexing="ij"), -1).reshape(-1, 3)
    bk = bk[np.lexsort((bk[:, 0], bk[:, 1], bk[:, 2]))].astype(np.int32)
    nb = len(bk)
    t = np.arange(512)
    g = bk[:, None, :] * 8 + np.stack([t & 7, (t >> 3) & 7, t >> 6], 1)[None, :, :]
    inside = np.all((g >= offset) & (g < offset + dims), axis=-1).reshape(-1)
    flags = np.where(inside, EXISTS | 1, 0).astype(np.uint8)
    speed = [np.where(inside, float(vel[i]), 0.0) for i in range(3)]
    cells = g.reshape(-1, 3)[inside]
    rng = np.random.default_rng(seed)
    pts = (np.repeat(cells, ppc, axis=0) + rng.uniform(-0.5, 0.5, size=(len(cells) * ppc, 3))) * h
    params = dict(h=(h, h, h), gravity=(0.0, 0.0, -9.81), box_lo=(-0.5 * h,) * 3, box_hi=((box - 0.5) * h,) * 3,
                  max_particles=int(len(pts) * 1.25) + (1 << 16), max_bricks=1 << 17)
    return {"keys": bk, "flags": flags, "u": speed[0], "v": speed[1], "w": speed[2], "particles": pts, "params": params,
            "fluid_cells": int(inside.sum())}
