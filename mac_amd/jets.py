"""Seeded jet set-ups for the time-step loop (brick format of include/mac_amd.h).

vertical_jet(): a box of nx x ny x nz cells (h per axis), open sides, solid
floor and a square inflow nozzle centred in the top layer.  Only the floor
and the nozzle exist at the start; fluid enters through the nozzle, whose
empty cells are refilled with jittered particles every step.
lateral_jet(): a horizontal nozzle on the x = 0 wall at mid-height, no floor.
breakup_jet(): a pulsed vertical jet falling into a pool at scale.
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


def breakup_jet(nx=512, ny=256, nz=256, h=0.01, nozzle=12, inflow=3.0, amp=0.05, period=40, margin=16, pool_depth=180,
                ppc=8, seed=0):
    """Large sparse jet: an nx x ny x nz box (h per axis) with a solid floor,
    a nozzle x nozzle inflow at the top centre at `inflow` m/s downward,
    pulsed by `amp` (relative) with a period of `period` steps from the seed
    (mac_sim_params.inflow_amp / inflow_period), falling as a column of fluid
    into a pool: the floor cells of [margin, n - margin) in x and y up to
    height pool_depth are Fluid, the column from the nozzle down to the pool
    is Fluid.  `ppc` particles per fluid cell at seeded uniform positions in
    the cell; particles leaving the box are deleted.  Returns the dict of
    fluid_block (keys, flags, u, v, w, particles, params, fluid_cells); the
    params carry the pulsation."""
    x0, y0 = (nx - nozzle) // 2, (ny - nozzle) // 2
    kmax = np.array([nx - 1, ny - 1, nz - 1])
    bk = np.stack(np.meshgrid(*[np.arange(0, kmax[i] // 8 + 1) for i in range(3)], indexing="ij"), -1).reshape(-1, 3)
    # bricks of the floor, the pool and the column (and their neighbours come with the first cell-type update)
    lo_pool = np.array([margin, margin, 0]) // 8
    hi_pool = np.array([nx - margin - 1, ny - margin - 1, pool_depth]) // 8
    in_pool = np.all((bk >= lo_pool) & (bk <= hi_pool), axis=1)
    in_floor = bk[:, 2] == 0
    lo_col = np.array([x0, y0, pool_depth]) // 8
    hi_col = np.array([x0 + nozzle - 1, y0 + nozzle - 1, nz - 1]) // 8
    in_col = np.all((bk >= lo_col) & (bk <= hi_col), axis=1)
    bk = bk[in_pool | in_floor | in_col]
    bk = bk[np.lexsort((bk[:, 0], bk[:, 1], bk[:, 2]))].astype(np.int32)
    nb = len(bk)
    t = np.arange(512)
    flags = np.zeros((nb, 512), np.uint8)
    w = np.zeros((nb, 512))
    fluid = np.zeros((nb, 512), bool)
    for i0 in range(0, nb, 4096):
        g = bk[i0:i0 + 4096, None, :] * 8 + np.stack([t & 7, (t >> 3) & 7, t >> 6], 1)[None, :, :]
        gx, gy, gz = g[..., 0], g[..., 1], g[..., 2]
        inside = (gx < nx) & (gy < ny) & (gz < nz)
        floor = inside & (gz == 0)
        pool = inside & (gz >= 1) & (gz <= pool_depth) & (gx >= margin) & (gx < nx - margin) & (gy >= margin) & (gy < ny - margin)
        col_xy = (gx >= x0) & (gx < x0 + nozzle) & (gy >= y0) & (gy < y0 + nozzle)
        col = inside & col_xy & (gz > pool_depth) & (gz < nz - 1)
        noz = inside & col_xy & (gz == nz - 1)
        f = np.zeros(gx.shape, np.uint8)
        f[floor] = SOLID | EXISTS | NODELETE
        # the floor face of the cells just above the floor is const (no flow)
        above = inside & (gz == 1) & ~pool
        f[above] = AIR | EXISTS | NODELETE | CONST[2]
        f[pool] = 1 | EXISTS
        f[pool & (gz == 1)] |= CONST[2]
        f[col] = 1 | EXISTS
        f[noz] = INFLOW | EXISTS | NODELETE | CONST[2]
        flags[i0:i0 + 4096] = f
        w[i0:i0 + 4096][col | noz] = -float(inflow)
        fluid[i0:i0 + 4096] = pool | col
    cells_b, cells_t = np.nonzero(fluid)
    nfl = len(cells_b)
    rng = np.random.default_rng(seed)
    pts = np.empty((nfl * ppc, 3))
    for i in range(3):
        ck = bk[cells_b, i] * 8 + ((cells_t >> (3 * i)) & 7)
        pts[:, i] = np.repeat(ck.astype(np.float64), ppc)
        pts[:, i] += rng.uniform(-0.5, 0.5, size=nfl * ppc)
        pts[:, i] *= h
    n = nb * 512
    params = dict(h=(h, h, h), gravity=(0.0, 0.0, -9.81), box_lo=(-0.5 * h,) * 3,
                  box_hi=((nx - 0.5) * h, (ny - 0.5) * h, (nz - 0.5) * h), max_particles=int(len(pts) * 1.25) + (1 << 16),
                  max_bricks=1 << 17, inflow_speed=(0.0, 0.0, -float(inflow)), inflow_amp=float(amp),
                  inflow_period=int(period))
    return {"keys": bk, "flags": flags.reshape(-1), "u": np.zeros(n), "v": np.zeros(n), "w": w.reshape(-1),
            "particles": pts, "params": params, "fluid_cells": int(nfl)}
