"""Seeded synthetic inputs in the brick format of include/mac_amd.h.

All draws come from a counter-based SplitMix64 keyed by (seed, stream,
index), so a field is bit-identical wherever and in whatever order it is
generated; normals use Box-Muller on two such uniforms.
"""
import numpy as np

AIR, FLUID, SOLID, INFLOW, INTERIOR, ZERO_PRES = 0, 1, 2, 3, 4, 5
CONST = (0x08, 0x10, 0x20)
EXISTS = 0x80

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, stream, idx):
    """SplitMix64 finaliser of (seed, stream, idx) -> uint64 (vectorised over idx)."""
    with np.errstate(over="ignore"):
        idx = np.asarray(idx, dtype=np.uint64)
        z = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + np.uint64(stream) * np.uint64(0x8CB92BA72F3D8DD7)
             + (idx + np.uint64(1)) * _G)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, idx):
    """Uniform doubles in (0, 1)."""
    v = splitmix64(seed, stream, idx) >> np.uint64(11)
    return (v.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed, stream, idx):
    idx = np.asarray(idx, dtype=np.uint64)
    u1 = uniform(seed, stream, idx * np.uint64(2))
    u2 = uniform(seed, stream, idx * np.uint64(2) + np.uint64(1))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def dense_to_bricks(a, nbx, nby, nbz):
    """(Z, Y, X) dense array -> (nb, 512) bricks, brick b = bx + nbx*(by + nby*bz),
    cell x + 8y + 64z."""
    a = np.asarray(a)
    b = a.reshape(nbz, 8, nby, 8, nbx, 8).transpose(0, 2, 4, 1, 3, 5)
    return np.ascontiguousarray(b.reshape(nbx * nby * nbz, 512))


def bricks_to_dense(b, nbx, nby, nbz):
    b = np.asarray(b).reshape(nbz, nby, nbx, 8, 8, 8).transpose(0, 3, 1, 4, 2, 5)
    return np.ascontiguousarray(b.reshape(nbz * 8, nby * 8, nbx * 8))


def box_keys(nbx, nby, nbz, origin=(0, 0, 0)):
    bz, by, bx = np.meshgrid(np.arange(nbz), np.arange(nby), np.arange(nbx), indexing="ij")
    k = np.stack([bx.ravel() + origin[0], by.ravel() + origin[1], bz.ravel() + origin[2]], axis=1)
    return np.ascontiguousarray(k.astype(np.int32))


def sphere_box(n, seed, nspheres=40, rmin=10.0, rmax=40.0, slab=30, velocity=True):
    """Projection benchmark input: n^3 box (n a multiple of 8) in 8^3 bricks,
    one-cell solid wall on the six domain faces (faces between the wall and
    the interior const, i.e. Neumann), fluid = floor slab z in [1, slab]
    union `nspheres` spheres of radius U(rmin, rmax) with centres uniform
    over the interior, the rest air.  Face velocities N(0,1) on every
    non-const face touching a fluid cell, 0 elsewhere.

    Returns dict(keys, flags, u, v, w, h, dims, fluid_fraction)."""
    assert n % 8 == 0
    nb = n // 8
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij", sparse=True)
    wall = (x == 0) | (x == n - 1) | (y == 0) | (y == n - 1) | (z == 0) | (z == n - 1)
    fluid = np.zeros((n, n, n), dtype=bool)
    if slab > 0:
        fluid[1:slab + 1, :, :] = True
    k = np.arange(nspheres, dtype=np.uint64)
    cx = 1.0 + uniform(seed, 1, 4 * k) * (n - 2)
    cy = 1.0 + uniform(seed, 1, 4 * k + np.uint64(1)) * (n - 2)
    cz = 1.0 + uniform(seed, 1, 4 * k + np.uint64(2)) * (n - 2)
    rr = rmin + uniform(seed, 1, 4 * k + np.uint64(3)) * (rmax - rmin)
    xc, yc, zc = x + 0.5, y + 0.5, z + 0.5
    for i in range(nspheres):
        r = rr[i]
        z0, z1 = max(0, int(cz[i] - r) - 1), min(n, int(cz[i] + r) + 2)
        y0, y1 = max(0, int(cy[i] - r) - 1), min(n, int(cy[i] + r) + 2)
        x0, x1 = max(0, int(cx[i] - r) - 1), min(n, int(cx[i] + r) + 2)
        d2 = ((xc[:, :, x0:x1] - cx[i]) ** 2 + (yc[:, y0:y1, :] - cy[i]) ** 2 + (zc[z0:z1, :, :] - cz[i]) ** 2)
        fluid[z0:z1, y0:y1, x0:x1] |= d2 <= r * r
    wall_full = np.broadcast_to(wall, (n, n, n))
    fluid &= ~wall_full
    mat = np.where(wall_full, SOLID, np.where(fluid, FLUID, AIR)).astype(np.uint8)
    flags = mat | np.uint8(EXISTS)
    cst = [(x == 1) | (x == n - 1), (y == 1) | (y == n - 1), (z == 1) | (z == n - 1)]
    for i in range(3):
        flags = flags | np.where(np.broadcast_to(cst[i], (n, n, n)), np.uint8(CONST[i]), np.uint8(0))
    vel = []
    if velocity:
        lin = (np.arange(n ** 3, dtype=np.uint64)).reshape(n, n, n)
        for i in range(3):
            lower = np.roll(fluid, 1, axis=2 - i)
            if i == 0:
                lower[:, :, 0] = False
            elif i == 1:
                lower[:, 0, :] = False
            else:
                lower[0, :, :] = False
            touch = (fluid | lower) & ~np.broadcast_to(cst[i], (n, n, n))
            comp = np.zeros((n, n, n), dtype=np.float64)
            ids = lin[touch] * np.uint64(3) + np.uint64(i)
            comp[touch] = normal(seed, 2, ids)
            vel.append(dense_to_bricks(comp, nb, nb, nb).reshape(-1))
    else:
        vel = [np.zeros(n ** 3) for _ in range(3)]
    return {
        "keys": box_keys(nb, nb, nb),
        "flags": dense_to_bricks(flags, nb, nb, nb).reshape(-1),
        "u": vel[0], "v": vel[1], "w": vel[2],
        "h": (1.0, 1.0, 1.0),
        "dims": (n, n, n),
        "fluid_fraction": float(fluid.sum()) / n ** 3,
    }


def random_grid(seed, nbx=2, nby=2, nbz=2, fluid_p=0.6, solid_p=0.05, const_p=0.1, drop_bricks=0, origin=(0, 0, 0)):
    """Small random grid for parity tests: random materials (Air, Fluid,
    Solid, Inflow, Interior, _0_Pres), random const bits, random missing
    cells and optionally missing bricks; N(0,1) velocities everywhere."""
    keys = box_keys(nbx, nby, nbz, origin)
    nb = keys.shape[0]
    ncell = nb * 512
    idx = np.arange(ncell, dtype=np.uint64)
    u = uniform(seed, 10, idx)
    mat = np.full(ncell, AIR, dtype=np.uint8)
    mat[u < fluid_p] = FLUID
    u2 = uniform(seed, 11, idx)
    mat[(u2 < solid_p)] = SOLID
    u3 = uniform(seed, 12, idx)
    sel = mat == FLUID
    mat[sel & (u3 < 0.05)] = INFLOW
    mat[sel & (u3 > 0.95)] = INTERIOR
    mat[(mat == AIR) & (u3 > 0.97)] = ZERO_PRES
    flags = mat | np.uint8(EXISTS)
    for i in range(3):
        c = uniform(seed, 13 + i, idx) < const_p
        flags = flags | np.where(c, np.uint8(CONST[i]), np.uint8(0))
    missing = uniform(seed, 16, idx) < 0.02
    flags[missing] = 0
    if drop_bricks:
        keep = np.ones(nb, dtype=bool)
        drop = (uniform(seed, 17, np.arange(nb, dtype=np.uint64)) * nb).astype(int)[:drop_bricks]
        keep[drop] = False
        keys = keys[keep]
        flags = flags.reshape(nb, 512)[keep].reshape(-1)
        nb = keys.shape[0]
        ncell = nb * 512
        idx = np.arange(ncell, dtype=np.uint64)
    vel = [normal(seed, 20 + i, idx) for i in range(3)]
    return {"keys": keys, "flags": flags, "u": vel[0], "v": vel[1], "w": vel[2], "h": (1.0, 1.0, 1.0)}


def line_1d_kat():
    """test/old/TestAlgorithms_Solve_Pressure.h:56-200 (test1) embedded in 3-D:
    cells -1 (air), 0 (fluid, u=2), 1 (fluid, u=1), 2 (air, u=0) along x,
    h = 1.  The y and z faces of the four cells are made Neumann by solid
    neighbours carrying const bits, so only the x coupling remains and the
    1-D answer holds: after the projection u(0) = u(1) = 1 (asserts :197-199)."""
    # bricks (-1,0,0) and (0,0,0); cells at y=z=1 (so y/z neighbours exist)
    keys = np.array([[-1, 0, 0], [0, 0, 0]], dtype=np.int32)
    flags = np.zeros((2, 512), dtype=np.uint8)
    u = np.zeros((2, 512))

    def cell(gx, gy, gz):
        b = 0 if gx < 0 else 1
        return b, (gx % 8) + 8 * gy + 64 * gz

    line = {-1: AIR, 0: FLUID, 1: FLUID, 2: AIR}
    speed = {-1: 0.0, 0: 2.0, 1: 1.0, 2: 0.0}
    for gx, m in line.items():
        b, t = cell(gx, 1, 1)
        flags[b, t] = m | EXISTS
        u[b, t] = speed[gx]
        # solid neighbours in y and z; the faces between them and the line cell const
        for (dy, dz) in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            bb, tt = cell(gx, 1 + dy, 1 + dz)
            flags[bb, tt] = SOLID | EXISTS
        flags[b, t] |= CONST[1] | CONST[2]                    # low y / low z faces of the line cell
        bb, tt = cell(gx, 2, 1)
        flags[bb, tt] |= CONST[1]                             # high y face = low y face of (gx,2,1)
        bb, tt = cell(gx, 1, 2)
        flags[bb, tt] |= CONST[2]
    # a face at x=3 (low x face of cell 3) so the air cell 2 has a +x neighbour
    return {"keys": keys, "flags": flags.reshape(-1), "u": u.reshape(-1), "v": np.zeros(1024), "w": np.zeros(1024),
            "h": (1.0, 1.0, 1.0), "cells": {gx: cell(gx, 1, 1) for gx in line}}
