"""ctypes binding of the CPU oracle (liboracle.so, built from mac_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never by the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p


def build():
    subprocess.check_call(["make", "-B", "-s", "-C", HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.orc_chunk_key.restype = ctypes.c_int
    lib.orc_chunk_key.argtypes = [ctypes.c_int]
    lib.orc_chunk_key_n.restype = ctypes.c_int
    lib.orc_chunk_key_n.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.orc_offset_neighbour.restype = ctypes.c_int
    lib.orc_offset_neighbour.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)]
    lib.orc_offset.restype = ctypes.c_int
    lib.orc_offset.argtypes = [ctypes.c_int] * 3
    lib.orc_assemble.restype = ctypes.c_int
    lib.orc_assemble.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P]
    lib.orc_divergence.restype = None
    lib.orc_divergence.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P]
    lib.orc_gradient_update.restype = None
    lib.orc_gradient_update.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P]
    lib.orc_cg_csr.restype = ctypes.c_int
    lib.orc_cg_csr.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_double, ctypes.c_int]
    lib.orc_project.restype = ctypes.c_int
    lib.orc_project.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    lib.orc_apply_A.restype = None
    lib.orc_apply_A.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P]
    lib.orc_mgpcg.restype = ctypes.c_int
    lib.orc_mgpcg.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                              ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                              ctypes.c_double, ctypes.c_int, _P]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def inv_h(h):
    return np.array([1.0 / float(x) for x in h], dtype=np.float64)


def chunk_key(k):
    return load().orc_chunk_key(int(k))


def chunk_key_n(k, n):
    return load().orc_chunk_key_n(int(k), int(n))


def offset_neighbour(n, dim, i, direction, off):
    """(wraps, neighbour offset) -- Offset::GetNeighbour with 0-based i."""
    r = ctypes.c_int()
    w = load().orc_offset_neighbour(int(n), int(dim), int(i), int(direction), int(off), ctypes.byref(r))
    return bool(w), r.value


def offset(kx, ky, kz):
    return load().orc_offset(int(kx), int(ky), int(kz))


def assemble(grid):
    """Reference CSR assembly; returns (n, offset, indice, value, row_of)."""
    keys = _c(grid["keys"], np.int32)
    flags = _c(grid["flags"], np.uint8)
    nb = keys.shape[0]
    ncell = nb * 512
    off = np.zeros(ncell + 1, dtype=np.int32)
    idx = np.zeros(7 * ncell, dtype=np.int32)
    val = np.zeros(7 * ncell, dtype=np.float64)
    row = np.zeros(ncell, dtype=np.int32)
    ih = inv_h(grid["h"])
    n = load().orc_assemble(nb, _p(keys), _p(flags), _p(ih), _p(off), _p(idx), _p(val), _p(row))
    nnz = off[n]
    return n, off[: n + 1].copy(), idx[:nnz].copy(), val[:nnz].copy(), row


def divergence(grid, u=None, v=None, w=None):
    keys = _c(grid["keys"], np.int32)
    flags = _c(grid["flags"], np.uint8)
    nb = keys.shape[0]
    u = _c(grid["u"] if u is None else u, np.float64)
    v = _c(grid["v"] if v is None else v, np.float64)
    w = _c(grid["w"] if w is None else w, np.float64)
    out = np.zeros(nb * 512)
    load().orc_divergence(nb, _p(keys), _p(flags), _p(inv_h(grid["h"])), _p(u), _p(v), _p(w), _p(out))
    return out


def gradient_update(grid, p, u, v, w):
    """In-place u -= grad p (copies returned)."""
    keys = _c(grid["keys"], np.int32)
    flags = _c(grid["flags"], np.uint8)
    u, v, w = (np.array(a, dtype=np.float64, copy=True) for a in (u, v, w))
    p = _c(p, np.float64)
    load().orc_gradient_update(keys.shape[0], _p(keys), _p(flags), _p(inv_h(grid["h"])), _p(p), _p(u), _p(v), _p(w))
    return u, v, w


def apply_A(grid, x):
    keys = _c(grid["keys"], np.int32)
    flags = _c(grid["flags"], np.uint8)
    x = _c(x, np.float64)
    y = np.zeros(keys.shape[0] * 512)
    load().orc_apply_A(keys.shape[0], _p(keys), _p(flags), _p(inv_h(grid["h"])), _p(x), _p(y))
    return y


def cg_csr(n, off, idx, val, b, rtol=1e-14, maxit=100000):
    x = np.zeros(n)
    it = load().orc_cg_csr(n, _p(_c(off, np.int32)), _p(_c(idx, np.int32)), _p(_c(val, np.float64)), _p(_c(b, np.float64)),
                           _p(x), rtol, maxit)
    return x, it


def project(grid, rtol=1e-14):
    """Reference projection (assembly + CG stand-in for UMFPACK + update).
    Returns dict(p, u, v, w, max_div, iters)."""
    keys = _c(grid["keys"], np.int32)
    flags = _c(grid["flags"], np.uint8)
    nb = keys.shape[0]
    u, v, w = (np.array(grid[k], dtype=np.float64, copy=True) for k in ("u", "v", "w"))
    p = np.zeros(nb * 512)
    h = _c(grid["h"], np.float64)
    md = ctypes.c_double()
    it = load().orc_project(nb, _p(keys), _p(flags), _p(h), _p(u), _p(v), _p(w), _p(p), rtol, ctypes.byref(md))
    return {"p": p, "u": u, "v": v, "w": w, "max_div": md.value, "iters": it}


def project_dirichlet(grid, viscosity, rtol=1e-14, real=None):
    """Viscous free-surface projection (Algorithms_Solve_Pressure_Dirichlet
    with Policy_Dirichlet_Boundary); real = (ur, vr, wr) for the air
    pressures, default the projected field.  Returns dict(p, u, v, w, max_div, iters)."""
    keys = _c(grid["keys"], np.int32)
    flags = _c(grid["flags"], np.uint8)
    nb = keys.shape[0]
    u, v, w = (np.array(grid[k], dtype=np.float64, copy=True) for k in ("u", "v", "w"))
    rr = [None, None, None] if real is None else [_c(a, np.float64) for a in real]
    p = np.zeros(nb * 512)
    h = _c(grid["h"], np.float64)
    md = ctypes.c_double()
    lib = load()
    lib.orc_project_dirichlet.argtypes = [ctypes.c_int] + [_P] * 9 + [ctypes.c_double, _P, ctypes.c_double, _P]
    it = lib.orc_project_dirichlet(nb, _p(keys), _p(flags), _p(h), _p(u), _p(v), _p(w),
                                   *[None if a is None else _p(a) for a in rr], float(viscosity), _p(p), rtol,
                                   ctypes.byref(md))
    return {"p": p, "u": u, "v": v, "w": w, "max_div": md.value, "iters": it}


def mgpcg(grid, rtol=1e-8, maxit=500, nu=2, omega=0.9, threads=0, smoother=1, cheb_lower=0.25, nu_coarse=2,
          cheb_lower_coarse=0.05, coarse_inner=4, coarse_damp=0.8):
    """CPU MG-PCG pressure (the same algorithm as the device solver).
    Returns dict(p, iters, ms_solve, ms_setup, true_rel_resid, levels, ok)."""
    keys = _c(grid["keys"], np.int32)
    flags = _c(grid["flags"], np.uint8)
    nb = keys.shape[0]
    u, v, w = (_c(grid[k], np.float64) for k in ("u", "v", "w"))
    p = np.zeros(nb * 512)
    out = np.zeros(5)
    h = _c(grid["h"], np.float64)
    rc = load().orc_mgpcg(nb, _p(keys), _p(flags), _p(h), _p(u), _p(v), _p(w), _p(p), rtol, maxit, nu, omega, smoother,
                          cheb_lower, nu_coarse, cheb_lower_coarse, coarse_inner, coarse_damp, threads, _p(out))
    return {"p": p, "iters": int(out[0]), "ms_solve": out[1], "ms_setup": out[2], "true_rel_resid": out[3],
            "levels": int(out[4]), "ok": rc == 0}


def mg_level(grid, level):
    """One level of the MG-PCG hierarchy of orc_mgpcg (a test hook): dims
    (n0, n1, n2), the stencil codes as [z, y, x] and the coarse operator's 7
    arrays (couplings -x,+x,-y,+y,-z,+z, diagonal) as [k, z, y, x]; level 0
    has no coefficient arrays (zeros).  dims is (0, 0, 0) past the top."""
    lib = load()
    lib.orc_mg_dump_request.restype = None
    lib.orc_mg_dump_request.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_longlong]
    dims = np.zeros(3, np.int32)
    lib.orc_mg_dump_request(int(level), _p(dims), None, None, 0)
    mgpcg(grid, rtol=1e-2, maxit=1)
    n = int(np.prod(dims.astype(np.int64)))
    if n == 0:
        return tuple(dims), None, None
    code = np.zeros(n, np.uint8)
    cf = np.zeros(7 * n)
    lib.orc_mg_dump_request(int(level), _p(dims), _p(code), _p(cf), n)
    mgpcg(grid, rtol=1e-2, maxit=1)
    shape = (int(dims[2]), int(dims[1]), int(dims[0]))
    return tuple(int(d) for d in dims), code.reshape(shape), cf.reshape((7,) + shape)


# ---------------------------------------------------------------- time step --
def corner_solve(h, signs, nb81):
    """The three-axis corner system (SolveMatrix / CreateMatrix) on a 3x3x3
    neighbourhood nb81[c, dz + 1, dy + 1, dx + 1] of face components."""
    lib = load()
    lib.orc_corner_solve.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P]
    hh = np.ascontiguousarray(h, np.float64)
    nb = np.ascontiguousarray(nb81, np.float64).reshape(81)
    res = np.zeros(3)
    lib.orc_corner_solve(_p(hh), int(signs[0]), int(signs[1]), int(signs[2]), _p(nb), _p(res))
    return res


class Sim:
    """CPU restatement of the time-step loop (sim_oracle.cpp)."""

    def __init__(self, h=(0.01, 0.01, 0.01), gravity=(0, 0, -9.81), viscosity=0.0, cfl=0.5, dt_min=0.0, dt_max=1e30,
                 layer_max=2, integrator=0, particle_rk=2, particles_per_inflow_cell=8, seed=0, box_lo=(0, 0, 0),
                 box_hi=(0, 0, 0), depth_interior=0, depth_remove=0, floor_axis=-1, floor_value=0.0, fill_cube_n=0,
                 inflow_speed=(0.0, 0.0, 0.0), inflow_amp=0.0, inflow_period=1, surface_bc=0, proj=None, **_ignored):
        lib = load()
        if not hasattr(lib, "_sim_bound"):
            lib.orc_sim_create.restype = ctypes.c_void_p
            lib.orc_sim_create.argtypes = [_P, _P, ctypes.c_ulonglong]
            lib.orc_sim_destroy.argtypes = [_P]
            lib.orc_sim_upload.argtypes = [_P, ctypes.c_int, _P, _P, _P, _P, _P, _P, ctypes.c_int, _P]
            lib.orc_sim_counts.argtypes = [_P, _P, _P]
            lib.orc_sim_download.argtypes = [_P] * 12
            lib.orc_sim_topology.argtypes = [_P]
            lib.orc_sim_compute_dt.restype = ctypes.c_double
            lib.orc_sim_compute_dt.argtypes = [_P]
            lib.orc_sim_forces.argtypes = [_P, ctypes.c_int]
            lib.orc_sim_advance.argtypes = [_P, ctypes.c_double, _P]
            lib.orc_sim_move.argtypes = [_P, ctypes.c_double]
            lib.orc_sim_interpolate.argtypes = [_P, ctypes.c_int, _P, _P]
            lib.orc_sim_extrapolate.argtypes = [_P, ctypes.c_int]
            lib.orc_sim_get_depth.argtypes = [_P, _P]
            lib.orc_sim_set_depth.argtypes = [_P, _P]
            lib.orc_sim_surface_bc.argtypes = [_P, ctypes.c_int, ctypes.c_int]
            lib.orc_sim_get_tick.restype = ctypes.c_ulonglong
            lib.orc_sim_get_tick.argtypes = [_P]
            lib.orc_sim_set_tick.argtypes = [_P, ctypes.c_ulonglong]
            lib.orc_sim_step.restype = ctypes.c_double
            lib.orc_sim_step.argtypes = [_P, _P]
            lib._sim_bound = True
        self.lib = lib
        prm = np.array(list(h) + list(gravity) + [viscosity, cfl, dt_min, dt_max] + list(box_lo) + list(box_hi)
                       + [floor_value, float((proj or {}).get("viscosity", 0.0))] + [float(x) for x in inflow_speed]
                       + [float(inflow_amp)], dtype=np.float64)
        iprm = np.array([layer_max, integrator, particle_rk, particles_per_inflow_cell, depth_interior, depth_remove,
                         floor_axis, fill_cube_n, inflow_period, surface_bc], dtype=np.int32)
        self.ptr = ctypes.c_void_p(lib.orc_sim_create(_p(prm), _p(iprm), seed))

    def __del__(self):
        try:
            self.lib.orc_sim_destroy(self.ptr)
        except Exception:
            pass

    def upload(self, keys, flags, layer=None, u=None, v=None, w=None, particles=None):
        keys = _c(keys, np.int32).reshape(-1, 3)
        flags = _c(flags, np.uint8)
        lay = None if layer is None else _c(layer, np.int8)
        uu = [None if a is None else _c(a, np.float64) for a in (u, v, w)]
        parts = np.zeros((0, 3)) if particles is None else _c(particles, np.float64).reshape(-1, 3)
        self.lib.orc_sim_upload(self.ptr, keys.shape[0], _p(keys), _p(flags), None if lay is None else _p(lay),
                                *[None if a is None else _p(a) for a in uu], parts.shape[0], _p(parts))

    def counts(self):
        nb, npart = ctypes.c_int(), ctypes.c_int()
        self.lib.orc_sim_counts(self.ptr, ctypes.byref(nb), ctypes.byref(npart))
        return nb.value, npart.value

    def download(self):
        nb, npart = self.counts()
        out = {"keys": np.zeros((nb, 3), np.int32), "flags": np.zeros(nb * 512, np.uint8),
               "layer": np.zeros(nb * 512, np.int8)}
        for k in ("u", "v", "w", "ax", "ay", "az", "p"):
            out[k] = np.zeros(nb * 512)
        out["particles"] = np.zeros((npart, 3))
        self.lib.orc_sim_download(self.ptr, *[_p(out[k]) for k in (
            "keys", "flags", "layer", "u", "v", "w", "ax", "ay", "az", "p", "particles")])
        return out

    def surface_bc(self, slot=0, dim=3):
        """The free-surface velocity boundary condition on speed slot `slot`
        (3 = the acceleration), in 3-D or (dim=2) on the x-y plane."""
        self.lib.orc_sim_surface_bc(self.ptr, slot, dim)

    def get_step(self):
        return int(self.lib.orc_sim_get_tick(self.ptr))

    def set_step(self, n):
        self.lib.orc_sim_set_tick(self.ptr, int(n))

    def get_depth(self):
        nb, _ = self.counts()
        d = np.zeros(nb * 512, np.int16)
        self.lib.orc_sim_get_depth(self.ptr, _p(d))
        return d

    def set_depth(self, depth):
        d = _c(depth, np.int16)
        assert d.size == self.counts()[0] * 512
        self.lib.orc_sim_set_depth(self.ptr, _p(d))

    def topology(self):
        self.lib.orc_sim_topology(self.ptr)

    def compute_dt(self):
        return self.lib.orc_sim_compute_dt(self.ptr)

    def forces(self, slot=0):
        self.lib.orc_sim_forces(self.ptr, slot)

    def advance(self, dt):
        md = ctypes.c_double()
        self.lib.orc_sim_advance(self.ptr, dt, ctypes.byref(md))
        return {"max_div": md.value}

    def move_particles(self, dt):
        self.lib.orc_sim_move(self.ptr, dt)

    def interpolate(self, pos):
        pos = _c(pos, np.float64).reshape(-1, 3)
        vel = np.zeros_like(pos)
        self.lib.orc_sim_interpolate(self.ptr, pos.shape[0], _p(pos), _p(vel))
        return vel

    def extrapolate(self, slot=0):
        self.lib.orc_sim_extrapolate(self.ptr, slot)

    def step(self):
        md = ctypes.c_double()
        dt = self.lib.orc_sim_step(self.ptr, ctypes.byref(md))
        return {"dt": dt, "max_div": md.value}
