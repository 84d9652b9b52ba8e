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


def mgpcg(grid, rtol=1e-8, maxit=500, nu=2, omega=0.9, threads=0):
    """CPU MG-PCG pressure (the same algorithm as the device solver).
    Returns dict(p, iters, ms_solve, ms_setup, true_rel_resid, levels, ok)."""
    keys = _c(grid["keys"], np.int32)
    flags = _c(grid["flags"], np.uint8)
    nb = keys.shape[0]
    u, v, w = (_c(grid[k], np.float64) for k in ("u", "v", "w"))
    p = np.zeros(nb * 512)
    out = np.zeros(5)
    h = _c(grid["h"], np.float64)
    rc = load().orc_mgpcg(nb, _p(keys), _p(flags), _p(h), _p(u), _p(v), _p(w), _p(p), rtol, maxit, nu, omega, threads,
                          _p(out))
    return {"p": p, "iters": int(out[0]), "ms_solve": out[1], "ms_setup": out[2], "true_rel_resid": out[3],
            "levels": int(out[4]), "ok": rc == 0}
