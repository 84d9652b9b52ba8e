"""mac_solve_csr_noalloc (the slot A solve on caller-owned device arrays and
the library's static scratch) against mac_solve_csr (host arrays, context
workspace) on the same systems, through the C-ABI: two calls in a row, the
second smaller, so the scratch's previous contents are in place when it
starts.  The kernels are shared and every reduction runs in a fixed order,
so the answers must agree bit for bit."""
import ctypes

import numpy as np
import pytest

from mac_amd import workloads as W
from mac_amd.lib import LinearSolver, MacError, load
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _box_system(nx, ny, nz, seed):
    nbx, nby, nbz = nx // 8, ny // 8, nz // 8
    mat = np.full((nz, ny, nx), W.AIR, np.uint8)
    mat[1:-1, 1:-1, 1:-1] = W.FLUID
    g = {"keys": W.box_keys(nbx, nby, nbz), "flags": W.dense_to_bricks(mat | W.EXISTS, nbx, nby, nbz).reshape(-1),
         "h": (0.01, 0.01, 0.01)}
    for i, k in enumerate(("u", "v", "w")):
        g[k] = W.normal(seed, i, np.arange(nx * ny * nz, dtype=np.uint64))
    n, off, idx, val, row = O.assemble(g)
    div = O.divergence(g)
    b = np.zeros(n)
    b[row[row >= 0]] = div[row >= 0]
    return n, off, idx, val, b


class _Hip:
    """Device buffers for the test through the HIP runtime the library links
    (the test allocates; the entry point under test does not)."""

    def __init__(self):
        import mac_amd.lib as L
        L.load()
        self.rt = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _noalloc(ls, n, off, idx, val, b, rtol):
    hip = _Hip()
    try:
        d_off = hip.put(np.asarray(off, np.int32))
        d_idx = hip.put(np.asarray(idx, np.int32))
        d_val = hip.put(np.asarray(val, np.float64))
        d_b = hip.put(np.asarray(b, np.float64))
        d_x = hip.put(np.full(n, np.nan))
        it = ctypes.c_int(-1)
        lib = load()
        rc = lib.mac_solve_csr_noalloc(ls.proj.ctx, n, int(off[n]), d_off, d_idx, d_val, d_b, d_x, rtol, ctypes.byref(it))
        if rc != 0:
            raise MacError(f"mac_solve_csr_noalloc {rc}: {lib.mac_last_error().decode()}")
        return hip.get(d_x, n, np.float64), it.value
    finally:
        hip.free()


@pytest.mark.parametrize("sizes", [((32, 32, 64), (16, 16, 16)), ((16, 16, 16), (16, 16, 8))],
                         ids=["grid_stride_then_one_workgroup", "one_workgroup_twice"])
def test_noalloc_matches_solve_csr(sizes):
    ls = LinearSolver(rtol=1e-13)
    for k, dims in enumerate(sizes):
        n, off, idx, val, b = _box_system(*dims, seed=20 + k)
        print(f"[case] noalloc call {k + 1}: {dims}, n = {n}, nnz = {int(off[n])}", flush=True)
        ref = ls.Solve_Linear(n, off, idx, val, b)
        x, iters = _noalloc(ls, n, off, idx, val, b, 1e-13)
        assert iters > 0
        assert np.array_equal(x, ref), np.abs(x - ref).max()
        xo, _ = O.cg_csr(n, off, idx, val, b, rtol=1e-14)
        assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


def test_noalloc_small_system_and_limits():
    ls = LinearSolver(rtol=1e-13)
    off3 = np.array([0, 2, 5, 7], dtype=np.int32)
    idx3 = np.array([0, 1, 0, 1, 2, 1, 2], dtype=np.int32)
    val3 = np.array([-2.0, 1.0, 1.0, -3.0, 1.0, 1.0, -2.0])
    b3 = np.array([1.0, 2.0, 3.0])
    x3, _ = _noalloc(ls, 3, off3, idx3, val3, b3, 1e-13)
    ref = ls.Solve_Linear(3, off3, idx3, val3, b3)
    assert np.allclose(x3, ref, rtol=1e-14, atol=1e-15)
    # non-symmetric, compressed columns (value[3c + r] is row r of column c)
    An = np.array([[2.0, -1.0, 0.5], [0.25, 3.0, -2.0], [1.0, 1.0, 1.0]])
    offn = np.array([0, 3, 6, 9], dtype=np.int32)
    idxn = np.array([0, 1, 2] * 3, dtype=np.int32)
    valn = An.T.reshape(-1).copy()
    xn, _ = _noalloc(ls, 3, offn, idxn, valn, b3, 1e-13)
    assert np.allclose(xn, ls.Solve_Linear(3, offn, idxn, valn, b3), rtol=1e-14, atol=1e-15)
    assert np.allclose(xn, np.linalg.solve(An, b3), rtol=1e-14, atol=1e-14)
    # beyond the scratch: a clear error before any device work
    lib = load()
    rc = lib.mac_solve_csr_noalloc(ls.proj.ctx, (1 << 18) + 1, 0, 1, 1, 1, 1, 1, 1e-8, None)
    assert rc != 0 and b"MAC_CSR_NOALLOC_MAX_N" in lib.mac_last_error()
