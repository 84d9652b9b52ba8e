"""Input/output files of tests/cpp/slots_test (the C++ slot adapters on a
stand-in grid), and its build line."""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "cpp", "slots_test.cpp")
BIN = os.path.join(HERE, "cpp", "slots_test")


def build():
    """g++ the adapter test against include/ and mac_amd/libmac_amd.so."""
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), SRC, "-o", BIN,
                           "-L" + os.path.join(ROOT, "mac_amd"), "-lmac_amd", "-Wl,-rpath,$ORIGIN/../../mac_amd"])


def write_input(path, g, n, off, idx, val, b, viscosity=0.0, accel=None):
    """accel = (ax, ay, az): slot B runs under the Force wrapper and projects
    them, the air pressures coming from g's u, v, w (the real speed)."""
    nb = g["keys"].shape[0]
    with open(path, "wb") as f:
        np.array([nb], np.int32).tofile(f)
        np.ascontiguousarray(g["keys"], np.int32).tofile(f)
        np.ascontiguousarray(g["flags"], np.uint8).tofile(f)
        for k in ("u", "v", "w"):
            np.ascontiguousarray(g[k], np.float64).tofile(f)
        np.ascontiguousarray(g["h"], np.float64).tofile(f)
        np.array([n], np.int32).tofile(f)
        np.ascontiguousarray(off[:n + 1], np.int32).tofile(f)
        nnz = int(off[n])
        np.array([nnz], np.int32).tofile(f)
        np.ascontiguousarray(idx[:nnz], np.int32).tofile(f)
        np.ascontiguousarray(val[:nnz], np.float64).tofile(f)
        np.ascontiguousarray(b[:n], np.float64).tofile(f)
        np.array([viscosity], np.float64).tofile(f)
        np.array([0 if accel is None else 1], np.int32).tofile(f)
        if accel is not None:
            for a in accel:
                np.ascontiguousarray(a, np.float64).tofile(f)


def read_output(path, nb, n):
    raw = open(path, "rb").read()
    m = nb * 512
    a = np.frombuffer(raw[: 8 * (4 * m + n)], np.float64)
    iters = int(np.frombuffer(raw[8 * (4 * m + n):], np.int32)[0])
    return {"p": a[:m], "u": a[m:2 * m], "v": a[2 * m:3 * m], "w": a[3 * m:4 * m], "x": a[4 * m:], "iters": iters}
