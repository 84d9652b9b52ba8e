"""Writes the projection golden vectors proj_*.npz.

Each fixture holds a small seeded grid (brick keys, flag bytes, face
velocities, spacing) and the expected results of the reference projection
(Algorithms_Solve_Pressure.h:44-221): the divergence, the pressure from a
direct sparse LU of the assembled system (scipy SuperLU standing in for
UMFPACK, Policy_Solve_Linear_Umfpack.h:56-70; both are exact factorisations),
and the face velocities after the gradient update.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mac_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def grids():
    yield "rand_seed1", W.random_grid(1)
    yield "rand_negorigin_dropped", W.random_grid(2, 3, 2, 2, origin=(-2, -1, 5), drop_bricks=2)
    g = W.sphere_box(24, 3, nspheres=3, rmin=3.0, rmax=6.0, slab=4)
    g["h"] = (0.01, 0.01, 0.01)
    yield "spheres24_h001", g
    g = W.random_grid(8, 2, 2, 1, fluid_p=0.85, const_p=0.05)
    g["h"] = (0.5, 0.25, 2.0)
    yield "rand_aniso_h", g
    yield "line_1d_kat", W.line_1d_kat()


def main():
    for name, g in grids():
        n, off, idx, val, row = O.assemble(g)
        A = sp.csc_matrix(sp.csr_matrix((val, idx, off), shape=(n, n)))
        div = O.divergence(g)
        b = np.zeros(n)
        sel = row >= 0
        b[row[sel]] = div[sel]
        x = spla.splu(A).solve(b) if n > 0 else np.zeros(0)
        p = np.zeros(row.size)
        p[sel] = x[row[sel]]
        u, v, w = O.gradient_update(g, p, g["u"], g["v"], g["w"])
        out = os.path.join(HERE, f"proj_{name}.npz")
        np.savez_compressed(out, keys=np.asarray(g["keys"], np.int32), flags=np.asarray(g["flags"], np.uint8),
                            u=g["u"], v=g["v"], w=g["w"], h=np.asarray(g["h"], np.float64), div=div, p=p,
                            u_out=u, v_out=v, w_out=w)
        print(out, n, os.path.getsize(out))


if __name__ == "__main__":
    main()
