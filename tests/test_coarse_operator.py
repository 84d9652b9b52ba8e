"""The oracle's coarse operator: on the coarse unknowns, level l+1 is the
Galerkin operator of the piecewise-constant aggregation, P0^T A_l P0 / 16
(P0 copies a coarse value onto its 8 children), with level 0 the reference's
own assembly (Algorithms_Solve_Pressure.h:103-133, negated).  Checked
against the brute-force sparse product on a small box with walls, a slab,
spheres and air, with h = 1 so that every coefficient is exact."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle.oracle as O
from mac_amd import workloads as W

DX = [-1, 1, 0, 0, 0, 0]
DY = [0, 0, -1, 1, 0, 0]
DZ = [0, 0, 0, 0, -1, 1]


def _grid(seed):
    g = W.sphere_box(32, seed, nspheres=6, rmin=3.0, rmax=6.0, slab=8)
    g["h"] = (1.0, 1.0, 1.0)
    return g


def _level0(g):
    """L0 = -A (reference assembly) over the unknowns, with each unknown's
    dense (x, y, z) in the brick bounding box."""
    n, off, idx, val, row = O.assemble(g)
    A = sp.csr_matrix((val[: off[n]], idx[: off[n]], off[: n + 1]), shape=(n, n))
    keys = np.asarray(g["keys"], np.int64)
    mn = keys.min(axis=0)
    t = np.arange(512)
    loc = np.stack([t & 7, (t >> 3) & 7, t >> 6], axis=1)
    xyz = np.zeros((n, 3), np.int64)
    for b in range(keys.shape[0]):
        r = row[b * 512:(b + 1) * 512]
        m = r >= 0
        xyz[r[m]] = 8 * (keys[b] - mn) + loc[m]
    return (-A).tocsr(), xyz


def _coarse_matrix(code, cf):
    """The level's operator over its unknowns from the dumped arrays."""
    dom = (code & 0x40) != 0
    zz, yy, xx = np.nonzero(dom)
    nc = zz.size
    ind = -np.ones(code.shape, np.int64)
    ind[zz, yy, xx] = np.arange(nc)
    rows, cols, vals = [np.arange(nc)], [np.arange(nc)], [cf[6][zz, yy, xx]]
    for k in range(6):
        z2, y2, x2 = zz + DZ[k], yy + DY[k], xx + DX[k]
        ok = (z2 >= 0) & (y2 >= 0) & (x2 >= 0) & (z2 < code.shape[0]) & (y2 < code.shape[1]) & (x2 < code.shape[2])
        j = -np.ones(nc, np.int64)
        j[ok] = ind[z2[ok], y2[ok], x2[ok]]
        m = j >= 0
        rows.append(np.nonzero(m)[0])
        cols.append(j[m])
        vals.append(-cf[k][zz[m], yy[m], xx[m]])
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nc, nc))
    return M, np.stack([xx, yy, zz], axis=1), ind


def _galerkin(Lf, xyz_f, ind_c):
    """P0^T Lf P0 / 16 on the coarse unknowns (fine unknowns whose parent is
    not a coarse unknown drop out of P0's range)."""
    par = xyz_f >> 1
    inside = (par[:, 0] < ind_c.shape[2]) & (par[:, 1] < ind_c.shape[1]) & (par[:, 2] < ind_c.shape[0])
    col = -np.ones(xyz_f.shape[0], np.int64)
    col[inside] = ind_c[par[inside, 2], par[inside, 1], par[inside, 0]]
    m = col >= 0
    nc = int(ind_c.max()) + 1
    P = sp.csr_matrix((np.ones(m.sum()), (np.nonzero(m)[0], col[m])), shape=(xyz_f.shape[0], nc))
    return (P.T @ Lf @ P / 16.0).tocsr()


@pytest.mark.parametrize("seed", [1, 2])
def test_coarse_levels_are_galerkin(seed):
    g = _grid(seed)
    Lf, xyz = _level0(g)
    checked = 0
    for level in (1, 2):
        dims, code, cf = O.mg_level(g, level)
        if code is None or not (code & 0x40).any():
            break
        Lc, xyz_c, ind_c = _coarse_matrix(code, cf)
        G = _galerkin(Lf, xyz, ind_c)
        diff = abs(G - Lc)
        scale = abs(Lc).max()
        assert diff.max() <= 1e-12 * scale, (level, diff.max(), scale)
        # not vacuous: the level has cells whose couplings are not the
        # rediscretised 1/h^2 of the level (strictly between 0 and it)
        ihc = 0.25 ** level
        vals = cf[:6][:, code & 0x40 != 0]
        assert ((vals > 0) & (vals < ihc * (1 - 1e-9))).any()
        Lf, xyz = Lc, xyz_c
        checked += 1
    assert checked == 2


def test_galerkin_check_catches_a_dropped_internal_coupling():
    """The same comparison fails for the operator a plausible mistake gives:
    the diagonal without the couplings between children subtracted."""
    g = _grid(1)
    Lf, xyz = _level0(g)
    dims, code, cf = O.mg_level(g, 1)
    bad = cf.copy()
    dom = code & 0x40 != 0
    # children's diagonals summed / 16 without the internal couplings:
    # every interior coarse cell's diagonal doubles (6/4 -> 12/4 of 1/h^2)
    bad[6][dom] = 2.0 * bad[6][dom]
    Lc, _, ind_c = _coarse_matrix(code, bad)
    G = _galerkin(Lf, xyz, ind_c)
    assert abs(G - Lc).max() > 0.1 * abs(Lc).max()
