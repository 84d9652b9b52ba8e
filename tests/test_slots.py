"""The C++ template-slot adapters (include/mac_amd_slots.hpp) driven on a
stand-in grid with the reference's grid calls: slot B
(Algorithms_Solve_Pressure_HIP) against the oracle's projection, slot A
(Policy_Solve_Linear_HIP) against the oracle's CG on the reference-assembled
CSR.  Bars: pressure and face velocities within 1e-8 relative, slot A within
1e-9; without a GPU both slots print the error and the program exits 0."""
import os
import subprocess

import numpy as np
import pytest

from mac_amd import workloads as W
from oracle import oracle as O
from tests import slots_io as S


def _binary():
    deps = [S.SRC, os.path.join(S.ROOT, "include", "mac_amd.h"), os.path.join(S.ROOT, "include", "mac_amd_slots.hpp")]
    if not os.path.exists(S.BIN) or os.path.getmtime(S.BIN) < max(os.path.getmtime(d) for d in deps):
        S.build()
    return S.BIN


def _case(tmp_path, g):
    # a reference grid stores speeds only for cells that exist
    ex = (g["flags"] & 0x80) > 0
    for k in ("u", "v", "w"):
        g[k] = np.where(ex, g[k], 0.0)
    n, off, idx, val, row = O.assemble(g)
    div = O.divergence(g)
    b = np.zeros(n)
    b[row[row >= 0]] = div[row >= 0]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    S.write_input(fin, g, n, off, idx, val, b)
    return n, off, idx, val, b, fin, fout


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def test_adapters_compile_and_report_missing_gpu(tmp_path):
    """No GPU here: the slots print the library's error and carry on, as the
    reference's solver policy does (Policy_Solve_Linear_Umfpack.h:15-21)."""
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be visible")
    g = W.random_grid(1)
    _, _, _, _, _, fin, fout = _case(tmp_path, g)
    r = subprocess.run([_binary(), fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "mac_amd error or warning :-5" in r.stdout


GRIDS = {"rand1": lambda: W.random_grid(1), "rand_neg_origin": lambda: W.random_grid(2, 3, 2, 2, origin=(-2, -1, 5)),
         "rand_dropped": lambda: W.random_grid(3, 3, 3, 2, drop_bricks=4)}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GRIDS))
def test_slots_match_oracle(tmp_path, name):
    g = GRIDS[name]()
    n, off, idx, val, b, fin, fout = _case(tmp_path, g)
    r = subprocess.run([_binary(), fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "error" not in r.stdout, r.stdout + r.stderr
    out = S.read_output(fout, g["keys"].shape[0], n)
    ref = O.project(g, rtol=1e-14)
    exists = (g["flags"] & 0x80) > 0
    assert rel(out["p"][exists], ref["p"][exists]) < 1e-8
    for k in ("u", "v", "w"):
        assert rel(out[k][exists], ref[k][exists]) < 1e-8, k
    xo, _ = O.cg_csr(n, off, idx, val, b, rtol=1e-14)
    assert rel(out["x"], xo) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GRIDS))
def test_viscous_slot_b_under_force_wrapper(tmp_path, name):
    """Jet3's composition (test/Jet3.cpp:626,633): the viscous slot B inside
    Algorithms_Solve_Pressure_Force projects the acceleration while the air
    pressures come from the real speed, Speed_GetRef(-1, false)
    (Policy_Dirichlet_Boundary.h:100-102).  The stand-in grid holds an
    acceleration that differs from its speed; the answer is the oracle's
    viscous projection of the acceleration with the speed as the real field."""
    g = GRIDS[name]()
    ex = (g["flags"] & 0x80) > 0
    for k in ("u", "v", "w"):
        g[k] = np.where(ex, g[k], 0.0)
    idxs = np.arange(g["flags"].size, dtype=np.uint64)
    acc = [np.where(ex, W.normal(17, 40 + i, idxs), 0.0) for i in range(3)]
    nu = 0.37
    n, off, idx, val, row = O.assemble(g)
    b = np.zeros(n)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    S.write_input(fin, g, n, off, idx, val, b, viscosity=nu, accel=acc)
    r = subprocess.run([_binary(), fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0 and "error" not in r.stdout, r.stdout + r.stderr
    out = S.read_output(fout, g["keys"].shape[0], n)
    ga = dict(g, u=acc[0], v=acc[1], w=acc[2])
    ref = O.project_dirichlet(ga, nu, rtol=1e-14, real=(g["u"], g["v"], g["w"]))
    wrong = O.project_dirichlet(ga, nu, rtol=1e-14)  # air pressures from the acceleration
    assert rel(out["p"][ex], ref["p"][ex]) < 1e-8
    for k in ("u", "v", "w"):
        assert rel(out[k][ex], ref[k][ex]) < 1e-8, k
    # the two fields give different air pressures, so the check tells them apart
    assert rel(wrong["p"][ex], ref["p"][ex]) > 1e-3
