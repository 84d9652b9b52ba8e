"""GPU parity of the projection (HIP library through the C-ABI) against the
CPU oracle on identical seeded inputs.

Bars: divergence and gradient update bit-exact (same operation order, no
FMA contraction on either side); operator within 1e-14 relative; pressure
and face velocities within 1e-8 relative of the oracle's direct-solve
stand-in (solves at rtol 1e-12); mixed precision within 1e-6 relative at
rtol 1e-10."""
import glob
import os

import numpy as np
import pytest

from mac_amd import workloads as W
from mac_amd.lib import MAC_FP64, MAC_MIXED, Projection
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRIDS = {
    "rand_2x2x2": lambda: W.random_grid(1),
    "rand_3x2x2_neg_origin": lambda: W.random_grid(2, 3, 2, 2, origin=(-2, -1, 5)),
    "rand_dropped_bricks": lambda: W.random_grid(3, 3, 3, 2, drop_bricks=4),
    "rand_dense_fluid": lambda: W.random_grid(4, 2, 3, 2, fluid_p=0.9, const_p=0.02),
    "single_brick": lambda: W.random_grid(5, 1, 1, 1),
    "line_1d_kat": W.line_1d_kat,
}


def dom_mask(g):
    f = g["flags"]
    return ((f & 0x80) > 0) & np.isin(f & 7, [1, 3, 4])


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make(g, **kw):
    pr = Projection(h=g["h"], **kw)
    pr.Init_Iteration(g["keys"], g["flags"], g["u"], g["v"], g["w"])
    return pr


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_divergence_bitexact(name):
    g = GRIDS[name]()
    pr = make(g)
    dg = pr.divergence()
    do = O.divergence(g)
    assert np.array_equal(dg, do), f"max diff {np.abs(dg - do).max()}"


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_operator(name):
    g = GRIDS[name]()
    pr = make(g)
    x = W.normal(7, 3, np.arange(g["flags"].size, dtype=np.uint64)) * dom_mask(g)
    yg = pr.apply_operator(x)
    yo = O.apply_A(g, x)
    assert rel(yg, yo) < 1e-14


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_gradient_bitexact(name):
    g = GRIDS[name]()
    pr = make(g)
    p = W.normal(9, 4, np.arange(g["flags"].size, dtype=np.uint64)) * dom_mask(g)
    pr.set_pressure(p)
    pr.gradient_update()
    ug, vg, wg = pr.get_velocity()
    uo, vo, wo = O.gradient_update(g, p, g["u"], g["v"], g["w"])
    assert np.array_equal(ug, uo) and np.array_equal(vg, vo) and np.array_equal(wg, wo)


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_project_fp64_vs_oracle(name):
    g = GRIDS[name]()
    ref = O.project(g, rtol=1e-14)
    pr = make(g, rtol=1e-12, precision=MAC_FP64)
    st = pr.Divergence_Projection()
    p = pr.get_pressure()
    u, v, w = pr.get_velocity()
    assert st["true_rel_resid"] <= 1e-12 * 1.0001
    assert rel(p, ref["p"]) < 1e-8, (rel(p, ref["p"]), st)
    for a, b in ((u, ref["u"]), (v, ref["v"]), (w, ref["w"])):
        assert rel(a, b) < 1e-8
    # max |div u| after the update equals the oracle's to the solve tolerance
    assert abs(st["max_div"] - ref["max_div"]) <= 1e-8 * max(1.0, ref["max_div"]) + 1e-9


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "proj_*.npz"))), ids=os.path.basename)
def test_superlu_golden_through_hip(path):
    """The committed golden systems (pressure from a direct sparse LU, SuperLU
    standing in for UMFPACK) pushed straight through the device projection:
    divergence bit-exact, pressure and corrected face velocities within 1e-8."""
    z = np.load(path)
    g = {k: z[k] for k in ("keys", "flags", "u", "v", "w")}
    g["h"] = tuple(float(x) for x in z["h"])
    pr = make(g, rtol=1e-12, precision=MAC_FP64)
    assert np.array_equal(pr.divergence(), z["div"])
    pr.Divergence_Projection()
    assert rel(pr.get_pressure(), z["p"]) < 1e-8
    for k, a in zip(("u", "v", "w"), pr.get_velocity()):
        assert rel(a, z[k + "_out"]) < 1e-8, k


def test_line_1d_known_answer():
    """test/old/TestAlgorithms_Solve_Pressure.h:56-200: u(0)=u(1)=1 after projection."""
    g = W.line_1d_kat()
    pr = make(g, rtol=1e-12)
    pr.Divergence_Projection()
    u, _, _ = pr.get_velocity()
    p = pr.get_pressure()
    for gx in (0, 1):
        b, t = g["cells"][gx]
        assert abs(u[b * 512 + t] - 1.0) < 1e-12
        assert abs(p[b * 512 + t] - 1.0) < 1e-12
    b, t = g["cells"][2]
    assert abs(u[b * 512 + t] - 1.0) < 1e-12


def test_uniform_flow_gives_zero_pressure():
    """test/old/TestSolvePressureUmfpack.h:271-336: a divergence-free uniform
    flow through the domain leaves p = 0 and u unchanged."""
    keys = W.box_keys(2, 2, 2)
    n = 16
    dense = np.full((n, n, n), W.FLUID | W.EXISTS, dtype=np.uint8)
    dense[[0, -1], :, :] = W.AIR | W.EXISTS
    dense[:, [0, -1], :] = W.AIR | W.EXISTS
    dense[:, :, [0, -1]] = W.AIR | W.EXISTS
    g = {"keys": keys, "flags": W.dense_to_bricks(dense, 2, 2, 2).reshape(-1), "h": (0.5, 0.25, 1.0)}
    n = g["flags"].size
    g["u"], g["v"], g["w"] = np.full(n, 0.3), np.full(n, -0.2), np.full(n, 1.5)
    pr = make(g, rtol=1e-12)
    st = pr.Divergence_Projection()
    assert st["iters"] == 0
    assert np.all(pr.get_pressure() == 0.0)
    u, v, w = pr.get_velocity()
    assert np.all(u == 0.3) and np.all(v == -0.2) and np.all(w == 1.5)


def small_spheres(n, seed):
    return W.sphere_box(n, seed, nspheres=6, rmin=3.0, rmax=10.0, slab=6)


def deep_pool():
    """The breakup jet's layout at 128x64x64: a pool 45 cells deep on a const
    floor, the falling column above it, N(0,1) face velocities on the fluid
    cells (seeded)."""
    from mac_amd import jets as J
    b = J.breakup_jet(nx=128, ny=64, nz=64, pool_depth=45, margin=4, nozzle=8, ppc=1, seed=0)
    g = {"keys": b["keys"], "flags": b["flags"], "h": (0.01, 0.01, 0.01)}
    fluid = dom_mask(g)
    rng = np.random.default_rng(0)
    for k in "uvw":
        a = np.zeros(g["flags"].size)
        a[fluid] = rng.standard_normal(int(fluid.sum()))
        g[k] = a
    return g


@pytest.mark.parametrize("precision", [MAC_FP64, MAC_MIXED])
def test_deep_pool_multigrid_size_independent(precision):
    """On the deep pool the device MG-PCG takes the oracle's iterations, 10
    with the Galerkin coarse operator (coarse operators rediscretised from the
    coarsened labels take 29 there), and reaches the tolerance."""
    g = deep_pool()
    cpu = O.mgpcg(g, rtol=1e-8)
    pr = make(g, rtol=1e-8, precision=precision)
    st = pr.Divergence_Projection()
    assert st["true_rel_resid"] <= 1e-8 * 1.0001
    assert abs(st["iters"] - cpu["iters"]) <= 1, (st["iters"], cpu["iters"])
    assert st["iters"] <= 12, st["iters"]
    assert rel(pr.get_pressure(), cpu["p"]) < 1e-6


@pytest.mark.parametrize("seed", [0, 1])
def test_sphere_box_48_fp64_vs_cpu_mg(seed):
    g = small_spheres(48, seed)
    cpu = O.mgpcg(g, rtol=1e-10)
    ref = O.project(g, rtol=1e-13)
    pr = make(g, rtol=1e-10, precision=MAC_FP64)
    st = pr.Divergence_Projection()
    p = pr.get_pressure()
    # same algorithm on the CPU: same iteration count up to rounding
    assert abs(st["iters"] - cpu["iters"]) <= 1, (st["iters"], cpu["iters"])
    assert st["levels"] == cpu["levels"]
    assert rel(p, ref["p"]) < 1e-7
    assert rel(p, cpu["p"]) < 1e-7


@pytest.mark.parametrize("seed", [0, 1])
def test_sphere_box_48_mixed(seed):
    g = small_spheres(48, seed)
    ref = O.project(g, rtol=1e-13)
    pr = make(g, rtol=1e-10, precision=MAC_MIXED)
    st = pr.Divergence_Projection()
    p = pr.get_pressure()
    assert st["true_rel_resid"] <= 1e-10 * 1.0001
    assert rel(p, ref["p"]) < 1e-6, rel(p, ref["p"])


def test_warm_start_previous_pressure():
    g = small_spheres(32, 3)
    pr = make(g, rtol=1e-10)
    pr.save_velocity()
    st0 = pr.Divergence_Projection()
    pr.restore_velocity()
    st1 = pr.Divergence_Projection(guess=True)  # exact previous answer as guess
    assert st1["iters"] < st0["iters"]
    assert st1["true_rel_resid"] <= 1e-10 * 1.0001


def test_solve_csr_slot_a():
    """Slot A: mac_solve_csr on the reference-assembled CSR vs the oracle's solve."""
    from mac_amd.lib import LinearSolver
    g = W.random_grid(6)
    n, off, idx, val, row = O.assemble(g)
    div = O.divergence(g)
    b = np.zeros(n)
    b[row[row >= 0]] = div[row >= 0]
    xo, _ = O.cg_csr(n, off, idx, val, b, rtol=1e-14)
    ls = LinearSolver(rtol=1e-13)
    ls.Solve_Linear_FactorizeMatrice(n, off, idx, val)
    xg = ls.Solve_Linear(n, off, idx, val, b)
    ls.Solve_Linear_Clean()
    assert rel(xg, xo) < 1e-9
    # n <= 3 host path (extrapolation corner systems)
    off3 = np.array([0, 2, 5, 7], dtype=np.int32)
    idx3 = np.array([0, 1, 0, 1, 2, 1, 2], dtype=np.int32)
    val3 = np.array([-2.0, 1.0, 1.0, -3.0, 1.0, 1.0, -2.0])
    b3 = np.array([1.0, 2.0, 3.0])
    x3 = ls.Solve_Linear(3, off3, idx3, val3, b3)
    A = np.array([[-2, 1, 0], [1, -3, 1], [0, 1, -2.0]])
    assert np.allclose(A @ x3, b3, rtol=0, atol=1e-14)
    # a non-symmetric 3x3 (the corner systems' shape) in compressed columns,
    # UMFPACK's layout: value[3c + r] is row r of column c
    An = np.array([[2.0, -1.0, 0.5], [0.25, 3.0, -2.0], [1.0, 1.0, 1.0]])
    offn = np.array([0, 3, 6, 9], dtype=np.int32)
    idxn = np.array([0, 1, 2] * 3, dtype=np.int32)
    xn = ls.Solve_Linear(3, offn, idxn, An.T.reshape(-1).copy(), b3)
    assert np.allclose(xn, np.linalg.solve(An, b3), rtol=1e-14, atol=1e-14)


def _box_system(nx, ny, nz, seed):
    """The reference-assembled pressure system of a fluid box in an air shell
    (nx x ny x nz cells, the outer cell layer air) with N(0,1) velocities."""
    nbx, nby, nbz = nx // 8, ny // 8, nz // 8
    mat = np.full((nz, ny, nx), W.AIR, np.uint8)
    mat[1:-1, 1:-1, 1:-1] = W.FLUID
    n_all = nx * ny * nz
    g = {"keys": W.box_keys(nbx, nby, nbz), "flags": W.dense_to_bricks(mat | W.EXISTS, nbx, nby, nbz).reshape(-1),
         "h": (0.01, 0.01, 0.01)}
    for i, k in enumerate(("u", "v", "w")):
        g[k] = W.normal(seed, i, np.arange(n_all, dtype=np.uint64))
    n, off, idx, val, row = O.assemble(g)
    div = O.divergence(g)
    b = np.zeros(n)
    b[row[row >= 0]] = div[row >= 0]
    return n, off, idx, val, b


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 32, 64)], ids=["one_workgroup", "grid_stride"])
def test_solve_csr_slot_a_paths(dims):
    """Both device paths of slot A: a jet-sized system (2,744 unknowns, one
    workgroup runs the whole solve) and the full 32 x 32 x 64 box of the
    small-jet configuration (55,800 unknowns, grid-stride launches), each to
    rtol 1e-13, against the oracle's CG; then a smaller system on the same
    context (the workspace is reused)."""
    import time

    from mac_amd.lib import LinearSolver
    n, off, idx, val, b = _box_system(*dims, seed=11)
    xo, _ = O.cg_csr(n, off, idx, val, b, rtol=1e-14)
    ls = LinearSolver(rtol=1e-13)
    ls.Solve_Linear(n, off, idx, val, b)  # first call sizes the workspace
    ts = []
    for _ in range(9):
        t0 = time.perf_counter()
        xg = ls.Solve_Linear(n, off, idx, val, b)
        ts.append((time.perf_counter() - t0) * 1e3)
    print(f"slot A n={n}: {np.median(ts):.2f} ms per solve to 1e-13, median of 9 (host arrays in and out)")
    assert rel(xg, xo) < 1e-10
    n2, off2, idx2, val2, b2 = _box_system(16, 16, 8, seed=12)
    x2 = ls.Solve_Linear(n2, off2, idx2, val2, b2)
    x2o, _ = O.cg_csr(n2, off2, idx2, val2, b2, rtol=1e-14)
    assert rel(x2, x2o) < 1e-10


@pytest.mark.parametrize("poison", ["nan", "huge"])
def test_guided_solve_falls_back_on_bad_guess(poison):
    """A guess whose first residual is worse than the zero start's (or not
    finite) is dropped: the solve restarts from zero and still converges."""
    g = small_spheres(32, 4)
    ref = O.project(g, rtol=1e-13)
    pr = make(g, rtol=1e-10)
    n = g["flags"].size
    bad = np.full(n, np.nan) if poison == "nan" else W.normal(11, 0, np.arange(n, dtype=np.uint64)) * 1e12
    pr.set_pressure(bad * dom_mask(g) if poison == "huge" else bad)
    st = pr.Divergence_Projection(guess=True)
    assert st["fallback_used"] == 1
    assert st["true_rel_resid"] <= 1e-10 * 1.0001
    assert rel(pr.get_pressure(), ref["p"]) < 1e-7


def test_guided_solve_falls_back_after_stall_iters():
    """stall_iters forces the fallback once a guided solve has not converged
    after that many iterations; the result equals the unguided solve's."""
    g = small_spheres(32, 5)
    pr = make(g, rtol=1e-10, check_every=1)
    pr.save_velocity()
    st0 = pr.Divergence_Projection()
    p0 = pr.get_pressure()
    pr.restore_velocity()
    pr.set_params(stall_iters=3)
    pr.set_pressure(0.5 * p0)  # a fair guess that still needs many iterations
    st1 = pr.Divergence_Projection(guess=True)
    assert st1["fallback_used"] == 1
    # the host polls at the stall iteration: exactly 3 iterations are dropped
    assert st1["iters"] == st0["iters"] + 3, (st0["iters"], st1["iters"])
    assert st1["true_rel_resid"] <= 1e-10 * 1.0001
    assert rel(pr.get_pressure(), p0) < 1e-8


@pytest.mark.parametrize("prec", [MAC_FP64, MAC_MIXED])
def test_graph_replays_match_eager_launches(prec):
    """The captured two-iteration graphs give bit-identical results to the
    same kernels launched one by one (event-timing mode), on every replay,
    with other solves and host round trips in between."""
    gs = [small_spheres(32, 7), small_spheres(32, 8)]
    ctx = {}
    for mode in ("eager", "graph"):
        for i, g in enumerate(gs):
            pr = make(g, rtol=1e-10, precision=prec)
            pr.enable_timing(mode == "eager")
            pr.save_velocity()
            ctx[(mode, i)] = pr
    outs = {}
    for rep in range(5):  # graphs start with the third solve on a grid
        for i in range(len(gs)):
            for mode in ("eager", "graph"):
                pr = ctx[(mode, i)]
                pr.restore_velocity()
                st = pr.Divergence_Projection()
                outs[(mode, i, rep)] = (st["iters"], pr.get_pressure(), np.concatenate(pr.get_velocity()))
    for i in range(len(gs)):
        for rep in range(5):
            ie, pe, ue = outs[("eager", i, rep)]
            ig, pg, ug = outs[("graph", i, rep)]
            assert ie == ig and np.array_equal(pe, pg) and np.array_equal(ue, ug), (i, rep)


@pytest.mark.parametrize("prec", [MAC_FP64, MAC_MIXED])
def test_coarse_tail_matches_launch_per_step(prec):
    """The last two V-cycle levels in one launch (k_vcycle_tail) against one
    launch per sweep, restriction, solve and prolongation: same iteration
    counts and the same answer to rounding (the steps are the same
    expressions; only the compiler's contraction choices may differ)."""
    for g, nl in ((small_spheres(32, 1), 3), (small_spheres(48, 0), 4)):
        res = {}
        for tail in (0, 1):
            pr = make(g, rtol=1e-10, precision=prec, coarse_tail=tail)
            pr.save_velocity()
            out = []
            for rep in range(3):  # the third solve runs as graphs
                pr.restore_velocity()
                st = pr.Divergence_Projection()
                assert st["levels"] == nl
                out.append((st["iters"], pr.get_pressure()))
            res[tail] = out
        for (i0, p0), (i1, p1) in zip(res[0], res[1]):
            assert i0 == i1
            assert rel(p1, p0) < (1e-12 if prec == MAC_FP64 else 1e-7), rel(p1, p0)


@pytest.mark.parametrize("prec", [MAC_FP64, MAC_MIXED])
@pytest.mark.parametrize("tail", [0, 1])
def test_fused_coarse_matches_separate_launches(prec, tail):
    """Restriction fused with the next level's first block step, prolongation
    fused with the first post-smoothing step (the prolongated face planes of
    the neighbour bricks formed in registers) against separate launches: same
    iteration counts and the same answer to rounding, on grids with missing
    and partly filled bricks, negative keys and 4-6 levels."""
    grids = [small_spheres(48, 0), small_spheres(64, 3),
             W.random_grid(3, 4, 4, 4, fluid_p=0.7, drop_bricks=9, origin=(-3, -1, 2))]
    for gi, g in enumerate(grids):
        res = {}
        for fuse in (0, 1):
            pr = make(g, rtol=1e-10, precision=prec, coarse_tail=tail, fuse_coarse=fuse)
            pr.save_velocity()
            out = []
            for rep in range(3):  # the third solve runs as graphs
                pr.restore_velocity()
                st = pr.Divergence_Projection()
                assert st["true_rel_resid"] <= 1e-10 * 1.0001
                out.append((st["iters"], pr.get_pressure()))
            res[fuse] = out
        for (i0, p0), (i1, p1) in zip(res[0], res[1]):
            assert i0 == i1, (gi, i0, i1)
            assert rel(p1, p0) < (1e-12 if prec == MAC_FP64 else 1e-7), (gi, rel(p1, p0))


@pytest.mark.parametrize("name", ["rand_2x2x2", "rand_dropped_bricks", "rand_dense_fluid", "spheres_32"])
@pytest.mark.parametrize("prec", [MAC_FP64, MAC_MIXED])
def test_viscous_air_pressure_vs_oracle(name, prec):
    """Algorithms_Solve_Pressure_Dirichlet with Policy_Dirichlet_Boundary on
    the device: the air pressures, the right side, the written pressure
    (air pressures included) and the updated velocities against the oracle;
    two solves in a row (the second reuses the context)."""
    g = small_spheres(32, 2) if name == "spheres_32" else GRIDS[name]()
    nu = 0.013
    ref = O.project_dirichlet(g, nu, rtol=1e-13)
    tol = 1e-8 if prec == MAC_FP64 else 1e-6
    pr = make(g, rtol=1e-12 if prec == MAC_FP64 else 1e-10, precision=prec, viscosity=nu)
    pr.save_velocity()
    for rep in range(2):
        pr.restore_velocity()
        pr.Divergence_Projection(guess=rep == 1)
        p = pr.get_pressure()
        u = np.concatenate(pr.get_velocity())
        assert rel(p, ref["p"]) < tol, (rep, rel(p, ref["p"]))
        assert rel(u, np.concatenate([ref[k] for k in ("u", "v", "w")])) < tol
        air = ((g["flags"] & 0x80) > 0) & ~dom_mask(g) & (ref["p"] != 0)
        assert air.any()
        assert np.allclose(p[air], ref["p"][air], rtol=1e-13, atol=0)


@pytest.mark.parametrize("prec", [MAC_FP64, MAC_MIXED])
def test_no_fluid_cells_leave_the_velocity(prec):
    """A grid with no Fluid, Inflow or Interior cell has no unknowns: no
    iteration, p = 0 and every face velocity unchanged, as in the oracle."""
    g = W.random_grid(9, 2, 2, 2, fluid_p=0.0)
    assert not dom_mask(g).any()
    ref = O.project(g, rtol=1e-14)
    pr = make(g, rtol=1e-12, precision=prec)
    st = pr.Divergence_Projection()
    assert st["iters"] == 0
    assert np.all(pr.get_pressure() == 0.0) and np.all(ref["p"] == 0.0)
    for a, b, k in zip(pr.get_velocity(), (ref["u"], ref["v"], ref["w"]), ("u", "v", "w")):
        assert np.array_equal(a, g[k]) and np.array_equal(b, g[k])
    pr.close()


def test_full_size_256_properties():
    """The projection benchmark input at its full size (256^3 box, slab + 40
    seeded spheres, 4.7M unknowns): FP64 and mixed solves to 1e-8 reach the
    tolerance on the true residual, leave max |div u| more than 1e5 times
    smaller than before, agree in pressure within 1e-6 (the mixed-precision
    bar), and a doubled velocity field gives a doubled pressure (linearity)."""
    g = W.sphere_box(256, 0)
    out = {}
    for prec in (MAC_FP64, MAC_MIXED):
        pr = make(g, rtol=1e-8, precision=prec)
        before = pr.max_divergence()
        st = pr.Divergence_Projection()
        assert st["true_rel_resid"] <= 1e-8 * 1.0001, st
        assert st["iters"] <= 12, st
        assert st["max_div"] < 1e-5 * before, (st["max_div"], before)
        out[prec] = pr.get_pressure()
        pr.close()
    assert rel(out[MAC_MIXED], out[MAC_FP64]) < 1e-6
    g2 = dict(g, u=2 * g["u"], v=2 * g["v"], w=2 * g["w"])
    pr = make(g2, rtol=1e-8, precision=MAC_FP64)
    pr.Divergence_Projection()
    assert rel(pr.get_pressure(), 2 * out[MAC_FP64]) < 1e-7
    pr.close()
