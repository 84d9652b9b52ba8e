"""GPU checks of the learned-guess plumbing: the brick grid -> network volume
packing (mac_unet_sample) against a numpy restatement of the same mapping and
scaling, the guess write-back (mac_unet_guess), the simulation's starting
iterate modes, and that a guided solve still reaches the tolerance.

Bars: packed inputs and targets within float32 rounding (2e-6 relative) of
numpy's FP64 values; guided solves at the same rtol as unguided ones, and
jets warm-started from the previous pressure within 1e-8 of the zero-start
run."""
import numpy as np
import pytest

from mac_amd import jets
from mac_amd import workloads as W
from mac_amd.lib import GUESS_PREVIOUS, GUESS_ZERO, MAC_FP64, MAC_MIXED, Projection, Sim, UNet

pytestmark = pytest.mark.gpu


def spheres(n, seed):
    return W.sphere_box(n, seed, nspheres=5, rmin=3.0, rmax=9.0, slab=5)


def project(g, **kw):
    pr = Projection(h=g["h"], rtol=1e-10, **kw)
    pr.Init_Iteration(g["keys"], g["flags"], g["u"], g["v"], g["w"])
    pr.save_velocity()
    return pr


def dense(g, a):
    nb = g["dims"][0] // 8
    return W.bricks_to_dense(np.asarray(a).reshape(-1, 512), nb, nb, nb)


def test_sample_packing_matches_numpy():
    g = spheres(32, 0)
    pr = project(g)
    div = pr.divergence()
    pr.restore_velocity()
    pr.project()
    p = pr.get_pressure()
    prev = W.normal(5, 1, np.arange(p.size, dtype=np.uint64))
    net = UNet(32, 32, 64, seed=0)
    x5, t, m = net.sample(pr, prev, out_scale=2.0)
    flags = g["flags"]
    fl = dense(g, flags)
    dom = (fl & 0x80 > 0) & np.isin(fl & 7, [1, 3, 4])
    f = -dense(g, div)
    s = np.sqrt(np.mean(f[dom] ** 2))
    hm = 1.0  # h = 1 on every axis
    vol = np.zeros((32, 32, 64, 5))
    vol[..., 3] = 1.0
    sub = vol[:, :, :32]
    sub[..., 0] = np.where(dom, f / s, 0.0)
    sub[..., 1] = dom
    air = ~dom & (fl & 0x80 > 0) & ((fl & 7) != 2)
    sub[..., 2] = air
    sub[..., 3] = ~dom & ~air
    sub[..., 4] = np.where(dom, dense(g, prev) * hm / s, 0.0)
    got = x5.reshape(32, 32, 64, 5)
    for c in range(5):
        err = np.abs(got[..., c] - vol[..., c]).max() / max(1.0, np.abs(vol[..., c]).max())
        assert err < 2e-6, (c, err)
    tgt = np.zeros((32, 32, 64))
    tgt[:, :, :32] = np.where(dom, dense(g, p) * hm / (s * 2.0), 0.0)
    assert np.abs(t.reshape(tgt.shape) - tgt).max() <= 2e-6 * np.abs(tgt).max()
    mk = np.zeros((32, 32, 64))
    mk[:, :, :32] = dom
    assert np.array_equal(m.reshape(mk.shape), mk)


def test_guess_writes_scaled_output_on_unknowns():
    """All-zero weights leave the head bias b as the output: the guess is
    b * out_scale * s / hm on the unknowns and 0 elsewhere."""
    g = spheres(32, 1)
    pr = project(g)
    div = pr.divergence()
    net = UNet(32, 32, 64, seed=0)
    params = np.zeros(net.nparams, np.float32)
    params[-1] = 0.75
    net.set_params(params)
    net.guess(pr, use_prev=False, out_scale=3.0)
    x = pr.get_pressure()
    fl = g["flags"]
    dom = ((fl & 0x80) > 0) & np.isin(fl & 7, [1, 3, 4])
    s = np.sqrt(np.mean(div[dom] ** 2))
    expect = np.where(dom, 0.75 * 3.0 * s, 0.0)
    assert np.allclose(x, expect, rtol=1e-6, atol=0)


def test_guess_over_bounding_box_matches_network_of_that_size():
    """A 32^3 grid under a 64 x 64 x 128 network: the guess runs over the
    grid's bounding box rounded up to what the kernels take (32, 32, 64) and
    equals, bit for bit, the guess of a network of exactly that volume with
    the same weights."""
    g = spheres(32, 3)
    big = UNet(64, 64, 128, seed=5)
    small = UNet(32, 32, 64, seed=5)
    small.set_params(big.get_params())
    xs = []
    for net in (big, small):
        pr = project(g)
        net.guess(pr, use_prev=False, out_scale=1.5)
        xs.append(pr.get_pressure())
        pr.close()
    assert np.abs(xs[1]).max() > 0
    assert np.array_equal(xs[0], xs[1])


@pytest.mark.parametrize("prec", [MAC_FP64, MAC_MIXED])
def test_untrained_guess_still_converges(prec):
    g = spheres(32, 2)
    pr = project(g, precision=prec)
    st0 = pr.project()
    p0 = pr.get_pressure()
    pr.restore_velocity()
    net = UNet(32, 32, 64, seed=7)
    net.guess(pr, use_prev=False, out_scale=1.0)
    st1 = pr.project(guess=True)
    assert st1["true_rel_resid"] <= 1e-10 * 1.0001
    assert np.linalg.norm(pr.get_pressure() - p0) <= 1e-7 * np.linalg.norm(p0)
    assert st0["fallback_used"] == 0


def test_sim_previous_pressure_warm_start():
    j = jets.vertical_jet(24, 24, 32, nozzle=4, inflow=1.0)
    out = {}
    for mode in (GUESS_ZERO, GUESS_PREVIOUS):
        sim = Sim(seed=3, proj={"rtol": 1e-12}, **j["params"])
        sim.set_guess(mode)
        sim.upload(j["keys"], j["flags"], u=j["u"], v=j["v"], w=j["w"])
        its = 0
        for _ in range(12):
            its += sum(sim.step()["pcg_iters"])
        out[mode] = (its, sim.download())
        sim.close()
    (i0, a), (i1, b) = out[GUESS_ZERO], out[GUESS_PREVIOUS]
    assert np.array_equal(a["keys"], b["keys"]) and np.array_equal(a["flags"], b["flags"])
    for k in ("u", "v", "w", "p"):
        assert np.linalg.norm(a[k] - b[k]) <= 1e-8 * max(1e-30, np.linalg.norm(a[k])), k
    assert i1 < i0, (i0, i1)


def test_sim_capture_pairs():
    j = jets.vertical_jet(24, 24, 32, nozzle=4, inflow=1.0)
    sim = Sim(seed=4, proj={"rtol": 1e-10}, **j["params"])
    sim.upload(j["keys"], j["flags"], u=j["u"], v=j["v"], w=j["w"])
    net = UNet(64, 64, 64, seed=0)
    for _ in range(6):
        sim.step()
    x5, t, m = net.capture(sim, 1.0)
    d = sim.download()
    assert m.sum() > 0 and np.isfinite(x5).all() and np.isfinite(t).all()
    # fluid mask channel == mask, masks partition the volume
    assert np.array_equal(x5[:, 1], m)
    assert np.array_equal(x5[:, 1] + x5[:, 2] + x5[:, 3], np.ones_like(m))
    # rhs channel has unit rms over the unknowns
    assert abs(np.sqrt(np.mean(x5[m > 0, 0].astype(np.float64) ** 2)) - 1.0) < 1e-5
    assert d["keys"].shape[0] > 0


def test_trained_guess_gain_on_held_out_inputs():
    """What the committed network gains over the method without it, on inputs
    it was not trained on (jet seeds 100-101, sphere-box seeds 500-502):
    fewer PCG iterations than the zero start over a jet run and on the
    projections, every solve at the same tolerance (fallback guards on)."""
    import os
    from mac_amd.train_guess import eval_jet, eval_projection, load_net
    w = os.path.join(os.path.dirname(__file__), "..", "mac_amd", "weights", "unet_guess.npz")
    net, osc = load_net(w, (64, 64, 64))
    zero = learned = 0
    for seed in (500, 501, 502):
        rng = np.random.default_rng(20_000 + seed)
        g = W.sphere_box(64, seed, nspheres=int(rng.integers(3, 9)), rmin=3.0, rmax=float(rng.uniform(8, 16)),
                         slab=int(rng.integers(0, 12)))
        r = eval_projection(g, net, osc, MAC_MIXED)
        print("[gain] sphere64", seed, r, flush=True)
        assert r["p_rel_diff"] < 1e-6
        assert r["guess_rel_err"] < 1.0  # closer to the answer than the zero start
        zero, learned = zero + r["iters_zero"], learned + r["iters_learned"]
    assert learned <= zero, (zero, learned)
    for seed in (100, 101):
        r = eval_jet(net, osc, seed, steps=30)
        print("[gain] jet48", seed, r, flush=True)
        assert r["learned"]["iters"] < r["zero"]["iters"]
        if r["learned"]["same_topology"]:
            assert r["learned"]["p_rel_diff"] < 1e-6


@pytest.mark.parametrize("n,steps", [(128, 40), (256, 20)])
def test_learned_guess_at_evaluation_sizes(n, steps):
    """Configuration 4's sizes: a held-out jet in an n^3 box for `steps`
    steps with the zero, previous-pressure and learned starting iterates
    (every projection to the same 1e-8; the learned start takes fewer
    iterations than the zero start and the runs end in the same state within
    1e-6), and the projection benchmark's input at n^3 (zero vs learned
    start, same answer within 1e-6)."""
    import os
    from mac_amd.train_guess import eval_jet, eval_projection, load_net
    w = os.path.join(os.path.dirname(__file__), "..", "mac_amd", "weights", "unet_guess.npz")
    net, osc = load_net(w, (n + 16, n + 16, n + 64))
    r = eval_jet(net, osc, 100, steps=steps, nx=n)
    print(f"[c4] jet {n}^3 {steps} steps:", {k: r[k]["iters"] for k in ("zero", "previous", "learned")}, flush=True)
    assert r["learned"]["iters"] < r["zero"]["iters"]
    for k in ("previous", "learned"):
        assert r[k]["same_topology"] and r[k]["p_rel_diff"] < 1e-6
    g = W.sphere_box(n, 0)
    p = eval_projection(g, net, osc, MAC_MIXED)
    print(f"[c4] benchmark input {n}^3:", p, flush=True)
    assert p["fallback"] == 0 and p["p_rel_diff"] < 1e-6 and p["guess_rel_err"] < 1.0
    net.close()
