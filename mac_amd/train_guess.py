"""Train the learned initial guess on the GPU and save its weights.

Training pairs come from the plain solver (FP64 MG-PCG to 1e-10):
  * jets: seeded vertical jets in a 48^3 box (nozzle 4-10 cells, inflow
    0.5-2.0 m/s), one pair every second time step (the step's last projection; the
    previous-pressure input is the pressure that projection started from);
  * spheres: seeded 64^3 free-surface boxes of the projection benchmark's
    family (floor slab + spheres, N(0,1) face velocities) with a zero
    previous pressure.
The grid's bounding box plus the one-brick margin the simulation may grow
fits the 64^3 network volume.  Pairs are generated and trained on as they
come (a replay buffer of recent pairs), with Adam on the masked mean-squared
error in output units; out_scale is the rms of the targets in the first
pairs.

    python -m mac_amd.train_guess --out mac_amd/weights/unet_guess.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import jets, workloads as W
from .lib import MAC_FP64, MAC_MIXED, Projection, Sim, UNet

VOL = (64, 64, 64)
# training seeds start here; the tests hold out jet seeds 100-101 and
# sphere-box seeds 500-502
JET_SEED0 = 1000
SPHERE_SEED0 = 10_000


def log(*a):
    print("[train]", *a, flush=True)


def jet_pairs(net, seed, steps, every=1):
    """Pairs from one seeded jet; yields (in5, target, mask) in out_scale 1 units."""
    rng = np.random.default_rng(10_000 + seed)
    nozzle = int(rng.integers(4, 11))
    inflow = float(rng.uniform(0.5, 2.0))
    j = jets.vertical_jet(48, 48, 48, h=0.01, nozzle=nozzle, inflow=inflow)
    sim = Sim(seed=seed, proj={"rtol": 1e-10}, **j["params"])
    sim.upload(j["keys"], j["flags"], u=j["u"], v=j["v"], w=j["w"])
    for k in range(steps):
        st = sim.step()
        if k >= 3 and k % every == 0 and st["nfluid"] > 64:
            yield net.capture(sim, 1.0)
    sim.close()


def sphere_pair(net, seed, n=64):
    rng = np.random.default_rng(20_000 + seed)
    g = W.sphere_box(n, seed, nspheres=int(rng.integers(3, 9)), rmin=3.0, rmax=float(rng.uniform(8, 16)),
                     slab=int(rng.integers(0, 12)))
    pr = Projection(h=g["h"], rtol=1e-10, precision=MAC_FP64)
    pr.Init_Iteration(g["keys"], g["flags"], g["u"], g["v"], g["w"])
    pr.project()
    out = net.sample(pr, None, 1.0)
    pr.close()
    return out


def save(path, net, out_scale, meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=net.get_params(), out_scale=np.float32(out_scale), shape=np.array(VOL, np.int32),
             meta=np.array(json.dumps(meta)))


def load_weights(path):
    """(params float32, out_scale, shape) from a weights file written by save()."""
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), float(z["out_scale"]), tuple(int(v) for v in z["shape"])


def load_net(path, shape):
    """A network of volume `shape` with the weights of `path`; returns (net, out_scale)."""
    params, out_scale, _ = load_weights(path)
    net = UNet(*shape)
    net.set_params(params)
    return net, out_scale


def eval_projection(g, net, out_scale, precision=MAC_FP64, rtol=1e-8, timed=0):
    """Zero start vs learned start on one projection input: iterations, the
    guess's relative residual ||div u - A x0|| / ||div u|| and error, and
    whether the guided solve fell back.  With timed = k > 0 also the median
    wall time of k solves each way (after the untimed solves that capture the
    solver's graphs), the learned one including its inference."""
    pr = Projection(h=g["h"], rtol=rtol, precision=precision, check_every=1)
    pr.Init_Iteration(g["keys"], g["flags"], g["u"], g["v"], g["w"])
    pr.save_velocity()
    times = {}
    if timed:
        for name, guided in (("zero", False), ("learned", True)):
            ts = []
            for k in range(3 + timed):
                pr.restore_velocity()
                pr.sync()
                t = time.perf_counter()
                if guided:
                    net.guess(pr, use_prev=False, out_scale=out_scale)
                pr.project(guess=guided)
                if k >= 3:
                    ts.append((time.perf_counter() - t) * 1e3)
            times[f"ms_{name}"] = float(np.median(ts))
        pr.restore_velocity()
    st0 = pr.project()
    p0 = pr.get_pressure()
    pr.restore_velocity()
    div = pr.divergence()
    net.guess(pr, use_prev=False, out_scale=out_scale)
    x0 = pr.get_pressure()
    # the reference's system A p = div u (apply_operator is that A)
    r0 = np.linalg.norm(div - pr.apply_operator(x0)) / np.linalg.norm(div)
    e0 = np.linalg.norm(x0 - p0) / np.linalg.norm(p0)
    pr.restore_velocity()
    pr.set_pressure(x0)
    st1 = pr.project(guess=True)
    p1 = pr.get_pressure()
    pr.close()
    return dict({"iters_zero": st0["iters"], "iters_learned": st1["iters"], "fallback": st1["fallback_used"],
                 "guess_rel_resid": float(r0), "guess_rel_err": float(e0),
                 "p_rel_diff": float(np.linalg.norm(p1 - p0) / np.linalg.norm(p0))}, **times)


def eval_jet(net, out_scale, seed, steps=60, nx=48):
    """Total PCG iterations of a seeded jet with a zero, previous-pressure and
    learned starting iterate, and the final state's difference to the zero-start run."""
    from .lib import GUESS_LEARNED, GUESS_PREVIOUS, GUESS_ZERO
    rng = np.random.default_rng(10_000 + seed)
    nozzle = int(rng.integers(4, 11))
    inflow = float(rng.uniform(0.5, 2.0))
    j = jets.vertical_jet(nx, nx, nx, h=0.01, nozzle=nozzle, inflow=inflow)
    out, states = {}, {}
    for name, mode in (("zero", GUESS_ZERO), ("previous", GUESS_PREVIOUS), ("learned", GUESS_LEARNED)):
        sim = Sim(seed=seed, proj={"rtol": 1e-8}, **j["params"])
        sim.set_guess(mode, net if mode == GUESS_LEARNED else None, out_scale)
        sim.upload(j["keys"], j["flags"], u=j["u"], v=j["v"], w=j["w"])
        its = 0
        ms_proj = 0.0
        t = time.time()
        for _ in range(steps):
            st = sim.step()
            its += sum(st["pcg_iters"])
            ms_proj += st["ms_project"]
        out[name] = {"iters": its, "s": time.time() - t, "ms_project": ms_proj, "nfluid_last": st["nfluid"]}
        states[name] = sim.download()
        sim.close()
    for name in ("previous", "learned"):
        a, b = states["zero"], states[name]
        same = a["keys"].shape == b["keys"].shape and np.array_equal(a["keys"], b["keys"])
        out[name]["same_topology"] = bool(same)
        if same:
            out[name]["p_rel_diff"] = float(np.linalg.norm(a["p"] - b["p"]) / max(1e-300, np.linalg.norm(a["p"])))
    out["nozzle"], out["inflow"] = nozzle, inflow
    return out


def eval_c4(weights, sizes=(128, 256), jet_steps=(150, 150), jet_seeds=(100, 101), box_seed=0, log_to=log):
    """The learned guess at the evaluation sizes: for each size n, held-out
    seeded vertical jets in an n^3 box (zero, previous-pressure and learned
    starting iterates, total PCG iterations and projection time over the
    steps; inference over the grid's bounding box counted in the learned
    time), and the projection benchmark's input at n^3 (zero vs learned start,
    mixed precision, median time of 5 solves each way)."""
    res = {}
    for n, steps in zip(sizes, jet_steps):
        # a simulated grid may grow a brick past the box on each side; every
        # guess runs over the grid's own bounding box anyway
        net, osc = load_net(weights, (n + 16, n + 16, n + 64))
        for seed in jet_seeds:
            r = eval_jet(net, osc, seed, steps=steps, nx=n)
            res[f"jet{n}_seed{seed}"] = r
            log_to(f"jet {n}^3 seed {seed}: " + json.dumps(r))
        g = W.sphere_box(n, box_seed)
        r = eval_projection(g, net, osc, MAC_MIXED, timed=5)
        res[f"spheres{n}_seed{box_seed}"] = r
        log_to(f"spheres {n}^3 seed {box_seed}: " + json.dumps(r))
        net.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "weights", "unet_guess.npz"))
    ap.add_argument("--jets", type=int, default=340)
    ap.add_argument("--jet-steps", type=int, default=120)
    ap.add_argument("--spheres", type=int, default=2000)
    ap.add_argument("--iters", type=int, default=40000, help="Adam steps in total")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--buffer", type=int, default=1000)
    ap.add_argument("--new-every", type=int, default=1, help="Adam steps per new pair while the sources last")
    ap.add_argument("--active-jets", type=int, default=8, help="jets stepped side by side")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--time-limit", type=float, default=1000.0, help="seconds; stops early and saves")
    ap.add_argument("--eval-c4", default=None, help="evaluate these weights at 128^3 and 256^3 (no training); "
                    "writes the results as JSON to --out")
    args = ap.parse_args(argv)

    if args.eval_c4:
        res = eval_c4(args.eval_c4)
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
        log(f"wrote {args.out}")
        return 0

    t0 = time.time()
    net = UNet(*VOL, seed=args.seed)
    rng = np.random.default_rng(args.seed)
    buf = []

    def sources():
        # at most active_jets jets at a time, a new seed when one ends; the
        # spheres interleaved so that both kinds arrive throughout
        nxt = 0
        js = []
        si = 0
        per_round = max(1, (args.spheres * args.active_jets) // max(1, args.jets * args.jet_steps // 2))
        while nxt < args.jets or js or si < args.spheres:
            while len(js) < args.active_jets and nxt < args.jets:
                js.append(jet_pairs(net, JET_SEED0 + nxt, args.jet_steps, every=2))
                nxt += 1
            for gen in list(js):
                try:
                    yield "jet", next(gen)
                except StopIteration:
                    js.remove(gen)
            for _ in range(per_round if (js or nxt < args.jets) else 1):
                if si < args.spheres:
                    yield "sphere", sphere_pair(net, SPHERE_SEED0 + si)
                    si += 1

    src = sources()
    # out_scale from the first pairs
    first = []
    for _ in range(16):
        first.append(next(src))
    tv = np.concatenate([p[1][p[2] > 0] for _, p in first])
    out_scale = float(np.sqrt(np.mean(tv.astype(np.float64) ** 2)))
    log(f"out_scale {out_scale:.4g} from {len(first)} pairs")
    for kind, (x5, t, m) in first:
        buf.append((kind, x5, t / out_scale, m))

    counts = {"jet": 0, "sphere": 0}
    losses = []
    exhausted = False
    it = 0
    while it < args.iters and time.time() - t0 < args.time_limit:
        # one new pair per new_every steps while the sources last
        if not exhausted and it % args.new_every == 0:
            try:
                kind, (x5, t, m) = next(src)
                buf.append((kind, x5, t / out_scale, m))
                counts[kind] += 1
                if len(buf) > args.buffer:
                    buf.pop(int(rng.integers(0, len(buf))))
            except StopIteration:
                exhausted = True
        kind, x5, t, m = buf[int(rng.integers(0, len(buf)))]
        lr = args.lr * (0.1 ** (it / max(1, args.iters)))
        losses.append(net.train_step(x5, t, m, lr))
        it += 1
        if it % 200 == 0:
            log(f"step {it}: loss {np.mean(losses[-200:]):.4g}, pairs jet {counts['jet']} sphere {counts['sphere']}, "
                f"{time.time() - t0:.0f}s")
    meta = {"iters": it, "pairs": counts, "final_loss": float(np.mean(losses[-100:])), "seconds": time.time() - t0,
            "lr": args.lr, "jets": args.jets, "spheres": args.spheres, "buffer": args.buffer,
            "loss_curve": [float(np.mean(losses[k:k + 500])) for k in range(0, len(losses), 2000)]}
    save(args.out, net, out_scale, meta)
    log(f"saved {args.out}: {meta}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
