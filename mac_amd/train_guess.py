"""Train the learned initial guess on the GPU and save its weights.

Training pairs come from the plain solver (FP64 MG-PCG to 1e-10):
  * jets: seeded vertical jets in a 48^3 box (nozzle 4-10 cells, inflow
    0.5-2.0 m/s), one pair per time step (the step's last projection; the
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
from .lib import MAC_FP64, Projection, Sim, UNet

VOL = (64, 64, 64)


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


def eval_projection(g, net, out_scale, precision=MAC_FP64, rtol=1e-8):
    """Zero start vs learned start on one projection input: iterations, the
    guess's relative residual ||div u - A x0|| / ||div u|| and error, and
    whether the guided solve fell back."""
    pr = Projection(h=g["h"], rtol=rtol, precision=precision, check_every=1)
    pr.Init_Iteration(g["keys"], g["flags"], g["u"], g["v"], g["w"])
    pr.save_velocity()
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
    return {"iters_zero": st0["iters"], "iters_learned": st1["iters"], "fallback": st1["fallback_used"],
            "guess_rel_resid": float(r0), "guess_rel_err": float(e0),
            "p_rel_diff": float(np.linalg.norm(p1 - p0) / np.linalg.norm(p0))}


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
        t = time.time()
        for _ in range(steps):
            its += sum(sim.step()["pcg_iters"])
        out[name] = {"iters": its, "s": time.time() - t}
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "weights", "unet_guess.npz"))
    ap.add_argument("--jets", type=int, default=12)
    ap.add_argument("--jet-steps", type=int, default=120)
    ap.add_argument("--spheres", type=int, default=240)
    ap.add_argument("--iters", type=int, default=3000, help="Adam steps in total")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--buffer", type=int, default=160)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--time-limit", type=float, default=480.0, help="seconds; stops early and saves")
    args = ap.parse_args(argv)

    t0 = time.time()
    net = UNet(*VOL, seed=args.seed)
    rng = np.random.default_rng(args.seed)
    buf = []

    def sources():
        js = [jet_pairs(net, s, args.jet_steps, every=2) for s in range(args.jets)]
        si = 0
        while js or si < args.spheres:
            for gen in list(js):
                try:
                    yield "jet", next(gen)
                except StopIteration:
                    js.remove(gen)
            for _ in range(2):
                if si < args.spheres:
                    yield "sphere", sphere_pair(net, si)
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
        # one new pair per 4 steps while the sources last
        if not exhausted and it % 4 == 0:
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
            "lr": args.lr, "jets": args.jets, "spheres": args.spheres}
    save(args.out, net, out_scale, meta)
    log(f"saved {args.out}: {meta}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
