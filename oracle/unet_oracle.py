"""FP64 numpy restatement of the learned-guess U-Net (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker for the HIP U-Net kernels
(mac_amd/csrc/unet*.{h,hip}); never used by the product.

Same architecture and parameter layout as the library: ten 3x3x3
convolutions (enc1a, enc1b, enc2a, enc2b, enc3a, enc3b, dec2a, dec2b, dec1a,
dec1b; weights [Cout][27 Cin] with K ordered (tap, cin), tap = 9(dz+1) +
3(dy+1) + (dx+1), then biases), the two 2x2x2 transposed convolutions up2,
up1 (weights [8][Cout][Cin], biases), and the 1x1x1 head (16 weights, 1 bias).
Arrays are channels-last (D, H, W, C).  With bf16=True every stored
activation and every convolution and transposed-convolution weight of the
forward pass is rounded to bfloat16 (round to nearest even) as the device
stores them.
"""
import numpy as np

CONV = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 32), (32, 32), (32, 16), (16, 16)]
NAMES = ["enc1a", "enc1b", "enc2a", "enc2b", "enc3a", "enc3b", "dec2a", "dec2b", "dec1a", "dec1b"]
UP = [("up2", 64, 32), ("up1", 32, 16)]


def layout():
    """name -> (offset, shape) of every parameter tensor in the flat vector."""
    out, off = {}, 0
    for name, (ci, co) in zip(NAMES, CONV):
        out[name + ".w"] = (off, (co, 27, ci))
        off += co * 27 * ci
        out[name + ".b"] = (off, (co,))
        off += co
    for name, ci, co in UP:
        out[name + ".w"] = (off, (8, co, ci))
        off += 8 * co * ci
        out[name + ".b"] = (off, (co,))
        off += co
    out["head.w"] = (off, (16,))
    off += 16
    out["head.b"] = (off, (1,))
    off += 1
    return out, off


def unpack(p):
    lay, n = layout()
    assert p.size == n
    return {k: np.asarray(p[o:o + int(np.prod(s))], np.float64).reshape(s) for k, (o, s) in lay.items()}


def pack(grads):
    lay, n = layout()
    out = np.zeros(n)
    for k, (o, s) in lay.items():
        out[o:o + int(np.prod(s))] = grads[k].reshape(-1)
    return out


def bf16(a):
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _taps():
    for tap in range(27):
        yield tap, tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1


def conv3(x, w, b):
    D, H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.broadcast_to(b, (D, H, W, w.shape[0])).copy()
    for tap, dz, dy, dx in _taps():
        xs = xp[1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        out += xs @ w[:, tap, :].T
    return out


def conv3_bwd(x, w, g):
    """gradients of sum(conv3(x,w,b) * g): (dx, dw, db)."""
    D, H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    for tap, dz, dy, dx in _taps():
        sl = (slice(1 + dz, 1 + dz + D), slice(1 + dy, 1 + dy + H), slice(1 + dx, 1 + dx + W))
        dw[:, tap, :] = np.einsum("zyxo,zyxi->oi", g, xp[sl])
        dxp[sl] += g @ w[:, tap, :]
    return dxp[1:-1, 1:-1, 1:-1], dw, g.sum(axis=(0, 1, 2))


def pool(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def pool_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def _octants(D, H, W):
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    return (x & 1) + 2 * (y & 1) + 4 * (z & 1)


def up(x, w, b):
    """out[2v + o] = b + W[o] x[v] (pre-activation)."""
    xf = np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)
    o = _octants(*xf.shape[:3])
    out = np.einsum("zyxi,zyxoi->zyxo", xf, w[o]) + b
    return out


def up_bwd(x, w, g):
    xf = np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)
    o = _octants(*xf.shape[:3])
    dw = np.zeros_like(w)
    for k in range(8):
        m = o == k
        dw[k] = np.einsum("no,ni->oi", g[m], xf[m])
    dxf = np.einsum("zyxo,zyxoi->zyxi", g, w[o])
    D, H, W, C = x.shape
    dx = dxf.reshape(D, 2, H, 2, W, 2, C).sum(axis=(1, 3, 5))
    return dx, dw, g.sum(axis=(0, 1, 2))


def forward(p, x5, shape, bf=False):
    """x5: (V, 5) inputs; returns (out (V,), cache)."""
    P = unpack(p)
    D, H, W = shape
    r = bf16 if bf else (lambda a: a)
    wr = bf16 if bf else (lambda a: a)
    x0 = np.zeros((D, H, W, 8))
    x0[..., :5] = np.asarray(x5, np.float64).reshape(D, H, W, 5)
    x0 = r(x0)
    c = {"x0": x0}

    def cv(name, x, relu=True):
        y = conv3(x, wr(P[name + ".w"]), P[name + ".b"])
        if relu:
            y = np.maximum(y, 0.0)
        return r(y)

    c["e1a"] = cv("enc1a", x0)
    c["e1"] = cv("enc1b", c["e1a"])
    c["p1"] = r(pool(c["e1"]))
    c["e2a"] = cv("enc2a", c["p1"])
    c["e2"] = cv("enc2b", c["e2a"])
    c["p2"] = r(pool(c["e2"]))
    c["e3a"] = cv("enc3a", c["p2"])
    c["e3"] = cv("enc3b", c["e3a"])
    c["u2"] = r(np.maximum(up(c["e3"], wr(P["up2.w"]), P["up2.b"]), 0.0))
    c["c2"] = np.concatenate([c["u2"], c["e2"]], axis=-1)
    c["d2a"] = cv("dec2a", c["c2"])
    c["d2"] = cv("dec2b", c["d2a"])
    c["u1"] = r(np.maximum(up(c["d2"], wr(P["up1.w"]), P["up1.b"]), 0.0))
    c["c1"] = np.concatenate([c["u1"], c["e1"]], axis=-1)
    c["d1a"] = cv("dec1a", c["c1"])
    c["d1"] = cv("dec1b", c["d1a"])
    out = c["d1"] @ P["head.w"] + P["head.b"][0]
    return out.reshape(-1), c


def backward(p, cache, gout, shape, bf=False):
    """FP64 gradients of sum(out * gout) w.r.t. the flat parameters.  With
    bf=True every stored activation gradient is rounded to bfloat16 where the
    device stores one (the head's output gradient, each data gradient of a
    convolution or transposed convolution, each sum of a skip gradient and a
    pooling gradient), and the convolutions' data gradients use the bfloat16
    weights, as the device does; parameter gradients are FP64 sums."""
    P = unpack(p)
    D, H, W = shape
    c = cache
    G = {}
    r = bf16 if bf else (lambda a: a)
    wr = bf16 if bf else (lambda a: a)
    go = np.asarray(gout, np.float64).reshape(D, H, W)
    G["head.w"] = np.einsum("zyx,zyxc->c", go, c["d1"])
    G["head.b"] = np.array([go.sum()])
    g = r(go[..., None] * P["head.w"])

    def cvb(name, x, y, g, need_dx=True):
        g = g * (y > 0)
        dx, dw, db = conv3_bwd(x, wr(P[name + ".w"]), g)
        G[name + ".w"], G[name + ".b"] = dw, db
        return r(dx)

    g = cvb("dec1b", c["d1a"], c["d1"], g)
    gc1 = cvb("dec1a", c["c1"], c["d1a"], g)
    gu1 = gc1[..., :16] * (c["u1"] > 0)
    gd2, G["up1.w"], G["up1.b"] = up_bwd(c["d2"], P["up1.w"], gu1)
    g = cvb("dec2b", c["d2a"], c["d2"], r(gd2))
    gc2 = cvb("dec2a", c["c2"], c["d2a"], g)
    gu2 = gc2[..., :32] * (c["u2"] > 0)
    ge3, G["up2.w"], G["up2.b"] = up_bwd(c["e3"], P["up2.w"], gu2)
    g = cvb("enc3b", c["e3a"], c["e3"], r(ge3))
    gp2 = cvb("enc3a", c["p2"], c["e3a"], g)
    ge2 = r(gc2[..., 32:] + pool_bwd(gp2))
    g = cvb("enc2b", c["e2a"], c["e2"], ge2)
    gp1 = cvb("enc2a", c["p1"], c["e2a"], g)
    ge1 = r(gc1[..., 16:] + pool_bwd(gp1))
    g = cvb("enc1b", c["e1a"], c["e1"], ge1)
    cvb("enc1a", c["x0"], c["e1a"], g)
    return pack(G)
