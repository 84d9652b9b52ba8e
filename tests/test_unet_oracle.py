"""CPU checks pinning the U-Net oracle (oracle/unet_oracle.py): the analytic
FP64 backward against central finite differences, and the convolution
against a direct per-voxel loop."""
import numpy as np

from oracle import unet_oracle as U

SHAPE = (4, 4, 64)


def _setup(seed=0):
    rng = np.random.default_rng(seed)
    lay, n = U.layout()
    p = np.zeros(n)
    for k, (o, s) in lay.items():
        fan = s[-1] * (27 if len(s) == 3 and s[1] == 27 else 1)
        p[o:o + int(np.prod(s))] = rng.normal(0, np.sqrt(2.0 / fan), int(np.prod(s)))
        if k.endswith(".b"):
            p[o:o + int(np.prod(s))] = rng.normal(0, 0.1, int(np.prod(s)))
    x5 = rng.normal(size=(np.prod(SHAPE), 5))
    return p, x5, rng


def test_layout_size_matches_library_formula():
    lay, n = U.layout()
    conv = sum(co * 27 * ci + co for ci, co in U.CONV)
    assert n == conv + 8 * 32 * 64 + 32 + 8 * 16 * 32 + 16 + 17


def test_conv3_matches_direct_loop():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(3, 4, 5, 2))
    w = rng.normal(size=(3, 27, 2))
    b = rng.normal(size=3)
    y = U.conv3(x, w, b)
    D, H, W, _ = x.shape
    ref = np.zeros((D, H, W, 3))
    for z in range(D):
        for yy in range(H):
            for xx in range(W):
                s = b.copy()
                for tap in range(27):
                    dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
                    if 0 <= z + dz < D and 0 <= yy + dy < H and 0 <= xx + dx < W:
                        s += w[:, tap, :] @ x[z + dz, yy + dy, xx + dx]
                ref[z, yy, xx] = s
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_backward_matches_finite_differences():
    p, x5, rng = _setup(2)
    gout = rng.normal(size=np.prod(SHAPE))
    out, cache = U.forward(p, x5, SHAPE)
    g = U.backward(p, cache, gout, SHAPE)
    lay, n = U.layout()
    eps = 1e-6
    # a few entries of every tensor
    for k, (o, s) in lay.items():
        for j in rng.choice(int(np.prod(s)), size=min(3, int(np.prod(s))), replace=False):
            idx = o + j
            pp, pm = p.copy(), p.copy()
            pp[idx] += eps
            pm[idx] -= eps
            fp = U.forward(pp, x5, SHAPE)[0] @ gout
            fm = U.forward(pm, x5, SHAPE)[0] @ gout
            fd = (fp - fm) / (2 * eps)
            assert abs(fd - g[idx]) <= 1e-5 * max(1.0, abs(fd)), (k, j, fd, g[idx])
