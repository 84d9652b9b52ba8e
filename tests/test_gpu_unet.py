"""GPU checks of the learned-guess U-Net (BF16 matrix-core convolutions, FP32
accumulation) against the FP64 numpy oracle.

Bars: the forward output within 2e-2 relative (L2) of the oracle run with the
same bfloat16 rounding of stored activations and convolution weights; every
parameter gradient within 1e-1 relative (L2) and cosine >= 0.995 of the
oracle's FP64 gradient on the same stored activations (the device keeps
activation gradients in bfloat16: about 2^-9 relative rounding per stored
tensor, compounded over the ~20 tensors between the loss and the first
layer)."""
import numpy as np
import pytest

from mac_amd.lib import UNet
from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu

SHAPE = (8, 8, 64)
# (D, H, W) with H a multiple of 16 on the first two levels: their
# convolutions take the LDS-tiled kernel, the third level the direct one
SHAPE_TILED = (4, 32, 64)


def _inputs(seed, shape=SHAPE):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(int(np.prod(shape)), 5)).astype(np.float32), rng


@pytest.mark.parametrize("seed,shape", [(0, SHAPE), (1, SHAPE), (2, SHAPE_TILED)])
def test_forward_matches_oracle(seed, shape):
    net = UNet(*shape, seed=seed)
    p = net.get_params().astype(np.float64)
    x5, _ = _inputs(seed, shape)
    out = net.forward(x5)
    ref, _ = U.forward(p, x5, shape, bf=True)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert err < 2e-2, err


@pytest.mark.parametrize("seed,shape", [(0, SHAPE), (1, SHAPE), (2, SHAPE_TILED)])
def test_backward_matches_oracle(seed, shape):
    net = UNet(*shape, seed=seed + 10)
    p = net.get_params().astype(np.float64)
    x5, rng = _inputs(seed, shape)
    net.forward(x5)
    gout = rng.normal(size=x5.shape[0]).astype(np.float32)
    g = net.backward(gout)
    _, cache = U.forward(p, x5, shape, bf=True)
    ref = U.backward(p, cache, gout.astype(np.float64), shape)
    lay, _ = U.layout()
    for k, (o, s) in lay.items():
        n = int(np.prod(s))
        a, b = g[o:o + n].astype(np.float64), ref[o:o + n]
        nb = np.linalg.norm(b)
        if nb == 0:
            assert np.linalg.norm(a) == 0
            continue
        rel = np.linalg.norm(a - b) / nb
        cos = a @ b / (np.linalg.norm(a) * nb)
        assert rel < 1e-1 and cos > 0.995, (k, rel, cos)


def test_training_reduces_loss():
    net = UNet(*SHAPE, seed=3)
    x5, rng = _inputs(3)
    # a smooth target the network can fit: a blend of the inputs
    tgt = (0.5 * x5[:, 0] - 0.3 * x5[:, 4] + 0.1).astype(np.float32)
    mask = (rng.uniform(size=x5.shape[0]) < 0.8).astype(np.float32)
    l0 = net.train_step(x5, tgt, mask, 1e-3)
    for _ in range(60):
        l1 = net.train_step(x5, tgt, mask, 1e-3)
    assert l1 < 0.5 * l0, (l0, l1)
