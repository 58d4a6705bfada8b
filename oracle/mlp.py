"""FP64 oracle of the control-variate network (test infrastructure only).

The reference has no learned component; this restates the network the GPU
path trains (64 -> 256 -> 256 -> 1, GELU) so its forward pass and gradients
can be checked.  ``bf16=True`` rounds every matrix-product operand to
bfloat16 (round to nearest even), which is what the GPU's MFMA tiles consume;
products of two bf16 values are exact in FP32, so the GPU then agrees with
this oracle to FP32 accumulation error.  Bias gradients are products with a
ones vector, as on the GPU.

Pinned by central finite differences of the loss (tests/test_oracle_mlp.py).
Parameter layout (flat, float): W1[nin,nh] b1[nh] W2[nh,nh] b2[nh] w3[nh] b3[1],
row-major, y = gelu(gelu(x W1 + b1) W2 + b2) w3 + b3.
"""
import math

import numpy as np
from scipy.special import erf


def sizes(nin=64, nh=256):
    return [(nin, nh), (nh,), (nh, nh), (nh,), (nh,), (1,)]


def unpack(theta, nin=64, nh=256):
    out, o = [], 0
    for sh in sizes(nin, nh):
        n = int(np.prod(sh))
        out.append(np.asarray(theta[o:o + n], dtype=np.float64).reshape(sh))
        o += n
    return out


def nparams(nin=64, nh=256):
    return sum(int(np.prod(s)) for s in sizes(nin, nh))


def bf16(x):
    """Round float values to bfloat16 (nearest even) and back to float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    r = ((u + np.uint64(0x7FFF) + lsb) >> np.uint64(16)) << np.uint64(16)
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _mm(a, b, rb):
    return (bf16(a) @ bf16(b)) if rb else a @ b


def forward(theta, x, nin=64, nh=256, rb=False):
    W1, b1, W2, b2, w3, b3 = unpack(theta, nin, nh)
    z1 = _mm(x, W1, rb) + b1
    h1 = gelu(z1)
    z2 = _mm(h1, W2, rb) + b2
    h2 = gelu(z2)
    y = _mm(h2, w3[:, None], rb)[:, 0] + b3[0]
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(theta, x, target, nin=64, nh=256, rb=False):
    """L = sum((y - t)^2) / (2n) and dL/dtheta (same flat layout)."""
    W1, b1, W2, b2, w3, b3 = unpack(theta, nin, nh)
    y, (x, z1, h1, z2, h2) = forward(theta, x, nin, nh, rb)
    n = x.shape[0]
    dy = (y - target) / n
    loss = 0.5 * np.sum((y - target) ** 2) / n
    gw3 = _mm(h2.T, dy[:, None], rb)[:, 0]
    ones = np.ones((1, n))
    gb3 = _mm(ones, dy[:, None], rb)[0]
    dz2 = _mm(dy[:, None], w3[None, :], rb) * gelu_grad(z2)
    gW2 = _mm(h1.T, dz2, rb)
    gb2 = _mm(ones, dz2, rb)[0]
    dz1 = _mm(dz2, W2.T, rb) * gelu_grad(z1)
    gW1 = _mm(x.T, dz1, rb)
    gb1 = _mm(ones, dz1, rb)[0]
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, gb3])
    return loss, g


def init(seed, nin=64, nh=256):
    """He-style initialisation used by the GPU trainer's host side."""
    rng = np.random.default_rng(seed)
    return np.concatenate([
        rng.standard_normal(nin * nh) * math.sqrt(2.0 / nin), np.zeros(nh),
        rng.standard_normal(nh * nh) * math.sqrt(2.0 / nh), np.zeros(nh),
        rng.standard_normal(nh) * math.sqrt(1.0 / nh), np.zeros(1)])
