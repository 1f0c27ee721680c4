"""TEST INFRASTRUCTURE ONLY: numpy restatement of the learned Kessler stand-in
(erf_amd/csrc/erfx_mlp.hip) -- 8 -> 128 -> 128 -> 4, exact (erf) GELU --
used to check the GPU's forward pass and gradients.

With bf16=True the values the GPU rounds to BF16 are rounded here too
(inputs, weights, both hidden activations, the hidden-layer error signal);
everything else is float64, so agreement with the GPU's FP32 accumulation is
to a stated tolerance, not bitwise.  With bf16=False it is the plain float64
network, whose backward pass tests/test_mlp_cpu.py checks against finite
differences.
"""
import math

import numpy as np

NI, NH, NO = 8, 128, 4
OFF_W1 = 0
OFF_B1 = OFF_W1 + NI * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NO
NPARAM = OFF_B3 + NO

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16 << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NI, NH), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(NH, NO), p[OFF_B3:NPARAM])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, x, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xb = r(x)
    z1 = b1 + xb @ r(W1)
    h1 = r(gelu(z1))
    z2 = b2 + h1 @ r(W2)
    h2 = r(gelu(z2))
    y = b3 + h2 @ r(W3)
    return y, (xb, z1, h1, z2, h2)


def loss_and_grad(p, x, t, bf=True):
    """Mean squared error over all n x 4 outputs and its gradient (NPARAM)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (xb, z1, h1, z2, h2) = forward(p, x, bf)
    n = x.shape[0]
    d = y - t
    loss = float((d * d).sum() / (n * NO))
    dy = 2.0 * d / (n * NO)
    dz2 = r((dy @ r(W3).T) * gelu_grad(z2))
    dz1 = (dz2 @ r(W2).T) * gelu_grad(z1)
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (xb.T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    g[OFF_W3:OFF_B3] = (h2.T @ dy).ravel()
    g[OFF_B3:NPARAM] = dy.sum(axis=0)
    return loss, g


def init_params(seed=0):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    for off, fi, fo in ((OFF_W1, NI, NH), (OFF_W2, NH, NH), (OFF_W3, NH, NO)):
        lim = math.sqrt(6.0 / (fi + fo))
        p[off:off + fi * fo] = rng.uniform(-lim, lim, fi * fo)
    return p
