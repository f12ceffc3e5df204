"""CPU restatement of the warm-start network (the checker of fhip_mlp).

TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module; the product
package never does.  The network has no counterpart in the reference (it is an
extension of the step path): a 3-128-128-13 perceptron with exact-erf GELU
hidden layers, the FP32 parameter layout of fhip_mlp_set_params (W1[128][3],
b1, W2[128][128], b2, W3[13][128], b3, row-major W[out][in]) and the loss
sum_o w_o (net(x)_o - y_o)^2 averaged over the samples.

``emulate=True`` computes in FP64 but rounds to BF16 (round to nearest even)
exactly where the device kernels feed an MFMA operand: the weights, the
inputs, the hidden activations, and the output and hidden gradients of the
backward pass; bias gradients and pre-activations stay unrounded, as on the
device.  ``emulate=False`` is the plain FP64 network, whose gradient is
pinned by central finite differences (tests/test_oracle_mlp.py); the BF16
rounding is pinned against torch's bfloat16 conversion.
"""
from __future__ import annotations

import math

import numpy as np

NIN, NH, NOUT = 3, 128, 13
OW1 = 0
OB1 = OW1 + NH * NIN
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NOUT * NH
NPARAM = OB3 + NOUT

_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(a) -> np.ndarray:
    """Round to BF16 (nearest even) via FP32, returned as FP64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return (p[OW1:OB1].reshape(NH, NIN), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3],
            p[OW3:OB3].reshape(NOUT, NH), p[OB3:NPARAM])


def _forward(params, x, emulate):
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(params)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    X0 = r(x)
    Z1 = X0 @ W1.T + b1
    A1 = r(gelu(Z1))
    Z2 = A1 @ W2.T + b2
    A2 = r(gelu(Z2))
    Z3 = A2 @ W3.T + b3
    return dict(W1=W1, W2=W2, W3=W3, X0=X0, Z1=Z1, A1=A1, Z2=Z2, A2=A2, Z3=Z3)


def forward(params, x, emulate=True) -> np.ndarray:
    return _forward(params, x, emulate)["Z3"]


def loss_and_gradient(params, x, y, w=None, emulate=True):
    """(loss, gradient in the parameter layout)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    f = _forward(params, x, emulate)
    n = x.shape[0]
    w = np.ones(NOUT) if w is None else np.asarray(w, dtype=np.float64)
    R = f["Z3"] - y
    loss = float(np.sum(w * R * R) / n)
    G3 = 2.0 * w * R / n
    G3b = r(G3)
    G2 = (G3b @ f["W3"]) * gelu_d(f["Z2"])
    G2b = r(G2)
    G1 = (G2b @ f["W2"]) * gelu_d(f["Z1"])
    G1b = r(G1)
    g = np.zeros(NPARAM)
    g[OW3:OB3] = (G3b.T @ f["A2"]).ravel()
    g[OB3:NPARAM] = G3.sum(axis=0)
    g[OW2:OB2] = (G2b.T @ f["A1"]).ravel()
    g[OB2:OW3] = G2.sum(axis=0)
    g[OW1:OB1] = (G1b.T @ f["X0"]).ravel()
    g[OB1:OW2] = G1.sum(axis=0)
    return loss, g


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights and zero biases (any generator: the device's own
    initialisation is not part of the check)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    for off, fo, fi in ((OW1, NH, NIN), (OW2, NH, NH), (OW3, NOUT, NH)):
        a = math.sqrt(6.0 / (fi + fo))
        p[off:off + fo * fi] = rng.uniform(-a, a, fo * fi)
    return p
