"""Pins of the warm-start network's CPU restatement (oracle/mlp.py).

The network has no counterpart in the reference, so its oracle is pinned by
mathematics: the FP64 gradient against central finite differences of the
loss in every parameter block, closed forms for a network whose last layer is
zero, and the BF16 rounding against torch's bfloat16 conversion (ties
included).  A planted mistake -- the backward pass through W2 without its
transpose -- must fail the finite-difference check.
"""
import numpy as np
import pytest

from oracle import mlp


def data(n=37, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (n, mlp.NIN)), rng.standard_normal((n, mlp.NOUT))


def fd_check(grad_fn, params, x, y, w, idx, h=1e-5):
    _, g = grad_fn(params, x, y, w)
    errs = []
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = mlp.loss_and_gradient(pp, x, y, w, emulate=False)
        lm, _ = mlp.loss_and_gradient(pm, x, y, w, emulate=False)
        fd = (lp - lm) / (2 * h)
        errs.append(abs(g[i] - fd) / max(abs(fd), 1e-3))
    return max(errs)


def sample_indices(rng):
    blocks = [(mlp.OW1, mlp.OB1), (mlp.OB1, mlp.OW2), (mlp.OW2, mlp.OB2), (mlp.OB2, mlp.OW3),
              (mlp.OW3, mlp.OB3), (mlp.OB3, mlp.NPARAM)]
    return [int(rng.integers(a, b)) for a, b in blocks for _ in range(6)]


def test_fp64_gradient_matches_finite_differences():
    x, y = data()
    p = mlp.init_params(5)
    p[mlp.OB1:mlp.OW2] = 0.1
    p[mlp.OB2:mlp.OW3] = -0.05
    w = np.linspace(0.5, 2.0, mlp.NOUT)
    idx = sample_indices(np.random.default_rng(11))
    err = fd_check(lambda *a: mlp.loss_and_gradient(*a, emulate=False), p, x, y, w, idx)
    assert err < 1e-6, err


def test_planted_transpose_mistake_fails():
    x, y = data()
    p = mlp.init_params(6)
    w = np.ones(mlp.NOUT)

    def wrong(params, x, y, w):
        f = mlp._forward(params, x, False)
        n = x.shape[0]
        R = f["Z3"] - y
        G3 = 2.0 * w * R / n
        G2 = (G3 @ f["W3"]) * mlp.gelu_d(f["Z2"])
        G1 = (G2 @ f["W2"].T) * mlp.gelu_d(f["Z1"])          # planted: W2^T instead of W2
        g = np.zeros(mlp.NPARAM)
        g[mlp.OW1:mlp.OB1] = (G1.T @ f["X0"]).ravel()
        g[mlp.OB1:mlp.OW2] = G1.sum(axis=0)
        return float(np.sum(w * R * R) / n), g

    idx = [mlp.OW1 + k for k in range(0, 30, 3)] + [mlp.OB1 + k for k in range(0, 30, 3)]
    assert fd_check(wrong, p, x, y, w, idx) > 1e-3


def test_zero_last_layer_closed_form():
    x, y = data(20)
    p = mlp.init_params(7)
    p[mlp.OW3:mlp.OB3] = 0.0
    b3 = np.linspace(-1, 1, mlp.NOUT)
    p[mlp.OB3:] = b3
    w = np.linspace(1.0, 3.0, mlp.NOUT)
    for emulate in (False, True):
        out = mlp.forward(p, x, emulate=emulate)
        assert np.allclose(out, b3[None, :], rtol=0, atol=0)
        loss, g = mlp.loss_and_gradient(p, x, y, w, emulate=emulate)
        assert loss == pytest.approx(np.sum(w * (b3 - y) ** 2) / len(x), rel=1e-14)
        assert np.allclose(g[mlp.OB3:], 2 * w * (b3 - y).mean(axis=0), rtol=1e-13, atol=1e-15)
        # nothing flows back through a zero last layer
        assert not np.any(g[:mlp.OW3])


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 4000)])
    # exact ties: low 16 bits 0x8000, with even and odd kept parts
    base = rng.integers(0x3F000000, 0x4F000000, 200, dtype=np.uint32)
    ties = ((base & 0xFFFF0000) | 0x8000).view(np.float32)
    v = np.concatenate([v, ties, -ties])
    ref = torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(v), ref)
