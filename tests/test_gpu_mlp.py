"""The warm-start network's MFMA kernels against the CPU oracle, and what the
trained network gains.

Forward outputs and full-batch gradients of fhip_mlp are compared with
oracle/mlp.py, which rounds to BF16 where the kernels feed MFMA operands
(the oracle is pinned by finite differences in tests/test_oracle_mlp.py).
Residual differences come from FP32 accumulation on the device against FP64
on the host, which can flip a BF16 rounding of an activation; the tolerances
below are two orders of magnitude under what a transposed operand or a
dropped term produces.  Training is deterministic (per-workgroup gradient
slots folded in a fixed order).  The last test trains the network on
equilibrated 12^3 replicas and runs a heating scan from the reference's
default start and from the warm start in one batch.
"""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

FWD_TOL = 2e-3      # max |y_gpu - y_oracle| / max |y_oracle|
GRAD_TOL = 2e-2     # per parameter block, normwise relative


def params_and_data(n, seed):
    rng = np.random.default_rng(seed)
    p = omlp.init_params(seed)
    p[omlp.OB1:omlp.OW2] = 0.1 * rng.standard_normal(omlp.NH)
    p[omlp.OB2:omlp.OW3] = 0.1 * rng.standard_normal(omlp.NH)
    p[omlp.OB3:] = 0.1 * rng.standard_normal(omlp.NOUT)
    p = p.astype(np.float32).astype(np.float64)
    x = rng.uniform(-1.5, 1.5, (n, omlp.NIN)).astype(np.float32).astype(np.float64)
    y = rng.standard_normal((n, omlp.NOUT)).astype(np.float32).astype(np.float64)
    return p, x, y


@pytest.mark.parametrize("n", [1, 16, 100, 1000])
def test_forward_matches_oracle(n):
    from feram_amd.warmstart import Mlp
    print(f"[gpu] mlp forward n={n}", flush=True)
    p, x, _ = params_and_data(n, 1)
    net = Mlp()
    net.params = p
    y = net.forward(x)
    yo = omlp.forward(p, x, emulate=True)
    err = np.max(np.abs(y - yo)) / np.max(np.abs(yo))
    assert err < FWD_TOL, err
    net.close()


@pytest.mark.parametrize("n", [16, 203])
def test_gradient_matches_oracle(n):
    from feram_amd.warmstart import Mlp
    print(f"[gpu] mlp gradient n={n}", flush=True)
    p, x, y = params_and_data(n, 2)
    w = np.linspace(0.5, 1.5, omlp.NOUT)
    net = Mlp()
    net.params = p
    loss, g = net.gradient(x, y, w)
    lo, go = omlp.loss_and_gradient(p, x, y, w, emulate=True)
    assert loss == pytest.approx(lo, rel=1e-3)
    blocks = {"W1": (omlp.OW1, omlp.OB1), "b1": (omlp.OB1, omlp.OW2), "W2": (omlp.OW2, omlp.OB2),
              "b2": (omlp.OB2, omlp.OW3), "W3": (omlp.OW3, omlp.OB3), "b3": (omlp.OB3, omlp.NPARAM)}
    for name, (a, b) in blocks.items():
        err = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        print(f"[gpu]   {name}: {err:.2e}", flush=True)
        assert err < GRAD_TOL, (name, err)
    # and against the exact FP64 gradient, at BF16 accuracy
    _, gx = omlp.loss_and_gradient(p, x, y, w, emulate=False)
    assert np.linalg.norm(g - gx) / np.linalg.norm(gx) < 5e-2
    net.close()


def test_training_fits_and_is_deterministic():
    from feram_amd.warmstart import Mlp
    print("[gpu] mlp training", flush=True)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (600, 3))
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], np.cos(x[:, 2]), x[:, 0] * x[:, 1] * x[:, 2]], axis=1)
    y = np.concatenate([t, t ** 2, np.tanh(3 * t), t[:, :1] - t[:, 1:2], t[:, 2:] + 0.5, t[:, :1] * 0.0 + 1.0,
                        t[:, 1:2] * t[:, 2:]], axis=1)
    assert y.shape[1] == omlp.NOUT
    a, b = Mlp(seed=9), Mlp(seed=9)
    ha = a.train(x, y, iters=1500, lr=3e-3)
    hb = b.train(x, y, iters=1500, lr=3e-3)
    assert np.array_equal(ha, hb) and np.array_equal(a.params, b.params)
    print(f"[gpu]   loss {ha[0]:.3e} -> {ha[-1]:.3e}", flush=True)
    assert ha[-1] < 2e-2 * ha[0]
    # the recorded loss is the loss of the parameters before each step
    lo, _ = omlp.loss_and_gradient(a.params.astype(np.float64), x, y, emulate=True)
    l_now, _ = a.gradient(x, y)
    assert l_now == pytest.approx(lo, rel=1e-2)
    a.close()
    b.close()


def test_warm_start_shortens_equilibration():
    from feram_amd.warmstart import WarmStart, equilibrium_samples, heating_scan
    print("[gpu] warm start: 96 training replicas 12^3, heating scan 8 T x 2", flush=True)
    rng = np.random.default_rng(12)
    cond = np.stack([rng.uniform(150, 450, 96), rng.uniform(0, 1e-3, 96), rng.uniform(-6, -4, 96)], axis=1)
    targets = equilibrium_samples(cond, L=12, n_therm=3000, n_avg=1000, seed=3)
    ws = WarmStart(seed=1)
    hist = ws.fit(cond, targets, iters=2000, lr=2e-3)
    print(f"[gpu]   training loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert hist[-1] < 0.5 * hist[0]
    temps = np.linspace(200, 450, 8)
    res = heating_scan(ws, temps, L=12, n_steps=3000, record=10)
    print(f"[gpu]   tau cold {res['cold']['tau']}", flush=True)
    print(f"[gpu]   tau warm {res['warm']['tau']}", flush=True)
    print(f"[gpu]   gain {res['gain']:.2f}", flush=True)
    uc, uw = np.array(res["cold"]["u"]), np.array(res["warm"]["u"])
    print(f"[gpu]   |u| cold {np.round(uc[:, 2], 3)}", flush=True)
    print(f"[gpu]   |u| warm {np.round(uw[:, 2], 3)}", flush=True)
    assert res["gain"] > 1.0
