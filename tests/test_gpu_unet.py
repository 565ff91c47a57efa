"""The learned Poisson stage on the GPU: forward output and parameter
gradients of the BF16/MFMA U-Net against the CPU oracle (which rounds to BF16
at the same points), training reduces the loss, and the learned FGMRES path
reaches the tolerance (with or without falling back to the plain V-cycle).

Tolerances: the device accumulates in FP32 in a different order and rounds
its stored BF16 activations independently, so single values can differ by a
BF16 ulp (2^-8 relative); normwise agreement is required to 2e-2 (forward)
and 5e-2 (gradients).
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from helpers import channel_pair, relerr, vfs

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import unet_oracle as uo  # noqa: E402

pytestmark = pytest.mark.gpu


def _nrel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _setup(levels=3, base=16):
    c, o, ctx = channel_pair(32, 16, 32, les=1)
    net = vfs.UNet(ctx, base=base, levels=levels, seed=7)
    orc = uo.UNetOracle(base, levels)
    rng = np.random.default_rng(4)
    shape = ctx.field_shape("B")
    b = np.zeros(shape)
    b[1:-1, 1:-1, 1:-1] = rng.standard_normal((shape[0] - 2, shape[1] - 2, shape[2] - 2))
    phi = np.zeros(shape)
    phi[1:-1, 1:-1, 1:-1] = rng.standard_normal((shape[0] - 2, shape[1] - 2, shape[2] - 2))
    ctx.upload("B", b)
    ctx.upload("PHI", phi)
    return c, o, ctx, net, orc, b, phi


def test_unet_forward_matches_oracle():
    c, o, ctx, net, orc, b, phi = _setup()
    params = net.params(0).astype(np.float64)
    net.set_output_scale(2.5)
    net.forward(ctx.device_ptr("B"), ctx.device_ptr("P"))
    y = ctx.download("P")
    out, cache = orc.forward(params, b[1:-1, 1:-1, 1:-1], bf16=True)
    want = out * cache["rms"] * 2.5
    assert np.all(y[0] == 0) and np.all(y[:, :, 0] == 0)
    assert _nrel(y[1:-1, 1:-1, 1:-1], want) <= 2e-2
    net.close(); ctx.close()


def test_unet_gradients_match_oracle():
    c, o, ctx, net, orc, b, phi = _setup()
    params = net.params(0).astype(np.float64)
    loss = net.gradients(ctx.device_ptr("B"), ctx.device_ptr("PHI"))
    g = net.params(1).astype(np.float64)
    lo, go, _ = orc.gradients(params, b[1:-1, 1:-1, 1:-1], phi[1:-1, 1:-1, 1:-1], 1.0, bf16=True)
    assert abs(loss - lo) <= 1e-2 * lo
    # layer by layer
    off = 0
    for sh in orc.shapes():
        n = int(np.prod(sh))
        assert _nrel(g[off:off + n], go[off:off + n]) <= 5e-2, (sh, _nrel(g[off:off + n], go[off:off + n]))
        off += n
    net.close(); ctx.close()


def test_unet_training_reduces_loss():
    c, o, ctx, net, orc, b, phi = _setup()
    # a real (b, phi) pair: one Poisson solve of the channel
    ctx.runge_kutta()
    ctx.poisson_rhs()
    ctx.upload("PHI", np.zeros(ctx.field_shape("PHI")))
    ctx.poisson_solve(vfs.default_solve_opts(tol=1e-10, use_guess=0))
    bd, pd = ctx.download("B"), ctx.download("PHI")
    scale = float(np.sqrt(np.mean(pd[1:-1, 1:-1, 1:-1] ** 2)) / np.sqrt(np.mean(bd[1:-1, 1:-1, 1:-1] ** 2)))
    net.set_output_scale(scale)
    losses = [net.train_step(ctx.device_ptr("B"), ctx.device_ptr("PHI"), 1e-3) for _ in range(40)]
    assert np.all(np.isfinite(losses))
    assert min(losses[-5:]) < 0.5 * losses[0], losses[::5]
    net.close(); ctx.close()


def test_learned_poisson_solve_reaches_tolerance():
    c, o, ctx, net, orc, b, phi = _setup()
    ctx.runge_kutta()
    ctx.poisson_rhs()
    ctx.upload("PHI", np.zeros(ctx.field_shape("PHI")))
    st = net.poisson_solve(vfs.default_solve_opts(tol=1e-8, use_guess=0))
    assert st.relres <= 1e-8
    x = ctx.download("PHI")
    r = ctx.download("B")
    # the oracle's operator agrees that x solves the system
    o.set("b", r)
    res = r - o.poisson_apply(x)
    assert np.linalg.norm(res) <= 1.01e-8 * np.linalg.norm(r)
    net.close(); ctx.close()
