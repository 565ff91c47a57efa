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


def test_k_marching_convolution_matches_tiled(monkeypatch):
    """The k-marching convolution (halo ring and weights in LDS) against the
    plane-by-plane one on a 4-level network over a grid whose i and j counts
    are not multiples of the tile: the same layer outputs up to FP32
    accumulation order (layers with more than one 16-channel chunk sum the
    K terms in another order), and the forward output matches the oracle."""
    c, o, ctx = channel_pair(40, 24, 24, les=1)
    with ctx:
        rng = np.random.default_rng(9)
        shape = ctx.field_shape("B")
        b = np.zeros(shape)
        b[1:-1, 1:-1, 1:-1] = rng.standard_normal((shape[0] - 2, shape[1] - 2, shape[2] - 2))
        ctx.upload("B", b)
        outs = []
        for km in ("1", "0"):
            monkeypatch.setenv("VFS_UNET_KMARCH", km)
            net = vfs.UNet(ctx, base=16, levels=4, seed=11)
            net.forward(ctx.device_ptr("B"), ctx.device_ptr("P"))
            outs.append(ctx.download("P"))
            params = net.params(0).astype(np.float64)
            net.close()
        monkeypatch.delenv("VFS_UNET_KMARCH")
        err = _nrel(outs[0], outs[1])
        print(f"[k-marching vs tiled] {err:.3e}", flush=True)
        assert err <= 1e-2
        orc = uo.UNetOracle(16, 4)
        out, cache = orc.forward(params, b[1:-1, 1:-1, 1:-1], bf16=True)
        assert _nrel(outs[0][1:-1, 1:-1, 1:-1], out * cache["rms"]) <= 2e-2


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


@pytest.mark.parametrize("direction", ["0", "1"], ids=["warm-start", "warm-start+direction"])
def test_learned_poisson_solve_reaches_tolerance(direction, monkeypatch):
    """The learned solve (warm start; optionally also the network as the
    first FGMRES direction) reaches 1e-8 under the oracle's operator."""
    monkeypatch.setenv("VFS_LEARNED_DIRECTION", direction)
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


@pytest.mark.parametrize("fill,direction,fallback", [(np.nan, "0", 1), (np.nan, "1", 1), (0.0, "1", 1), (0.0, "0", 0)],
                         ids=["nan-weights", "nan-weights+direction", "zero-weights+direction", "zero-weights"])
def test_learned_solve_falls_back(fill, direction, fallback, monkeypatch):
    """A network that cannot help triggers the fallback: after three
    iterations the residual has not halved, the solve restarts from the
    original guess with the plain V-cycle, reports fallback = 1 and still
    reaches the tolerance under the oracle's operator.  Non-finite weights
    poison the warm start (fallback either way); all-zero weights give a zero
    warm start (alpha = 0: the guess is kept and the V-cycle FGMRES converges
    on its own, no fallback) and, as the first direction, a zero direction
    (fallback)."""
    monkeypatch.setenv("VFS_LEARNED_DIRECTION", direction)
    c, o, ctx, net, orc, b, phi = _setup()
    p = net.params(0)
    p[:] = fill
    net.set_params(p)
    ctx.runge_kutta()
    ctx.poisson_rhs()
    ctx.upload("PHI", np.zeros(ctx.field_shape("PHI")))
    for fp32 in (1, 0):
        ctx.upload("PHI", np.zeros(ctx.field_shape("PHI")))
        st = net.poisson_solve(vfs.default_solve_opts(tol=1e-8, mg_fp32=fp32, use_guess=0))
        assert st.fallback == fallback and st.relres <= 1e-8
        x, r = ctx.download("PHI"), ctx.download("B")
        assert np.all(np.isfinite(x))
        o.set("b", r)
        assert np.linalg.norm(r - o.poisson_apply(x)) <= 1.01e-8 * np.linalg.norm(r)
    net.close(); ctx.close()


def test_learned_stage_on_held_out_seed():
    """The shipped network on the 256x192x256 channel of a seed that neither
    training nor validation used, three steps into the run: the warm start
    x0 = g + alpha net(b - A g) from the previous step's correction g lowers
    the residual below that of g alone, and the learned FGMRES reaches 1e-8
    without falling back.  Prints the residual ratio and the iterations and
    times of the solves with and without the network."""
    import importlib

    learned = importlib.import_module("vfs-wind_amd.learned")
    ctx = learned.channel_context(256, 192, 256, learned.HELD_OUT_SEED)
    with ctx:
        opts = vfs.default_solve_opts(tol=1e-8)
        for _ in range(3):
            ctx.fractional_step(opts)
        # the next step's right-hand side; PHI holds the previous correction (the guess)
        ctx.calc_k_flux()
        ctx.contra2cart()
        ctx.eddy_viscosity()
        ctx.pressure_gradient()
        ctx.runge_kutta()
        ctx.poisson_rhs()
        n = int(np.prod(ctx.field_shape("PHI")))
        b, g = ctx.download("B"), ctx.download("PHI")

        def apply(x):
            ctx.upload("PHI", x)
            ctx.poisson_apply_dev(ctx.device_ptr("PHI"), ctx.device_ptr("RHS"))
            return ctx.download("RHS").reshape(-1)[:n].reshape(x.shape)

        r = b - apply(g)
        net = learned.make_net(ctx)
        ctx.upload("B", r)
        net.forward(ctx.device_ptr("B"), ctx.device_ptr("PHI"))
        y = ctx.download("PHI")
        Ay = apply(y)
        alpha = float(np.vdot(r, Ay) / np.vdot(Ay, Ay))
        ratio = float(np.linalg.norm(r - alpha * Ay) / np.linalg.norm(r))
        res = {}
        for name, nn in (("fp32_mg", None), ("fp32_mg_net", net)):
            ctx.upload("B", b)
            ctx.upload("PHI", g)
            ctx.sync()
            o = vfs.default_solve_opts(tol=1e-8, mg_fp32=1, use_guess=1)
            st = nn.poisson_solve(o) if nn is not None else ctx.poisson_solve(o)
            res[name] = (st.iterations, round(st.ms, 2), st.fallback, st.relres)
        print(f"[held-out] warm-start residual ratio {ratio:.3e} (alpha {alpha:.3e}); solves {res}", flush=True)
        assert ratio < 1.0
        assert res["fp32_mg_net"][2] == 0 and res["fp32_mg_net"][3] <= 1e-8
        net.close()
