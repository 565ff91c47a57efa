"""Training and use of the learned Poisson stage (the BF16 U-Net of unet.hip).

Training pairs are what the solver itself produces during a run: after every
fractional step of the seeded channel, the Poisson right-hand side B and the
converged correction PHI are read back and random windows of them become
training samples.  Training runs on the device through the library's own
forward/backward/Adam kernels on a small context whose interior is the window
size; the trained parameters are resolution independent (a convolutional
network) and are loaded into a network on the full grid for inference.

    python vfs-wind_amd/learned.py --nx 256 --ny 192 --nz 256 --steps 100 --crops 20 \\
        --epochs 3 --out vfs-wind_amd/weights/unet_channel.npz
"""
from __future__ import annotations

import argparse
import importlib
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
DEFAULT_WEIGHTS = PKG / "weights" / "unet_channel.npz"


def _vfs():
    return importlib.import_module("vfs-wind_amd")


def _cases():
    return importlib.import_module("vfs-wind_amd.cases")


def channel_context(nx, ny, nz, seed, vfs=None):
    """Context on the seeded channel with dt = 0.3 dt_CFL1, ready to step."""
    vfs = vfs or _vfs()
    cases = _cases()
    c = cases.channel(nx, ny, nz, seed=seed)
    p = cases.params_for(vfs, c, dt=1e-3)
    ctx = vfs.Context(p)
    ctx.upload("COOR", c.coor)
    ctx.form_metrics()
    ctx.mean_k_area()
    ctx.upload("UCONT", c.ucont)
    ctx.contra2cart()
    ctx.upload("UCAT_OLD", ctx.download("UCAT"))
    p.dt = 0.3 * ctx.min_dt()
    ctx.set_params(p)
    return ctx


def collect_pairs(nx, ny, nz, steps, crops, window, seeds, log=print):
    """(b, phi) windows (interior cells, float32) from solver runs."""
    vfs = _vfs()
    wk, wj, wi = window
    B, PHI = [], []
    rng = np.random.default_rng(12345)
    for seed in seeds:
        ctx = channel_context(nx, ny, nz, seed, vfs)
        opts = vfs.default_solve_opts(tol=1e-8)
        for s in range(steps):
            ctx.fractional_step(opts)
            b = ctx.download("B")[1:-1, 1:-1, 1:-1]
            phi = ctx.download("PHI")[1:-1, 1:-1, 1:-1]
            for _ in range(crops):
                k0 = rng.integers(0, b.shape[0] - wk + 1)
                j0 = rng.integers(0, b.shape[1] - wj + 1)
                i0 = rng.integers(0, b.shape[2] - wi + 1)
                bb = b[k0:k0 + wk, j0:j0 + wj, i0:i0 + wi]
                pp = phi[k0:k0 + wk, j0:j0 + wj, i0:i0 + wi]
                B.append(bb.astype(np.float32))
                PHI.append((pp - pp.mean()).astype(np.float32))
            if (s + 1) % 10 == 0:
                log(f"seed {seed}: {s + 1} steps, {len(B)} pairs")
        ctx.close()
    return np.stack(B), np.stack(PHI)


def _window_context(window, vfs):
    wk, wj, wi = window
    p = vfs.make_params(wi + 2, wj + 2, wk + 2, per=(0, 0, 0), bctype=(1, 1, 1, 1, 1, 1), les=0)
    return vfs.Context(p)


def train(B, PHI, *, base=16, levels=4, epochs=3, lr=1e-3, seed=20261202, log=print):
    """Train on (b, phi) windows; returns (params float32, out_scale, losses)."""
    vfs = _vfs()
    window = B.shape[1:]
    ctx = _window_context(window, vfs)
    net = vfs.UNet(ctx, base=base, levels=levels, seed=seed)
    rb = np.sqrt((B.astype(np.float64) ** 2).mean(axis=(1, 2, 3)))
    rp = np.sqrt((PHI.astype(np.float64) ** 2).mean(axis=(1, 2, 3)))
    out_scale = float(np.median(rp / np.maximum(rb, 1e-300)))
    net.set_output_scale(out_scale)
    shape = ctx.field_shape("B")
    bf = np.zeros(shape)
    pf = np.zeros(shape)
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.time()
    for ep in range(epochs):
        order = rng.permutation(len(B))
        ep_loss = []
        for q in order:
            bf[1:-1, 1:-1, 1:-1] = B[q]
            pf[1:-1, 1:-1, 1:-1] = PHI[q]
            ctx.upload("B", bf)
            ctx.upload("PHI", pf)
            ep_loss.append(net.train_step(ctx.device_ptr("B"), ctx.device_ptr("PHI"), lr))
        losses.append(float(np.mean(ep_loss)))
        log(f"epoch {ep}: mean loss {losses[-1]:.4e} ({time.time() - t0:.1f} s)")
    params = net.params(0)
    net.close()
    ctx.close()
    return params, out_scale, losses


def save(path, params, out_scale, base, levels, info=""):
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(path, params=params.astype(np.float32), out_scale=np.float64(out_scale), base=np.int32(base),
                        levels=np.int32(levels), info=np.array(info))


def load(path=DEFAULT_WEIGHTS):
    """(params, out_scale, base, levels) or None when no trained weights exist."""
    p = Path(path)
    if not p.exists():
        return None
    z = np.load(p, allow_pickle=False)
    return z["params"], float(z["out_scale"]), int(z["base"]), int(z["levels"])


def make_net(ctx, path=DEFAULT_WEIGHTS):
    """A UNet on ctx's grid with the trained parameters (None when absent)."""
    w = load(path)
    if w is None:
        return None
    params, out_scale, base, levels = w
    vfs = _vfs()
    net = vfs.UNet(ctx, base=base, levels=levels)
    net.set_params(params)
    net.set_output_scale(out_scale)
    return net


def main(argv=None):
    if str(PKG.parent) not in sys.path:
        sys.path.insert(0, str(PKG.parent))
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--crops", type=int, default=20)
    ap.add_argument("--seeds", type=int, nargs="+", default=[20261202])
    ap.add_argument("--window", type=int, nargs=3, default=[64, 48, 64], help="k j i window (each divisible by 8)")
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--base", type=int, default=16)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    B, PHI = collect_pairs(a.nx, a.ny, a.nz, a.steps, a.crops, tuple(a.window), a.seeds, log)
    log(f"{len(B)} training pairs of {tuple(a.window)} cells")
    params, out_scale, losses = train(B, PHI, base=a.base, levels=a.levels, epochs=a.epochs, lr=a.lr, log=log)
    info = (f"channel {a.nx}x{a.ny}x{a.nz}, seeds {a.seeds}, {a.steps} steps x {a.crops} windows "
            f"{tuple(a.window)}, {a.epochs} epochs, lr {a.lr}, losses {losses}")
    save(a.out, params, out_scale, a.base, a.levels, info)
    log(f"saved {a.out}: {info}")


if __name__ == "__main__":
    main()
