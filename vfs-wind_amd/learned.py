"""Training and use of the learned Poisson stage (the BF16 U-Net of unet.hip).

Training pairs are what the solver itself produces during a run: after every
fractional step of a seeded channel, the residual r = B - A phi_prev of the
solve's initial guess (the previous step's correction) and the correction
e = PHI - phi_prev the solve added are read back and random windows of them
become training samples -- the learned stage is applied to such residuals.
Training, validation and held-out evaluation use disjoint seeds (TRAIN_SEEDS,
VALIDATION_SEED, HELD_OUT_SEED and the benchmark's seed).  Training runs on
the device through the library's own forward/backward/Adam kernels on a small
context whose interior is the window size; the trained parameters are
resolution independent (a convolutional network) and are loaded into a
network on the full grid for inference.

    python vfs-wind_amd/learned.py --nx 256 --ny 192 --nz 256 --steps 20 --crops 10 \\
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


# Seeds of the training runs, of the validation run and of the held-out runs the
# tests and the benchmark evaluate on (cases.SEED): pairwise disjoint.
TRAIN_SEEDS = tuple(range(1001, 1011))
VALIDATION_SEED = 2001
HELD_OUT_SEED = 3001


def collect_pairs(nx, ny, nz, steps, crops, window, seeds, log=print, rng_seed=12345):
    """(r, e) windows (interior cells, float32) from solver runs: after every
    fractional step, r = B - A phi_prev is the residual of the initial guess
    the step's solve started from (the previous step's correction, zero at the
    first step) and e = PHI - phi_prev the correction that solve added; the
    learned stage is applied to exactly such residuals."""
    vfs = _vfs()
    wk, wj, wi = window
    R, E = [], []
    rng = np.random.default_rng(rng_seed)
    for seed in seeds:
        ctx = channel_context(nx, ny, nz, seed, vfs)
        opts = vfs.default_solve_opts(tol=1e-8)
        n = int(np.prod(ctx.field_shape("PHI")))
        for s in range(steps):
            prev = ctx.download("PHI")
            ctx.fractional_step(opts)
            b, phi = ctx.download("B"), ctx.download("PHI")
            ctx.upload("PHI", prev)
            ctx.poisson_apply_dev(ctx.device_ptr("PHI"), ctx.device_ptr("RHS"))
            r = (b - ctx.download("RHS").reshape(-1)[:n].reshape(b.shape))[1:-1, 1:-1, 1:-1]
            e = (phi - prev)[1:-1, 1:-1, 1:-1]
            ctx.upload("PHI", phi)
            for _ in range(crops):
                k0 = rng.integers(0, r.shape[0] - wk + 1)
                j0 = rng.integers(0, r.shape[1] - wj + 1)
                i0 = rng.integers(0, r.shape[2] - wi + 1)
                rr = r[k0:k0 + wk, j0:j0 + wj, i0:i0 + wi]
                ee = e[k0:k0 + wk, j0:j0 + wj, i0:i0 + wi]
                R.append(rr.astype(np.float32))
                E.append((ee - ee.mean()).astype(np.float32))
            if (s + 1) % 10 == 0:
                log(f"seed {seed}: {s + 1} steps, {len(R)} pairs")
        ctx.close()
    return np.stack(R), np.stack(E)


def _window_context(window, vfs):
    wk, wj, wi = window
    p = vfs.make_params(wi + 2, wj + 2, wk + 2, per=(0, 0, 0), bctype=(1, 1, 1, 1, 1, 1), les=0)
    return vfs.Context(p)


def validation_loss(net, ctx, B, PHI):
    """Mean normalised loss of the network on (b, phi) windows (no update)."""
    shape = ctx.field_shape("B")
    bf, pf = np.zeros(shape), np.zeros(shape)
    out = []
    for q in range(len(B)):
        bf[1:-1, 1:-1, 1:-1] = B[q]
        pf[1:-1, 1:-1, 1:-1] = PHI[q]
        ctx.upload("B", bf)
        ctx.upload("PHI", pf)
        out.append(net.gradients(ctx.device_ptr("B"), ctx.device_ptr("PHI")))
    return float(np.mean(out))


def train(B, PHI, *, base=16, levels=4, epochs=3, lr=1e-3, seed=20261202, log=print, val=None):
    """Train on (b, phi) windows; returns (params float32, out_scale, losses,
    validation losses); `val` = (B, PHI) windows of a validation run."""
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
    losses, val_losses = [], []
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
        if val is not None:
            val_losses.append(validation_loss(net, ctx, *val))
        log(f"epoch {ep}: mean loss {losses[-1]:.4e}"
            + (f", validation {val_losses[-1]:.4e}" if val is not None else "") + f" ({time.time() - t0:.1f} s)")
    params = net.params(0)
    net.close()
    ctx.close()
    return params, out_scale, losses, val_losses


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


def time_inference(nx, ny, nz, reps, log=print):
    """ms per forward pass of the shipped network (random weights when none are
    shipped) on the seeded channel's Poisson right-hand side; the BF16 MFMA
    operations of one pass (the convolutions as dense contractions)."""
    vfs = _vfs()
    ctx = channel_context(nx, ny, nz, HELD_OUT_SEED, vfs)
    ctx.fractional_step(vfs.default_solve_opts(tol=1e-8))
    net = make_net(ctx) or vfs.UNet(ctx, base=16, levels=4)
    b, y = ctx.device_ptr("B"), ctx.device_ptr("RHS")
    net.forward(b, y)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(b, y)
    ctx.sync()
    ms = (time.perf_counter() - t0) / reps * 1e3
    levels, base = net.levels, net.base
    pts = nx * ny * nz
    flop = 0.0
    for l in range(levels):  # 27 taps x cin x cout multiply-adds per point
        c_out = base << l
        c_in = 16 if l == 0 else base << (l - 1)
        flop += 2 * 27 * c_in * c_out * (pts >> (3 * l))
        if l < levels - 1:
            c_up = base << (levels - 1) if l == levels - 2 else base << (l + 1)
            flop += 2 * 27 * (c_up + c_out) * c_out * (pts >> (3 * l))
    log(f"forward pass {ms:.3f} ms, {flop / 1e12:.3f} TFLOP (BF16, first layer padded to 16 channels), "
        f"{flop / (ms * 1e-3) / 1e12:.1f} TFLOP/s")
    net.close()
    ctx.close()
    return ms, flop


def main(argv=None):
    if str(PKG.parent) not in sys.path:
        sys.path.insert(0, str(PKG.parent))
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--crops", type=int, default=10)
    ap.add_argument("--seeds", type=int, nargs="+", default=list(TRAIN_SEEDS))
    ap.add_argument("--val-steps", type=int, default=5)
    ap.add_argument("--window", type=int, nargs=3, default=[64, 48, 64], help="k j i window (each divisible by 8)")
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--base", type=int, default=16)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    ap.add_argument("--time-inference", type=int, default=0, metavar="REPS",
                    help="only time REPS forward passes on the nx x ny x nz channel")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.time_inference:
        time_inference(a.nx, a.ny, a.nz, a.time_inference, log)
        return
    held_out = {VALIDATION_SEED, HELD_OUT_SEED, _cases().SEED}
    if held_out & set(a.seeds):
        raise SystemExit(f"training seeds must not include the validation / held-out seeds {sorted(held_out)}")
    B, PHI = collect_pairs(a.nx, a.ny, a.nz, a.steps, a.crops, tuple(a.window), a.seeds, log)
    log(f"{len(B)} training pairs of {tuple(a.window)} cells")
    val = collect_pairs(a.nx, a.ny, a.nz, a.val_steps, a.crops, tuple(a.window), [VALIDATION_SEED], log, rng_seed=54321)
    params, out_scale, losses, val_losses = train(B, PHI, base=a.base, levels=a.levels, epochs=a.epochs, lr=a.lr, log=log,
                                                  val=val)
    info = (f"channel {a.nx}x{a.ny}x{a.nz}, training seeds {a.seeds}, validation seed {VALIDATION_SEED}, "
            f"{a.steps} steps x {a.crops} windows {tuple(a.window)} of (r, e) pairs, {a.epochs} epochs, lr {a.lr}, "
            f"losses {losses}, validation losses {val_losses}")
    save(a.out, params, out_scale, a.base, a.levels, info)
    log(f"saved {a.out}: {info}")


if __name__ == "__main__":
    main()
