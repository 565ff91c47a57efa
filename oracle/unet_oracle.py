"""CPU restatement of the learned Poisson stage's U-Net (forward and backward).

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP network's output
and gradients.  There is no reference counterpart (the learned stage is new);
the backward pass here is pinned by finite differences in FP64
(tests/test_unet_oracle_cpu.py).

Architecture (same as vfs-wind_amd/csrc/unet.hip): interior cells
(nk, nj, ni) channels-last; level l has base * 2^l channels; encoder: 3x3x3
conv + ReLU per level, 2x2x2 average pooling between levels; decoder:
nearest upsampling, concat [upsampled | encoder skip], 3x3x3 conv + ReLU;
output 1x1x1 conv.  Input: channel 0 = b / rms(b), 15 zero channels.  With
bf16=True the tensors are rounded to BF16 (nearest even) at the same points
the device stores them: weights, input, every conv/pool output, the masked
gradients fed to the backward convolutions.
"""
from __future__ import annotations

import numpy as np


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, np.uint64(0x7FC00000), r).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def _taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def _shift(x, dk, dj, di):
    """y[k, j, i] = x[k + dk, j + dj, i + di] with zero padding."""
    y = np.zeros_like(x)
    nk, nj, ni = x.shape[:3]
    ks, ke = max(0, -dk), min(nk, nk - dk)
    js, je = max(0, -dj), min(nj, nj - dj)
    is_, ie = max(0, -di), min(ni, ni - di)
    y[ks:ke, js:je, is_:ie] = x[ks + dk:ke + dk, js + dj:je + dj, is_ + di:ie + di]
    return y


def conv3(x, w, b=None):
    """x (nk, nj, ni, cin), w (27, cin, cout) -> (nk, nj, ni, cout)."""
    out = np.zeros(x.shape[:3] + (w.shape[2],), dtype=np.float64)
    for t, (dk, dj, di) in enumerate(_taps()):
        out += _shift(x, dk, dj, di).astype(np.float64) @ w[t].astype(np.float64)
    if b is not None:
        out += b
    return out


def conv3_data_grad(g, w):
    """dX[p][ci] = sum_tap sum_co g[p - tap][co] w[tap][ci][co]."""
    out = np.zeros(g.shape[:3] + (w.shape[1],), dtype=np.float64)
    for t, (dk, dj, di) in enumerate(_taps()):
        out += _shift(g, -dk, -dj, -di).astype(np.float64) @ w[t].astype(np.float64).T
    return out


def conv3_weight_grad(x, g):
    """dW[tap][ci][co] = sum_p x[p + tap][ci] g[p][co]."""
    cin, cout = x.shape[3], g.shape[3]
    dw = np.zeros((27, cin, cout))
    g2 = g.reshape(-1, cout).astype(np.float64)
    for t, (dk, dj, di) in enumerate(_taps()):
        dw[t] = _shift(x, dk, dj, di).reshape(-1, cin).astype(np.float64).T @ g2
    return dw


def pool(x):
    nk, nj, ni, c = x.shape
    return x.reshape(nk // 2, 2, nj // 2, 2, ni // 2, 2, c).mean(axis=(1, 3, 5))


def upsample(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


class UNetOracle:
    """Parameters in the device's flattened order: for each encoder level
    (0..L-1) then each decoder level (0..L-2): W (27, cin, cout), bias (cout);
    then the output weights (base) and bias (1)."""

    def __init__(self, base: int, levels: int):
        self.base, self.L = base, levels
        self.enc = [(16 if l == 0 else base << (l - 1), base << l) for l in range(levels)]
        self.dec = {}
        for l in range(levels - 2, -1, -1):
            cu = (base << (levels - 1)) if l == levels - 2 else (base << (l + 1))
            self.dec[l] = (cu + (base << l), base << l)

    def shapes(self):
        s = []
        for cin, cout in self.enc:
            s += [(27, cin, cout), (cout,)]
        for l in range(self.L - 1):
            cin, cout = self.dec[l]
            s += [(27, cin, cout), (cout,)]
        s += [(self.base,), (1,)]
        return s

    def unflatten(self, flat):
        out, off = [], 0
        for sh in self.shapes():
            n = int(np.prod(sh))
            out.append(np.asarray(flat[off:off + n], dtype=np.float64).reshape(sh))
            off += n
        return out

    def flatten(self, parts):
        return np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts])

    def forward(self, params, b, bf16=True):
        """b: interior right-hand side (nk, nj, ni). Returns (out, cache)."""
        r = lambda a: to_bf16(a).astype(np.float64) if bf16 else np.asarray(a, dtype=np.float64)
        P = self.unflatten(params)
        L = self.L
        ew = [(r(P[2 * l]), P[2 * l + 1]) for l in range(L)]
        dw = {}
        for l in range(L - 1):
            dw[l] = (r(P[2 * L + 2 * l]), P[2 * L + 2 * l + 1])
        wo, bo = P[-2], P[-1]
        rms = np.sqrt(np.mean(b * b))
        x0 = np.zeros(b.shape + (16,))
        x0[..., 0] = b / rms
        x0 = r(x0)
        e, pl, cat, d = {}, {}, {}, {}
        e[0] = r(np.maximum(conv3(x0, *ew[0]), 0))
        for l in range(1, L):
            pl[l] = r(pool(e[l - 1]))
            e[l] = r(np.maximum(conv3(pl[l], *ew[l]), 0))
        for l in range(L - 2, -1, -1):
            coarse = e[L - 1] if l == L - 2 else d[l + 1]
            cat[l] = np.concatenate([upsample(coarse), e[l]], axis=3)
            d[l] = r(np.maximum(conv3(cat[l], *dw[l]), 0))
        out = d[0] @ wo + bo[0]
        cache = dict(x0=x0, e=e, pl=pl, cat=cat, d=d, ew=ew, dw=dw, wo=wo, rms=rms)
        return out, cache

    def gradients(self, params, b, phi, out_scale, bf16=True):
        """Loss = mean((out - phi / (rms(b) out_scale))^2) and its parameter gradient."""
        r = lambda a: to_bf16(a).astype(np.float64) if bf16 else np.asarray(a, dtype=np.float64)
        out, c = self.forward(params, b, bf16)
        L = self.L
        t = phi / (c["rms"] * out_scale)
        n = out.size
        loss = float(np.mean((out - t) ** 2))
        gout = 2.0 * (out - t) / n
        grads_enc = [None] * L
        grads_dec = {}
        gwo = np.einsum("kji,kjic->c", gout, c["d"][0])
        gbo = np.array([gout.sum()])
        gd = {0: gout[..., None] * c["wo"]}
        ge = {l: np.zeros_like(c["e"][l]) for l in range(L)}
        for l in range(0, L - 1):
            gm = r(np.where(c["d"][l] > 0, gd[l], 0.0))
            W, _ = c["dw"][l]
            grads_dec[l] = (conv3_weight_grad(c["cat"][l], gm), gm.sum(axis=(0, 1, 2)))
            gcat = conv3_data_grad(gm, W)
            cs = c["e"][l].shape[3]
            cu = gcat.shape[3] - cs
            ge[l] += gcat[..., cu:]
            gc = gcat[..., :cu]
            nk, nj, ni = gc.shape[:3]
            gcoarse = gc.reshape(nk // 2, 2, nj // 2, 2, ni // 2, 2, cu).sum(axis=(1, 3, 5))
            if l == L - 2:
                ge[L - 1] = ge[L - 1] + gcoarse
            else:
                gd[l + 1] = gcoarse
        for l in range(L - 1, -1, -1):
            gm = r(np.where(c["e"][l] > 0, ge[l], 0.0))
            W, _ = c["ew"][l]
            X = c["x0"] if l == 0 else c["pl"][l]
            grads_enc[l] = (conv3_weight_grad(X, gm), gm.sum(axis=(0, 1, 2)))
            if l == 0:
                break
            gp = conv3_data_grad(gm, W)
            ge[l - 1] = ge[l - 1] + upsample(gp) * 0.125
        parts = []
        for l in range(L):
            parts += list(grads_enc[l])
        for l in range(L - 1):
            parts += list(grads_dec[l])
        parts += [gwo, gbo]
        return loss, self.flatten(parts), out
