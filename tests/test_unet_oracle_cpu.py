"""Pins of the U-Net oracle (CPU): the convolution against a direct loop
restatement of 3D cross-correlation with zero padding, and the analytic
backward pass against central finite differences in FP64."""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import unet_oracle as uo  # noqa: E402


def test_conv3_matches_direct_correlation():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 5, 6, 3))
    w = rng.standard_normal((27, 3, 2))
    y = uo.conv3(x, w)
    # y[k,j,i,o] = sum over the 3x3x3 window (tap t = 9(dk+1) + 3(dj+1) + di+1)
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    ref = np.zeros_like(y)
    for k in range(4):
        for j in range(5):
            for i in range(6):
                win = xp[k:k + 3, j:j + 3, i:i + 3].reshape(27, 3)
                ref[k, j, i] = np.einsum("tc,tco->o", win, w)
    assert np.allclose(y, ref, atol=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.001953125, 1.005859375, -3.140625, 0.0], dtype=np.float32)
    r = uo.to_bf16(v)
    # the BF16 spacing at 1 is 2^-7: 1 + 2^-8 is a tie and goes to the even 1.0,
    # 1 + 2^-9 rounds down, 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert list(r) == [1.0, 1.0, 1.0, 1.0078125, -3.140625, 0.0]


def test_backward_matches_finite_differences():
    net = uo.UNetOracle(base=16, levels=2)
    rng = np.random.default_rng(3)
    shapes = net.shapes()
    params = np.concatenate([0.2 * rng.standard_normal(int(np.prod(s))) for s in shapes])
    b = rng.standard_normal((4, 4, 4))
    phi = rng.standard_normal((4, 4, 4))
    loss, g, _ = net.gradients(params, b, phi, out_scale=1.0, bf16=False)
    idx = rng.choice(params.size, 40, replace=False)
    h = 1e-6
    for q in idx:
        p1, p2 = params.copy(), params.copy()
        p1[q] += h
        p2[q] -= h
        l1, _, _ = net.gradients(p1, b, phi, 1.0, bf16=False)
        l2, _, _ = net.gradients(p2, b, phi, 1.0, bf16=False)
        fd = (l1 - l2) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (q, fd, g[q])
