"""Closed-form pins of the oracle's momentum path (no GPU).

On a Cartesian grid with spacings (hx, hy, hz) the metrics are constant, so
Formfunction_2's central second-order fluxes reduce to textbook difference
operators whose action on a Fourier mode is known exactly:

* a velocity component v = A sin(k x_a) carried by a uniform velocity U along
  the same axis a, in a fully periodic box, has the discrete rate
  dv/dt = lambda v with lambda = -i U sin(k h_a)/h_a + nu (2 cos(k h_a) - 2)/h_a^2
  acting on exp(i k x_a) (central advection, three-point Laplacian);
* one step of the reference's low-storage RK4 (stage weights 1/4, 1/3, 1/2, 1)
  on that linear problem multiplies exp(i k x) by 1 + z + z^2/2 + z^3/6 + z^4/24,
  z = lambda dt;
* a linear pressure p = G.x on an affine grid gives the face gradients
  dP = (face area vector) . G exactly;
* the periodic mean-flux forcing adds (mean_k_flux - inlet_flux) hx hy /
  (dt mean_k_area) to dP.z, so one forcing-only Euler step restores the
  target flux;
* Calc_Minimum_dt of a uniform flow is min_c h_c / |u_c|.

A sign or scale error on the convective or viscous flux, a wrong 1/ren, an
RK stage weight out of place, a transposed metric product or a dropped
forcing term each changes one of these by far more than round-off.
"""
import numpy as np
import pytest

import pyoracle as po

H = (0.1, 0.07, 0.13)


def cartesian_box(n=(16, 12, 16), h=H, per=(1, 1, 1), bctype=(100,) * 6, **kw):
    nx, ny, nz = n
    mx, my, mz = nx + 2, ny + 2, nz + 2
    coor = np.zeros((mz, my, mx, 3))
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coor[: nz + 1, : ny + 1, : nx + 1, 0] = I * h[0]
    coor[: nz + 1, : ny + 1, : nx + 1, 1] = J * h[1]
    coor[: nz + 1, : ny + 1, : nx + 1, 2] = K * h[2]
    p = po.make_params(mx, my, mz, per=per, bctype=bctype, **kw)
    return p, po.Oracle(p, coor)


def cell_coordinate(o, axis, h):
    """Centre coordinate along `axis` of every DA cell, ghosts included
    (cell n spans nodes n-1..n, so its centre is (n - 1/2) h)."""
    idx = np.arange(o.shape[2 - axis], dtype=float)
    shape = [1, 1, 1]
    shape[2 - axis] = -1
    return np.broadcast_to(((idx - 0.5) * h).reshape(shape), o.shape)


def area(h, axis):
    return h[(axis + 1) % 3] * h[(axis + 2) % 3]


def mode_case(a, b, U, nu, les=0, laplacian=1, nu_t=0.0, m=2, amp=0.3):
    """Component b = amp sin(k x_a) carried by U along axis a (b != a)."""
    n = (16, 12, 16)
    p, o = cartesian_box(n, H, les=les, laplacian=laplacian, ren=1.0 / nu, cs2=0.0)
    N = n[a]
    k = 2 * np.pi * m / (N * H[a])
    x = cell_coordinate(o, a, H[a])
    v = amp * np.sin(k * x)
    ucont = np.zeros(o.shape + (3,))
    ucont[..., a] = U * area(H, a)
    ucont[..., b] = v * area(H, b)
    o.set("ucont", ucont)
    o.contra2cart()
    o.set("ucat_old", o["ucat"])
    if les:
        o.set("nu_t", np.full(o.shape, nu_t))
    lam = -1j * U * np.sin(k * H[a]) / H[a] + (nu + nu_t) * (2 * np.cos(k * H[a]) - 2) / H[a] ** 2
    return p, o, k, x, lam


AXES = [(0, 1), (2, 0), (1, 2), (0, 2)]
INNER = (slice(1, -1),) * 3


@pytest.mark.parametrize("a,b", AXES)
@pytest.mark.parametrize("laplacian", [1, 0])
def test_formfunction_fourier_mode(a, b, laplacian):
    """rhs_b / area_b = A Im(lambda exp(i k x)) at every cell; rhs_a = 0."""
    U, nu, amp = 0.7, 0.05, 0.3
    p, o, k, x, lam = mode_case(a, b, U, nu, laplacian=laplacian, amp=amp)
    o.set("rhs", np.zeros(o.shape + (3,)))
    o.formfunction2(1.0)
    want = amp * np.imag(lam * np.exp(1j * k * x))
    got = o["rhs"][..., b] / area(H, b)
    np.testing.assert_allclose(got[INNER], want[INNER], rtol=0, atol=1e-12 * np.abs(want).max())
    assert np.abs(o["rhs"][..., a][INNER]).max() < 1e-12 * np.abs(want).max()
    # the advection part alone (nu -> 0) is odd in U: the sign of the convective flux
    p2, o2, _, _, lam2 = mode_case(a, b, -U, 1e-300, laplacian=laplacian, amp=amp)
    o2.set("rhs", np.zeros(o2.shape + (3,)))
    o2.formfunction2(1.0)
    want2 = amp * (U * np.sin(k * H[a]) / H[a]) * np.cos(k * x)
    np.testing.assert_allclose(o2["rhs"][..., b][INNER] / area(H, b), want2[INNER], rtol=0,
                               atol=1e-12 * np.abs(want2).max())


def test_formfunction_eddy_viscosity_scale():
    """With les on, a uniform nu_t adds to 1/ren in the viscous flux."""
    U, nu, nu_t, amp = 0.4, 0.02, 0.03, 0.5
    p, o, k, x, lam = mode_case(0, 2, U, nu, les=1, nu_t=nu_t, amp=amp)
    o.set("rhs", np.zeros(o.shape + (3,)))
    o.formfunction2(1.0)
    want = amp * np.imag(lam * np.exp(1j * k * x))
    np.testing.assert_allclose(o["rhs"][..., 2][INNER] / area(H, 2), want[INNER], rtol=0,
                               atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("a,b", AXES[:3])
def test_runge_kutta_amplification(a, b):
    """One RK4 step multiplies exp(i k x) by the degree-4 Taylor polynomial of
    exp(z), z = lambda dt (the stage weights 1/4, 1/3, 1/2, 1 give exactly
    that polynomial for a linear operator); the carrier component stays U."""
    U, nu, amp, dt = 0.7, 0.05, 0.3, 0.1
    p, o, k, x, lam = mode_case(a, b, U, nu, amp=amp)
    p.dt, p.st = dt, 1.0
    o.set("dp", np.zeros(o.shape + (3,)))
    o.set("nu_t", np.zeros(o.shape))
    o.runge_kutta()
    z = lam * dt
    assert abs(z) > 0.4  # every power of z matters at this step size
    g = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    want = amp * np.imag(g * np.exp(1j * k * x))
    got = o["ucont"][..., b] / area(H, b)
    np.testing.assert_allclose(got[INNER], want[INNER], rtol=0, atol=1e-12 * amp)
    np.testing.assert_allclose(o["ucont"][..., a][INNER], U * area(H, a), rtol=1e-13)
    # an explicit Euler step (a different polynomial) is far away
    euler = amp * np.imag((1 + z) * np.exp(1j * k * x))
    assert np.abs(got[INNER] - euler[INNER]).max() > 1e-3 * amp


def affine_box(n=(7, 6, 8)):
    """Nodes x = A (i, j, k) with a full (skewed) matrix A: every metric is an
    exact constant on this grid."""
    nx, ny, nz = n
    mx, my, mz = nx + 2, ny + 2, nz + 2
    A = np.array([[0.11, 0.02, -0.015], [0.01, 0.09, 0.02], [-0.012, 0.018, 0.13]])
    K, J, I = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    xi = np.stack([I, J, K], -1).astype(float)
    coor = np.zeros((mz, my, mx, 3))
    coor[: nz + 1, : ny + 1, : nx + 1] = (xi @ A.T)[: nz + 1, : ny + 1, : nx + 1]
    p = po.make_params(mx, my, mz, per=(0, 0, 0), bctype=(1, 1, 1, 1, 1, 1))
    o = po.Oracle(p, coor)
    cent = (xi - 0.5) @ A.T  # centres of every DA cell, ghosts extrapolated
    return p, o, cent


def test_pressure_gradient_linear_pressure():
    """p = G.x: dP.x = icsi.G, dP.y = jeta.G, dP.z = kzet.G on every interior face,
    zero on the boundary faces."""
    p, o, cent = affine_box()
    G = np.array([0.8, -1.3, 0.45])
    o.set("p", cent @ G)
    o.pressure_gradient()
    dp = o["dp"]
    mz, my, mx = o.shape
    for c, f, sl in ((0, "icsi", (slice(1, -1), slice(1, -1), slice(1, mx - 2))),
                     (1, "jeta", (slice(1, -1), slice(1, my - 2), slice(1, -1))),
                     (2, "kzet", (slice(1, mz - 2), slice(1, -1), slice(1, -1)))):
        want = o.m[f][sl] @ G
        np.testing.assert_allclose(dp[sl + (c,)], want, rtol=1e-12, atol=1e-14)
    assert np.all(dp[:, :, mx - 2, 0] == 0) and np.all(dp[:, my - 2, :, 1] == 0) and np.all(dp[mz - 2, :, :, 2] == 0)


def test_mean_flux_forcing():
    """Periodic-k channel on an orthogonal grid with stretched y: Calc_k_Flux and
    mean_k_area of a uniform stream, the forcing term of dP.z, and one
    forcing-only Euler step landing on the target flux."""
    nx, ny, nz = 6, 8, 5
    hx, hz = 0.2, 0.15
    yj = np.tanh(2.0 * (2 * np.arange(ny + 1) / ny - 1)) / np.tanh(2.0)
    mx, my, mz = nx + 2, ny + 2, nz + 2
    coor = np.zeros((mz, my, mx, 3))
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coor[: nz + 1, : ny + 1, : nx + 1, 0] = I * hx
    coor[: nz + 1, : ny + 1, : nx + 1, 1] = yj[J]
    coor[: nz + 1, : ny + 1, : nx + 1, 2] = K * hz
    dt, W, target = 0.01, 1.3, 2.0
    p = po.make_params(mx, my, mz, per=(1, 0, 1), bctype=(100, 100, 1, 1, 100, 100), dt=dt)
    o = po.Oracle(p, coor)
    hy = np.diff(yj)
    A = nx * hx * (yj[-1] - yj[0])
    assert o.mean_k_area() == pytest.approx(A, rel=1e-13)
    ucont = np.zeros(o.shape + (3,))
    ucont[..., 2] = W * hx * np.concatenate([[0], hy, [0]])[None, :, None]
    o.set("ucont", ucont)
    flux = o.calc_k_flux()
    assert flux == pytest.approx(W * A, rel=1e-13)
    p.inlet_flux = target
    o.set("p", np.zeros(o.shape))
    o.pressure_gradient()
    dpz = o["dp"][1:-1, 1:-1, 1:-1, 2]
    want = (flux - target) * hx * hy[None, :, None] / (dt * A)
    np.testing.assert_allclose(dpz, np.broadcast_to(want, dpz.shape), rtol=1e-12)
    assert np.abs(o["dp"][..., :2]).max() == 0
    o.set("ucont", ucont - dt * o["dp"])
    assert o.calc_k_flux() == pytest.approx(target, rel=1e-13)


def test_min_dt_uniform_flow():
    """Calc_Minimum_dt of a uniform flow on a Cartesian grid = min_c h_c/|u_c|."""
    p, o = cartesian_box((6, 5, 7), H, per=(0, 0, 0), bctype=(1,) * 6)
    for u in (np.array([0.3, -0.2, 1.1]), np.array([-2.0, 0.5, 0.25])):
        ucont = np.zeros(o.shape + (3,))
        for c in range(3):
            ucont[..., c] = u[c] * area(H, c)
        o.set("ucont", ucont)
        assert o.min_dt() == pytest.approx(min(H[c] / abs(u[c]) for c in range(3)), rel=1e-13)
