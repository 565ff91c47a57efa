"""Pins of the FP64 CPU oracle by closed forms, invariants and library
routines (no GPU).  Each check is chosen so that a plausible restatement
mistake (a dropped term, a wrong sign or index, a transposed metric) fails.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import cases, channel_pair


def cartesian(nx, ny, nz, h=(0.1, 0.2, 0.05), per=(0, 0, 0), bctype=(1, 1, 1, 1, 1, 1), **kw):
    mx, my, mz = nx + 2, ny + 2, nz + 2
    coor = np.zeros((mz, my, mx, 3))
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coor[: nz + 1, : ny + 1, : nx + 1, 0] = I * h[0]
    coor[: nz + 1, : ny + 1, : nx + 1, 1] = J * h[1]
    coor[: nz + 1, : ny + 1, : nx + 1, 2] = K * h[2]
    p = po.make_params(mx, my, mz, per=per, bctype=bctype, **kw)
    return p, po.Oracle(p, coor)


def test_metrics_closed_form_cartesian():
    hx, hy, hz = 0.1, 0.2, 0.05
    p, o = cartesian(6, 5, 7, (hx, hy, hz))
    s = (slice(1, -1),) * 3
    csi, eta, zet = o.m["csi"][s], o.m["eta"][s], o.m["zet"][s]
    np.testing.assert_allclose(csi[..., 0], hy * hz, rtol=1e-14)
    np.testing.assert_allclose(eta[..., 1], hx * hz, rtol=1e-14)
    np.testing.assert_allclose(zet[..., 2], hx * hy, rtol=1e-14)
    for v, d in ((csi, (1, 2)), (eta, (0, 2)), (zet, (0, 1))):
        assert np.abs(v[..., list(d)]).max() < 1e-15
    np.testing.assert_allclose(o.m["aj"][s], 1 / (hx * hy * hz), rtol=1e-12)
    for f, ax, area in (("icsi", 0, hy * hz), ("jeta", 1, hx * hz), ("kzet", 2, hx * hy)):
        np.testing.assert_allclose(o.m[f][s][..., ax], area, rtol=1e-12)
        np.testing.assert_allclose(o.m[f[0] + "aj"][s], 1 / (hx * hy * hz), rtol=1e-12)


def test_metrics_closed_form_sheared():
    """y = eta * hy * (1 + a * xi): the cell-averaged edge vectors of cell
    (i, j, k) are exact, giving csi = (hy hz (1 + a (i - 1/2)), 0, 0),
    eta = (-hz hy a (j - 1/2), hz hx, 0), zet = (0, 0, hx hy (1 + a (i - 1/2)))
    and aj = 1 / (hx hy hz (1 + a (i - 1/2))).  The node-plane formulas of
    the non-NEWMETRIC build would put i (not i - 1/2) here, so this separates
    the two; face metrics are the means of the two cells, ghosts copy their
    neighbours in the order i-, i+, j-, j+, k-, k+."""
    hx, hy, hz, a = 0.1, 0.2, 0.05, 0.15
    nx, ny, nz = 6, 5, 7
    mx, my, mz = nx + 2, ny + 2, nz + 2
    coor = np.zeros((mz, my, mx, 3))
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coor[: nz + 1, : ny + 1, : nx + 1, 0] = I * hx
    coor[: nz + 1, : ny + 1, : nx + 1, 1] = J * hy * (1 + a * I)
    coor[: nz + 1, : ny + 1, : nx + 1, 2] = K * hz
    p = po.make_params(mx, my, mz, per=(0, 0, 0), bctype=(1, 1, 1, 1, 1, 1))
    o = po.Oracle(p, coor)
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    ic = np.clip(i, 1, mx - 2).astype(float)
    jc = np.clip(j, 1, my - 2).astype(float)
    s = 1 + a * (ic - 0.5)
    exp = {
        "csi": np.stack([hy * hz * s, 0 * s, 0 * s], -1),
        "eta": np.stack([-hz * hy * a * (jc - 0.5), hz * hx + 0 * s, 0 * s], -1),
        "zet": np.stack([0 * s, 0 * s, hx * hy * s], -1),
    }
    for f, e in exp.items():  # every cell, ghosts included (mirrored)
        np.testing.assert_allclose(o.m[f], e, rtol=1e-13, atol=1e-16, err_msg=f)
    np.testing.assert_allclose(o.m["aj"], 1 / (hx * hy * hz * s), rtol=1e-13)
    si = slice(1, -1)
    # I faces 0..mx-2: mean of cells i and i+1
    fi = 1 + a * (0.5 * (np.clip(np.arange(mx - 1), 1, mx - 2) + np.clip(np.arange(1, mx), 1, mx - 2)) - 0.5)
    np.testing.assert_allclose(o.m["icsi"][si, si, : mx - 1, 0], np.broadcast_to(hy * hz * fi, (nz, ny, mx - 1)),
                               rtol=1e-13)
    vol = 0.5 * (1 / o.m["aj"][si, si, : mx - 1] + 1 / o.m["aj"][si, si, 1:])
    np.testing.assert_allclose(o.m["iaj"][si, si, : mx - 1], 1 / vol, rtol=1e-13)
    # J faces 0..my-2: eta.x is the mean of the two cells' -hz hy a (j - 1/2)
    fj = 0.5 * (np.clip(np.arange(my - 1), 1, my - 2) + np.clip(np.arange(1, my), 1, my - 2)) - 0.5
    np.testing.assert_allclose(o.m["jeta"][si, : my - 1, si, 0],
                               np.broadcast_to((-hz * hy * a * fj)[None, :, None], (nz, my - 1, nx)), rtol=1e-13)
    # K faces: zet is k independent
    np.testing.assert_allclose(o.m["kzet"][: mz - 1, si, si], o.m["zet"][: mz - 1, si, si], rtol=1e-13, atol=1e-16)


def test_contra2cart_recovers_uniform_flow():
    """Ucont = u . (node-plane area vector) on any grid gives back u exactly."""
    u = np.array([0.3, -0.2, 1.1])
    # Contra2Cart averages the fluxes of faces i-1 and i, which equal
    # u . csi of the cell only on an affine grid: use the Cartesian box
    p, o = cartesian(5, 4, 6, per=(1, 0, 1), bctype=(100, 100, 1, 1, 100, 100))
    ucont = np.zeros(o["ucont"].shape)
    ucont[..., 0] = o.m["csi"] @ u
    ucont[..., 1] = o.m["eta"] @ u
    ucont[..., 2] = o.m["zet"] @ u
    o.set("ucont", ucont)
    o.contra2cart()
    np.testing.assert_allclose(o["ucat"][1:-1, 1:-1, 1:-1], np.broadcast_to(u, o["ucat"][1:-1, 1:-1, 1:-1].shape),
                               rtol=1e-13, atol=1e-15)
    # no-slip ghost rows mirror the first interior cell with a sign flip
    np.testing.assert_allclose(o["ucat"][1:-1, 0, 1:-1], -o["ucat"][1:-1, 1, 1:-1], rtol=0, atol=0)


@pytest.mark.parametrize("case", ["channel", "box"])
def test_contra2cart_idempotent(case):
    """A second Contra2Cart on an unchanged Ucont rewrites the same Ucat
    (ghosts included): the device RungeKutta relies on this to drop the
    stage-start call that repeats the previous stage's closing one."""
    c = cases.channel(24, 16, 20) if case == "channel" else cases.box_flow(24, 20, 32)
    p = po.make_params(*c.da, per=c.per, bctype=c.bctype, inlet_flux=c.inlet_flux)
    o = po.Oracle(p, c.coor)
    o.set("ucont", c.ucont)
    o.contra2cart()
    o.set("ucat_old", 0.9 * o["ucat"])
    rng = np.random.default_rng(1)
    o.set("ucont", c.ucont * (1 + 0.01 * rng.standard_normal(c.ucont.shape)))
    o.contra2cart()
    u1 = o["ucat"].copy()
    o.contra2cart()
    assert np.array_equal(o["ucat"], u1)


def test_eddy_viscosity_linear_field():
    """u = A x on a Cartesian grid: central differences are exact, so
    nu_t = Cs^2 * V^(2/3) * sqrt(2 S:S) with S = sym(A)."""
    hx, hy, hz = 0.1, 0.2, 0.05
    p, o = cartesian(6, 6, 6, (hx, hy, hz), les=1, cs2=0.01)
    A = np.array([[0.3, 0.1, -0.2], [0.05, -0.4, 0.2], [0.7, 0.0, 0.1]])
    cent = o.m["cent"]
    ucat = np.einsum("ab,kjib->kjia", A, cent)
    # ghost cells: extrapolate the linear field so one-sided and central agree
    X = np.zeros(cent.shape)
    K, J, I = np.meshgrid(np.arange(o.shape[0]), np.arange(o.shape[1]), np.arange(o.shape[2]), indexing="ij")
    X[..., 0], X[..., 1], X[..., 2] = (I - 0.5) * hx, (J - 0.5) * hy, (K - 0.5) * hz
    ucat = np.einsum("ab,kjib->kjia", A, X)
    o.set("ucat", ucat)
    o.eddy_viscosity()
    S = 0.5 * (A + A.T)
    expect = 0.01 * (hx * hy * hz) ** (2 / 3) * np.sqrt(2 * np.sum(S * S))
    np.testing.assert_allclose(o["nu_t"][1:-1, 1:-1, 1:-1], expect, rtol=1e-12)


def test_formfunction_uniform_flow_is_steady():
    """A uniform flow on the periodic Cartesian box has zero convective and
    viscous divergence: Formfunction_2 returns zero (to round-off)."""
    p, o = cartesian(6, 5, 6, per=(1, 1, 1), bctype=(100,) * 6, les=1, cs2=0.01)
    u = np.array([0.3, -0.2, 1.1])
    ucont = np.zeros(o["ucont"].shape)
    ucont[..., 0] = o.m["csi"] @ u
    ucont[..., 1] = o.m["eta"] @ u
    ucont[..., 2] = o.m["zet"] @ u
    o.set("ucont", ucont)
    o.contra2cart()
    o.eddy_viscosity()
    for so in (1, 0):
        p.second_order = so
        o.set("rhs", np.zeros(o["rhs"].shape))
        o.formfunction2(1.0)
        assert np.abs(o["rhs"]).max() < 1e-14


def test_formfunction_conservation_periodic_box():
    """Flux form: on a fully periodic Cartesian box every face flux enters two
    cells with opposite signs, so the RHS summed over all cells vanishes for
    any velocity field (convective + viscous, both schemes)."""
    p, o = cartesian(6, 5, 7, per=(1, 1, 1), bctype=(100,) * 6, les=1, cs2=0.01, ren=50.0)
    rng = np.random.default_rng(1)
    ucont = rng.standard_normal(o["ucont"].shape)
    for so in (1, 0):
        p.second_order = so
        o.set("ucont", ucont)
        o.contra2cart()
        o.eddy_viscosity()
        o.set("rhs", np.zeros(o["rhs"].shape))
        o.formfunction2(1.0)
        r = o["rhs"][1:-1, 1:-1, 1:-1]
        for comp in range(3):
            assert abs(r[..., comp].sum()) < 1e-12 * np.abs(r[..., comp]).sum()


def test_poisson_operator_reduces_to_7_point():
    """On a Cartesian grid the 19-point operator is the 7-point Laplacian
    with face coefficients area/spacing, Neumann at walls."""
    hx, hy, hz = 0.1, 0.2, 0.05
    p, o = cartesian(5, 4, 6, (hx, hy, hz), per=(1, 0, 1), bctype=(100, 100, 1, 1, 100, 100))
    rng = np.random.default_rng(0)
    x = np.zeros(o.shape)
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal((6, 4, 5))
    y = o.poisson_apply(x)
    xi = x[1:-1, 1:-1, 1:-1]
    gx, gy, gz = hy * hz / hx, hx * hz / hy, hx * hy / hz
    lap = gx * (np.roll(xi, -1, 2) - 2 * xi + np.roll(xi, 1, 2)) + gz * (np.roll(xi, -1, 0) - 2 * xi + np.roll(xi, 1, 0))
    yy = np.zeros_like(xi)
    yy[:, 1:, :] += gy * (xi[:, :-1, :] - xi[:, 1:, :])
    yy[:, :-1, :] += gy * (xi[:, 1:, :] - xi[:, :-1, :])
    np.testing.assert_allclose(y[1:-1, 1:-1, 1:-1], lap + yy, rtol=1e-12, atol=1e-13)


def test_poisson_constant_nullspace_and_solution():
    """A.1 = 0 on the curvilinear channel, the RHS is made consistent (sum b = 0),
    and on a tiny grid the solver agrees with a dense least-squares solve."""
    c = cases.channel(6, 4, 6)
    p = po.make_params(*c.da, per=c.per, bctype=c.bctype, dt=0.01)
    o = po.Oracle(p, c.coor)
    one = np.zeros(o.shape)
    one[1:-1, 1:-1, 1:-1] = 1.0
    assert np.abs(o.poisson_apply(one)).max() < 1e-13
    o.set("ucont", c.ucont)
    o.poisson_rhs()
    b = o["b"]
    assert abs(b.sum()) < 1e-12 * np.abs(b).sum()
    # dense matrix by probing the operator with unit vectors
    idx = [(k, j, i) for k in range(1, o.shape[0] - 1) for j in range(1, o.shape[1] - 1) for i in range(1, o.shape[2] - 1)]
    n = len(idx)
    A = np.zeros((n, n))
    for col, (k, j, i) in enumerate(idx):
        e = np.zeros(o.shape)
        e[k, j, i] = 1.0
        y = o.poisson_apply(e)
        A[:, col] = [y[q] for q in idx]
    bb = np.array([b[q] for q in idx])
    xs = np.linalg.lstsq(A, bb, rcond=None)[0]
    xs -= xs.mean()
    o.set("phi", np.zeros(o.shape))
    it, rel = o.poisson_solve(tol=1e-12, levels=1)
    assert rel <= 1e-12
    phi = np.array([o["phi"][q] for q in idx])
    np.testing.assert_allclose(phi, xs, rtol=0, atol=1e-9 * np.abs(xs).max())


def test_projection_divergence_matches_residual():
    """After the solve and the projection, max|div U| * aj is the residual
    of the Poisson equation times dt: it falls with the solver tolerance."""
    c, o, _ = channel_pair(16, 12, 16, with_gpu=False)
    o.calc_k_flux()
    o.pressure_gradient()
    o.runge_kutta()
    d0 = o.divergence()
    o.poisson_rhs()
    o.set("phi", np.zeros(o.shape))
    o.poisson_solve(tol=1e-12)
    o.update_pressure()
    o.projection()
    d1 = o.divergence()
    assert d1 < 1e-9 * d0 and d1 < 1e-10


def test_dfuncs_partition_of_unity():
    lib = po.load()
    for f in (lib.orc_dfunc_2h, lib.orc_dfunc_s3h, lib.orc_dfunc_s4h):
        for r in np.linspace(-0.5, 0.5, 11):
            s = sum(f(r - j) for j in range(-4, 5))
            assert abs(s - 1.0) < 5e-8
    # Gaussian kernel integrates to one (trapezoid on a fine grid)
    r = np.linspace(-20, 20, 40001)
    v = np.array([lib.orc_dfunc_exp(x, 2.0) for x in r])
    assert abs(np.trapezoid(v, r) - 1.0) < 1e-7


def test_snes_residual_time_term_and_consistency():
    """FormFunction_SNES: at u = Ucont_o with RHS_o = F(Ucont_o) the residual is
    F(Ucont_o) - dP (the two halves add up); a perturbation d changes it by
    -d/dt to leading order (sign and scale of the time term)."""
    from helpers import channel_pair
    c, o, _ = channel_pair(12, 8, 10, with_gpu=False)
    o.eddy_viscosity()
    o.pressure_gradient()
    uo = o["ucont"].copy()
    o.set("rhs", np.zeros(o.shape + (3,)))
    o.formfunction2(1.0)
    rhso = o["rhs"].copy()
    r = o.formfunction_snes(uo, uo, rhso)
    want = rhso - o["dp"]
    inner = np.zeros(o.shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    assert np.allclose(r[inner], want[inner], rtol=1e-12, atol=1e-12 * np.abs(want).max())
    o.p.dt = dt = 1e-6 * o.p.dt  # the time term dominates: R(u+d) - R(u) = -d/dt + O(|d|)
    r = o.formfunction_snes(uo, uo, rhso)
    rng = np.random.default_rng(2)
    d = np.zeros_like(uo)
    d[2:-2, 2:-3, 2:-2] = 1e-7 * rng.standard_normal(d[2:-2, 2:-3, 2:-2].shape)
    r2 = o.formfunction_snes(uo + d, uo, rhso)
    lin = (r2 - r) * (-dt)
    sel = np.abs(d) > 0
    assert np.allclose(lin[sel], d[sel], rtol=0, atol=1e-3 * np.abs(d).max())


def test_oracle_thread_count_invariance():
    """The oracle's OpenMP loops write disjoint cells and its reductions run
    in a fixed order: one fractional step (les 2, Poisson to 1e-10) gives the
    same bits on 1 thread and on 4."""
    out = []
    before = po.set_threads()
    try:
        for n in (1, 4):
            po.set_threads(n)
            c, o, _ = channel_pair(20, 14, 18, les=2, with_gpu=False)
            o.p.ti = 5  # past the first steps, so the dynamic coefficient is computed
            it, md, rel = o.fractional_step(tol=1e-10)
            out.append((it, md, rel, o["ucont"].copy(), o["p"].copy(), o["nu_t"].copy()))
    finally:
        po.set_threads(before)
    a, b = out
    assert a[:3] == b[:3]
    for x, y in zip(a[3:], b[3:]):
        assert np.array_equal(x, y)
