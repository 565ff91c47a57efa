"""FormBCS and Contra2Cart boundary codes beyond walls, slip and the k+ outlet
on the GPU against the oracle (bcs.c:1745-2340, rhs.c:593-660): moving lid
(2), Couette walls (12), j+ outflow (4), far field (6) on every face or with a
wall mix, the characteristic outlet (8) before and after its first two steps,
and the k- outlet (4); then one RK4 momentum step (RungeKutta with its
FormBCS calls, rhs.c:1586-1708) on the same codes.  Tolerance 1e-12 normwise
relative (the boundary sums reduce in another order)."""
import numpy as np
import pytest

from helpers import cases, relerr, vfs
import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = {
    "lid": (1, 1, 1, 2, 1, 1),
    "couette": (1, 1, 12, 12, 10, 10),
    "j-outflow": (10, 10, 1, 4, 10, 10),
    "far-all": (6, 6, 6, 6, 6, 6),
    "far-mix": (6, 6, 1, 10, 5, 6),
    "char-outlet": (10, 10, 10, 10, 5, 8),
    "k-outlet": (10, 10, 1, 10, 4, 10),
    "interface-symmetry": (0, 3, 1, 10, 5, 4),
}


def pair(bctype, ti, nx=18, ny=14, nz=20):
    c = cases.box_flow(nx, ny, nz, bump=0.2)
    mx, my, mz = c.da
    fin = 0.7 * nx * ny
    op = po.make_params(mx, my, mz, per=(0, 0, 0), bctype=bctype, les=0, dt=2e-3, inlet_flux=-1.0, flux_in_sum=fin)
    op.ti = ti
    o = po.Oracle(op, c.coor)
    gp = vfs.make_params(mx, my, mz, per=(0, 0, 0), bctype=bctype, les=0, dt=2e-3, inlet_flux=-1.0, flux_in_sum=fin)
    ctx = vfs.Context(gp)
    ctx.upload("COOR", c.coor)
    ctx.form_metrics()
    ctx.set_time_index(ti, 0, 1)
    rng = np.random.default_rng(8)
    u = c.ucont + 0.03 * rng.standard_normal(c.ucont.shape) * np.abs(c.ucont).max()
    o.set("ucont", u)
    o.contra2cart()
    o.set("ucat_old", o["ucat"])
    ctx.upload("UCONT", u)
    ctx.contra2cart()
    ctx.upload("UCAT_OLD", ctx.download("UCAT"))
    return o, ctx


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("ti", [0, 5])
def test_form_bcs_codes(name, ti):
    o, ctx = pair(CASES[name], ti)
    with ctx:
        assert relerr(ctx.download("UCAT"), o["ucat"]) <= TOL
        for scale in (0, 1):
            o.form_bcs(scale)
            ctx.form_bcs(scale)
            assert relerr(ctx.download("UCONT"), o["ucont"]) <= TOL, (name, scale)
            assert relerr(ctx.download("UCAT"), o["ucat"]) <= TOL, (name, scale)


@pytest.mark.parametrize("name", list(CASES))
def test_runge_kutta_with_codes(name):
    o, ctx = pair(CASES[name], 5)
    with ctx:
        o.pressure_gradient()
        ctx.pressure_gradient()
        o.runge_kutta()
        ctx.runge_kutta()
        assert relerr(ctx.download("UCONT"), o["ucont"]) <= TOL, name
        assert relerr(ctx.download("UCAT"), o["ucat"]) <= TOL, name


def test_unsupported_codes_are_refused():
    c = cases.box_flow(8, 8, 8)
    mx, my, mz = c.da
    for bc in ((-1, 1, 1, 1, 5, 4), (11, 10, 10, 10, 5, 4), (10, 10, 13, 10, 5, 4), (10, 10, 10, 10, 5, 2 + 10 * 0 + 7)):
        with pytest.raises(vfs.VFSError, match="not supported"):
            vfs.Context(vfs.make_params(mx, my, mz, per=(0, 0, 0), bctype=bc, les=0))
