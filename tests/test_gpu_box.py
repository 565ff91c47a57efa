"""The actuator-line configuration on the GPU (small size): uniform-inflow box
with slip sides, inlet (5) and outlet (4); InflowFlux, RK4 and a full
fractional step against the oracle, and the -imp 4 momentum solve with the
rotor's F_eul, judged by the oracle's residual at the GPU root."""
import importlib

import numpy as np
import pytest

import pyoracle as po
from helpers import relerr, vfs

cases = importlib.import_module("vfs-wind_amd.cases")
pytestmark = pytest.mark.gpu


def _pair(nx=24, ny=20, nz=32, les=1):
    c = cases.box_flow(nx, ny, nz)
    op = po.make_params(*c.da, per=c.per, bctype=c.bctype, ren=c.ren, dt=1e-3, inlet_flux=c.inlet_flux, les=les)
    o = po.Oracle(op, c.coor)
    o.set("ucont", c.ucont)
    o.contra2cart()
    o.set("ucat_old", o["ucat"])
    op.dt = 0.3 * o.min_dt()
    gp = cases.params_for(vfs, c, op.dt, les=les)
    ctx = vfs.Context(gp)
    ctx.upload("COOR", c.coor)
    ctx.form_metrics()
    ctx.upload("UCONT", c.ucont)
    ctx.contra2cart()
    ctx.upload("UCAT_OLD", ctx.download("UCAT"))
    return c, o, ctx


def test_inflow_flux_and_rk():
    c, o, ctx = _pair()
    assert relerr(ctx.download("UCAT"), o["ucat"]) <= 1e-12
    fo = o.inflow_flux()
    fg = ctx.inflow_flux()
    assert abs(fg - fo) <= 1e-12 * abs(fo)
    assert relerr(ctx.download("UCONT"), o["ucont"]) <= 1e-12
    assert relerr(ctx.download("UCAT"), o["ucat"]) <= 1e-12
    # the inlet area is the box cross-section: uniform unit inflow carries it
    assert fo == pytest.approx(24 * 20, rel=1e-2)
    o.eddy_viscosity(); ctx.eddy_viscosity()
    o.pressure_gradient(); ctx.pressure_gradient()
    o.runge_kutta(); ctx.runge_kutta()
    assert relerr(ctx.download("UCONT"), o["ucont"]) <= 1e-12
    assert relerr(ctx.download("UCAT"), o["ucat"]) <= 1e-12
    ctx.close()


def test_box_fractional_step():
    c, o, ctx = _pair()
    it, md, rel = o.fractional_step(tol=1e-10)
    st = ctx.fractional_step(vfs.default_solve_opts(tol=1e-10))
    assert st.poisson_relres <= 1e-10 and rel <= 1e-10
    assert relerr(ctx.download("UCONT"), o["ucont"]) <= 1e-7
    assert st.maxdiv <= 1e-8 and md <= 1e-8
    ctx.close()


def test_box_cn_with_rotor():
    c, o, ctx = _pair(32, 32, 32)
    ts = cases.synthetic_turbine(8.0, (16.1, 15.9, 12.3), deltafunc=7, halfwidth=2.0)
    alm = vfs.ALM(ctx, ts.turbine_struct(vfs), ts.nodes, ts.foils)
    alm.step(o.p.dt)
    feul = ctx.download("FEUL")
    assert np.abs(feul).max() > 0
    ctx.eddy_viscosity(); o.eddy_viscosity()
    ctx.pressure_gradient(); o.pressure_gradient()
    ctx.store_old_state()
    uo = ctx.download("UCONT_O")
    rhso = ctx.download("RHS_O")
    tol = 1e-6
    st = ctx.implicit_matrixfree(vfs.default_newton_opts(rtol=tol, use_feul=1))
    assert st.fnorm <= tol * st.fnorm0
    u = ctx.download("UCONT")
    # the GPU solve starts with InflowFlux, which sets the inlet fluxes and the
    # inlet ghost velocity that every residual evaluation reads: do the same
    # on the oracle, whose residual at the starting state is the GPU's fnorm0
    o.set("ucat_old", ctx.download("UCAT_OLD"))
    o.inflow_flux()
    r0 = o.formfunction_snes(o["ucont"].copy(), uo, rhso, feul)
    assert abs(np.linalg.norm(r0) - st.fnorm0) <= 1e-12 * st.fnorm0
    # the oracle's residual (with the same F_eul) vanishes at the GPU root; the
    # post-solve outlet fix rewrites the z-flux at k = mz-2 and the ghost
    # velocities behind it, which rows k >= mz-4 read through the QUICK
    # stencil, so those rows are left out
    r = o.formfunction_snes(u, uo, rhso, feul)
    keep = slice(0, r.shape[0] - 4)
    assert np.linalg.norm(r[keep]) <= 2 * tol * np.linalg.norm(r0[keep])
    alm.close()
    ctx.close()


@pytest.mark.parametrize("path", ["short", "long"])
@pytest.mark.parametrize("cells", [(24, 20, 32), (23, 20, 31)], ids=["24x20x32", "odd-23x20x31"])
def test_box_poisson_one_pass_forms_bitwise(monkeypatch, cells, path):
    """On the box (no periodic direction, so the compact smoother's i and k
    neighbours at the walls are ghosts; with an odd i count one colour has a
    line fewer per row): the FP32 V-cycle with the finest level in the
    colour-split layout, in the natural layout (VFS_MG_SPLIT=0), and the
    two-pass natural form (VFS_ZEBRA_FUSED=0, VFS_MG_FUSED_RESTRICT=0) give
    the same solve, bit for bit (each with the same line segments)."""
    out = []
    variants = ({}, {"VFS_MG_SPLIT": "0"}, {"VFS_ZEBRA_FUSED": "0", "VFS_MG_FUSED_RESTRICT": "0"})
    for env in variants:
        env = {"VFS_SMOOTH_COMPACT": "1", "VFS_LINE_PATH": path, **env}
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        c, o, ctx = _pair(*cells)
        try:
            ctx.poisson_rhs()
            ctx.upload("PHI", np.zeros(o.shape))
            st = ctx.poisson_solve(vfs.default_solve_opts(tol=1e-10, use_guess=0, maxit=400))
            assert st.relres <= 1e-10
            out.append((st.iterations, ctx.download("PHI")))
        finally:
            ctx.close()
            for k in env:
                monkeypatch.delenv(k)
    for its, phi in out[1:]:
        assert its == out[0][0] and np.array_equal(phi, out[0][1])
