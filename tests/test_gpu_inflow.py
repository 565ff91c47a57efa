"""InflowFlux's inlet profiles and inflow planes on the GPU against the oracle
(bcs.c:247-626, bcs.c:29-95): each supported -inlet profile on a
slip/inlet/outlet box whose inlet plane spans [-1/2, 1/2]^2, the -inlet 100
plane with and without a held flux, and whole fractional steps with the
pseudo-periodic inlet and with -inlet 100.  Tolerance 1e-12 normwise
relative (the inlet area and flux sums reduce in another order)."""
import numpy as np
import pytest

from helpers import cases, relerr, vfs
import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12


def centred_pair(nx=20, ny=16, nz=24, inletprofile=1, inlet_flux=-1.0, pseudo=0, bump=0.15):
    c = cases.box_flow(nx, ny, nz, bump=bump)
    IM, JM, KM = nx + 1, ny + 1, nz + 1
    coor = c.coor.copy()
    coor[:KM, :JM, :IM, 0] = coor[:KM, :JM, :IM, 0] / nx - 0.5
    coor[:KM, :JM, :IM, 1] = coor[:KM, :JM, :IM, 1] / ny - 0.5
    coor[:KM, :JM, :IM, 2] = coor[:KM, :JM, :IM, 2] / nz
    mx, my, mz = c.da
    op = po.make_params(mx, my, mz, per=(0, 0, 0), bctype=c.bctype, les=0, dt=1e-3, inlet_flux=inlet_flux,
                        inletprofile=inletprofile, pseudo=pseudo)
    o = po.Oracle(op, coor)
    gp = vfs.make_params(mx, my, mz, per=(0, 0, 0), bctype=c.bctype, les=0, dt=1e-3, inlet_flux=inlet_flux,
                         inletprofile=inletprofile, pseudo=pseudo)
    ctx = vfs.Context(gp)
    ctx.upload("COOR", coor)
    ctx.form_metrics()
    # fluxes of a smooth field on this grid, then Ucat from them on both sides
    rng = np.random.default_rng(9)
    u = o["ucont"].copy()
    u[1:-1, 1:-1, :-1, 0] = 0.05 * rng.standard_normal((nz, ny, mx - 1)) / (ny * nz)
    u[1:-1, :-1, 1:-1, 1] = 0.05 * rng.standard_normal((nz, my - 1, nx)) / (nx * nz)
    u[:-1, 1:-1, 1:-1, 2] = (1 + 0.05 * rng.standard_normal((mz - 1, ny, nx))) / (nx * ny)
    u[:, :, 0, 0] = u[:, :, mx - 2, 0] = 0
    u[:, 0, :, 1] = u[:, my - 2, :, 1] = 0
    o.set("ucont", u)
    o.contra2cart()
    o.set("ucat_old", o["ucat"])
    ctx.upload("UCONT", u)
    ctx.contra2cart()
    ctx.upload("UCAT_OLD", ctx.download("UCAT"))
    return o, ctx


@pytest.mark.parametrize("profile,flux", [(1, -1.0), (1, 0.9), (2, 0.9), (0, -1.0), (3, -1.0), (10, -1.0), (12, -1.0),
                                          (14, -1.0), (14, 0.6), (15, -1.0), (21, -1.0)])
def test_inlet_profiles(profile, flux):
    o, ctx = centred_pair(inletprofile=profile, inlet_flux=flux)
    with ctx:
        f_cpu = o.inflow_flux()
        f_gpu = ctx.inflow_flux()
        assert abs(f_gpu - f_cpu) <= TOL * max(abs(f_cpu), 1e-300)
        assert relerr(ctx.download("UCONT"), o["ucont"]) <= TOL
        assert relerr(ctx.download("UCAT"), o["ucat"]) <= TOL
        if flux > 0 and profile in (1, 2, 14):
            assert abs(f_gpu - flux) <= 1e-12 * flux


@pytest.mark.parametrize("flux", [-1.0, 0.9])
def test_plane_inlet(flux):
    o, ctx = centred_pair(inletprofile=100, inlet_flux=flux)
    with ctx:
        rng = np.random.default_rng(3)
        my, mx = o.shape[1], o.shape[2]
        plane = np.zeros((my, mx, 3))
        plane[..., 2] = 1.0
        plane += 0.1 * rng.standard_normal(plane.shape)
        o.set_inflow_plane(plane)
        ctx.inflow_set_plane(plane)
        assert np.array_equal(ctx.inflow_get_plane(), plane)
        f_cpu = o.inflow_flux()
        f_gpu = ctx.inflow_flux()
        assert abs(f_gpu - f_cpu) <= TOL * abs(f_cpu)
        assert relerr(ctx.download("UCONT"), o["ucont"]) <= TOL
        assert relerr(ctx.download("UCAT"), o["ucat"]) <= TOL


@pytest.mark.parametrize("mode", ["pseudo", "plane"])
def test_fractional_steps_with_inflow_planes(mode):
    """Two RK4 steps: -pseudo 1 (the plane is Ucat of plane (6 mz) / 10 at the
    start of each step) and -inlet 100 (a fixed plane set on both sides)."""
    pseudo = 1 if mode == "pseudo" else 0
    o, ctx = centred_pair(inletprofile=100 if mode == "plane" else 1, inlet_flux=1.0, pseudo=pseudo)
    with ctx:
        if mode == "plane":
            my, mx = o.shape[1], o.shape[2]
            plane = np.zeros((my, mx, 3))
            plane[..., 2] = 1.0 + 0.05 * np.random.default_rng(1).standard_normal((my, mx))
            o.set_inflow_plane(plane)
            ctx.inflow_set_plane(plane)
        o.mean_k_area()
        ctx.mean_k_area()
        for s in range(2):
            st = ctx.fractional_step(vfs.default_solve_opts(tol=1e-11))
            it, md, rel = o.fractional_step(tol=1e-11)
            assert st.poisson_relres <= 1e-11 and rel <= 1e-11
        err = relerr(ctx.download("UCONT"), o["ucont"])
        assert err <= 1e-8, err


def test_unsupported_profiles_are_refused():
    c = cases.box_flow(8, 8, 8)
    mx, my, mz = c.da
    for prof in (11, 13, 16, 20, 101):
        with pytest.raises(vfs.VFSError, match="not supported"):
            vfs.Context(vfs.make_params(mx, my, mz, per=(0, 0, 0), bctype=c.bctype, les=0, inletprofile=prof))
    with pytest.raises(vfs.VFSError, match="-pseudo"):
        vfs.Context(vfs.make_params(mx, my, mz, per=(1, 0, 1), bctype=(100, 100, 1, 1, 100, 100), pseudo=1))
