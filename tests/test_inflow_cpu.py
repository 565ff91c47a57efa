"""InflowFlux's inlet profiles and inflow planes (bcs.c:247-626, rhs2.c:716-862)
on the CPU: the oracle pinned by closed forms, and the C ABI's inflow-section
writer and reader (host code) pinned by their round trip.

Oracle pins, on a Cartesian box whose inlet plane spans [-1/2, 1/2]^2:
- a uniform inlet with -flux carries exactly that flux;
- the pipe profile (-inlet 12, uin = 2 (1 - 4 r^2)) carries the midpoint-rule
  integral of the profile over the inlet cells, with r taken at the centres
  of the cells' faces;
- an inflow plane (-inlet 100) of uniform velocity U gives
  ucont.z = (U + ucat(1)) / 2 . zet, and with -flux the held flux;
- the pseudo-periodic plane is Ucat of plane (6 mz) / 10.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from helpers import cases, vfs

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import pyoracle as po  # noqa: E402


def centred_box(nx=12, ny=10, nz=8, inletprofile=1, inlet_flux=-1.0, pseudo=0):
    """Box of nx x ny x nz cells spanning [-1/2, 1/2]^2 x [0, 1] (unit length),
    slip sides, inlet on k-, outlet on k+."""
    c = cases.box_flow(nx, ny, nz, h=1.0, bump=0.0)
    IM, JM, KM = nx + 1, ny + 1, nz + 1
    coor = c.coor.copy()
    coor[:KM, :JM, :IM, 0] = coor[:KM, :JM, :IM, 0] / nx - 0.5
    coor[:KM, :JM, :IM, 1] = coor[:KM, :JM, :IM, 1] / ny - 0.5
    coor[:KM, :JM, :IM, 2] = coor[:KM, :JM, :IM, 2] / nz
    mx, my, mz = c.da
    p = po.make_params(mx, my, mz, per=(0, 0, 0), bctype=c.bctype, les=0, inlet_flux=inlet_flux,
                       inletprofile=inletprofile, pseudo=pseudo)
    o = po.Oracle(p, coor)
    rng = np.random.default_rng(5)
    u = np.zeros(o["ucat"].shape)
    u[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, (nz, ny, nx, 3))
    o.set("ucat", u)
    return c, coor, o


def test_uniform_inlet_carries_the_given_flux():
    c, coor, o = centred_box(inletprofile=1, inlet_flux=0.37)
    assert abs(o.inflow_flux() - 0.37) <= 1e-14


def test_pipe_profile_midpoint_integral():
    nx, ny = 12, 10
    c, coor, o = centred_box(nx, ny, inletprofile=12)
    flux = o.inflow_flux()
    # face centres of the inlet cells (equal to the cell centres in x and y on this grid)
    xc = (np.arange(nx) + 0.5) / nx - 0.5
    yc = (np.arange(ny) + 0.5) / ny - 0.5
    R2 = xc[None, :] ** 2 + yc[:, None] ** 2
    want = float((2 * (1 - 4 * R2) * (1.0 / nx) * (1.0 / ny)).sum())
    assert abs(flux - want) <= 1e-13
    # and cell by cell: ucont.z / |zet| is the profile
    uz = o["ucont"][0, 1:-1, 1:-1, 2] * nx * ny
    assert np.allclose(uz, 2 * (1 - 4 * R2), rtol=0, atol=1e-13)


@pytest.mark.parametrize("inlet_flux", [-1.0, 0.8])
def test_plane_inlet(inlet_flux):
    nx, ny = 12, 10
    c, coor, o = centred_box(nx, ny, inletprofile=100, inlet_flux=inlet_flux)
    U = np.array([0.1, -0.2, 1.3])
    plane = np.broadcast_to(U, (ny + 2, nx + 2, 3)).copy()
    o.set_inflow_plane(plane)
    u1 = o["ucat"][1].copy()
    flux = o.inflow_flux()
    assert np.array_equal(o["ucat"][0, 1:-1, 1:-1], np.broadcast_to(U, (ny, nx, 3)))
    uz = o["ucont"][0, 1:-1, 1:-1, 2]
    if inlet_flux < 0:
        # zet of an inlet face of this grid is (0, 0, 1 / (nx ny))
        want = 0.5 * (U[2] + u1[1:-1, 1:-1, 2]) / (nx * ny)
        assert np.allclose(uz, want, rtol=0, atol=1e-15)
        assert abs(flux - uz.sum()) <= 1e-14
    else:
        assert abs(flux - inlet_flux) <= 1e-14


def test_pseudo_periodic_plane_is_the_kplane_velocity():
    c, coor, o = centred_box(nz=10)
    mz = o.shape[0]
    assert np.array_equal(o.pseudo_periodic_plane(), o["ucat"][(6 * mz) // 10])


def test_inflow_sections_round_trip(tmp_path):
    """Planes written by save_inflow_section at ti = 1..250 (files of 100
    planes) are read back by read_inflow_section at the same ti, after a
    restart, and through the recycling period."""
    mx, my, dt, period = 6, 5, 0.001, 100
    plane = lambda v: np.full((my, mx, 3), float(v))
    for ti in range(1, 251):
        vfs.inflow_write_section(tmp_path, ti, 1, period, 0, dt, plane(ti))
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["inflow_000001_dt=0.001.dat", "inflow_000101_dt=0.001.dat", "inflow_000201_dt=0.001.dat"]
    assert (tmp_path / names[0]).stat().st_size == 100 * mx * my * 24
    (tmp_path / "inflow").mkdir()
    for n in names:
        (tmp_path / n).rename(tmp_path / "inflow" / n)
    # fresh start: ti = 0 reads the first plane, ti = 1 is zero, then plane ti
    r = vfs.InflowReader(tmp_path, dt, period, 0, 20000, 0, mx, my)
    assert np.all(r.next(0) == 1) and np.all(r.next(1) == 0)
    for ti in range(2, 251):
        assert np.all(r.next(ti) == ti), ti
    r.close()
    # a restart at ti = 101 opens the second file
    r = vfs.InflowReader(tmp_path, dt, period, 0, 20000, 101, mx, my)
    for ti in range(101, 231):
        assert np.all(r.next(ti) == ti), ti
    r.close()
    # a restart at ti = 150 positions itself inside the second file; files are
    # switched only after tistart + 90, so past its end the planes are zero
    r = vfs.InflowReader(tmp_path, dt, period, 0, 20000, 150, mx, my)
    for ti in range(150, 215):
        assert np.all(r.next(ti) == (ti if ti <= 200 else 0)), ti
    r.close()
    # recycling every 200 steps: ti = 201.. reads the planes of ti - 200
    r = vfs.InflowReader(tmp_path, dt, period, 0, 200, 0, mx, my)
    for ti in range(0, 251):
        got = r.next(ti)
        assert np.all(got == (1 if ti == 0 else 0 if ti == 1 else ti - 200 if ti > 200 else ti)), ti
    r.close()
    with pytest.raises(vfs.VFSError, match="cannot open the inflow file"):
        vfs.InflowReader(tmp_path, 0.002, period, 0, 20000, 0, mx, my).next(0)
