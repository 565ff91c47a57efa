"""FormBCS / Contra2Cart boundary codes beyond walls, slip and the k+ outlet
(bcs.c:1745-2340, rhs.c:593-660), oracle pins on the CPU.

On a Cartesian box of unit spacing (face area vectors along the axes):
- a moving lid (code 2, j+) puts the wall velocity (1, 0, 0) halfway between
  ghost and interior cells, a Couette wall (code 12) puts (0, 0, 1) there, and
  a j+ outflow (code 4) copies the interior velocity into the ghost cell;
- with every face far field (code 6) the flux-difference correction balances
  the inflow and outflow faces exactly;
- the characteristic outlet (code 8, k+) carries the inlet flux FluxInSum
  from the third step on, and leaves the fluxes alone before;
- the k- outlet (code 4) carries FluxInSum.
"""
import numpy as np
import pytest

from helpers import cases
import pyoracle as po


def box(bctype, nx=10, ny=8, nz=12, per=(0, 0, 0), seed=3, flux_in_sum=0.0, ti=0):
    c = cases.box_flow(nx, ny, nz, bump=0.0)
    mx, my, mz = c.da
    p = po.make_params(mx, my, mz, per=per, bctype=bctype, les=0, flux_in_sum=flux_in_sum)
    p.ti = ti
    o = po.Oracle(p, c.coor)
    rng = np.random.default_rng(seed)
    u = c.ucont + 0.05 * rng.standard_normal(c.ucont.shape)
    o.set("ucont", u)
    o.contra2cart()
    o.set("ucat_old", o["ucat"])
    return o


def interior(a):
    return a[1:-1, :, 1:-1]


def test_lid_couette_and_j_outflow_ghosts():
    o = box((1, 1, 1, 2, 1, 1))
    o.contra2cart()
    u = o["ucat"]
    wall = 0.5 * (u[1:-1, -1, 1:-1] + u[1:-1, -2, 1:-1])
    assert np.allclose(wall, [1.0, 0.0, 0.0], rtol=0, atol=1e-14)
    o = box((1, 1, 12, 12, 1, 1))
    o.contra2cart()
    u = o["ucat"]
    for gj, ij in ((0, 1), (-1, -2)):
        wall = 0.5 * (u[1:-1, gj, 1:-1] + u[1:-1, ij, 1:-1])
        assert np.allclose(wall, [0.0, 0.0, 1.0], rtol=0, atol=1e-14)
    o = box((10, 10, 1, 4, 10, 10))
    o.contra2cart()
    u = o["ucat"]
    assert np.array_equal(u[1:-1, -1, 1:-1], u[1:-1, -2, 1:-1])


def _face_fluxes(o):
    u = o["ucont"]
    mz, my, mx = o.shape
    inn = u[1:-1, 1:-1, 0, 0].sum() + u[1:-1, 0, 1:-1, 1].sum() + u[0, 1:-1, 1:-1, 2].sum()
    out = u[1:-1, 1:-1, mx - 2, 0].sum() + u[1:-1, my - 2, 1:-1, 1].sum() + u[mz - 2, 1:-1, 1:-1, 2].sum()
    return inn, out


def test_far_field_balances_the_boundary_flux():
    o = box((6, 6, 6, 6, 6, 6), seed=4)
    u = o["ucont"].copy()
    u[..., 2] *= 1.3  # more out than in after the face fluxes are rebuilt from a non-uniform field
    u[: o.shape[0] // 2, ..., 2] *= 0.5
    o.set("ucont", u)
    o.contra2cart()
    o.form_bcs(0)
    inn, out = _face_fluxes(o)
    assert abs(inn - out) <= 1e-13 * max(abs(inn), 1.0)


@pytest.mark.parametrize("ti", [0, 1, 2, 7])
def test_characteristic_outlet_carries_the_inflow(ti):
    fin = 0.8 * 10 * 8
    o = box((10, 10, 10, 10, 5, 8), flux_in_sum=fin, ti=ti)
    before = o["ucont"][-2, 1:-1, 1:-1, 2].sum()
    o.form_bcs(0)
    after = o["ucont"][-2, 1:-1, 1:-1, 2].sum()
    if ti >= 2:
        assert abs(after - fin) <= 1e-13 * fin
    else:
        assert after == before


def test_k_minus_outlet_carries_the_inflow():
    fin = 0.6 * 10 * 8
    o = box((10, 10, 10, 10, 4, 10), flux_in_sum=fin)
    o.form_bcs(0)
    assert abs(o["ucont"][0, 1:-1, 1:-1, 2].sum() - fin) <= 1e-13 * fin
