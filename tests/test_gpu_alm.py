"""GPU parity of the actuator-line path (rotor_model.c) against the FP64 oracle.

Seeded non-orthogonal box with the benchmark rotor at its centre; each stage
of the rotor_model 6 block of Flow_Solver (solvers.c:535-563) is compared with
the oracle on the same inputs.  Integer boxes must match exactly; FP64 values
within 1e-12 relative (normwise) -- the device's sqrt/asin/exp/acos may differ
from glibc's by an ulp.
"""
import importlib
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
import pyoracle as po  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relerr(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d


def _setup(rotor_model, deltafunc, halfwidth, nx=40, ny=36, nz=44):
    vfs = importlib.import_module("vfs-wind_amd")
    cases = importlib.import_module("vfs-wind_amd.cases")
    b = cases.box(nx, ny, nz, h=1.0, bump=0.3)
    mx, my, mz = b["da"]
    hub = (nx / 2 + 0.13, ny / 2 - 0.21, nz / 2 + 0.37)
    ts = cases.synthetic_turbine(8.0, hub, rotor_model=rotor_model, deltafunc=deltafunc, halfwidth=halfwidth,
                                 r_nacelle=1.1)
    p = vfs.make_params(mx, my, mz, per=(0, 0, 0), bctype=(10, 10, 10, 10, 5, 4), les=0)
    ctx = vfs.Context(p)
    ctx.upload("COOR", b["coor"])
    ctx.form_metrics()
    ctx.upload("UCAT", b["ucat"])
    op = po.make_params(mx, my, mz, per=(0, 0, 0), bctype=(10, 10, 10, 10, 5, 4), les=0)
    o = po.Oracle(op, b["coor"])
    o.set("ucat", b["ucat"])
    t = po.make_turbine(ts.nx, ts.ny, ts.nz, ts.xc, ts.yc, ts.zc, ts.angvel, ts.pitch, ts.r_nacelle, ts.refangle)
    alm = vfs.ALM(ctx, ts.turbine_struct(vfs), ts.nodes, ts.foils)
    return vfs, ctx, o, t, ts, alm


@pytest.mark.parametrize("rotor_model,deltafunc,halfwidth", [(6, 7, 2.0), (3, 10, 4.0), (6, 0, 4.0)],
                         ids=["model6-gauss", "model3-s4h", "model6-2h"])
def test_alm_stages(rotor_model, deltafunc, halfwidth):
    vfs, ctx, o, t, ts, alm = _setup(rotor_model, deltafunc, halfwidth)
    nb, nvb = ts.nodes.shape[:2]
    foils = po.Foils(ts.foils)
    dt = 0.05
    # airfoil_ACL at the initial positions
    cent0, _ = po.alm_elements(ts.nodes)
    tw, ch = po.alm_twist_chord(t, cent0, foils)
    assert relerr(alm.get("TWIST"), tw) <= TOL and relerr(alm.get("CHORD"), ch) <= TOL
    for step in range(2):
        # rotor_Rot
        alm.rotate(dt)
        x, u = po.alm_rotate(t, dt, ts.nodes if step == 0 else x)
        assert relerr(alm.get("NODES"), x.reshape(-1, 3)) <= TOL
        assert relerr(alm.get("NODE_VEL"), u.reshape(-1, 3)) <= TOL
        cent, dA = po.alm_elements(x)
        assert relerr(alm.get("CENT"), cent) <= TOL and relerr(alm.get("DA"), dA) <= TOL
        # Pre_process (integer boxes: exact)
        alm.preprocess()
        box = o.alm_preprocess(alm.get("CENT"))
        assert np.array_equal(alm.get("BOX").astype(np.int32), box)
        if rotor_model == 6:
            ref = po.alm_refline(t, x)
            rcent, _ = po.alm_elements(ref)
            assert relerr(alm.get("REF_CENT"), rcent) <= TOL
            rbox = o.alm_preprocess(alm.get("REF_CENT"))
            assert np.array_equal(alm.get("REF_BOX").astype(np.int32), rbox)
            scent, sbox = alm.get("REF_CENT"), rbox
        else:
            scent, sbox = alm.get("CENT"), box
        # Calc_U_lagr
        alm.u_lagr()
        ul = alm.get("U_LAGR")
        assert relerr(ul, o.alm_ulagr(scent, sbox)) <= TOL
        # Calc_F_lagr_ACL on the device's U_lagr
        alm.f_lagr()
        f, aoa, cl, cd = po.alm_flagr(t, nb, nvb, alm.get("CENT"), ul, alm.get("NODE_VEL"), alm.get("TWIST"),
                                      alm.get("CHORD"), foils)
        assert relerr(alm.get("F_LAGR"), f) <= TOL
        assert relerr(alm.get("AOA"), aoa) <= TOL
        assert relerr(alm.get("CL"), cl) <= TOL and relerr(alm.get("CD"), cd) <= TOL
        # Calc_F_eul
        alm.f_eul()
        feul = o.alm_feul(alm.get("CENT"), box, alm.get("F_LAGR"), alm.get("DA"), df=deltafunc, halfwidth=halfwidth)
        got = ctx.download("FEUL")
        assert np.abs(got).max() > 0
        assert relerr(got, feul) <= TOL
    alm.close()
    ctx.close()


def test_alm_step_matches_stages():
    vfs, ctx, o, t, ts, alm = _setup(6, 7, 2.0)
    alm.step(0.05)
    feul_step = ctx.download("FEUL")
    f_step = alm.get("F_LAGR")
    alm.close()
    alm2 = vfs.ALM(ctx, ts.turbine_struct(vfs), ts.nodes, ts.foils)
    alm2.rotate(0.05); alm2.preprocess(); alm2.u_lagr(); alm2.f_lagr(); alm2.f_eul()
    assert np.array_equal(ctx.download("FEUL"), feul_step)
    assert np.array_equal(alm2.get("F_LAGR"), f_step)
    alm2.close()
    ctx.close()


def test_alm_uref_disk():
    """Uref_ACL on the reference's own Urefdata000 disk, scaled into the box."""
    vfs, ctx, o, t, ts, alm = _setup(6, 7, 2.0)
    tb = importlib.import_module("vfs-wind_amd.turbine")
    deck = ROOT / "tests" / "golden" / "clipper_tsr5"
    # disk of radius 8 (R = 48 scaled by 6), 1R upstream of the hub
    x, tris, cent, dA = tb.read_ucd_disk(deck / "Urefdata000", 6.0, (ts.xc, ts.yc, ts.zc), (0, 0, 1), 48.0, 6.0,
                                         loc_refvel=0.5)
    alm.set_disk(cent, dA)
    uref = alm.uref()
    box = o.alm_preprocess(cent)
    u = o.alm_ulagr(cent, box)
    want = float((u[:, 2] * dA).sum() / dA.sum())
    assert abs(uref - want) <= TOL * abs(want)
    assert 0.9 < uref < 1.1  # the inflow is 1 + 5% noise
    alm.close()
    ctx.close()
