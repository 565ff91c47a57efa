"""Checkpoints through the C ABI: Ucont_P_Binary_Output / Ucont_P_Binary_Input
(main.c:376-770) on a device context.  The written files hold the context's
fields bitwise in PETSc's big-endian Vec format; a fresh context restarted
from them holds the same fields, Ucont_o = Ucont and Ucat_old = Ucat, zero
blanking without immersed bodies, and steps on like the original (the
pressure correction's initial guess is not checkpointed, so the next solves
start from different guesses: agreement to the solve tolerance)."""
import struct

import numpy as np
import pytest

from helpers import cases, channel_pair, relerr, vfs

pytestmark = pytest.mark.gpu


def test_checkpoint_round_trip_and_restart(tmp_path):
    c, o, ctx = channel_pair(24, 16, 20, with_gpu=True)
    opts = vfs.default_solve_opts(tol=1e-10)
    ctx.fractional_step(opts)
    nv = np.zeros(ctx.field_shape("NVERT"))
    nv[5:8, 4:6, 6:9] = 1.0
    ctx.upload("NVERT", nv)
    ctx.checkpoint_write(tmp_path, 12, 0)
    names = {"vfield": "UCONT", "ufield": "UCAT", "pfield": "P", "nvfield": "NVERT"}
    saved = {f: ctx.download(f) for f in names.values()}
    for stem, field in names.items():
        path = tmp_path / f"{stem}000012_0.dat"
        raw = path.read_bytes()
        n = saved[field].size
        assert raw[:8] == struct.pack(">ii", 1211214, n)
        assert np.array_equal(np.frombuffer(raw[8:], ">f8"), saved[field].ravel())
    ctx.upload("NVERT", np.zeros_like(nv))

    ctx2 = vfs.Context(ctx.params)
    ctx2.upload("COOR", c.coor)
    ctx2.form_metrics()
    ctx2.mean_k_area()
    for immersed in (1, 0):
        ctx2.checkpoint_read(tmp_path, 12, 0, immersed=immersed)
        for field in ("UCONT", "UCAT", "P"):
            assert np.array_equal(ctx2.download(field), saved[field]), field
        assert np.array_equal(ctx2.download("UCONT_O"), saved["UCONT"])
        assert np.array_equal(ctx2.download("UCAT_OLD"), saved["UCAT"])
        assert np.array_equal(ctx2.download("NVERT"), saved["NVERT"] if immersed else np.zeros_like(nv))
    with pytest.raises(vfs.VFSError, match="cannot open"):
        ctx2.checkpoint_read(tmp_path, 13, 0)
    # a restart steps on like the original
    ctx.upload("PHI", np.zeros(ctx.field_shape("PHI")))
    ctx.upload("UCAT_OLD", saved["UCAT"])
    for s in range(2):
        a = ctx.fractional_step(opts)
        b = ctx2.fractional_step(opts)
        assert a.poisson_relres <= 1e-10 and b.poisson_relres <= 1e-10
    assert relerr(ctx2.download("UCONT"), ctx.download("UCONT")) <= 1e-8
    ctx2.close()
    ctx.close()


def test_checkpoint_read_rejects_other_grid(tmp_path):
    c, o, ctx = channel_pair(16, 12, 16)
    ctx.checkpoint_write(tmp_path, 0, 1)
    small = cases.channel(12, 12, 16)
    ctx2 = vfs.Context(cases.params_for(vfs, small, 1e-3))
    with pytest.raises(vfs.VFSError, match="expected"):
        ctx2.checkpoint_read(tmp_path, 0, 1)
    ctx2.close()
    ctx.close()
