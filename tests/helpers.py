"""Shared builders for the parity tests: one seeded case fed identically to the
FP64 CPU oracle and to the HIP library."""
from __future__ import annotations

import importlib

import numpy as np

import pyoracle as po

vfs = importlib.import_module("vfs-wind_amd")
cases = importlib.import_module("vfs-wind_amd.cases")


def relerr(a: np.ndarray, b: np.ndarray) -> float:
    """Normwise relative difference max|a-b| / max|b| (0 when both vanish)."""
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    if den == 0.0:
        return num
    return num / den


def channel_pair(nx=32, ny=24, nz=32, les=1, second_order=1, laplacian=1, with_gpu=True, seed=cases.SEED,
                 max_cs=0.5, homo=(0, 0, 0)):
    """Oracle and (optionally) GPU context on the same channel case, metrics
    formed, initial Ucat from Contra2Cart, dt = 0.3 dt_CFL1."""
    c = cases.channel(nx, ny, nz, seed=seed)
    op = po.make_params(*c.da, per=c.per, bctype=c.bctype, ren=c.ren, dt=1e-3, inlet_flux=c.inlet_flux,
                        les=les, second_order=second_order, laplacian=laplacian, max_cs=max_cs, homo=homo)
    o = po.Oracle(op, c.coor)
    o.set("ucont", c.ucont)
    o.contra2cart()
    o.set("ucat_old", o["ucat"])
    dt = 0.3 * o.min_dt()
    op.dt = dt
    o.mean_k_area()
    ctx = None
    if with_gpu:
        gp = cases.params_for(vfs, c, dt, les=les, second_order=second_order, laplacian=laplacian, max_cs=max_cs,
                              homo=homo)
        ctx = vfs.Context(gp)
        ctx.upload("COOR", c.coor)
        ctx.form_metrics()
        ctx.mean_k_area()
        ctx.upload("UCONT", c.ucont)
        ctx.contra2cart()
        ctx.upload("UCAT_OLD", ctx.download("UCAT"))
    return c, o, ctx
