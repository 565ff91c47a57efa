"""Actuator-line turbine in a uniform-inflow box, run for a number of steps.

The configuration of the reference's Clipper rotor deck restated on a
synthetic rotor (Test_08_ClipperTurbine/TSR5.0: bcs "10 10 10 10 5 4",
-rotor_modeled 6, -imp 4): a Cartesian-curvilinear box of nx x ny x nz cells
(unit spacing, node displacement `bump`), slip walls on the i and j faces,
uniform inlet U = 1 on k- (plus seeded noise inside), outlet on k+; one
3-blade rotor of radius R cells at the box centre, axis +z, tip-speed ratio
7 at a fixed angular velocity, 41 nodes (40 elements) per blade, CL = 2 pi
alpha clipped at +-1.2, CD = 0.01, forces spread with the Gaussian kernel
(-deltafunc 7, halfwidth 2).  Each step: the rotor block (rotate,
Pre_process, U_lagr, F_lagr, F_eul), then the fractional step with the
Crank-Nicolson momentum solve (or RK4 with --imp 0), the Poisson solve to
1e-8, the pressure update and the projection.

    python vfs-wind_amd/turbine_run.py --nx 384 --ny 192 --nz 256 --steps 100

Per-step lines go to stderr, one JSON summary line to stdout.
"""
from __future__ import annotations

import argparse
import importlib
import json
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent


def main(argv=None):
    if str(PKG.parent) not in sys.path:
        sys.path.insert(0, str(PKG.parent))
    vfs = importlib.import_module("vfs-wind_amd")
    cases = importlib.import_module("vfs-wind_amd.cases")
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nx", type=int, default=384)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--radius", type=float, default=32.0, help="rotor radius in cells")
    ap.add_argument("--tsr", type=float, default=7.0)
    ap.add_argument("--bump", type=float, default=0.3)
    ap.add_argument("--cfl", type=float, default=0.3, help="dt as a fraction of the CFL-1 step")
    ap.add_argument("--imp", type=int, default=4, choices=(0, 4), help="0: RK4, 4: Crank-Nicolson Newton-Krylov")
    ap.add_argument("--newton-rtol", type=float, default=1e-4)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    c = cases.box_flow(a.nx, a.ny, a.nz, bump=a.bump)
    p = cases.params_for(vfs, c, dt=1e-3)
    ctx = vfs.Context(p)
    ctx.upload("COOR", c.coor)
    ctx.form_metrics()
    ctx.upload("UCONT", c.ucont)
    ctx.contra2cart()
    ctx.upload("UCAT_OLD", ctx.download("UCAT"))
    p.dt = a.cfl * ctx.min_dt()
    ctx.set_params(p)
    hub = (0.5 * a.nx, 0.5 * a.ny, 0.5 * a.nz)
    ts = cases.synthetic_turbine(a.radius, hub, tsr=a.tsr, deltafunc=7, halfwidth=2.0)
    alm = vfs.ALM(ctx, ts.turbine_struct(vfs), ts.nodes, ts.foils)
    axis = np.array([0.0, 0.0, 1.0])
    log(f"case {a.nx}x{a.ny}x{a.nz} built in {time.perf_counter() - t0:.1f} s; dt = {p.dt:.4e}; "
        f"{alm.n_elmt} line elements")
    nopts = vfs.default_newton_opts(rtol=a.newton_rtol, use_feul=1)
    sopts = vfs.default_solve_opts(tol=1e-8)
    rows = []
    ctx.sync()
    t_run = time.perf_counter()
    for s in range(a.steps):
        ta = time.perf_counter()
        alm.step(p.dt)
        ctx.sync()
        ms_alm = (time.perf_counter() - ta) * 1e3
        st = ctx.fractional_step_cn(nopts, sopts) if a.imp == 4 else ctx.fractional_step(sopts)
        f = alm.get("F_LAGR")
        da = alm.get("DA")
        axial = float((f @ axis * da).sum())
        rows.append(dict(newton=st.newton_iterations, gmres=st.linear_iterations, poisson=st.poisson_iterations,
                         relres=st.poisson_relres, maxdiv=st.maxdiv, ms_alm=ms_alm, ms_momentum=st.ms_momentum,
                         ms_poisson=st.ms_poisson, ms_total=st.ms_total + ms_alm, axial=axial))
        log(f"step {s}: newton {st.newton_iterations} gmres {st.linear_iterations} poisson {st.poisson_iterations} "
            f"relres {st.poisson_relres:.2e} maxdiv {st.maxdiv:.2e} alm {ms_alm:.1f} ms momentum "
            f"{st.ms_momentum:.1f} ms poisson {st.ms_poisson:.1f} ms axial force {axial:.4e}")
    ctx.sync()
    wall = time.perf_counter() - t_run
    alm.close()
    ctx.close()
    cells = a.nx * a.ny * a.nz
    mean = lambda k: float(np.mean([r[k] for r in rows]))
    out = {
        "workload": f"actuator-line box {a.nx}x{a.ny}x{a.nz}, bcs 10 10 10 10 5 4, rotor R={a.radius:g} TSR {a.tsr:g}, "
                    f"Gaussian spreading eps=2, {'CN Newton-Krylov' if a.imp == 4 else 'RK4'} + Poisson 1e-8",
        "steps": a.steps, "cells": cells, "dt": p.dt,
        "ms_per_step": wall / a.steps * 1e3, "mcell_steps_per_s": cells * a.steps / wall / 1e6,
        "ms_alm": mean("ms_alm"), "ms_momentum": mean("ms_momentum"), "ms_poisson": mean("ms_poisson"),
        "newton_iterations_mean": mean("newton"), "gmres_iterations_mean": mean("gmres"),
        "poisson_iterations_mean": mean("poisson"), "poisson_relres_worst": max(r["relres"] for r in rows),
        "maxdiv_worst": max(r["maxdiv"] for r in rows), "axial_force_last": rows[-1]["axial"],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
