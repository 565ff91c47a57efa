"""Poisson preconditioners side by side on the seeded curvilinear channel:
FGMRES with the FP32 geometric V-cycle, the FP64 one, and the algebraic
multigrid (the role of the reference's BoomerAMG, -poisson 1), each solving
the same right-hand side (one momentum stage, then PoissonRHS) from a zero
guess.  The AMG's host set-up falls in its first solve.

    python vfs-wind_amd/poisson_compare.py --nx 64 --ny 48 --nz 64 --reps 3
"""
import argparse
import importlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=48)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-8)
    a = ap.parse_args(argv)
    vfs = importlib.import_module("vfs-wind_amd")
    cases = importlib.import_module("vfs-wind_amd.cases")
    c = cases.channel(a.nx, a.ny, a.nz)
    ctx = vfs.Context(cases.params_for(vfs, c, 1e-3, les=0))
    out = {}
    with ctx:
        ctx.upload("COOR", c.coor)
        ctx.form_metrics()
        ctx.mean_k_area()
        ctx.upload("UCONT", c.ucont)
        ctx.contra2cart()
        ctx.upload("UCAT_OLD", ctx.download("UCAT"))
        ctx.runge_kutta()
        ctx.poisson_rhs()
        for name, kw in (("gmg_fp32", {}), ("gmg_fp64", {"mg_fp32": 0}), ("amg_fp64", {"amg": 1})):
            runs = []
            for _ in range(a.reps):
                ctx.upload("PHI", np.zeros(ctx.shape))
                st = ctx.poisson_solve(vfs.default_solve_opts(tol=a.tol, use_guess=0, maxit=400, **kw))
                runs.append((st.iterations, st.ms, st.relres))
                print(f"[{name}] iterations {st.iterations} relres {st.relres:.2e} ms {st.ms:.2f}", flush=True)
            out[name] = {"iterations": runs[-1][0], "first_ms": round(runs[0][1], 3),
                         "ms": round(float(np.median([r[1] for r in runs[1:]] or [runs[0][1]])), 3),
                         "relres": max(r[2] for r in runs)}
    print(json.dumps({"grid": [a.nx, a.ny, a.nz], "tol": a.tol, "solves": out}))


if __name__ == "__main__":
    main()
