#!/usr/bin/env python3
"""Benchmark of the fractional-step core on one MI355X.

Workload (one "step" = one Flow_Solver time step, solvers.c:313-759, -imp 0):
seeded curvilinear channel 256x192x256 cells (12,582,912), periodic x/z,
no-slip walls, LES (Cs = 0.1), central second-order convection, RK4 momentum,
Poisson to a 1e-8 relative residual with FGMRES(20) + FP32 V-cycle, pressure
update, projection and divergence check, dt = 0.3 dt_CFL1 held fixed.

Prints one JSON line: Mcell-steps/s, the roofline entry of the dominant kernel
(HIP-event timing of its launches inside the timed region) and the FP64 CPU
oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per unit of each timed kernel (every input read once, every
# output written once, FP64 unless stated) and the units one launch covers:
# "point" = every DA point (the launch grid), "cell" = interior cells,
# "line point" = interior points of one smoother colour, (mz-2)(my-2)ceil((mx-2)/2).
ALGO = {
    # One RK stage's Formfunction_2 with the stage update fused (Ucont = a (Rhs - dP) + pUcont), in the
    # reference's stored-metric accounting: Formfunction_2 reads Ucont 3 + Ucat 3 + nu_t 1 + Nvert 1 +
    # cell metrics 9 + face metrics 30 and writes Rhs 3 (400 B/cell); the fused update adds pUcont 3 +
    # dP 3 in and Ucont 3 out (72 B/cell).  Timed as one region: face pass, its periodic fix, both cell passes.
    "formfunction2": (400 + 72, "cell"),
    # the face pass as it runs on this workload (second order, face metrics rebuilt from the cell
    # metrics, flux divergence formed in the same kernel): ucont 24 + ucat 24 + nu_t 8 + nvert 8 +
    # cell metrics 80; csi.Fp, eta.Fp, zet.Fp out 24
    "face_fluxes": (24 + 24 + 8 + 8 + 80 + 24, "point"),
    # (the path with flux planes) 18 fluxes 144; Fp out 24
    "cell_fp": (144 + 24, "point"),
    # (after the fused face pass) csi.Fp, eta.Fp, zet.Fp 24 + iaj/jaj/kaj 24 + nvert 8 + dP 24 + pUcont 24;
    # Ucont out 24
    "cell_rhs": (24 + 24 + 8 + 24 + 24 + 24, "point"),
    # ucont 24 + csi/eta/zet 72 + nvert 8; ucat out 24 (interior cells)
    "contra2cart": (24 + 72 + 8 + 24, "cell"),
    # ucat 24 + csi/eta/zet 72 + aj 8 + nvert 8; nu_t out 8
    "eddy_viscosity": (24 + 72 + 8 + 8 + 8, "point"),
    # p 8 + nvert 8 + 3 face metric families 240; dP out 24
    "pressure_gradient": (8 + 8 + 240 + 24, "point"),
    # FP32 smoother: 16 off-line coefficient / diagonal ratios as FP16 32 + diagonal 4 + b 4 + x 4;
    # line rhs out 4
    "line_rhs_fp32": (32 + 4 + 4 + 4 + 4, "line point"),
    # FP32: Thomas factors lo, 1/pivot, upper 12 + line rhs 4; x out 4
    "line_solve_fp32": (12 + 4 + 4, "line point"),
    # compact FP32 smoother (weak cross couplings): four axis ratios as FP16 8 + diagonal 4 + b 4 + x 4;
    # line rhs out 4
    "line_rhs_compact_fp32": (8 + 4 + 4 + 4 + 4, "line point"),
    # the same smoother in one pass (right-hand side formed inside the line solve): four axis ratios
    # 8 + diagonal 4 + b 4 + x 4 + Thomas factors 12; x out 4
    "line_relax_compact_fp32": (8 + 4 + 4 + 4 + 12 + 4, "line point"),
    # FP64: 19 coefficients 152 + x 8 + row mask 1; y out 8
    "poisson_apply_fp64": (152 + 8 + 1 + 8, "point"),
}


# HBM bytes per launch of the finest-level launch, from rocprofv3 FETCH_SIZE (x2, calibrated on
# copy kernels of known size) + WRITE_SIZE on this workload at 256x192x256
TRAFFIC_SOURCE = ("not measured by this run: rocprofv3 --pmc FETCH_SIZE (x2.0) + WRITE_SIZE of the same kernel on "
                  "this workload in a separate run, profiles/hbm_traffic_channel_256x192x256_pmc.md")
MEASURED_TRAFFIC = {
    "formfunction2": 4.200e9,
    "face_fluxes": 2.423e9,
    "cell_rhs": 1.723e9,
    "line_rhs_fp32": 0.382e9,
    "line_rhs_compact_fp32": 0.208e9,
    "line_relax_compact_fp32": 0.230e9,
    "poisson_apply_fp64": 2.504e9,
    "contra2cart": 1.880e9,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        cc = "gcc"
    return model, cc


def cpu_baseline(c, timed_steps=1):
    """The restated reference CPU path: the FP64 oracle (OpenMP, fixed-order
    reductions) built -O3 -march=native -fopenmp on this host, on the same
    channel as the GPU run.  One untimed step (it builds the multigrid
    hierarchy), then `timed_steps` steps on every OpenMP thread and one step
    on one thread.  Mcell-steps/s of full steps (RK4 + FGMRES/MG to 1e-8 +
    projection)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    flags = "-O3 -march=native -fopenmp -ffp-contract=off"
    try:
        path = po.build_native()
    except (OSError, subprocess.CalledProcessError) as e:
        log(f"native oracle build failed ({e}); using the portable build")
        path, flags = None, "-O3 -fopenmp -ffp-contract=off"
    po.load(path)
    threads = po.set_threads(0)
    nx, ny, nz = c.cells
    p = po.make_params(*c.da, per=c.per, bctype=c.bctype, ren=c.ren, dt=1e-3, inlet_flux=c.inlet_flux)
    o = po.Oracle(p, c.coor)
    o.set("ucont", c.ucont)
    o.contra2cart()
    o.set("ucat_old", o["ucat"])
    p.dt = 0.3 * o.min_dt()
    o.mean_k_area()
    o.fractional_step(tol=1e-8)  # hierarchy set-up; the first solve starts from zero
    cells = nx * ny * nz
    t0 = time.perf_counter()
    for _ in range(timed_steps):
        o.fractional_step(tol=1e-8)
    t_all = (time.perf_counter() - t0) / timed_steps
    log(f"cpu baseline: {threads} threads {t_all:.2f} s/step")
    po.set_threads(1)
    t0 = time.perf_counter()
    o.fractional_step(tol=1e-8)
    t_one = time.perf_counter() - t0
    po.set_threads(threads)
    log(f"cpu baseline: 1 thread {t_one:.2f} s/step")
    model, cc = _host_info()
    return {"value": round(cells / t_all / 1e6, 4), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"restated reference CPU path (FP64 oracle), channel {nx}x{ny}x{nz}, {timed_steps} full "
                      f"step(s) (RK4 + FGMRES/MG 1e-8 + projection) on {threads} threads after one untimed step: "
                      f"{t_all:.2f} s/step; one step on 1 thread: {t_one:.2f} s",
            "one_core_value": round(cells / t_one / 1e6, 4),
            "cpu_model": model, "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "compiler": cc, "flags": flags}


def poisson_comparison(vfs, ctx, reps=3):
    """ms per Poisson solve on the same right-hand side: plain FP64 V-cycle,
    FP32 V-cycle, FP32 V-cycle + learned warm start / first stage.  The
    right-hand side is that of one more step of the run (momentum stage, then
    PoissonRHS); the previous step's correction is the initial guess, as in
    the time loop, and each variant is also run from a zero guess.  For the
    first solve of each variant the correction is also applied (pressure
    update and projection of the same pre-projection Ucont), and the pressure,
    velocity and divergence are compared with the all-FP64 variant."""
    learned = importlib.import_module("vfs-wind_amd.learned")
    ctx.calc_k_flux()
    if ctx.params.les:
        ctx.contra2cart()
        ctx.eddy_viscosity()
    ctx.pressure_gradient()
    ctx.runge_kutta()
    ctx.poisson_rhs()
    guess = ctx.download("PHI")
    zero = np.zeros_like(guess)
    ucont_pre, p_pre = ctx.download("UCONT"), ctx.download("P")
    net = learned.make_net(ctx)
    variants = [("fp64_mg", 0, None), ("fp32_mg", 1, None)]
    if net is not None:
        variants.append(("fp32_mg_net", 1, net))
    out = {}
    fields = {}
    for start, g in (("previous_guess", guess), ("zero_guess", zero)):
        res = {}
        for name, fp32, nn in variants:
            ms, its, fb, rel = [], [], [], []
            for rep in range(reps):
                ctx.upload("PHI", g)
                ctx.sync()
                opts = vfs.default_solve_opts(tol=1e-8, mg_fp32=fp32, use_guess=1)
                st = nn.poisson_solve(opts) if nn is not None else ctx.poisson_solve(opts)
                ms.append(st.ms); its.append(st.iterations); fb.append(st.fallback); rel.append(st.relres)
                if rep == 0 and start == "previous_guess":
                    ctx.update_pressure()
                    ctx.projection()
                    fields[name] = (ctx.download("P"), ctx.download("UCONT"), ctx.divergence())
                    ctx.upload("UCONT", ucont_pre)
                    ctx.upload("P", p_pre)
                    ctx.contra2cart()
            res[name] = {"ms": round(float(np.median(ms)), 3), "iterations": int(np.median(its)),
                         "fallback": int(max(fb)), "relres": float(max(rel))}
        base = res["fp64_mg"]["ms"]
        for name in res:
            res[name]["speedup_vs_fp64_mg"] = round(base / res[name]["ms"], 3)
        out[start] = res
    # errors against the all-FP64 variant, after the pressure update and projection
    p64, u64, d64 = fields["fp64_mg"]
    err = {"fp64_mg": {"maxdiv": d64}}
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    for name, (p_, u_, d_) in fields.items():
        if name != "fp64_mg":
            err[name] = {"p_rel_vs_fp64": rel(p_, p64), "ucont_rel_vs_fp64": rel(u_, u64), "maxdiv": d_}
    out["errors_after_projection"] = err
    if net is not None:
        net.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=192)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--mg-fp64", action="store_true", help="plain FP64 V-cycle instead of FP32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-poisson-comparison", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for {args.gpus})", file=sys.stderr)
        sys.exit(2)

    vfs = importlib.import_module("vfs-wind_amd")
    cases = importlib.import_module("vfs-wind_amd.cases")
    log(f"building case {args.nx}x{args.ny}x{args.nz}")
    c = cases.channel(args.nx, args.ny, args.nz)
    params = cases.params_for(vfs, c, dt=1e-3)
    ctx = vfs.Context(params)
    ctx.upload("COOR", c.coor)
    ctx.form_metrics()
    ctx.mean_k_area()
    ctx.upload("UCONT", c.ucont)
    ctx.contra2cart()
    ctx.upload("UCAT_OLD", ctx.download("UCAT"))
    params.dt = 0.3 * ctx.min_dt()
    ctx.set_params(params)
    opts = vfs.default_solve_opts(tol=1e-8, mg_fp32=0 if args.mg_fp64 else 1)
    log(f"dt = {params.dt:.4e}; warmup {args.warmup} steps")
    for s in range(args.warmup):
        st = ctx.fractional_step(opts)
        log(f"warmup step {s}: poisson its {st.poisson_iterations} relres {st.poisson_relres:.2e} "
            f"maxdiv {st.maxdiv:.2e} ms {st.ms_total:.1f}")
    ctx.sync()
    # inside the timed region only the roofline entry's region carries HIP events
    ctx.profile(True, only=["formfunction2"])
    its, worst = [], 0.0
    t0 = time.perf_counter()
    for s in range(args.steps):
        st = ctx.fractional_step(opts)
        its.append(st.poisson_iterations)
        worst = max(worst, st.poisson_relres)
        log(f"step {s}: poisson its {st.poisson_iterations} relres {st.poisson_relres:.2e} maxdiv {st.maxdiv:.2e} "
            f"momentum {st.ms_momentum:.1f} ms poisson {st.ms_poisson:.1f} ms")
    ctx.sync()
    elapsed = time.perf_counter() - t0
    cells = args.nx * args.ny * args.nz
    ms_step = elapsed / args.steps * 1e3
    value = cells * args.steps / elapsed / 1e6

    units = {"point": params.mx * params.my * params.mz, "cell": cells,
             "line point": (params.mz - 2) * (params.my - 2) * ((params.mx - 1) // 2)}

    def read_kernels():
        out = {}
        for name, (bpu, unit) in ALGO.items():
            ms, cnt = ctx.profile_read(name)
            if cnt:
                avg = ms / cnt
                out[name] = {"total_ms": ms, "launches": cnt, "avg_ms": avg,
                             "GBs": bpu * units[unit] / (avg * 1e-3) / 1e9}
        return out

    kern = read_kernels()  # the roofline entry, timed inside the timed region
    # every timed kernel, on two more (untimed) steps
    ctx.profile(True)
    for s in range(2):
        ctx.fractional_step(opts)
    ctx.sync()
    kern_all = read_kernels()
    ctx.profile(False)
    # the roofline entry: one RK stage's Formfunction_2 (the largest share of the step); otherwise the
    # kernel with the largest total time
    dom = "formfunction2" if "formfunction2" in kern else (max(kern, key=lambda k: kern[k]["total_ms"]) if kern else None)
    roofline = None
    if dom:
        a = kern[dom]["GBs"]
        bpu, unit = ALGO[dom]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(a / HBM_PEAK_GBS, 4),
                    "traffic": MEASURED_TRAFFIC.get(dom) if (args.nx, args.ny, args.nz) == (256, 192, 256) else None,
                    "traffic_source": TRAFFIC_SOURCE if (args.nx, args.ny, args.nz) == (256, 192, 256) else None,
                    "bytes_per_unit": bpu, "unit_of_work": unit, "units_per_launch": units[unit],
                    "avg_launch_ms": round(kern[dom]["avg_ms"], 4)}
    log("kernels (2 untimed steps, every kernel timed): " +
        json.dumps({k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in kern_all.items()}))
    pois = None
    if not args.no_poisson_comparison:
        log("Poisson solve comparison (same right-hand side)")
        pois = poisson_comparison(vfs, ctx)
        log("poisson: " + json.dumps(pois))
    ctx.close()

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (FP64 oracle, bounded sample)")
        cpu = cpu_baseline(c)

    out = {
        "metric": "Mcell-steps/s at 1e-8 Poisson resid; ms/Poisson solve plain MG vs FP32+net",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"curvilinear channel {args.nx}x{args.ny}x{args.nz} cells, periodic x/z, no-slip walls, "
                        f"LES Cs=0.1, central 2nd order, RK4 + Poisson FGMRES(20)/"
                        f"{'FP64' if args.mg_fp64 else 'FP32'} V-cycle to 1e-8 + projection, dt=0.3 dt_CFL1",
            "cells": cells,
            "poisson_iterations_mean": sum(its) / len(its),
            "poisson_relres_worst": worst,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "poisson_ms_per_solve": pois,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
