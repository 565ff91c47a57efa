"""Readers of the actuator-line input files and the turbine set-up.

They mirror what the reference reads before the time loop:
  Turbine.inp   main.c:2042-2101       (header line, 16 numbers per turbine)
  acldata000    ACL_read_ucd           rotor_model.c:683-830
  FOILnn        airfoil_ACL            rotor_model.c:2028-2048
  CLnn, CDnn    airfoil_ACL            rotor_model.c:2050-2087
  Urefdata000   disk_read_ucd          rotor_model.c:54-365 (OneDmZ = 1)
Lengths are divided by the same reference lengths the reference divides by.
"""
from __future__ import annotations

import importlib
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

TURBINE_KEYS = ["nx_tb", "ny_tb", "nz_tb", "x_c", "y_c", "z_c", "indf_axis", "Tipspeedratio", "J_rotation",
                "r_rotor", "CP_max", "TSR_max", "angvel_fixed", "Torque_generator", "pitch", "CT"]


def _numbers(text: str):
    for tok in text.split():
        yield tok


def read_turbine_inp(path, n_turbines: int = 1, chact_leng: float = 1.0) -> list[dict]:
    """Turbine.inp (main.c:2042-2101): axis normalised, hub centre divided by -chact_leng."""
    lines = Path(path).read_text().splitlines()
    toks = " ".join(lines[1:]).split()
    out = []
    for t in range(n_turbines):
        v = [float(x) for x in toks[16 * t:16 * (t + 1)]]
        if len(v) != 16:
            raise ValueError(f"{path}: turbine {t} has {len(v)} of 16 numbers")
        d = dict(zip(TURBINE_KEYS, v))
        rr = math.sqrt(d["nx_tb"] ** 2 + d["ny_tb"] ** 2 + d["nz_tb"] ** 2)
        d["nx_tb"], d["ny_tb"], d["nz_tb"] = d["nx_tb"] / rr, d["ny_tb"] / rr, d["nz_tb"] / rr
        d["x_c"], d["y_c"], d["z_c"] = d["x_c"] / chact_leng, d["y_c"] / chact_leng, d["z_c"] / chact_leng
        out.append(d)
    return out


def read_acldata(path, num_blade: int, reflength: float, hub=(0.0, 0.0, 0.0)) -> np.ndarray:
    """acldata000 (rotor_model.c:686-830): per blade a node count then x y z
    rows; divided by reflength and moved to the hub.  Returns (num_blade, n_v, 3)."""
    toks = Path(path).read_text().split()
    pos = 0
    n_v = int(toks[pos]); pos += 1
    nodes = np.empty((num_blade, n_v, 3))
    for b in range(num_blade):
        if b != 0:
            pos += 1  # the next blade's count line
        for j in range(n_v):
            for c in range(3):
                nodes[b, j, c] = float(toks[pos]) / reflength + hub[c]
                pos += 1
    return nodes


def _skip2(path):
    lines = Path(path).read_text().splitlines()
    return " ".join(lines[2:]).split()


def read_foil(path, reflength_wt: float):
    """FOILnn: two header lines, a count, then r twist chord rows (r, chord / reflength_wt)."""
    toks = _skip2(path)
    n = int(toks[0])
    v = np.array([float(x) for x in toks[1:1 + 3 * n]]).reshape(n, 3)
    return v[:, 0] / reflength_wt, v[:, 1].copy(), v[:, 2] / reflength_wt


def read_polar(path):
    """CLnn / CDnn: two header lines, a count, then angle (deg) value rows."""
    toks = _skip2(path)
    n = int(toks[0])
    v = np.array([float(x) for x in toks[1:1 + 2 * n]]).reshape(n, 2)
    return v[:, 0].copy(), v[:, 1].copy()


def read_foils(directory, num_foiltype: int, reflength_wt: float) -> list[dict]:
    d = Path(directory)
    out = []
    for f in range(num_foiltype):
        r, tw, ch = read_foil(d / f"FOIL{f:02d}", reflength_wt)
        acl, cl = read_polar(d / f"CL{f:02d}")
        acd, cd = read_polar(d / f"CD{f:02d}")
        out.append({"r": r, "twist": tw, "chord": ch, "ang_cl": acl, "cl": cl, "ang_cd": acd, "cd": cd})
    return out


def read_ucd_disk(path, reflength: float, hub, axis, r_rotor: float, reflength_wt: float, loc_refvel: float = 1.0,
                  one_d_upstream: bool = True):
    """Urefdata000 (disk_read_ucd, rotor_model.c:54-365): UCD triangles; nodes
    divided by reflength, moved to the hub and, for the U_ref disk, moved
    loc_refvel * 2R upstream along the axis.  Returns (nodes, tris, cent, dA)."""
    lines = Path(path).read_text().splitlines()
    toks = " ".join(lines[3:]).split()
    n_v, n_e = int(toks[0]), int(toks[1])
    pos = 5
    x = np.empty((n_v, 3))
    R = r_rotor / reflength_wt
    for i in range(n_v):
        for c in range(3):
            x[i, c] = float(toks[pos + 1 + c]) / reflength + hub[c]
        pos += 4
        if one_d_upstream:
            rr = loc_refvel * 2.0 * R
            x[i] -= rr * np.asarray(axis)
    tris = np.empty((n_e, 3), dtype=np.int64)
    for e in range(n_e):
        tris[e] = [int(toks[pos + 3]) - 1, int(toks[pos + 4]) - 1, int(toks[pos + 5]) - 1]
        pos += 6
    d12 = x[tris[:, 1]] - x[tris[:, 0]]
    d13 = x[tris[:, 2]] - x[tris[:, 0]]
    nf = np.stack([d12[:, 1] * d13[:, 2] - d12[:, 2] * d13[:, 1],
                   -d12[:, 0] * d13[:, 2] + d12[:, 2] * d13[:, 0],
                   d12[:, 0] * d13[:, 1] - d12[:, 1] * d13[:, 0]], axis=1)
    dA = np.sqrt((nf ** 2).sum(axis=1)) / 2.0
    cent = (x[tris[:, 0]] + x[tris[:, 1]] + x[tris[:, 2]]) / 3.0
    return x, tris, cent, dA


@dataclass
class TurbineSetup:
    """Everything vfs_alm_create takes for one turbine."""

    nx: float
    ny: float
    nz: float
    xc: float
    yc: float
    zc: float
    angvel: float
    pitch: float
    r_nacelle: float
    refangle: float
    rotor_model: int
    deltafunc: int
    halfwidth: float
    nodes: np.ndarray                    # (num_blade, n_v, 3)
    foils: list = field(default_factory=list)

    def turbine_struct(self, vfs=None):
        vfs = vfs or importlib.import_module("vfs-wind_amd")
        t = vfs.Turbine()
        for k in ("nx", "ny", "nz", "xc", "yc", "zc", "angvel", "pitch", "r_nacelle", "refangle", "halfwidth"):
            setattr(t, k, float(getattr(self, k)))
        t.rotor_model, t.deltafunc = int(self.rotor_model), int(self.deltafunc)
        return t


def load_deck(directory, *, num_blade=3, num_foiltype=1, reflength_wt=1.0, r_nacelle=0.0, refangle=20.0,
              rotor_model=6, deltafunc=10, halfwidth=4.0, chact_leng=1.0) -> TurbineSetup:
    """Read a turbine deck directory (Turbine.inp, acldata000, FOIL/CL/CD) the way
    the reference's rotor_model 6 set-up does (main.c:2229-2238), with
    -fixturbineangvel (angvel = angvel_fixed)."""
    d = Path(directory)
    t = read_turbine_inp(d / "Turbine.inp", 1, chact_leng)[0]
    nodes = read_acldata(d / "acldata000", num_blade, reflength_wt, (t["x_c"], t["y_c"], t["z_c"]))
    foils = read_foils(d, num_foiltype, reflength_wt)
    return TurbineSetup(t["nx_tb"], t["ny_tb"], t["nz_tb"], t["x_c"], t["y_c"], t["z_c"], t["angvel_fixed"], t["pitch"],
                        r_nacelle / reflength_wt, refangle, rotor_model, deltafunc, halfwidth, nodes, foils)
