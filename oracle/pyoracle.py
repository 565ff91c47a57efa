"""ctypes binding of the FP64 CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liborc.so"


class OrcParams(C.Structure):
    _fields_ = [
        ("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
        ("per_i", C.c_int), ("per_j", C.c_int), ("per_k", C.c_int),
        ("bctype", C.c_int * 6),
        ("les", C.c_int), ("second_order", C.c_int), ("laplacian", C.c_int),
        ("ren", C.c_double), ("dt", C.c_double), ("st", C.c_double), ("cs2", C.c_double),
        ("inlet_flux", C.c_double), ("mean_k_area", C.c_double), ("mean_k_flux", C.c_double),
        ("poisson_threshold", C.c_double), ("flux_in_sum", C.c_double),
    ]


METRIC_NAMES = ["csi", "eta", "zet", "aj", "icsi", "ieta", "izet", "iaj", "jcsi", "jeta", "jzet", "jaj",
                "kcsi", "keta", "kzet", "kaj", "cent", "nvert"]


class OrcMetrics(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in METRIC_NAMES] + [("coor", C.POINTER(C.c_double))]


_lib = None


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR), "-B" if force else "all"], check=True,
                       capture_output=True, text=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    P, M, D = C.POINTER(OrcParams), C.POINTER(OrcMetrics), C.POINTER(C.c_double)
    sig = {
        "orc_form_metrics": ([P, D, M], None),
        "orc_contra2cart": ([P, M, D, D, D], None),
        "orc_form_bcs": ([P, M, D, D, D, C.c_int], None),
        "orc_eddy_viscosity": ([P, M, D, D], None),
        "orc_calc_k_flux": ([P, M, D], C.c_double),
        "orc_mean_k_area": ([P, M], C.c_double),
        "orc_pressure_gradient": ([P, M, D, D], None),
        "orc_formfunction2": ([P, M, D, D, D, D, C.c_double], None),
        "orc_runge_kutta": ([P, M, D, D, D, D, D, D], None),
        "orc_min_dt": ([P, M, D], C.c_double),
        "orc_poisson_rhs": ([P, M, D, D], None),
        "orc_poisson_apply": ([P, M, D, D], None),
        "orc_poisson_solve": ([P, M, D, D, C.c_double, C.c_int, C.c_int, C.c_int, D], C.c_int),
        "orc_remove_mean": ([P, M, D], None),
        "orc_update_pressure": ([P, M, D, D], None),
        "orc_projection": ([P, M, D, D, D, D], None),
        "orc_divergence": ([P, M, D], C.c_double),
        "orc_fractional_step": ([P, M, D, D, D, D, D, D, D, C.c_double, D, D], C.c_int),
        "orc_periodic_fix": ([P, D, C.c_int], None),
        "orc_dfunc_2h": ([C.c_double], C.c_double),
        "orc_dfunc_s3h": ([C.c_double], C.c_double),
        "orc_dfunc_s4h": ([C.c_double], C.c_double),
        "orc_dfunc_exp": ([C.c_double, C.c_double], C.c_double),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(mx, my, mz, *, per=(1, 0, 1), bctype=(100, 100, 1, 1, 100, 100), les=1, second_order=1,
                laplacian=1, ren=2800.0, dt=1e-3, st=1.0, cs2=0.01, inlet_flux=-1.0, poisson_threshold=0.1,
                flux_in_sum=0.0) -> OrcParams:
    p = OrcParams()
    p.mx, p.my, p.mz = mx, my, mz
    p.per_i, p.per_j, p.per_k = per
    for q in range(6):
        p.bctype[q] = int(bctype[q])
    p.les, p.second_order, p.laplacian = les, second_order, laplacian
    p.ren, p.dt, p.st, p.cs2 = ren, dt, st, cs2
    p.inlet_flux, p.poisson_threshold, p.flux_in_sum = inlet_flux, poisson_threshold, flux_in_sum
    p.mean_k_area, p.mean_k_flux = 0.0, 0.0
    return p


class Oracle:
    """Host arrays of one DMDA block and the restated reference functions."""

    def __init__(self, params: OrcParams, coor: np.ndarray, nvert: np.ndarray | None = None):
        self.lib = load()
        self.p = params
        self.shape = (params.mz, params.my, params.mx)
        n3 = self.shape + (3,)
        self.coor = np.ascontiguousarray(coor, dtype=np.float64)
        self.m = {}
        for name in METRIC_NAMES:
            vec = name not in ("aj", "iaj", "jaj", "kaj", "nvert")
            self.m[name] = np.zeros(n3 if vec else self.shape)
        if nvert is not None:
            self.m["nvert"][...] = nvert
        self.cm = OrcMetrics(**{n: _ptr(self.m[n]) for n in METRIC_NAMES}, coor=_ptr(self.coor))
        self.lib.orc_form_metrics(C.byref(self.p), _ptr(self.coor), C.byref(self.cm))
        self.fields = {
            "ucont": np.zeros(n3), "ucat": np.zeros(n3), "ucat_old": np.zeros(n3), "nu_t": np.zeros(self.shape),
            "p": np.zeros(self.shape), "phi": np.zeros(self.shape), "dp": np.zeros(n3), "rhs": np.zeros(n3),
            "b": np.zeros(self.shape),
        }

    def __getitem__(self, k):
        return self.fields[k]

    def set(self, name: str, arr: np.ndarray):
        self.fields[name][...] = arr

    def _f(self, name):
        return _ptr(self.fields[name])

    def mean_k_area(self):
        self.p.mean_k_area = self.lib.orc_mean_k_area(C.byref(self.p), C.byref(self.cm))
        return self.p.mean_k_area

    def calc_k_flux(self):
        self.p.mean_k_flux = self.lib.orc_calc_k_flux(C.byref(self.p), C.byref(self.cm), self._f("ucont"))
        return self.p.mean_k_flux

    def contra2cart(self):
        self.lib.orc_contra2cart(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("ucat"),
                                 self._f("ucat_old"))

    def form_bcs(self, outflow_scale=0):
        self.lib.orc_form_bcs(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("ucat"),
                              self._f("ucat_old"), int(outflow_scale))

    def eddy_viscosity(self):
        self.lib.orc_eddy_viscosity(C.byref(self.p), C.byref(self.cm), self._f("ucat"), self._f("nu_t"))

    def pressure_gradient(self):
        self.lib.orc_pressure_gradient(C.byref(self.p), C.byref(self.cm), self._f("p"), self._f("dp"))

    def formfunction2(self, scale=1.0):
        self.lib.orc_formfunction2(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("ucat"),
                                   self._f("nu_t"), self._f("rhs"), scale)

    def runge_kutta(self):
        self.lib.orc_runge_kutta(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("ucat"),
                                 self._f("ucat_old"), self._f("nu_t"), self._f("dp"),
                                 C.cast(None, C.POINTER(C.c_double)))

    def min_dt(self):
        return self.lib.orc_min_dt(C.byref(self.p), C.byref(self.cm), self._f("ucont"))

    def poisson_rhs(self):
        self.lib.orc_poisson_rhs(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("b"))

    def poisson_apply(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.shape)
        self.lib.orc_poisson_apply(C.byref(self.p), C.byref(self.cm), _ptr(x), _ptr(y))
        return y

    def poisson_solve(self, tol=1e-8, restart=20, maxit=400, levels=6):
        rel = C.c_double()
        it = self.lib.orc_poisson_solve(C.byref(self.p), C.byref(self.cm), self._f("b"), self._f("phi"), tol,
                                        restart, maxit, levels, C.byref(rel))
        return it, rel.value

    def update_pressure(self):
        self.lib.orc_update_pressure(C.byref(self.p), C.byref(self.cm), self._f("p"), self._f("phi"))

    def projection(self):
        self.lib.orc_projection(C.byref(self.p), C.byref(self.cm), self._f("phi"), self._f("ucont"),
                                self._f("ucat"), self._f("ucat_old"))

    def divergence(self):
        return self.lib.orc_divergence(C.byref(self.p), C.byref(self.cm), self._f("ucont"))

    def fractional_step(self, tol=1e-8):
        md, rel = C.c_double(), C.c_double()
        it = self.lib.orc_fractional_step(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("ucat"),
                                          self._f("ucat_old"), self._f("nu_t"), self._f("p"), self._f("phi"),
                                          self._f("dp"), tol, C.byref(md), C.byref(rel))
        return it, md.value, rel.value
