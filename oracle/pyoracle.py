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
        ("max_cs", C.c_double), ("i_homo", C.c_int), ("j_homo", C.c_int), ("k_homo", C.c_int),
        ("inletprofile", C.c_int), ("pseudo_periodic", C.c_int), ("inflow_plane", C.c_void_p), ("ti", C.c_int),
    ]


METRIC_NAMES = ["csi", "eta", "zet", "aj", "icsi", "ieta", "izet", "iaj", "jcsi", "jeta", "jzet", "jaj",
                "kcsi", "keta", "kzet", "kaj", "cent", "nvert"]


class OrcMetrics(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in METRIC_NAMES] + [("coor", C.POINTER(C.c_double))]


class OrcTurbine(C.Structure):
    _fields_ = [("nx", C.c_double), ("ny", C.c_double), ("nz", C.c_double), ("xc", C.c_double), ("yc", C.c_double),
                ("zc", C.c_double), ("angvel", C.c_double), ("pitch", C.c_double), ("r_nacelle", C.c_double),
                ("refangle", C.c_double), ("shen", C.c_int), ("prandtl", C.c_int), ("shen1", C.c_int),
                ("r_rotor", C.c_double), ("tsr", C.c_double), ("num_blade", C.c_int), ("a_shen", C.c_double),
                ("b_shen", C.c_double), ("c_shen", C.c_double), ("correction_shen1", C.c_double),
                ("correction3d_cl", C.c_int), ("c0_cl", C.c_double), ("smoothforce", C.c_int)]


class OrcFoil(C.Structure):
    _fields_ = [("n_ac", C.c_int), ("r_ac", C.POINTER(C.c_double)), ("twist", C.POINTER(C.c_double)),
                ("chord", C.POINTER(C.c_double)), ("n_cl", C.c_int), ("ang_cl", C.POINTER(C.c_double)),
                ("cl", C.POINTER(C.c_double)), ("n_cd", C.c_int), ("ang_cd", C.POINTER(C.c_double)),
                ("cd", C.POINTER(C.c_double)), ("r_beg", C.c_double), ("r_end", C.c_double)]


_lib = None


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR), "-B" if force else "all"], check=True,
                       capture_output=True, text=True)
    return LIB_PATH


def build_native() -> Path:
    """Builds liborc_native.so (-march=native) for the host this runs on."""
    path = ORACLE_DIR / "liborc_native.so"
    subprocess.run(["make", "-C", str(ORACLE_DIR), "-B", path.name], check=True, capture_output=True, text=True)
    return path


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """The oracle library (liborc.so unless `path` names another build of the
    same sources); the first call fixes which one the process uses."""
    global _lib
    if _lib is not None:
        return _lib
    if path is None:
        if not LIB_PATH.exists():
            build()
        path = LIB_PATH
    lib = C.CDLL(str(path))
    P, M, D = C.POINTER(OrcParams), C.POINTER(OrcMetrics), C.POINTER(C.c_double)
    sig = {
        "orc_form_metrics": ([P, D, M], None),
        "orc_contra2cart": ([P, M, D, D, D], None),
        "orc_form_bcs": ([P, M, D, D, D, C.c_int], None),
        "orc_eddy_viscosity": ([P, M, D, D], None),
        "orc_eddy_viscosity_cs": ([P, M, D, D, D], None),
        "orc_smagorinsky_constant": ([P, M, D, C.c_int, C.c_int, C.c_int, D], None),
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
        "orc_inlet_area": ([P, M], C.c_double),
        "orc_inflow_flux": ([P, M, D, D], C.c_double),
        "orc_pseudo_periodic_plane": ([P, D, D], None),
        "orc_formfunction_snes": ([P, M, D, D, D, D, D, D, D, D, D, D], None),
        "orc_set_threads": ([C.c_int], C.c_int),
        "orc_arbitrary_rotate": ([D, C.c_double, D, D], None),
        "orc_alm_elements": ([C.c_int, C.c_int, D, D, D], None),
        "orc_alm_rotate": ([C.POINTER(OrcTurbine), C.c_double, C.c_int, D, D], None),
        "orc_alm_refline": ([C.POINTER(OrcTurbine), C.c_int, D, D], None),
        "orc_alm_twist_chord": ([C.POINTER(OrcTurbine), C.c_int, D, C.POINTER(OrcFoil), C.c_int, D, D], None),
        "orc_alm_preprocess": ([P, M, C.c_int, D, C.POINTER(C.c_int)], None),
        "orc_alm_ulagr": ([P, M, D, C.c_int, D, C.POINTER(C.c_int), D], None),
        "orc_alm_flagr": ([C.POINTER(OrcTurbine), C.c_int, C.c_int, D, D, D, D, D, C.POINTER(OrcFoil), C.c_int,
                           D, D, D, D], None),
        "orc_alm_feul": ([P, M, C.c_int, D, C.POINTER(C.c_int), D, D, C.c_int, C.c_double, D], None),
        "orc_alm_loads": ([C.POINTER(OrcTurbine), C.c_int, D, D, D, D, D], None),
        "orc_turbine_angvel": ([C.c_int, D, C.c_double, C.c_double, C.c_int, C.c_int, D], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def set_threads(n: int = 0) -> int:
    """OpenMP team size of the oracle (n > 0 sets it); returns the current one."""
    return load().orc_set_threads(int(n))


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(mx, my, mz, *, per=(1, 0, 1), bctype=(100, 100, 1, 1, 100, 100), les=1, second_order=1,
                laplacian=1, ren=2800.0, dt=1e-3, st=1.0, cs2=0.01, inlet_flux=-1.0, poisson_threshold=0.1,
                flux_in_sum=0.0, max_cs=0.5, homo=(0, 0, 0), inletprofile=1, pseudo=0) -> OrcParams:
    p = OrcParams()
    p.mx, p.my, p.mz = mx, my, mz
    p.per_i, p.per_j, p.per_k = per
    for q in range(6):
        p.bctype[q] = int(bctype[q])
    p.les, p.second_order, p.laplacian = les, second_order, laplacian
    p.ren, p.dt, p.st, p.cs2 = ren, dt, st, cs2
    p.inlet_flux, p.poisson_threshold, p.flux_in_sum = inlet_flux, poisson_threshold, flux_in_sum
    p.mean_k_area, p.mean_k_flux = 0.0, 0.0
    p.max_cs = max_cs
    p.i_homo, p.j_homo, p.k_homo = (int(h) for h in homo)
    p.inletprofile, p.pseudo_periodic = int(inletprofile), int(pseudo)
    p.inflow_plane = None
    p.ti = 0
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

    def eddy_viscosity(self, cs: np.ndarray | None = None):
        if cs is None:
            self.lib.orc_eddy_viscosity(C.byref(self.p), C.byref(self.cm), self._f("ucat"), self._f("nu_t"))
        else:
            cs = np.ascontiguousarray(cs, dtype=np.float64)
            self.lib.orc_eddy_viscosity_cs(C.byref(self.p), C.byref(self.cm), self._f("ucat"), _ptr(cs),
                                           self._f("nu_t"))

    def smagorinsky_constant(self, ti=2, tistart=0, rstart=0) -> np.ndarray:
        """Compute_Smagorinsky_Constant_1 on the current Ucat: lCs per point."""
        cs = np.zeros(self.shape)
        self.lib.orc_smagorinsky_constant(C.byref(self.p), C.byref(self.cm), self._f("ucat"), int(ti), int(tistart),
                                          int(rstart), _ptr(cs))
        return cs

    def pressure_gradient(self):
        self.lib.orc_pressure_gradient(C.byref(self.p), C.byref(self.cm), self._f("p"), self._f("dp"))

    def formfunction2(self, scale=1.0):
        self.lib.orc_formfunction2(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("ucat"),
                                   self._f("nu_t"), self._f("rhs"), scale)

    def runge_kutta(self):
        self.lib.orc_runge_kutta(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("ucat"),
                                 self._f("ucat_old"), self._f("nu_t"), self._f("dp"),
                                 C.cast(None, C.POINTER(C.c_double)))

    def set_inflow_plane(self, plane: np.ndarray | None):
        """user->ucont_plane: (my, mx, 3) Cartesian velocity of the inlet plane."""
        if plane is None:
            self._plane, self.p.inflow_plane = None, None
            return
        self._plane = np.ascontiguousarray(plane, dtype=np.float64).reshape(self.shape[1], self.shape[2], 3)
        self.p.inflow_plane = self._plane.ctypes.data

    def pseudo_periodic_plane(self) -> np.ndarray:
        out = np.zeros((self.shape[1], self.shape[2], 3))
        self.lib.orc_pseudo_periodic_plane(C.byref(self.p), self._f("ucat"), _ptr(out))
        return out

    def inflow_flux(self):
        return self.lib.orc_inflow_flux(C.byref(self.p), C.byref(self.cm), self._f("ucont"), self._f("ucat"))

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


# ---- actuator line ---------------------------------------------------------
def _iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def make_turbine(nx, ny, nz, xc, yc, zc, angvel, pitch=0.0, r_nacelle=0.0, refangle=20.0, shen=0, prandtl=0, shen1=0,
                 r_rotor=0.0, tsr=0.0, num_blade=3, a_shen=0.125, b_shen=21.0, c_shen=0.1, correction_shen1=1.0,
                 correction3d_cl=0, c0_cl=1.0, smoothforce=0) -> OrcTurbine:
    """Turbine parameters; the correction options default to main.c:161, 224-238."""
    rr = float(np.sqrt(nx * nx + ny * ny + nz * nz))
    return OrcTurbine(nx / rr, ny / rr, nz / rr, xc, yc, zc, angvel, pitch, r_nacelle, refangle, shen, prandtl, shen1,
                      r_rotor, tsr, num_blade, a_shen, b_shen, c_shen, correction_shen1, correction3d_cl, c0_cl,
                      smoothforce)


def alm_loads(t: OrcTurbine, cent, flagr, ulagr, dA):
    """Calc_forces_ACL sums: (A_Sum, F_Sum, P_Sum, U_mean, Fx, Fy, Fz, M_Sum)."""
    lib = load()
    cent, flagr, ulagr, dA = (np.ascontiguousarray(a, dtype=np.float64) for a in (cent, flagr, ulagr, dA))
    out = np.zeros(8)
    lib.orc_alm_loads(C.byref(t), int(dA.size), _ptr(cent), _ptr(flagr), _ptr(ulagr), _ptr(dA), _ptr(out))
    return out


class Foils:
    """ctypes array of OrcFoil that keeps its numpy tables alive."""

    def __init__(self, foils: list):
        self.keep = []
        self.arr = (OrcFoil * max(1, len(foils)))()
        self.n = len(foils)
        for q, f in enumerate(foils):
            a = [np.ascontiguousarray(f[k], dtype=np.float64) for k in ("r", "twist", "chord", "ang_cl", "cl", "ang_cd", "cd")]
            self.keep += a
            o = self.arr[q]
            o.n_ac, o.r_ac, o.twist, o.chord = len(a[0]), _ptr(a[0]), _ptr(a[1]), _ptr(a[2])
            o.n_cl, o.ang_cl, o.cl = len(a[3]), _ptr(a[3]), _ptr(a[4])
            o.n_cd, o.ang_cd, o.cd = len(a[5]), _ptr(a[5]), _ptr(a[6])
            o.r_beg, o.r_end = float(a[0][0]), float(a[0][-1])


def alm_elements(nodes: np.ndarray):
    """nodes (nb, nvb, 3) -> element centres (n, 3) and lengths (n,)."""
    lib = load()
    nb, nvb = nodes.shape[:2]
    x = np.ascontiguousarray(nodes, dtype=np.float64)
    cent, dA = np.zeros((nb * (nvb - 1), 3)), np.zeros(nb * (nvb - 1))
    lib.orc_alm_elements(nb, nvb, _ptr(x), _ptr(cent), _ptr(dA))
    return cent, dA


def alm_rotate(t: OrcTurbine, dt: float, nodes: np.ndarray):
    """rotor_Rot node loop: returns rotated nodes and node velocities."""
    lib = load()
    x = np.ascontiguousarray(nodes, dtype=np.float64).copy()
    u = np.zeros_like(x)
    lib.orc_alm_rotate(C.byref(t), dt, x.size // 3, _ptr(x), _ptr(u))
    return x, u


def alm_refline(t: OrcTurbine, nodes: np.ndarray):
    lib = load()
    x = np.ascontiguousarray(nodes, dtype=np.float64)
    r = np.zeros_like(x)
    lib.orc_alm_refline(C.byref(t), x.size // 3, _ptr(x), _ptr(r))
    return r


def alm_twist_chord(t: OrcTurbine, cent: np.ndarray, foils: Foils):
    lib = load()
    c = np.ascontiguousarray(cent, dtype=np.float64)
    tw, ch = np.zeros(len(c)), np.zeros(len(c))
    lib.orc_alm_twist_chord(C.byref(t), len(c), _ptr(c), foils.arr, foils.n, _ptr(tw), _ptr(ch))
    return tw, ch


def alm_flagr(t: OrcTurbine, nb: int, nvb: int, cent, ulagr, unode, twist, chord, foils: Foils, cl=None, cd=None):
    lib = load()
    n = nb * (nvb - 1)
    cent, ulagr, unode, twist, chord = (np.ascontiguousarray(a, dtype=np.float64) for a in (cent, ulagr, unode, twist, chord))
    aoa = np.zeros(n)
    cl = np.zeros(n) if cl is None else np.ascontiguousarray(cl, dtype=np.float64).copy()
    cd = np.zeros(n) if cd is None else np.ascontiguousarray(cd, dtype=np.float64).copy()
    f = np.zeros((n, 3))
    lib.orc_alm_flagr(C.byref(t), nb, nvb, _ptr(cent), _ptr(ulagr), _ptr(unode), _ptr(twist), _ptr(chord), foils.arr,
                      foils.n, _ptr(aoa), _ptr(cl), _ptr(cd), _ptr(f))
    return f, aoa, cl, cd


def _snes_methods():
    def formfunction_snes(self, u, ucont_o, rhs_o, feul=None):
        """FormFunction_SNES at u; leaves ucont/ucat at the wall-zeroed u."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        uo = np.ascontiguousarray(ucont_o, dtype=np.float64)
        ro = np.ascontiguousarray(rhs_o, dtype=np.float64)
        fe = None if feul is None else np.ascontiguousarray(feul, dtype=np.float64)
        r = np.zeros(self.shape + (3,))
        self.lib.orc_formfunction_snes(C.byref(self.p), C.byref(self.cm), _ptr(u), _ptr(uo), _ptr(ro), self._f("dp"),
                                       None if fe is None else _ptr(fe), self._f("ucont"), self._f("ucat"),
                                       self._f("ucat_old"), self._f("nu_t"), _ptr(r))
        return r

    Oracle.formfunction_snes = formfunction_snes


_snes_methods()


def _alm_methods():
    def alm_preprocess(self, cent):
        c = np.ascontiguousarray(cent, dtype=np.float64)
        box = np.zeros((len(c), 6), dtype=np.int32)
        self.lib.orc_alm_preprocess(C.byref(self.p), C.byref(self.cm), len(c), _ptr(c), _iptr(box))
        return box

    def alm_ulagr(self, cent, box):
        c = np.ascontiguousarray(cent, dtype=np.float64)
        b = np.ascontiguousarray(box, dtype=np.int32)
        u = np.zeros((len(c), 3))
        self.lib.orc_alm_ulagr(C.byref(self.p), C.byref(self.cm), self._f("ucat"), len(c), _ptr(c), _iptr(b), _ptr(u))
        return u

    def alm_feul(self, cent, box, flagr, dA, df=10, halfwidth=4.0):
        c = np.ascontiguousarray(cent, dtype=np.float64)
        b = np.ascontiguousarray(box, dtype=np.int32)
        f = np.ascontiguousarray(flagr, dtype=np.float64)
        a = np.ascontiguousarray(dA, dtype=np.float64)
        feul = np.zeros(self.shape + (3,))
        self.lib.orc_alm_feul(C.byref(self.p), C.byref(self.cm), len(c), _ptr(c), _iptr(b), _ptr(f), _ptr(a), int(df),
                              float(halfwidth), _ptr(feul))
        return feul

    Oracle.alm_preprocess = alm_preprocess
    Oracle.alm_ulagr = alm_ulagr
    Oracle.alm_feul = alm_feul


_alm_methods()


def turbine_angvel(mode, prm, dt, uref, ti, tistart, state):
    """Calc_turbineangvel: state = [angvel, Tipspeedratio, Torque_aero, Torque_generator,
    ang_axis] updated and returned; prm as orc_turbine_angvel documents."""
    lib = load()
    prm = np.ascontiguousarray(prm, dtype=np.float64)
    st = np.ascontiguousarray(state, dtype=np.float64).copy()
    lib.orc_turbine_angvel(int(mode), _ptr(prm), float(dt), float(uref), int(ti), int(tistart), _ptr(st))
    return st
