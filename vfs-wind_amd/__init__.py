"""vfs-wind_amd: MI355X (gfx950) fractional-step core of VFS-Wind.

Python binding of the C ABI in include/vfs_wind.h (libvfswind.so, built by
build_lib.py).  The binding only moves host arrays in and out of a device
context and calls the ABI; every computation runs in the HIP library.  It
raises at import time if the library is missing: there is no CPU fallback.

The package directory name contains a hyphen, so import it with
``importlib.import_module("vfs-wind_amd")``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libvfswind.so"

VFS_OK, VFS_EINVAL, VFS_EHIP, VFS_ENOMEM, VFS_ENOCONV = 0, 1, 2, 3, 4

# vfs_field enumeration (include/vfs_wind.h)
FIELDS = [
    "COOR", "NVERT", "CSI", "ETA", "ZET", "AJ", "ICSI", "IETA", "IZET", "IAJ",
    "JCSI", "JETA", "JZET", "JAJ", "KCSI", "KETA", "KZET", "KAJ", "CENT",
    "UCONT", "UCAT", "UCAT_OLD", "NU_T", "P", "PHI", "DP", "RHS", "B", "FEUL", "UCONT_O", "RHS_O", "CS",
]
FIELD = {name: i for i, name in enumerate(FIELDS)}
VECTOR_FIELDS = {"COOR", "CSI", "ETA", "ZET", "ICSI", "IETA", "IZET", "JCSI", "JETA", "JZET",
                 "KCSI", "KETA", "KZET", "CENT", "UCONT", "UCAT", "UCAT_OLD", "DP", "RHS", "FEUL",
                 "UCONT_O", "RHS_O"}
METRIC_FIELDS = ["CSI", "ETA", "ZET", "AJ", "ICSI", "IETA", "IZET", "IAJ", "JCSI", "JETA", "JZET",
                 "JAJ", "KCSI", "KETA", "KZET", "KAJ", "CENT"]


class Params(C.Structure):
    """vfs_params: every option a hot-path kernel reads."""

    _fields_ = [
        ("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
        ("per_i", C.c_int), ("per_j", C.c_int), ("per_k", C.c_int),
        ("bctype", C.c_int * 6),
        ("les", C.c_int), ("second_order", C.c_int), ("laplacian", C.c_int),
        ("ren", C.c_double), ("dt", C.c_double), ("st", C.c_double),
        ("cs2", C.c_double), ("inlet_flux", C.c_double),
        ("poisson_threshold", C.c_double), ("flux_in_sum", C.c_double),
        ("max_cs", C.c_double), ("dynamic_freq", C.c_int),
        ("i_homo_filter", C.c_int), ("j_homo_filter", C.c_int), ("k_homo_filter", C.c_int),
        ("inletprofile", C.c_int), ("pseudo_periodic", C.c_int),
    ]


class SolveOpts(C.Structure):
    _fields_ = [
        ("tol", C.c_double), ("restart", C.c_int), ("maxit", C.c_int), ("levels", C.c_int),
        ("pre", C.c_int), ("post", C.c_int), ("coarse_sweeps", C.c_int),
        ("mg_fp32", C.c_int), ("use_guess", C.c_int), ("amg", C.c_int), ("amg_thresh", C.c_double),
        ("deep_level", C.c_int), ("deep_pre", C.c_int), ("deep_post", C.c_int),
    ]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("relres", C.c_double), ("ms", C.c_double), ("fallback", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [
        ("poisson_iterations", C.c_int), ("poisson_relres", C.c_double), ("maxdiv", C.c_double),
        ("ms_momentum", C.c_double), ("ms_poisson", C.c_double), ("ms_total", C.c_double),
        ("newton_iterations", C.c_int), ("linear_iterations", C.c_int),
    ]


class NewtonOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("restart", C.c_int), ("max_newton", C.c_int), ("max_linear", C.c_int),
                ("use_feul", C.c_int)]


class NewtonStats(C.Structure):
    _fields_ = [("newton_iterations", C.c_int), ("linear_iterations", C.c_int), ("function_evals", C.c_int),
                ("fnorm0", C.c_double), ("fnorm", C.c_double), ("ms", C.c_double)]


class Turbine(C.Structure):
    """vfs_turbine: one actuator-line turbine (Turbine.inp + control options)."""

    _fields_ = [
        ("nx", C.c_double), ("ny", C.c_double), ("nz", C.c_double),
        ("xc", C.c_double), ("yc", C.c_double), ("zc", C.c_double),
        ("angvel", C.c_double), ("pitch", C.c_double), ("r_nacelle", C.c_double), ("refangle", C.c_double),
        ("rotor_model", C.c_int), ("deltafunc", C.c_int), ("halfwidth", C.c_double),
        ("shen", C.c_int), ("prandtl", C.c_int), ("shen1", C.c_int), ("r_rotor", C.c_double), ("tsr", C.c_double),
        ("num_blade", C.c_int), ("a_shen", C.c_double), ("b_shen", C.c_double), ("c_shen", C.c_double),
        ("correction_shen1", C.c_double), ("correction3d_cl", C.c_int), ("c0_cl", C.c_double), ("smoothforce", C.c_int),
    ]

    def __init__(self, *args, **kw):
        # the correction options' defaults of main.c:161, 224-238
        base = {"num_blade": 3, "a_shen": 0.125, "b_shen": 21.0, "c_shen": 0.1, "correction_shen1": 1.0, "c0_cl": 1.0}
        base.update(kw)
        super().__init__(*args, **base)


class TurbineControl(C.Structure):
    """vfs_turbine_control: Calc_turbineangvel's mode and parameters."""

    _fields_ = [
        ("mode", C.c_int), ("angvel_fixed", C.c_double), ("tsr", C.c_double), ("r_rotor", C.c_double),
        ("reflength_wt", C.c_double), ("refvel_wt", C.c_double), ("refvel_cfd", C.c_double), ("rho_air", C.c_double),
        ("J_rotation", C.c_double), ("CP_max", C.c_double), ("TSR_max", C.c_double), ("X_control", C.c_double),
        ("fixturbinegeneratortorque", C.c_int), ("torque_generator", C.c_double),
    ]


class Foil(C.Structure):
    """vfs_foil: radial twist/chord table and CL/CD polars of one foil type."""

    _fields_ = [
        ("n_ac", C.c_int), ("r_ac", C.c_void_p), ("twist", C.c_void_p), ("chord", C.c_void_p),
        ("n_cl", C.c_int), ("ang_cl", C.c_void_p), ("cl", C.c_void_p),
        ("n_cd", C.c_int), ("ang_cd", C.c_void_p), ("cd", C.c_void_p),
    ]


# vfs_alm_array enumeration and the number of values per element / node
ALM_ARRAYS = ["NODES", "NODE_VEL", "CENT", "REF_CENT", "DA", "TWIST", "CHORD", "U_LAGR", "F_LAGR",
              "AOA", "CL", "CD", "BOX", "REF_BOX"]
_ALM_SHAPE = {"NODES": ("node", 3), "NODE_VEL": ("node", 3), "CENT": ("elmt", 3), "REF_CENT": ("elmt", 3),
              "DA": ("elmt", 1), "TWIST": ("elmt", 1), "CHORD": ("elmt", 1), "U_LAGR": ("elmt", 3),
              "F_LAGR": ("elmt", 3), "AOA": ("elmt", 1), "CL": ("elmt", 1), "CD": ("elmt", 1),
              "BOX": ("elmt", 6), "REF_BOX": ("elmt", 6)}


# entry points and their argument types
_SIGS = {
    "vfs_last_error": ([], C.c_char_p),
    "vfs_ctx_create": ([C.POINTER(Params), C.POINTER(C.c_void_p)], C.c_int),
    "vfs_ctx_destroy": ([C.c_void_p], C.c_int),
    "vfs_set_params": ([C.c_void_p, C.POINTER(Params)], C.c_int),
    "vfs_get_params": ([C.c_void_p, C.POINTER(Params)], C.c_int),
    "vfs_field_size": ([C.c_void_p, C.c_int], C.c_size_t),
    "vfs_upload": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "vfs_download": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "vfs_field_device_ptr": ([C.c_void_p, C.c_int], C.c_void_p),
    "vfs_field_written": ([C.c_void_p, C.c_int], C.c_int),
    "vfs_sync": ([C.c_void_p], C.c_int),
    "vfs_form_metrics": ([C.c_void_p], C.c_int),
    "vfs_mean_k_area": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "vfs_contra2cart": ([C.c_void_p], C.c_int),
    "vfs_form_bcs": ([C.c_void_p, C.c_int], C.c_int),
    "vfs_eddy_viscosity": ([C.c_void_p], C.c_int),
    "vfs_smagorinsky_constant": ([C.c_void_p, C.c_int, C.c_int, C.c_int], C.c_int),
    "vfs_set_time_index": ([C.c_void_p, C.c_int, C.c_int, C.c_int], C.c_int),
    "vfs_inflow_flux": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "vfs_calc_k_flux": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "vfs_pressure_gradient": ([C.c_void_p], C.c_int),
    "vfs_formfunction2": ([C.c_void_p, C.c_double], C.c_int),
    "vfs_runge_kutta": ([C.c_void_p], C.c_int),
    "vfs_min_dt": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "vfs_poisson_rhs": ([C.c_void_p], C.c_int),
    "vfs_poisson_apply": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "vfs_poisson_solve": ([C.c_void_p, C.POINTER(SolveOpts), C.POINTER(SolveStats)], C.c_int),
    "vfs_update_pressure": ([C.c_void_p], C.c_int),
    "vfs_projection": ([C.c_void_p], C.c_int),
    "vfs_divergence": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "vfs_divergence_noalloc": ([C.POINTER(Params), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "vfs_fractional_step": ([C.c_void_p, C.POINTER(SolveOpts), C.POINTER(StepStats)], C.c_int),
    "vfs_alm_create": ([C.c_void_p, C.POINTER(Turbine), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                        C.POINTER(C.c_void_p)], C.c_int),
    "vfs_alm_destroy": ([C.c_void_p], C.c_int),
    "vfs_alm_counts": ([C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "vfs_alm_rotate": ([C.c_void_p, C.c_double], C.c_int),
    "vfs_alm_preprocess": ([C.c_void_p], C.c_int),
    "vfs_alm_u_lagr": ([C.c_void_p], C.c_int),
    "vfs_alm_f_lagr": ([C.c_void_p], C.c_int),
    "vfs_alm_f_eul": ([C.c_void_p], C.c_int),
    "vfs_alm_step": ([C.c_void_p, C.c_double], C.c_int),
    "vfs_alm_get": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "vfs_alm_set_disk": ([C.c_void_p, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
    "vfs_alm_uref": ([C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "vfs_alm_loads": ([C.c_void_p, C.c_void_p], C.c_int),
    "vfs_alm_turbine_angvel": ([C.c_void_p, C.POINTER(TurbineControl), C.c_double, C.c_double, C.c_int, C.c_int,
                                C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "vfs_formfunction_snes": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_int], C.c_int),
    "vfs_store_old_state": ([C.c_void_p], C.c_int),
    "vfs_implicit_matrixfree": ([C.c_void_p, C.POINTER(NewtonOpts), C.POINTER(NewtonStats)], C.c_int),
    "vfs_fractional_step_cn": ([C.c_void_p, C.POINTER(NewtonOpts), C.POINTER(SolveOpts), C.POINTER(StepStats)], C.c_int),
    "vfs_unet_create": ([C.c_void_p, C.c_int, C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)], C.c_int),
    "vfs_unet_destroy": ([C.c_void_p], C.c_int),
    "vfs_unet_set_output_scale": ([C.c_void_p, C.c_double], C.c_int),
    "vfs_unet_forward": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "vfs_unet_gradients": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)], C.c_int),
    "vfs_unet_train_step": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double)], C.c_int),
    "vfs_unet_param_count": ([C.c_void_p], C.c_longlong),
    "vfs_unet_get_params": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "vfs_unet_set_params": ([C.c_void_p, C.c_void_p], C.c_int),
    "vfs_poisson_solve_learned": ([C.c_void_p, C.c_void_p, C.POINTER(SolveOpts), C.POINTER(SolveStats)], C.c_int),
    "vfs_inflow_set_plane": ([C.c_void_p, C.c_void_p], C.c_int),
    "vfs_inflow_get_plane": ([C.c_void_p, C.c_void_p], C.c_int),
    "vfs_pseudo_periodic_plane": ([C.c_void_p], C.c_int),
    "vfs_inflow_write_section": ([C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                  C.c_void_p], C.c_int),
    "vfs_inflow_reader_open": ([C.c_char_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.POINTER(C.c_void_p)], C.c_int),
    "vfs_inflow_reader_next": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "vfs_inflow_reader_close": ([C.c_void_p], C.c_int),
    "vfs_amg_host_build": ([C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int,
                            C.POINTER(C.c_void_p)], C.c_int),
    "vfs_amg_host_levels": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "vfs_amg_host_matrix": ([C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_longlong), C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "vfs_amg_host_splitting": ([C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "vfs_amg_host_colouring": ([C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_void_p, C.c_void_p], C.c_int),
    "vfs_amg_host_free": ([C.c_void_p], C.c_int),
    "vfs_vec_write": ([C.c_char_p, C.c_void_p, C.c_longlong], C.c_int),
    "vfs_vec_read": ([C.c_char_p, C.c_void_p, C.c_longlong], C.c_int),
    "vfs_vec_length": ([C.c_char_p, C.POINTER(C.c_longlong)], C.c_int),
    "vfs_checkpoint_write": ([C.c_void_p, C.c_char_p, C.c_int, C.c_int], C.c_int),
    "vfs_checkpoint_read": ([C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int], C.c_int),
    "vfs_grid_dims": ([C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
                      C.c_int),
    "vfs_grid_read": ([C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_void_p],
                      C.c_int),
    "vfs_profile": ([C.c_void_p, C.c_int], C.c_int),
    "vfs_profile_select": ([C.c_void_p, C.c_uint], C.c_int),
    "vfs_profile_read": ([C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                          C.POINTER(C.c_char_p)], C.c_int),
}

KERNEL_IDS = {"face_fluxes": 0, "cell_rhs": 1, "line_rhs_fp32": 2, "line_rhs_fp64": 3,
              "poisson_apply_fp64": 4, "contra2cart": 5, "pressure_gradient": 6, "eddy_viscosity": 7,
              "line_solve_fp32": 8, "line_solve_fp64": 9, "cell_fp": 10, "formfunction2": 11,
              "line_rhs_compact_fp32": 12, "line_relax_compact_fp32": 13}


class VFSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vfs error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libvfswind.so; raises if it was not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with vfs-wind_amd/build_lib.py (no CPU fallback exists)")
    lib = C.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _raise(lib, rc):
    if rc != VFS_OK:
        raise VFSError(rc, lib.vfs_last_error().decode())


# ---- file formats (host side; no device needed) -------------------------------
GRID_FORMATS = {"text": 0, "binary": 1, "xyz": 2}


def vec_write(path, arr: np.ndarray):
    """Writes a PETSc binary Vec file (VecView's format) holding arr's values."""
    lib = load_library()
    a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
    _raise(lib, lib.vfs_vec_write(os.fsencode(path), a.ctypes.data, a.size))


def vec_length(path) -> int:
    lib = load_library()
    n = C.c_longlong()
    _raise(lib, lib.vfs_vec_length(os.fsencode(path), C.byref(n)))
    return n.value


def vec_read(path, shape=None) -> np.ndarray:
    """Reads a PETSc binary Vec file; `shape` (if given) must hold exactly its values."""
    lib = load_library()
    n = vec_length(path) if shape is None else int(np.prod(shape))
    out = np.empty(n)
    _raise(lib, lib.vfs_vec_read(os.fsencode(path), out.ctypes.data, n))
    return out if shape is None else out.reshape(shape)


def inflow_write_section(directory, ti: int, tistart: int, period: int, minus: int, dt: float, plane: np.ndarray):
    """save_inflow_section: appends a (my, mx, 3) plane to directory/inflow_%06d_dt=%g.dat."""
    lib = load_library()
    a = np.ascontiguousarray(plane, dtype=np.float64)
    my, mx = a.shape[:2]
    _raise(lib, lib.vfs_inflow_write_section(os.fsencode(directory), ti, tistart, period, minus, dt, mx, my,
                                             a.ctypes.data))


class InflowReader:
    """read_inflow_section's reader: planes of path_inflow/inflow/inflow_%06d_dt=%g.dat."""

    def __init__(self, path_inflow, dt_inflow: float, period: int, minus: int, recycle: int, tistart: int, mx: int,
                 my: int):
        self.lib = load_library()
        self.shape = (my, mx, 3)
        h = C.c_void_p()
        _raise(self.lib, self.lib.vfs_inflow_reader_open(os.fsencode(path_inflow), dt_inflow, period, minus, recycle,
                                                         tistart, mx, my, C.byref(h)))
        self.h = h

    def next(self, ti: int) -> np.ndarray:
        out = np.empty(self.shape)
        _raise(self.lib, self.lib.vfs_inflow_reader_next(self.h, ti, out.ctypes.data))
        return out

    def close(self):
        if self.h:
            self.lib.vfs_inflow_reader_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AmgHost:
    """The AMG setup (PMIS, FF1, Galerkin) run on the host on a CSR matrix
    (scipy.sparse-like: indptr, indices, data); levels' operators, splittings
    and interpolations are returned as (indptr, indices, data, shape)."""

    def __init__(self, indptr, indices, data, theta=0.75, max_levels=25, max_coarse=0):
        self.lib = load_library()
        ip = np.ascontiguousarray(indptr, dtype=np.int32)
        ix = np.ascontiguousarray(indices, dtype=np.int32)
        dv = np.ascontiguousarray(data, dtype=np.float64)
        h = C.c_void_p()
        _raise(self.lib, self.lib.vfs_amg_host_build(len(ip) - 1, ip.ctypes.data, ix.ctypes.data, dv.ctypes.data, theta,
                                                     max_levels, max_coarse, C.byref(h)))
        self.h = h
        nl = C.c_int()
        _raise(self.lib, self.lib.vfs_amg_host_levels(self.h, C.byref(nl)))
        self.levels = nl.value

    def matrix(self, level: int, which: int = 0):
        r, c, z = C.c_int(), C.c_int(), C.c_longlong()
        _raise(self.lib, self.lib.vfs_amg_host_matrix(self.h, level, which, C.byref(r), C.byref(c), C.byref(z), None, None,
                                                      None))
        ptr = np.empty(r.value + 1, np.int32)
        col = np.empty(z.value, np.int32)
        val = np.empty(z.value)
        _raise(self.lib, self.lib.vfs_amg_host_matrix(self.h, level, which, None, None, None, ptr.ctypes.data,
                                                      col.ctypes.data, val.ctypes.data))
        return ptr, col, val, (r.value, c.value)

    def splitting(self, level: int) -> np.ndarray:
        n = self.matrix(level, 0)[3][0]
        out = np.empty(n, np.int32)
        _raise(self.lib, self.lib.vfs_amg_host_splitting(self.h, level, out.ctypes.data))
        return out

    def colouring(self, level: int):
        """(rows grouped by colour, colour offsets) of the smoother on a level."""
        n = self.matrix(level, 0)[3][0]
        nc = C.c_int()
        _raise(self.lib, self.lib.vfs_amg_host_colouring(self.h, level, C.byref(nc), None, None))
        rows = np.empty(n, np.int32)
        off = np.empty(nc.value + 1, np.int32)
        _raise(self.lib, self.lib.vfs_amg_host_colouring(self.h, level, C.byref(nc), rows.ctypes.data, off.ctypes.data))
        return rows, off

    def close(self):
        if self.h:
            self.lib.vfs_amg_host_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grid_dims(path, fmt: str = "text", block: int = 0):
    """(IM, JM, KM) node counts of one block of a grid.dat / xyz.dat file."""
    lib = load_library()
    im, jm, km = C.c_int(), C.c_int(), C.c_int()
    _raise(lib, lib.vfs_grid_dims(os.fsencode(path), GRID_FORMATS[fmt], block, C.byref(im), C.byref(jm), C.byref(km)))
    return im.value, jm.value, km.value


def grid_read(path, fmt: str = "text", block: int = 0, cl: float = 1.0, L_dim: float = 1.0) -> np.ndarray:
    """Node coordinates of one block as the DA coordinate array (KM+1, JM+1, IM+1, 3)."""
    lib = load_library()
    im, jm, km = grid_dims(path, fmt, block)
    out = np.zeros((km + 1, jm + 1, im + 1, 3))
    _raise(lib, lib.vfs_grid_read(os.fsencode(path), GRID_FORMATS[fmt], block, cl, L_dim, im + 1, jm + 1, km + 1,
                                  out.ctypes.data))
    return out


def default_newton_opts(rtol=1e-4, restart=30, max_newton=50, max_linear=1000, use_feul=0) -> NewtonOpts:
    o = NewtonOpts()
    o.rtol, o.restart, o.max_newton, o.max_linear, o.use_feul = rtol, restart, max_newton, max_linear, use_feul
    return o


def make_params(mx, my, mz, *, per=(1, 0, 1), bctype=(100, 100, 1, 1, 100, 100), les=1, second_order=1,
                laplacian=1, ren=2800.0, dt=1e-3, st=1.0, cs2=0.01, inlet_flux=-1.0,
                poisson_threshold=0.1, flux_in_sum=0.0, max_cs=0.5, dynamic_freq=1, homo=(0, 0, 0), inletprofile=1,
                pseudo=0) -> Params:
    p = Params()
    p.mx, p.my, p.mz = mx, my, mz
    p.per_i, p.per_j, p.per_k = per
    for q in range(6):
        p.bctype[q] = int(bctype[q])
    p.les, p.second_order, p.laplacian = les, second_order, laplacian
    p.ren, p.dt, p.st, p.cs2 = ren, dt, st, cs2
    p.inlet_flux, p.poisson_threshold, p.flux_in_sum = inlet_flux, poisson_threshold, flux_in_sum
    p.max_cs, p.dynamic_freq = max_cs, dynamic_freq
    p.i_homo_filter, p.j_homo_filter, p.k_homo_filter = (int(h) for h in homo)
    p.inletprofile, p.pseudo_periodic = int(inletprofile), int(pseudo)
    return p


def default_solve_opts(tol=1e-8, restart=20, maxit=200, levels=6, pre=2, post=3, coarse_sweeps=20, mg_fp32=1,
                       use_guess=1, amg=0, amg_thresh=0.75, deep_level=0, deep_pre=1, deep_post=2) -> SolveOpts:
    o = SolveOpts()
    o.tol, o.restart, o.maxit, o.levels = tol, restart, maxit, levels
    o.pre, o.post, o.coarse_sweeps, o.mg_fp32, o.use_guess = pre, post, coarse_sweeps, mg_fp32, use_guess
    o.amg, o.amg_thresh = amg, amg_thresh
    o.deep_level, o.deep_pre, o.deep_post = deep_level, deep_pre, deep_post
    return o


class Context:
    """A device-resident DMDA block (the reference's UserCtx for the hot path)."""

    def __init__(self, params: Params, lib: C.CDLL | None = None):
        self.lib = lib or load_library()
        self.params = params
        h = C.c_void_p()
        self._check(self.lib.vfs_ctx_create(C.byref(params), C.byref(h)))
        self.h = h
        self.shape = (params.mz, params.my, params.mx)

    # -- plumbing --------------------------------------------------------------
    def _check(self, rc: int):
        if rc != VFS_OK:
            raise VFSError(rc, self.lib.vfs_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.vfs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def field_shape(self, name: str):
        return self.shape + ((3,) if name in VECTOR_FIELDS else ())

    def upload(self, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.field_shape(name):
            raise ValueError(f"{name}: expected shape {self.field_shape(name)}, got {a.shape}")
        self._check(self.lib.vfs_upload(self.h, FIELD[name], a.ctypes.data))

    def download(self, name: str) -> np.ndarray:
        a = np.empty(self.field_shape(name), dtype=np.float64)
        self._check(self.lib.vfs_download(self.h, FIELD[name], a.ctypes.data))
        return a

    def device_ptr(self, name: str) -> int:
        return int(self.lib.vfs_field_device_ptr(self.h, FIELD[name]) or 0)

    def field_written(self, name: str):
        """Declare a write through a kept device pointer (vfs_field_written)."""
        self._check(self.lib.vfs_field_written(self.h, FIELD[name]))

    def set_params(self, params: Params):
        self._check(self.lib.vfs_set_params(self.h, C.byref(params)))
        self.params = params

    def sync(self):
        self._check(self.lib.vfs_sync(self.h))

    # -- reference entry points ---------------------------------------------
    def form_metrics(self):
        self._check(self.lib.vfs_form_metrics(self.h))

    def mean_k_area(self) -> float:
        v = C.c_double()
        self._check(self.lib.vfs_mean_k_area(self.h, C.byref(v)))
        return v.value

    def contra2cart(self):
        self._check(self.lib.vfs_contra2cart(self.h))

    def form_bcs(self, outflow_scale: int = 0):
        self._check(self.lib.vfs_form_bcs(self.h, int(outflow_scale)))

    def inflow_flux(self) -> float:
        v = C.c_double()
        self._check(self.lib.vfs_inflow_flux(self.h, C.byref(v)))
        return v.value

    def eddy_viscosity(self):
        self._check(self.lib.vfs_eddy_viscosity(self.h))

    def smagorinsky_constant(self, ti: int = 2, tistart: int = 0, rstart: int = 0):
        self._check(self.lib.vfs_smagorinsky_constant(self.h, int(ti), int(tistart), int(rstart)))

    def set_time_index(self, ti: int, tistart: int = 0, rstart: int = 0):
        self._check(self.lib.vfs_set_time_index(self.h, int(ti), int(tistart), int(rstart)))

    def calc_k_flux(self) -> float:
        v = C.c_double()
        self._check(self.lib.vfs_calc_k_flux(self.h, C.byref(v)))
        return v.value

    def pressure_gradient(self):
        self._check(self.lib.vfs_pressure_gradient(self.h))

    def formfunction2(self, scale: float = 1.0):
        self._check(self.lib.vfs_formfunction2(self.h, C.c_double(scale)))

    def runge_kutta(self):
        self._check(self.lib.vfs_runge_kutta(self.h))

    def min_dt(self) -> float:
        v = C.c_double()
        self._check(self.lib.vfs_min_dt(self.h, C.byref(v)))
        return v.value

    def poisson_rhs(self):
        self._check(self.lib.vfs_poisson_rhs(self.h))

    def poisson_apply_dev(self, x_ptr: int, y_ptr: int):
        self._check(self.lib.vfs_poisson_apply(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr)))

    def poisson_solve(self, opts: SolveOpts | None = None, check: bool = True) -> SolveStats:
        st = SolveStats()
        rc = self.lib.vfs_poisson_solve(self.h, C.byref(opts or default_solve_opts()), C.byref(st))
        if check or rc != VFS_ENOCONV:
            self._check(rc)
        return st

    def update_pressure(self):
        self._check(self.lib.vfs_update_pressure(self.h))

    def projection(self):
        self._check(self.lib.vfs_projection(self.h))

    def divergence(self) -> float:
        v = C.c_double()
        self._check(self.lib.vfs_divergence(self.h, C.byref(v)))
        return v.value

    def profile(self, enable: bool, only=None):
        """Start (clearing earlier records) or stop HIP-event timing; `only`
        restricts it to the named kernels (KERNEL_IDS)."""
        if enable and only:
            mask = 0
            for k in only:
                mask |= 1 << KERNEL_IDS[k]
            self._check(self.lib.vfs_profile_select(self.h, mask))
        else:
            self._check(self.lib.vfs_profile(self.h, int(enable)))

    def profile_read(self, kernel: str):
        """(total ms, launch count) recorded for a kernel since profile(True)."""
        ms, cnt, name = C.c_double(), C.c_longlong(), C.c_char_p()
        self._check(self.lib.vfs_profile_read(self.h, KERNEL_IDS[kernel], C.byref(ms), C.byref(cnt), C.byref(name)))
        return ms.value, cnt.value

    def fractional_step(self, opts: SolveOpts | None = None) -> StepStats:
        st = StepStats()
        self._check(self.lib.vfs_fractional_step(self.h, C.byref(opts or default_solve_opts()), C.byref(st)))
        return st

    def fractional_step_cn(self, nopts: NewtonOpts | None = None, opts: SolveOpts | None = None) -> StepStats:
        st = StepStats()
        self._check(self.lib.vfs_fractional_step_cn(self.h, C.byref(nopts or default_newton_opts()),
                                                    C.byref(opts or default_solve_opts()), C.byref(st)))
        return st

    # -- inflow planes ------------------------------------------------------------
    def inflow_set_plane(self, plane: np.ndarray):
        a = np.ascontiguousarray(plane, dtype=np.float64)
        assert a.size == 3 * self.params.mx * self.params.my
        self._check(self.lib.vfs_inflow_set_plane(self.h, a.ctypes.data))

    def inflow_get_plane(self) -> np.ndarray:
        out = np.empty((self.params.my, self.params.mx, 3))
        self._check(self.lib.vfs_inflow_get_plane(self.h, out.ctypes.data))
        return out

    def pseudo_periodic_plane(self):
        self._check(self.lib.vfs_pseudo_periodic_plane(self.h))

    # -- checkpoints (vfield/ufield/pfield/nvfield) -----------------------------
    def checkpoint_write(self, directory, ti: int, block: int = 0):
        self._check(self.lib.vfs_checkpoint_write(self.h, os.fsencode(directory), ti, block))

    def checkpoint_read(self, directory, ti: int, block: int = 0, immersed: int = 0):
        self._check(self.lib.vfs_checkpoint_read(self.h, os.fsencode(directory), ti, block, immersed))

    # -- implicit momentum (-imp 4) ---------------------------------------------
    def store_old_state(self):
        self._check(self.lib.vfs_store_old_state(self.h))

    def formfunction_snes(self, u_dev: int, r_dev: int, use_feul: int = 0):
        self._check(self.lib.vfs_formfunction_snes(self.h, u_dev, r_dev, int(use_feul)))

    def implicit_matrixfree(self, nopts: NewtonOpts | None = None, check: bool = True) -> NewtonStats:
        st = NewtonStats()
        rc = self.lib.vfs_implicit_matrixfree(self.h, C.byref(nopts or default_newton_opts()), C.byref(st))
        if check:
            self._check(rc)
        return st


class ALM:
    """An actuator-line turbine attached to a Context (the reference's IBMNodes
    + FSInfo + ACL of one turbine on the rotor_model 3/6 path)."""

    def __init__(self, ctx: Context, turbine: Turbine, nodes: np.ndarray, foils: list):
        self.ctx, self.lib = ctx, ctx.lib
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        if nodes.ndim != 3 or nodes.shape[2] != 3:
            raise ValueError("nodes: expected (num_blade, nodes_per_blade, 3)")
        nb, nvb = nodes.shape[:2]
        self._keep = []
        fs = (Foil * max(1, len(foils)))()
        for q, f in enumerate(foils):
            arrs = [np.ascontiguousarray(f[k], dtype=np.float64) for k in ("r", "twist", "chord", "ang_cl", "cl", "ang_cd", "cd")]
            self._keep += arrs
            fs[q].n_ac, fs[q].r_ac, fs[q].twist, fs[q].chord = len(arrs[0]), arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data
            fs[q].n_cl, fs[q].ang_cl, fs[q].cl = len(arrs[3]), arrs[3].ctypes.data, arrs[4].ctypes.data
            fs[q].n_cd, fs[q].ang_cd, fs[q].cd = len(arrs[5]), arrs[5].ctypes.data, arrs[6].ctypes.data
        h = C.c_void_p()
        ctx._check(self.lib.vfs_alm_create(ctx.h, C.byref(turbine), nb, nvb, nodes.ctypes.data, len(foils), fs, C.byref(h)))
        self.h = h
        ne, nv = C.c_int(), C.c_int()
        ctx._check(self.lib.vfs_alm_counts(h, C.byref(ne), C.byref(nv)))
        self.n_elmt, self.n_nodes = ne.value, nv.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.vfs_alm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rotate(self, dt: float):
        self.ctx._check(self.lib.vfs_alm_rotate(self.h, float(dt)))

    def preprocess(self):
        self.ctx._check(self.lib.vfs_alm_preprocess(self.h))

    def u_lagr(self):
        self.ctx._check(self.lib.vfs_alm_u_lagr(self.h))

    def f_lagr(self):
        self.ctx._check(self.lib.vfs_alm_f_lagr(self.h))

    def f_eul(self):
        self.ctx._check(self.lib.vfs_alm_f_eul(self.h))

    def step(self, dt: float):
        self.ctx._check(self.lib.vfs_alm_step(self.h, float(dt)))

    def get(self, name: str) -> np.ndarray:
        kind, w = _ALM_SHAPE[name]
        n = self.n_elmt if kind == "elmt" else self.n_nodes
        a = np.empty((n, w) if w > 1 else (n,), dtype=np.float64)
        self.ctx._check(self.lib.vfs_alm_get(self.h, ALM_ARRAYS.index(name), a.ctypes.data))
        return a

    def set_disk(self, cent: np.ndarray, dA: np.ndarray):
        cent = np.ascontiguousarray(cent, dtype=np.float64)
        dA = np.ascontiguousarray(dA, dtype=np.float64)
        self.ctx._check(self.lib.vfs_alm_set_disk(self.h, len(dA), cent.ctypes.data, dA.ctypes.data))

    def uref(self) -> float:
        v = C.c_double()
        self.ctx._check(self.lib.vfs_alm_uref(self.h, C.byref(v)))
        return v.value

    def loads(self) -> np.ndarray:
        """Calc_forces_ACL: (A_Sum, F_Sum, P_Sum, U_mean, Fx, Fy, Fz, M_Sum)."""
        out = np.zeros(8)
        self.ctx._check(self.lib.vfs_alm_loads(self.h, out.ctypes.data))
        return out

    def turbine_angvel(self, control: "TurbineControl", dt: float, uref: float, ti: int, tistart: int = 0):
        """Calc_turbineangvel: (angular velocity, tip speed ratio) after the update."""
        w, t = C.c_double(), C.c_double()
        self.ctx._check(self.lib.vfs_alm_turbine_angvel(self.h, C.byref(control), float(dt), float(uref), int(ti),
                                                        int(tistart), C.byref(w), C.byref(t)))
        return w.value, t.value


class UNet:
    """The learned Poisson stage on a Context's grid (BF16 U-Net on the matrix cores)."""

    def __init__(self, ctx: Context, base: int = 16, levels: int = 4, seed: int = 20261202):
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        ctx._check(self.lib.vfs_unet_create(ctx.h, base, levels, seed, C.byref(h)))
        self.h = h
        self.base, self.levels = base, levels
        self.n_params = int(self.lib.vfs_unet_param_count(h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.vfs_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_output_scale(self, s: float):
        self.ctx._check(self.lib.vfs_unet_set_output_scale(self.h, float(s)))

    def forward(self, b_dev: int, y_dev: int):
        self.ctx._check(self.lib.vfs_unet_forward(self.h, b_dev, y_dev))

    def gradients(self, b_dev: int, phi_dev: int) -> float:
        v = C.c_double()
        self.ctx._check(self.lib.vfs_unet_gradients(self.h, b_dev, phi_dev, C.byref(v)))
        return v.value

    def train_step(self, b_dev: int, phi_dev: int, lr: float) -> float:
        v = C.c_double()
        self.ctx._check(self.lib.vfs_unet_train_step(self.h, b_dev, phi_dev, float(lr), C.byref(v)))
        return v.value

    def params(self, which: int = 0) -> np.ndarray:
        a = np.empty(self.n_params, dtype=np.float32)
        self.ctx._check(self.lib.vfs_unet_get_params(self.h, int(which), a.ctypes.data))
        return a

    def set_params(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {a.size}")
        self.ctx._check(self.lib.vfs_unet_set_params(self.h, a.ctypes.data))

    def poisson_solve(self, opts: SolveOpts | None = None, check: bool = True) -> SolveStats:
        st = SolveStats()
        rc = self.lib.vfs_poisson_solve_learned(self.ctx.h, self.h, C.byref(opts or default_solve_opts()), C.byref(st))
        if check:
            self.ctx._check(rc)
        return st
