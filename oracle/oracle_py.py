"""ctypes wrapper of the CPU oracle (liboracle_pbsm3d.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from chm_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_pbsm3d.so")

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.oracle_create.argtypes = [C.POINTER(vp), C.POINTER(abi.Mesh), C.POINTER(abi.Config)]
        L.oracle_step.argtypes = [vp, C.POINTER(abi.Forcing), C.c_int32, C.POINTER(abi.Out), C.POINTER(abi.Stats)]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_destroy.restype = None
        L.oracle_set_solver.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.oracle_set_replay.argtypes = [vp, dp, dp]
        L.oracle_system_nnz.argtypes = [vp, C.c_int]
        L.oracle_system_nnz.restype = C.c_int64
        L.oracle_export_system.argtypes = [vp, C.c_int, ip, ip, dp, dp]
        L.oracle_export_face_trace.argtypes = [vp] + [dp] * 6
        L.oracle_export_cell_trace.argtypes = [vp, dp, dp]
        L.oracle_get_state.argtypes = [vp, dp, dp]
        L.oracle_set_state.argtypes = [vp, dp, dp]
        L.oracle_last_error.argtypes = [vp]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_face_geometry.argtypes = [C.c_int64, dp, ip, dp] + [dp] * 7
        L.oracle_csr_nnz.argtypes = [C.c_int64, C.c_int, ip]
        L.oracle_csr_nnz.restype = C.c_int64
        L.oracle_csr_pattern.argtypes = [C.c_int64, C.c_int, ip, ip, ip]
        L.oracle_solve_csr.argtypes = [C.c_int64, ip, ip, dp, dp, dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_int64, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.oracle_dense_solve.argtypes = [C.c_int64, ip, ip, dp, dp, dp]
        L.oracle_precond_apply.argtypes = [C.c_int64, ip, ip, dp, C.c_int, C.c_int64, C.c_int, dp, dp]
        L.oracle_face_colors.argtypes = [C.c_int64, ip, ip]
        L.oracle_fetch_nearest.argtypes = [C.c_int64, dp, dp, C.c_int64, dp, dp, ip]
        L.oracle_fetch_nearest.restype = None
        L.oracle_fetch.argtypes = [C.c_int64, dp, dp, dp, dp, C.POINTER(abi.FetchConfig), dp, dp]
        L.oracle_fetch.restype = None
        L.oracle_expint_e1.argtypes = [C.c_double]
        L.oracle_expint_e1.restype = C.c_double
        L.oracle_scale_wind.argtypes = [C.c_int64, ip, dp, dp, dp, dp, C.POINTER(abi.WindConfig), dp, dp, dp]
        L.oracle_scale_wind.restype = None
        L.oracle_snow_init.argtypes = [C.c_int64, dp, C.POINTER(abi.SnowConfig), dp]
        L.oracle_snow_init.restype = None
        L.oracle_snow_drift.argtypes = [C.c_int64, C.POINTER(abi.SnowConfig), dp, dp]
        L.oracle_snow_drift.restype = None
        L.oracle_config_default.argtypes = [C.POINTER(abi.Config)]
        L.oracle_config_default.restype = None
        L.oracle_max_threads.restype = C.c_int
        L.oracle_set_ilut.argtypes = [C.c_double, C.c_double]
        L.oracle_set_ilut.restype = None
        L.oracle_std_dry_air_density.argtypes = [C.c_double, C.c_double]
        L.oracle_std_dry_air_density.restype = C.c_double
        L.oracle_is_nan.argtypes = [C.c_double]
        L.oracle_is_nan.restype = C.c_int
        L.oracle_ustar_coupled.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_int)]
        L.oracle_ustar_coupled.restype = C.c_double
        _lib = L
    return _lib


def face_geometry(mesh) -> abi.Geometry:
    return abi.geometry_args(lib().oracle_face_geometry, mesh)


def csr_pattern(n_faces: int, n_layer: int, neigh: np.ndarray):
    keep = abi.Keep()
    nb = np.ascontiguousarray(neigh, dtype=np.int32).reshape(-1)
    nnz = lib().oracle_csr_nnz(n_faces, n_layer, keep.iptr(nb))
    rp = np.zeros(n_faces * n_layer + 1, dtype=np.int32)
    col = np.zeros(nnz, dtype=np.int32)
    lib().oracle_csr_pattern(n_faces, n_layer, keep.iptr(nb), keep.iptr(rp), keep.iptr(col))
    return rp, col


def oracle_config(**overrides) -> abi.Config:
    """Reference defaults with the reference's solver (GMRES(30) + ILUT)."""
    c = abi.Config()
    lib().oracle_config_default(C.byref(c))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


class Oracle:
    """One PBSM3D module instance on the CPU (init in the constructor, run in step())."""

    def __init__(self, geom: abi.Geometry, cfg: abi.Config | None = None, direct: bool = False, threads: int = 1):
        self.cfg = cfg if cfg is not None else oracle_config()
        self.F = geom.n_faces
        self.L = int(self.cfg.nLayer)
        self._keep = abi.Keep()
        mesh = geom.struct(self._keep)
        self._h = C.c_void_p()
        rc = lib().oracle_create(C.byref(self._h), C.byref(mesh), C.byref(self.cfg))
        if rc != 0:
            raise ValueError(f"oracle_create failed ({rc})")
        lib().oracle_set_solver(self._h, 1 if direct else 0, threads, -1, -1)

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = C.c_void_p()

    __del__ = close

    def set_solver(self, direct=None, threads=None, precond=-1, dep_precond=-1):
        lib().oracle_set_solver(self._h, int(bool(direct)), int(threads or 1), precond, dep_precond)

    def set_replay(self, conc=None, qdep=None):
        """The following steps take conc / qdep (another solver's solutions) instead of
        solving; None: solve as usual."""
        keep = abi.Keep()
        lib().oracle_set_replay(self._h, keep.dptr(conc), keep.dptr(qdep))

    def step(self, forcing: dict, dt: int = 900, want_fields: bool = True):
        keep = abi.Keep()
        f = abi.forcing_struct(forcing, keep)
        outs = {k: np.zeros(self.F) for k in abi.OUT_FIELDS}
        if want_fields:
            outs["conc"] = np.zeros(self.F * self.L)
            outs["qdep"] = np.zeros(self.F)
        o = abi.out_struct(outs, keep)
        st = abi.Stats()
        rc = lib().oracle_step(self._h, C.byref(f), dt, C.byref(o), C.byref(st))
        if rc != 0:
            raise RuntimeError(f"oracle_step failed ({rc}): {lib().oracle_last_error(self._h).decode()}")
        return outs, st.as_dict()

    def system(self, which: int = 0):
        L = lib()
        nnz = L.oracle_system_nnz(self._h, which)
        n = self.F * (self.L if which == 0 else 1)
        rp = np.zeros(n + 1, dtype=np.int32)
        col = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        rhs = np.zeros(n)
        keep = abi.Keep()
        L.oracle_export_system(self._h, which, keep.iptr(rp), keep.iptr(col), keep.dptr(val), keep.dptr(rhs))
        return rp, col, val, rhs

    def face_trace(self) -> dict:
        names = ("ustar", "hs", "c_salt", "z0", "saltation", "uth")
        out = {k: np.zeros(self.F) for k in names}
        keep = abi.Keep()
        lib().oracle_export_face_trace(self._h, *(keep.dptr(out[k]) for k in names))
        return out

    def cell_trace(self) -> dict:
        out = {k: np.zeros(self.F * self.L) for k in ("u_z", "csubl")}
        keep = abi.Keep()
        lib().oracle_export_cell_trace(self._h, keep.dptr(out["u_z"]), keep.dptr(out["csubl"]))
        return out

    def get_state(self):
        a, b = np.zeros(self.F), np.zeros(self.F)
        keep = abi.Keep()
        lib().oracle_get_state(self._h, keep.dptr(a), keep.dptr(b))
        return a, b

    def set_state(self, sum_drift, sum_subl):
        keep = abi.Keep()
        lib().oracle_set_state(self._h, keep.dptr(sum_drift), keep.dptr(sum_subl))


def solve_csr(rp, col, val, b, precond=abi.PC_ILUT, restart=30, maxit=1000, tol=1e-8, F=0, L=1, x0=None):
    keep = abi.Keep()
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it, res = C.c_int(0), C.c_double(0)
    rc = lib().oracle_solve_csr(n, keep.iptr(rp), keep.iptr(col), keep.dptr(val), keep.dptr(b), keep.dptr(x),
                                precond, restart, maxit, tol, F, L, C.byref(it), C.byref(res))
    return x, it.value, res.value, rc


def fetch_nearest(cx, cy, qx, qy):
    """Index of the centroid nearest to each query point (brute force, ties to the lower index)."""
    keep = abi.Keep()
    qx = np.ascontiguousarray(qx, dtype=np.float64)
    idx = np.zeros(len(qx), dtype=np.int32)
    lib().oracle_fetch_nearest(len(cx), keep.dptr(cx), keep.dptr(cy), len(qx), keep.dptr(qx), keep.dptr(qy),
                               keep.iptr(idx))
    return idx


def fetch(cx, cy, cz, canopy, cfg, vw_dir):
    """fetchr's fetch per face (canopy None: no vegetation)."""
    keep = abi.Keep()
    out = np.zeros(len(cx))
    lib().oracle_fetch(len(cx), keep.dptr(cx), keep.dptr(cy), keep.dptr(cz),
                       keep.dptr(canopy) if canopy is not None else None, C.byref(cfg), keep.dptr(vw_dir),
                       keep.dptr(out))
    return out


def std_dry_air_density(altitude: float, temperature_k: float) -> float:
    return lib().oracle_std_dry_air_density(float(altitude), float(temperature_k))


def is_nan(v: float) -> bool:
    return lib().oracle_is_nan(float(v)) == 1


def ustar_coupled(u2: float, lam: float):
    """(u*, bracketed) of the z0_ustar_coupling root search for one face."""
    ok = C.c_int(0)
    r = lib().oracle_ustar_coupled(float(u2), float(lam), C.byref(ok))
    return r, ok.value == 1


def expint_e1(x: float) -> float:
    return lib().oracle_expint_e1(float(x))


def scale_wind(neigh, cx, cy, canopy, lai, cfg, U_R, snowdepthavg=None):
    """scale_wind_vert's U_2m_above_srf per face (domain mode when cfg.smooth)."""
    keep = abi.Keep()
    out = np.zeros(len(cx))
    lib().oracle_scale_wind(len(cx), keep.iptr(np.asarray(neigh).reshape(-1)), keep.dptr(cx), keep.dptr(cy),
                            keep.dptr(canopy), keep.dptr(lai), C.byref(cfg), keep.dptr(U_R), keep.dptr(snowdepthavg),
                            keep.dptr(out))
    return out


def snow_init(n_faces: int, swe0, cfg) -> np.ndarray:
    """snobal's initial snowcover state [len(abi.SNOW_FIELDS) x F] (swe0 None: no initial condition)."""
    keep = abi.Keep()
    st = np.zeros((len(abi.SNOW_FIELDS), n_faces))
    lib().oracle_snow_init(n_faces, keep.dptr(swe0) if swe0 is not None else None, C.byref(cfg), keep.dptr(st))
    return st


def snow_drift(state, cfg, drift_mass) -> np.ndarray:
    """The snowcover state after one step's drift_mass (a new array)."""
    keep = abi.Keep()
    st = np.array(state, dtype=np.float64, copy=True, order="C")
    lib().oracle_snow_drift(st.shape[1], C.byref(cfg), keep.dptr(np.ascontiguousarray(drift_mass, dtype=np.float64)),
                            st.ctypes.data_as(C.POINTER(C.c_double)))
    return st


def face_colors(neigh: np.ndarray):
    """Face colours of the multicolour D-ILU and their count (neigh: [F x 3] or flat)."""
    nb = np.ascontiguousarray(neigh, dtype=np.int32).reshape(-1)
    F = nb.size // 3
    col = np.zeros(F, dtype=np.int32)
    keep = abi.Keep()
    nc = lib().oracle_face_colors(F, keep.iptr(nb), keep.iptr(col))
    return col, nc


def dense_solve(rp, col, val, b):
    keep = abi.Keep()
    n = len(rp) - 1
    x = np.zeros(n)
    rc = lib().oracle_dense_solve(n, keep.iptr(rp), keep.iptr(col), keep.dptr(val), keep.dptr(b), keep.dptr(x))
    if rc != 0:
        raise RuntimeError("singular")
    return x


def precond_apply(rp, col, val, r, precond, F=0, L=1):
    keep = abi.Keep()
    n = len(rp) - 1
    z = np.zeros(n)
    lib().oracle_precond_apply(n, keep.iptr(rp), keep.iptr(col), keep.dptr(val), precond, F, L,
                               keep.dptr(np.ascontiguousarray(r, dtype=np.float64)), keep.dptr(z))
    return z
