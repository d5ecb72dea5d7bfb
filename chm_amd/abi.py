"""ctypes mirror of include/pbsm3d.h and numpy helpers to fill it.

The structs hold raw pointers into numpy arrays (host) or torch tensors
(device); :class:`Keep` keeps every referenced buffer alive for as long as the
struct is in use.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

PBSM3D_OK = 0
PBSM3D_EINVAL = -1
PBSM3D_EHIP = -2
PBSM3D_ENOCONV = -3
PBSM3D_ENAN = -4
PBSM3D_ENOMEM = -5

FP64 = 0
MIXED = 1

PC_NONE, PC_JACOBI, PC_COLUMN, PC_ILU0, PC_ILUT, PC_MCILU0 = range(6)
WS_ZERO, WS_LEARNED = 0, 1
KRYLOV_GMRES, KRYLOV_BICGSTAB, KRYLOV_BICGSTAB_SPLIT = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint8)


class Mesh(C.Structure):
    _fields_ = [
        ("n_faces", C.c_int64),
        ("neigh", _ip),
        ("normal_x", _dp),
        ("normal_y", _dp),
        ("edge_length", _dp),
        ("area", _dp),
        ("center_x", _dp),
        ("center_y", _dp),
        ("center_z", _dp),
        ("canopy_height", _dp),
        ("lai", _dp),
        ("stalk_number", _dp),
        ("stalk_diameter", _dp),
        ("is_water", _up),
    ]


class Config(C.Structure):
    _fields_ = [
        ("nLayer", C.c_int32),
        ("do_fixed_settling", C.c_int32),
        ("settling_velocity", C.c_double),
        ("do_sublimation", C.c_int32),
        ("do_lateral_diff", C.c_int32),
        ("smooth_coeff", C.c_double),
        ("min_sd_trans", C.c_double),
        ("cutoff", C.c_double),
        ("snow_diffusion_const", C.c_double),
        ("enable_veg", C.c_int32),
        ("use_tanh_fetch", C.c_int32),
        ("use_exp_fetch", C.c_int32),
        ("use_R94_lambda", C.c_int32),
        ("z0_snow", C.c_double),
        ("settling_scale", C.c_double),
        ("tol", C.c_double),
        ("max_iters", C.c_int32),
        ("restart", C.c_int32),
        ("precision", C.c_int32),
        ("precond", C.c_int32),
        ("warm_start", C.c_int32),
        ("dep_precond", C.c_int32),
        ("ortho_passes", C.c_int32),
        ("block_size", C.c_int32),
        ("krylov", C.c_int32),
        ("inner_rtol", C.c_double),
        ("use_PomLi_probability", C.c_int32),
        ("z0_ustar_coupling", C.c_int32),
        ("rouault_diffusion_coef", C.c_int32),
        ("coarse_tiles", C.c_int32),
    ]


class Forcing(C.Structure):
    _fields_ = [
        ("U_R", _dp),
        ("U_2m_above_srf", _dp),
        ("vw_dir", _dp),
        ("swe", _dp),
        ("t", _dp),
        ("rh", _dp),
        ("snowdepthavg", _dp),
        ("fetch", _dp),
        ("p_snow_hours", _dp),
        ("on_device", C.c_int32),
    ]


class Out(C.Structure):
    _fields_ = [
        ("Qsalt", _dp),
        ("Qsusp", _dp),
        ("Qsubl", _dp),
        ("Qsubl_mass", _dp),
        ("sum_subl", _dp),
        ("drift_mass", _dp),
        ("sum_drift", _dp),
        ("pbsm_more_than_avail", _dp),
        ("blowingsnow_probability", _dp),
        ("conc", _dp),
        ("qdep", _dp),
        ("on_device", C.c_int32),
    ]


class FetchConfig(C.Structure):
    """pbsm3d_fetch_config of include/pbsm3d_fetch.h."""

    _fields_ = [
        ("steps", C.c_int32),
        ("incl_veg", C.c_int32),
        ("max_distance", C.c_double),
        ("I", C.c_double),
    ]


class WindConfig(C.Structure):
    """pbsm3d_wind_config of include/pbsm3d_wind.h."""

    _fields_ = [
        ("ignore_canopy", C.c_int32),
        ("smooth", C.c_int32),
        ("z0_snow", C.c_double),
    ]


class SnowConfig(C.Structure):
    """pbsm3d_snow_config of include/pbsm3d_snow.h."""

    _fields_ = [
        ("drift_density", C.c_double),
        ("max_active_layer", C.c_double),
        ("ic_rho", C.c_double),
        ("max_h2o_vol", C.c_double),
        ("mass_threshold", C.c_double),
    ]


# snowcover state fields (include/pbsm3d_snow.h), field-major [SNOW_FIELDS x F]
SNOW_FIELDS = ("z_s", "m_s", "rho", "layer_count", "z_s_0", "z_s_l", "m_s_0", "m_s_l", "T_s", "T_s_0", "T_s_l",
               "cc_s", "cc_s_0", "cc_s_l", "h2o_total", "h2o_vol", "h2o", "h2o_max", "h2o_sat")


class Stats(C.Structure):
    _fields_ = [
        ("suspension_present", C.c_int32),
        ("deposition_present", C.c_int32),
        ("susp_iters", C.c_int32),
        ("dep_iters", C.c_int32),
        ("susp_resid", C.c_double),
        ("dep_resid", C.c_double),
        ("susp_rhs_max", C.c_double),
        ("dep_rhs_max", C.c_double),
        ("susp_refinements", C.c_int32),
        ("susp_restarts", C.c_int32),
        ("ms_assembly", C.c_double),
        ("ms_susp_solve", C.c_double),
        ("ms_post", C.c_double),
        ("ms_dep_solve", C.c_double),
        ("ms_total", C.c_double),
        ("susp_resid0", C.c_double),
        ("susp_cold_fallback", C.c_int32),
        ("reserved2", C.c_int32),
        ("susp_resid_coarse", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if not name.startswith("reserved")}


OUT_FIELDS = ("Qsalt", "Qsusp", "Qsubl", "Qsubl_mass", "sum_subl", "drift_mass", "sum_drift", "pbsm_more_than_avail",
              "blowingsnow_probability")
FORCING_FIELDS = ("U_R", "U_2m_above_srf", "vw_dir", "swe", "t", "rh", "snowdepthavg", "fetch", "p_snow_hours")


def default_config(**overrides) -> Config:
    """config.PBSM3D defaults of the reference (PBSM3D.cpp:123-145, 223-258)."""
    c = Config(
        nLayer=10,
        do_fixed_settling=0,
        settling_velocity=0.5,
        do_sublimation=1,
        do_lateral_diff=1,
        smooth_coeff=820.0,
        min_sd_trans=0.1,
        cutoff=0.3,
        snow_diffusion_const=0.3,
        enable_veg=1,
        use_tanh_fetch=1,
        use_exp_fetch=0,
        use_R94_lambda=1,
        z0_snow=0.01,
        settling_scale=1.0,
        tol=1e-8,
        max_iters=1000,
        restart=30,
        precision=FP64,
        precond=PC_MCILU0,
        warm_start=WS_ZERO,
        dep_precond=PC_MCILU0,
        ortho_passes=1,
        block_size=256,
        krylov=KRYLOV_GMRES,
        inner_rtol=0.0,
        use_PomLi_probability=0,
        z0_ustar_coupling=0,
        rouault_diffusion_coef=0,
    )
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Keep:
    """Holds references to every buffer a ctypes struct points into."""

    def __init__(self):
        self.refs = []

    def dptr(self, a):
        if a is None:
            return None
        if hasattr(a, "data_ptr"):  # torch tensor
            self.refs.append(a)
            return C.cast(C.c_void_p(a.data_ptr()), _dp)
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.refs.append(a)
        return a.ctypes.data_as(_dp)

    def iptr(self, a):
        if a is None:
            return None
        if hasattr(a, "data_ptr"):
            self.refs.append(a)
            return C.cast(C.c_void_p(a.data_ptr()), _ip)
        a = np.ascontiguousarray(a, dtype=np.int32)
        self.refs.append(a)
        return a.ctypes.data_as(_ip)

    def uptr(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.uint8)
        self.refs.append(a)
        return a.ctypes.data_as(_up)


class Geometry:
    """Per-face geometry in the ABI's layout (edges face-major, [3*i + j])."""

    def __init__(self, neigh, normal_x, normal_y, edge_length, area, center_x, center_y, center_z,
                 canopy_height=None, lai=None, stalk_number=None, stalk_diameter=None, is_water=None):
        self.neigh = np.ascontiguousarray(neigh, dtype=np.int32).reshape(-1)
        self.normal_x = np.ascontiguousarray(normal_x, dtype=np.float64).reshape(-1)
        self.normal_y = np.ascontiguousarray(normal_y, dtype=np.float64).reshape(-1)
        self.edge_length = np.ascontiguousarray(edge_length, dtype=np.float64).reshape(-1)
        self.area = np.ascontiguousarray(area, dtype=np.float64)
        self.center_x = np.ascontiguousarray(center_x, dtype=np.float64)
        self.center_y = np.ascontiguousarray(center_y, dtype=np.float64)
        self.center_z = np.ascontiguousarray(center_z, dtype=np.float64)
        self.canopy_height = canopy_height
        self.lai = lai
        self.stalk_number = stalk_number
        self.stalk_diameter = stalk_diameter
        self.is_water = is_water

    @property
    def n_faces(self) -> int:
        return int(self.area.shape[0])

    def struct(self, keep: Keep) -> Mesh:
        return Mesh(
            n_faces=self.n_faces,
            neigh=keep.iptr(self.neigh),
            normal_x=keep.dptr(self.normal_x),
            normal_y=keep.dptr(self.normal_y),
            edge_length=keep.dptr(self.edge_length),
            area=keep.dptr(self.area),
            center_x=keep.dptr(self.center_x),
            center_y=keep.dptr(self.center_y),
            center_z=keep.dptr(self.center_z),
            canopy_height=keep.dptr(self.canopy_height),
            lai=keep.dptr(self.lai),
            stalk_number=keep.dptr(self.stalk_number),
            stalk_diameter=keep.dptr(self.stalk_diameter),
            is_water=keep.uptr(self.is_water),
        )


def forcing_struct(forcing: dict, keep: Keep, on_device: bool = False) -> Forcing:
    return Forcing(
        **{k: keep.dptr(forcing.get(k)) for k in FORCING_FIELDS},
        on_device=1 if on_device else 0,
    )


def out_struct(outs: dict, keep: Keep, on_device: bool = False) -> Out:
    kw = {k: keep.dptr(outs.get(k)) for k in OUT_FIELDS + ("conc", "qdep")}
    return Out(**kw, on_device=1 if on_device else 0)


def geometry_args(lib_fn, mesh) -> Geometry:
    """Call a ``*_face_geometry`` C function on a :class:`chm_amd.meshgen.Mesh`."""
    f = mesh.n_faces
    out = {k: np.zeros(3 * f) for k in ("normal_x", "normal_y", "edge_length")}
    out.update({k: np.zeros(f) for k in ("area", "center_x", "center_y", "center_z")})
    keep = Keep()
    vx = np.ascontiguousarray(mesh.vertex, dtype=np.float64)
    el = np.ascontiguousarray(mesh.elem, dtype=np.int32)
    rc = lib_fn(
        C.c_int64(f), keep.dptr(vx), keep.iptr(el), keep.dptr(mesh.area),
        *(keep.dptr(out[k]) for k in ("normal_x", "normal_y", "edge_length", "area", "center_x", "center_y", "center_z")),
    )
    if rc != 0:
        raise RuntimeError(f"face geometry failed: {rc}")
    return Geometry(mesh.neigh, **out)
