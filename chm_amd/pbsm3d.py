"""Python binding of the MI355X PBSM3D library (ctypes over include/pbsm3d.h).

The library is the only compute path: if libpbsm3d.so is missing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbsm3d.so")

EXPORTED = (
    "pbsm3d_abi_version",
    "pbsm3d_config_default",
    "pbsm3d_face_geometry",
    "pbsm3d_create",
    "pbsm3d_step",
    "pbsm3d_system_nnz",
    "pbsm3d_export_system",
    "pbsm3d_get_state",
    "pbsm3d_set_state",
    "pbsm3d_destroy",
    "pbsm3d_last_error",
    "pbsm3d_profile_enable",
    "pbsm3d_profile_read",
    "pbsm3d_profile_classes",
    "pbsm3d_profile_sample",
    "pbsm3d_spmv_tuning",
    "pbsm3d_spmv_select",
    "pbsm3d_face_colors",
    "pbsm3d_precond_tuning",
    "pbsm3d_precond_select",
    "pbsm3d_bicg_restart",
    "pbsm3d_fetch_config_default",
    "pbsm3d_fetch_create",
    "pbsm3d_fetch_run",
    "pbsm3d_fetch_nearest",
    "pbsm3d_fetch_destroy",
    "pbsm3d_fetch_last_error",
    "pbsm3d_batch_create",
    "pbsm3d_batch_step",
    "pbsm3d_batch_set_warm_start",
    "pbsm3d_batch_use_graphs",
    "pbsm3d_batch_info",
    "pbsm3d_batch_get_state",
    "pbsm3d_batch_set_state",
    "pbsm3d_batch_training_pairs",
    "pbsm3d_batch_destroy",
    "pbsm3d_batch_last_error",
    "pbsm3d_wind_config_default",
    "pbsm3d_wind_create",
    "pbsm3d_wind_run",
    "pbsm3d_wind_run_noalloc",
    "pbsm3d_wind_destroy",
    "pbsm3d_wind_last_error",
    "pbsm3d_snow_config_default",
    "pbsm3d_snow_create",
    "pbsm3d_snow_drift",
    "pbsm3d_snow_outputs",
    "pbsm3d_snow_get_state",
    "pbsm3d_snow_set_state",
    "pbsm3d_snow_destroy",
    "pbsm3d_snow_last_error",
    "pbsm3d_mlp_create",
    "pbsm3d_mlp_destroy",
    "pbsm3d_mlp_param_count",
    "pbsm3d_mlp_get_params",
    "pbsm3d_mlp_set_params",
    "pbsm3d_mlp_set_normalization",
    "pbsm3d_mlp_train",
    "pbsm3d_mlp_predict",
    "pbsm3d_mlp_predict_path",
    "pbsm3d_mlp_gradients",
    "pbsm3d_mlp_last_error",
    "pbsm3d_training_pairs",
    "pbsm3d_set_warm_start",
    "pbsm3d_coarse_system",
)

N_FEATURES, N_TARGETS = 16, 11

# launch shapes of the position-order SpMV: v % 5 = 0 / 4 one thread per row (blocks of
# 256 / 128), 1 / 2 / 3 one thread per face column (blocks of 256 / 128 / 64); v >= 5 with the
# XCD workgroup mapping
SPMV_VARIANTS = 10
PRECOND_MODES = 8
PBSM3D_WIND_NOALLOC_MAX_FACES = 4194304  # include/pbsm3d_wind.h

KERNEL_CLASSES = ("assemble", "spmv64", "spmv32", "precond", "ortho", "residual", "deposition", "warmstart",
                  "factor")

_lib = None


class Pbsm3dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"PBSM3D error {code}: {msg}")
        self.code = code


def lib():
    """Load libpbsm3d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python -m chm_amd.build)")
        # One HIP runtime per process: torch ships its own libamdhip64.so.7, and a
        # process that uses both torch device tensors and this library must resolve
        # the library's libamdhip64.so.7 to that same copy, so torch loads first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.pbsm3d_abi_version.restype = C.c_int
        L.pbsm3d_config_default.argtypes = [C.POINTER(abi.Config)]
        L.pbsm3d_config_default.restype = None
        L.pbsm3d_face_geometry.argtypes = [C.c_int64, dp, ip, dp] + [dp] * 7
        L.pbsm3d_create.argtypes = [C.POINTER(vp), C.POINTER(abi.Mesh), C.POINTER(abi.Config), C.c_int]
        L.pbsm3d_step.argtypes = [vp, C.POINTER(abi.Forcing), C.c_int32, C.POINTER(abi.Out), C.POINTER(abi.Stats)]
        L.pbsm3d_system_nnz.argtypes = [vp, C.c_int]
        L.pbsm3d_system_nnz.restype = C.c_int64
        L.pbsm3d_export_system.argtypes = [vp, C.c_int, ip, ip, dp, dp]
        L.pbsm3d_get_state.argtypes = [vp, dp, dp]
        L.pbsm3d_set_state.argtypes = [vp, dp, dp]
        L.pbsm3d_destroy.argtypes = [vp]
        L.pbsm3d_destroy.restype = None
        L.pbsm3d_last_error.argtypes = [vp]
        L.pbsm3d_last_error.restype = C.c_char_p
        L.pbsm3d_profile_enable.argtypes = [vp, C.c_int]
        L.pbsm3d_profile_read.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), dp, C.c_int]
        L.pbsm3d_profile_classes.argtypes = [vp, C.c_uint32]
        L.pbsm3d_profile_sample.argtypes = [vp, C.c_int]
        L.pbsm3d_spmv_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_int32]
        L.pbsm3d_spmv_select.argtypes = [vp, C.c_int, C.c_int]
        L.pbsm3d_face_colors.argtypes = [C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.pbsm3d_precond_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_float)]
        L.pbsm3d_precond_select.argtypes = [vp, C.c_int, C.c_int]
        L.pbsm3d_bicg_restart.argtypes = [vp, C.c_double]
        L.pbsm3d_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(abi.Mesh), C.POINTER(abi.Config),
                                          C.c_int]
        L.pbsm3d_batch_step.argtypes = [vp, C.POINTER(abi.Forcing), C.c_int32, C.POINTER(abi.Out),
                                        C.POINTER(abi.Stats)]
        L.pbsm3d_batch_set_warm_start.argtypes = [vp, vp]
        L.pbsm3d_batch_use_graphs.argtypes = [vp, C.c_int]
        L.pbsm3d_batch_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.pbsm3d_batch_get_state.argtypes = [vp, C.c_int32, dp, dp]
        L.pbsm3d_batch_set_state.argtypes = [vp, C.c_int32, dp, dp]
        L.pbsm3d_batch_training_pairs.argtypes = [vp, C.c_int32, vp, vp]
        L.pbsm3d_batch_destroy.argtypes = [vp]
        L.pbsm3d_batch_destroy.restype = None
        L.pbsm3d_batch_last_error.argtypes = [vp]
        L.pbsm3d_batch_last_error.restype = C.c_char_p
        L.pbsm3d_fetch_config_default.argtypes = [C.POINTER(abi.FetchConfig)]
        L.pbsm3d_fetch_config_default.restype = None
        L.pbsm3d_fetch_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int64, dp, dp, dp, dp,
                                          C.POINTER(abi.FetchConfig)]
        L.pbsm3d_fetch_run.argtypes = [vp, vp, vp, C.c_int]
        L.pbsm3d_fetch_nearest.argtypes = [vp, C.c_int64, dp, dp, C.POINTER(C.c_int32)]
        L.pbsm3d_fetch_destroy.argtypes = [vp]
        L.pbsm3d_fetch_destroy.restype = None
        L.pbsm3d_fetch_last_error.argtypes = [vp]
        L.pbsm3d_fetch_last_error.restype = C.c_char_p
        L.pbsm3d_wind_config_default.argtypes = [C.POINTER(abi.WindConfig)]
        L.pbsm3d_wind_config_default.restype = None
        L.pbsm3d_wind_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int64, C.POINTER(C.c_int32), dp, dp, dp, dp,
                                         C.POINTER(abi.WindConfig)]
        L.pbsm3d_wind_run.argtypes = [vp, vp, vp, vp, C.c_int]
        L.pbsm3d_wind_run_noalloc.argtypes = [C.c_int, C.c_int64, vp, vp, vp, vp, vp, C.POINTER(abi.WindConfig), vp,
                                              vp, vp]
        L.pbsm3d_wind_destroy.argtypes = [vp]
        L.pbsm3d_wind_destroy.restype = None
        L.pbsm3d_wind_last_error.argtypes = [vp]
        L.pbsm3d_wind_last_error.restype = C.c_char_p
        L.pbsm3d_snow_config_default.argtypes = [C.POINTER(abi.SnowConfig)]
        L.pbsm3d_snow_config_default.restype = None
        L.pbsm3d_snow_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int64, dp, C.POINTER(abi.SnowConfig)]
        L.pbsm3d_snow_drift.argtypes = [vp, vp, C.c_int]
        L.pbsm3d_snow_outputs.argtypes = [vp, vp, vp, C.c_int]
        L.pbsm3d_snow_get_state.argtypes = [vp, dp]
        L.pbsm3d_snow_set_state.argtypes = [vp, dp]
        L.pbsm3d_snow_destroy.argtypes = [vp]
        L.pbsm3d_snow_destroy.restype = None
        L.pbsm3d_snow_last_error.argtypes = [vp]
        L.pbsm3d_snow_last_error.restype = C.c_char_p
        fp = C.POINTER(C.c_float)
        L.pbsm3d_mlp_create.argtypes = [C.POINTER(vp), C.c_int, C.c_uint64]
        L.pbsm3d_mlp_destroy.argtypes = [vp]
        L.pbsm3d_mlp_destroy.restype = None
        L.pbsm3d_mlp_param_count.restype = C.c_int64
        L.pbsm3d_mlp_get_params.argtypes = [vp, fp]
        L.pbsm3d_mlp_set_params.argtypes = [vp, fp]
        L.pbsm3d_mlp_set_normalization.argtypes = [vp, fp]
        L.pbsm3d_mlp_train.argtypes = [vp, vp, vp, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, dp]
        L.pbsm3d_mlp_predict.argtypes = [vp, vp, vp, C.c_int64]
        L.pbsm3d_mlp_predict_path.argtypes = [vp, vp, vp, C.c_int64, C.c_int]
        L.pbsm3d_mlp_gradients.argtypes = [vp, fp, fp, C.c_int64, fp, dp]
        L.pbsm3d_mlp_last_error.argtypes = [vp]
        L.pbsm3d_mlp_last_error.restype = C.c_char_p
        L.pbsm3d_training_pairs.argtypes = [vp, vp, vp]
        L.pbsm3d_set_warm_start.argtypes = [vp, vp]
        L.pbsm3d_coarse_system.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _check(rc: int, h=None):
    if rc != 0:
        msg = lib().pbsm3d_last_error(h).decode(errors="replace")
        raise Pbsm3dError(rc, msg)


def face_geometry(mesh) -> abi.Geometry:
    """Per-face geometry of a :class:`chm_amd.meshgen.Mesh` through the library's face API."""
    return abi.geometry_args(lib().pbsm3d_face_geometry, mesh)


def face_colors(neigh) -> tuple:
    """Face colours of the multicolour D-ILU preconditioner (host computation) and their count."""
    nb = np.ascontiguousarray(neigh, dtype=np.int32).reshape(-1)
    F = nb.size // 3
    col = np.zeros(F, dtype=np.int32)
    nc = lib().pbsm3d_face_colors(F, nb.ctypes.data_as(C.POINTER(C.c_int32)), col.ctypes.data_as(C.POINTER(C.c_int32)))
    if nc < 0:
        raise Pbsm3dError(nc, "pbsm3d_face_colors failed")
    return col, nc


def default_config(**overrides) -> abi.Config:
    c = abi.Config()
    lib().pbsm3d_config_default(C.byref(c))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Context:
    """One PBSM3D instance on one GPU: ``init`` in the constructor, ``run`` in :meth:`step`."""

    def __init__(self, geom: abi.Geometry, cfg: abi.Config | None = None, device: int = 0):
        self.cfg = cfg if cfg is not None else default_config()
        self.F = geom.n_faces
        self.L = int(self.cfg.nLayer)
        keep = abi.Keep()
        mesh = geom.struct(keep)
        self._h = C.c_void_p()
        _check(lib().pbsm3d_create(C.byref(self._h), C.byref(mesh), C.byref(self.cfg), device))
        self.last_stats: dict = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().pbsm3d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, forcing: dict, dt: int = 900, outs: dict | None = None, want_fields: bool = False,
             on_device: bool = False) -> tuple[dict, dict]:
        """Run one timestep.

        ``forcing`` maps the reference's variable names to arrays of F values
        (numpy host arrays, or torch device tensors with ``on_device=True``).
        Returns (outputs, stats); outputs are numpy arrays unless ``outs``
        supplies device tensors (then written in place).
        """
        keep = abi.Keep()
        f = abi.forcing_struct(forcing, keep, on_device=on_device)
        out_on_dev = False
        if outs is None:
            outs = {k: np.zeros(self.F) for k in abi.OUT_FIELDS}
            if want_fields:
                outs["conc"] = np.zeros(self.F * self.L)
                outs["qdep"] = np.zeros(self.F)
        else:
            out_on_dev = any(hasattr(v, "data_ptr") for v in outs.values())
        o = abi.out_struct(outs, keep, on_device=out_on_dev)
        st = abi.Stats()
        _check(lib().pbsm3d_step(self._h, C.byref(f), int(dt), C.byref(o), C.byref(st)), self._h)
        self.last_stats = st.as_dict()
        return outs, self.last_stats

    def system(self, which: int = 0):
        """(row_ptr, col, val, rhs) of the last assembled suspension (0) or deposition (1) system."""
        L = lib()
        nnz = L.pbsm3d_system_nnz(self._h, which)
        n = self.F * (self.L if which == 0 else 1)
        rp = np.zeros(n + 1, dtype=np.int32)
        col = np.zeros(nnz, dtype=np.int32)
        val = np.zeros(nnz)
        rhs = np.zeros(n)
        keep = abi.Keep()
        _check(L.pbsm3d_export_system(self._h, which, keep.iptr(rp), keep.iptr(col), keep.dptr(val), keep.dptr(rhs)),
               self._h)
        return rp, col, val, rhs

    def profile(self, enable: bool = True, classes=None, every: int = 1):
        """Turn HIP-event timing of the kernel classes on or off; classes: the names
        (KERNEL_CLASSES) to bracket, None = all; every: bracket only every every-th launch
        of each class."""
        mask = (1 << len(KERNEL_CLASSES)) - 1 if classes is None else sum(
            1 << KERNEL_CLASSES.index(n) for n in classes)
        _check(lib().pbsm3d_profile_classes(self._h, mask), self._h)
        _check(lib().pbsm3d_profile_sample(self._h, every), self._h)
        _check(lib().pbsm3d_profile_enable(self._h, 1 if enable else 0), self._h)

    def profile_read(self, reset: bool = True) -> dict:
        """{class: (launches, device ms)} since the last reset."""
        out = {}
        for k, name in enumerate(KERNEL_CLASSES):
            n, ms = C.c_int64(0), C.c_double(0)
            _check(lib().pbsm3d_profile_read(self._h, k, C.byref(n), C.byref(ms), 1 if reset else 0), self._h)
            out[name] = (int(n.value), float(ms.value))
        return out

    def spmv_tuning(self, which: int = 0) -> dict:
        """Launch shape the SpMV tuner chose for the FP64 (which=0) or FP32 (1) operator and
        the average time of every shape tried (ms); chosen = -1 before the first solve."""
        n = SPMV_VARIANTS
        chosen, ms = C.c_int32(-1), (C.c_float * n)()
        _check(lib().pbsm3d_spmv_tuning(self._h, which, C.byref(chosen), ms, n), self._h)
        return {"chosen": int(chosen.value), "ms": [round(float(v), 5) for v in ms]}

    def precond_tuning(self, which: int = 0) -> dict:
        """The D-ILU sweep mode of the FP64 (which=0) / FP32 (1) solve -- bit 0 the XCD
        workgroup mapping; split operator only: bit 1 non-temporal factor loads, bit 2
        layer-parallel sweeps -- and the tuner's average apply time of each mode (ms,
        0 = not tried)."""
        mode, ms = C.c_int32(-1), (C.c_float * PRECOND_MODES)()
        _check(lib().pbsm3d_precond_tuning(self._h, which, C.byref(mode), ms), self._h)
        return {"mode": int(mode.value), "ms": [round(float(v), 5) for v in ms]}

    def precond_select(self, which: int, mode: int):
        """Fix the D-ILU sweep mode of the FP64 (0) / FP32 (1) solve (0 .. 1, or 0 .. 7 with
        the split operator); -1 re-tunes."""
        _check(lib().pbsm3d_precond_select(self._h, which, mode), self._h)

    def bicg_restart(self, eps: float):
        """Shadow-residual restart threshold of the FP32 BiCGStab runs (default 1e-5, 0: off);
        stats["susp_restarts"] counts the restarts of a step."""
        _check(lib().pbsm3d_bicg_restart(self._h, float(eps)), self._h)

    def spmv_select(self, which: int, variant: int):
        """Fix the SpMV launch shape of the FP64 (0) / FP32 (1) operator; -1 re-tunes."""
        _check(lib().pbsm3d_spmv_select(self._h, which, variant), self._h)

    def training_pairs(self, X, Y):
        """Write features [F x 16] and solved targets [F x 11] of the last step into
        float32 device tensors X, Y (needs nLayer = 10)."""
        _check(lib().pbsm3d_training_pairs(self._h, C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr())), self._h)

    def set_warm_start(self, net: "Mlp | None"):
        self._warm = net  # keep the network alive while attached
        _check(lib().pbsm3d_set_warm_start(self._h, net._h if net is not None else None), self._h)

    def coarse_system(self):
        """The learned coarse space of the last step (cfg.coarse_tiles = T > 0): tile of each
        face [F], the basis w [L*F] (reference order) and the Galerkin matrix G [T x T]."""
        T = int(self.cfg.coarse_tiles)
        tile = np.zeros(self.F, dtype=np.int32)
        w = np.zeros(self.F * self.L)
        G = np.zeros(T * T)
        _check(lib().pbsm3d_coarse_system(self._h, tile.ctypes.data_as(C.POINTER(C.c_int32)),
                                           w.ctypes.data_as(C.POINTER(C.c_double)),
                                           G.ctypes.data_as(C.POINTER(C.c_double))), self._h)
        return tile, w, G.reshape(T, T)

    def get_state(self):
        a, b = np.zeros(self.F), np.zeros(self.F)
        keep = abi.Keep()
        _check(lib().pbsm3d_get_state(self._h, keep.dptr(a), keep.dptr(b)), self._h)
        return a, b

    def set_state(self, sum_drift, sum_subl):
        keep = abi.Keep()
        _check(lib().pbsm3d_set_state(self._h, keep.dptr(sum_drift), keep.dptr(sum_subl)), self._h)


class Mlp:
    """The warm-start network (16 -> 128 -> 128 -> 128 -> 11) living on one GPU."""

    def __init__(self, device: int = 0, seed: int = 0):
        self._h = C.c_void_p()
        rc = lib().pbsm3d_mlp_create(C.byref(self._h), device, seed)
        if rc != 0:
            raise Pbsm3dError(rc, "pbsm3d_mlp_create failed")

    def _ck(self, rc):
        if rc != 0:
            raise Pbsm3dError(rc, lib().pbsm3d_mlp_last_error(self._h).decode(errors="replace"))

    def close(self):
        if getattr(self, "_h", None):
            lib().pbsm3d_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def param_count() -> int:
        return int(lib().pbsm3d_mlp_param_count())

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.param_count(), dtype=np.float32)
        self._ck(lib().pbsm3d_mlp_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._ck(lib().pbsm3d_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def set_normalization(self, x_mean, x_std, y_mean, y_std):
        v = np.concatenate([x_mean, x_std, y_mean, y_std]).astype(np.float32)
        assert v.size == 2 * N_FEATURES + 2 * N_TARGETS
        self._ck(lib().pbsm3d_mlp_set_normalization(self._h, v.ctypes.data_as(C.POINTER(C.c_float))))

    @staticmethod
    def _f32(*ts):
        import torch

        for t in ts:
            if t.dtype != torch.float32 or not t.is_contiguous() or not t.is_cuda:
                raise TypeError("network arrays must be contiguous float32 device tensors")

    def train(self, X, Y, epochs: int = 4, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> float:
        """X [n x 16], Y [n x 11] float32 device tensors."""
        self._f32(X, Y)
        loss = C.c_double(0)
        self._ck(lib().pbsm3d_mlp_train(self._h, C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()), X.shape[0],
                                        epochs, batch, lr, seed, C.byref(loss)))
        return loss.value

    def predict(self, X, Y, path: int = 0):
        """Y = net(X) on device tensors; path 0 default, 1 fused one-launch kernel, 2 layer by layer."""
        self._f32(X, Y)
        self._ck(lib().pbsm3d_mlp_predict_path(self._h, C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()),
                                               X.shape[0], path))

    def gradients(self, X: np.ndarray, Y: np.ndarray):
        """(loss, flat gradient) of the normalised MSE on one batch of host arrays."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.zeros(self.param_count(), dtype=np.float32)
        loss = C.c_double(0)
        fp = C.POINTER(C.c_float)
        self._ck(lib().pbsm3d_mlp_gradients(self._h, X.ctypes.data_as(fp), Y.ctypes.data_as(fp), X.shape[0],
                                            g.ctypes.data_as(fp), C.byref(loss)))
        return loss.value, g


def fetch_config(**overrides) -> abi.FetchConfig:
    c = abi.FetchConfig()
    lib().pbsm3d_fetch_config_default(C.byref(c))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Fetch:
    """Upwind fetch on the GPU (the fetchr module's computation) for fixed face centroids.

    canopy_height: per-face canopy heights, or None when the mesh carries no vegetation."""

    def __init__(self, center_x, center_y, center_z, canopy_height=None, cfg: abi.FetchConfig | None = None,
                 device: int = 0):
        L = lib()
        self._keep = abi.Keep()
        self.n_faces = int(len(center_x))
        cfg = cfg if cfg is not None else fetch_config()
        h = C.c_void_p()
        k = self._keep
        rc = L.pbsm3d_fetch_create(C.byref(h), device, self.n_faces, k.dptr(center_x), k.dptr(center_y),
                                   k.dptr(center_z), k.dptr(canopy_height) if canopy_height is not None else None,
                                   C.byref(cfg))
        if rc != 0:
            raise Pbsm3dError(rc, L.pbsm3d_fetch_last_error(None).decode())
        self._h = h

    def run(self, vw_dir):
        """fetch per face for wind directions vw_dir: numpy (host) or a float64 CUDA tensor (device)."""
        L = lib()
        if hasattr(vw_dir, "data_ptr"):
            import torch

            out = torch.empty_like(vw_dir)
            rc = L.pbsm3d_fetch_run(self._h, C.c_void_p(vw_dir.data_ptr()), C.c_void_p(out.data_ptr()), 1)
        else:
            vw = np.ascontiguousarray(vw_dir, dtype=np.float64)
            out = np.zeros(self.n_faces)
            rc = L.pbsm3d_fetch_run(self._h, vw.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), 0)
        if rc != 0:
            raise Pbsm3dError(rc, L.pbsm3d_fetch_last_error(self._h).decode())
        return out

    def nearest(self, qx, qy):
        L = lib()
        qx = np.ascontiguousarray(qx, dtype=np.float64)
        qy = np.ascontiguousarray(qy, dtype=np.float64)
        idx = np.zeros(len(qx), dtype=np.int32)
        rc = L.pbsm3d_fetch_nearest(self._h, len(qx), qx.ctypes.data_as(C.POINTER(C.c_double)),
                                    qy.ctypes.data_as(C.POINTER(C.c_double)), idx.ctypes.data_as(C.POINTER(C.c_int32)))
        if rc != 0:
            raise Pbsm3dError(rc, L.pbsm3d_fetch_last_error(self._h).decode())
        return idx

    def close(self):
        if getattr(self, "_h", None):
            lib().pbsm3d_fetch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Batch:
    """Parameter-ensemble members advanced together (include/pbsm3d_batch.h): members of
    equal nLayer share one union system whose sweeps and BiCGStab kernels are launched
    once for all of them, with per-member BiCGStab scalars and stop flags."""

    def __init__(self, geoms, cfgs, device: int = 0):
        L = lib()
        self.M = len(geoms)
        if len(cfgs) != self.M:
            raise ValueError("one config per member")
        self._keep = abi.Keep()
        meshes = (abi.Mesh * self.M)(*[g.struct(self._keep) for g in geoms])
        self.cfgs = (abi.Config * self.M)(*cfgs)
        self.F = [g.n_faces for g in geoms]
        self.L = [int(c.nLayer) for c in cfgs]
        h = C.c_void_p()
        rc = L.pbsm3d_batch_create(C.byref(h), self.M, meshes, self.cfgs, device)
        if rc != 0:
            raise Pbsm3dError(rc, L.pbsm3d_batch_last_error(None).decode(errors="replace"))
        self._h = h

    def _check(self, rc):
        if rc != 0:
            raise Pbsm3dError(rc, lib().pbsm3d_batch_last_error(self._h).decode(errors="replace"))

    def step(self, forcings, dt: int = 900, outs=None, want_fields: bool = False, on_device: bool = False):
        """One timestep of every member: forcings[m] as Context.step takes it; returns
        (outputs per member, stats per member)."""
        keep = abi.Keep()
        fs = (abi.Forcing * self.M)(*[abi.forcing_struct(f, keep, on_device=on_device) for f in forcings])
        if outs is None:
            outs = []
            for m in range(self.M):
                o = {k: np.zeros(self.F[m]) for k in abi.OUT_FIELDS}
                if want_fields:
                    o["conc"] = np.zeros(self.F[m] * self.L[m])
                    o["qdep"] = np.zeros(self.F[m])
                outs.append(o)
        os_ = (abi.Out * self.M)(*[abi.out_struct(o, keep, on_device=any(hasattr(v, "data_ptr") for v in o.values()))
                                   for o in outs])
        st = (abi.Stats * self.M)()
        self._check(lib().pbsm3d_batch_step(self._h, fs, int(dt), os_, st))
        return outs, [s.as_dict() for s in st]

    def set_warm_start(self, net):
        self._check(lib().pbsm3d_batch_set_warm_start(self._h, net._h if net is not None else None))

    def use_graphs(self, enable: bool = True):
        """Replay the BiCGStab iterations from captured hipGraphs (default) or launch them."""
        self._check(lib().pbsm3d_batch_use_graphs(self._h, 1 if enable else 0))

    def info(self) -> dict:
        u, n, it = C.c_int32(0), C.c_int64(0), C.c_int64(0)
        self._check(lib().pbsm3d_batch_info(self._h, C.byref(u), C.byref(n), C.byref(it)))
        return {"unions": int(u.value), "launches": int(n.value), "iterations": int(it.value)}

    def get_state(self, m: int):
        a, b = np.zeros(self.F[m]), np.zeros(self.F[m])
        keep = abi.Keep()
        self._check(lib().pbsm3d_batch_get_state(self._h, m, keep.dptr(a), keep.dptr(b)))
        return a, b

    def training_pairs(self, m: int, X, Y):
        """Member m's features [F_m x 16] and solved targets [F_m x 11] of the last step
        into float32 device tensors X, Y (nLayer 10)."""
        self._check(lib().pbsm3d_batch_training_pairs(self._h, m, C.c_void_p(X.data_ptr()),
                                                       C.c_void_p(Y.data_ptr())))

    def set_state(self, m: int, sum_drift, sum_subl):
        keep = abi.Keep()
        self._check(lib().pbsm3d_batch_set_state(self._h, m, keep.dptr(sum_drift), keep.dptr(sum_subl)))

    def close(self):
        if getattr(self, "_h", None):
            lib().pbsm3d_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wind_config(**overrides) -> abi.WindConfig:
    c = abi.WindConfig()
    lib().pbsm3d_wind_config_default(C.byref(c))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Wind:
    """Vertical wind scaling on the GPU (the scale_wind_vert module's computation):
    U_2m_above_srf from U_R and snowdepthavg."""

    def __init__(self, neigh, center_x, center_y, canopy_height=None, lai=None, cfg: abi.WindConfig | None = None,
                 device: int = 0):
        L = lib()
        self._keep = abi.Keep()
        k = self._keep
        self.n_faces = int(len(center_x))
        cfg = cfg if cfg is not None else wind_config()
        h = C.c_void_p()
        rc = L.pbsm3d_wind_create(C.byref(h), device, self.n_faces, k.iptr(np.asarray(neigh).reshape(-1)),
                                  k.dptr(center_x), k.dptr(center_y), k.dptr(canopy_height), k.dptr(lai),
                                  C.byref(cfg))
        if rc != 0:
            raise Pbsm3dError(rc, L.pbsm3d_wind_last_error(None).decode())
        self._h = h

    def run(self, U_R, snowdepthavg=None):
        """numpy arrays (host) or float64 CUDA tensors (device); snowdepthavg None: absent."""
        L = lib()
        if hasattr(U_R, "data_ptr"):
            import torch

            out = torch.empty_like(U_R)
            sd = C.c_void_p(snowdepthavg.data_ptr()) if snowdepthavg is not None else None
            rc = L.pbsm3d_wind_run(self._h, C.c_void_p(U_R.data_ptr()), sd, C.c_void_p(out.data_ptr()), 1)
        else:
            u = np.ascontiguousarray(U_R, dtype=np.float64)
            sdv = np.ascontiguousarray(snowdepthavg, dtype=np.float64) if snowdepthavg is not None else None
            out = np.zeros(self.n_faces)
            rc = L.pbsm3d_wind_run(self._h, u.ctypes.data_as(C.c_void_p),
                                   sdv.ctypes.data_as(C.c_void_p) if sdv is not None else None,
                                   out.ctypes.data_as(C.c_void_p), 0)
        if rc != 0:
            raise Pbsm3dError(rc, L.pbsm3d_wind_last_error(self._h).decode())
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().pbsm3d_wind_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def wind_run_noalloc(neigh, center_x, center_y, U_R, out, snowdepthavg=None, canopy_height=None, lai=None,
                     cfg: abi.WindConfig | None = None, device: int = 0) -> None:
    """pbsm3d_wind_run_noalloc on device tensors (neigh int32 [F, 3], the rest float64 [F]):
    writes U_2m_above_srf into ``out``; the library allocates nothing."""
    L = lib()
    cfg = cfg if cfg is not None else wind_config()

    def p(t):
        return C.c_void_p(t.data_ptr()) if t is not None else None

    rc = L.pbsm3d_wind_run_noalloc(device, int(center_x.shape[0]), p(neigh), p(center_x), p(center_y),
                                   p(canopy_height), p(lai), C.byref(cfg), p(U_R), p(snowdepthavg), p(out))
    if rc != 0:
        raise Pbsm3dError(rc, L.pbsm3d_wind_last_error(None).decode())


def snow_config(**overrides) -> abi.SnowConfig:
    c = abi.SnowConfig()
    lib().pbsm3d_snow_config_default(C.byref(c))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Snow:
    """Snowcover update from PBSM3D's drift_mass on the GPU (the drift consumer of the
    snobal module); provides swe and snowdepthavg."""

    def __init__(self, n_faces: int, swe0=None, cfg: abi.SnowConfig | None = None, device: int = 0):
        L = lib()
        self._keep = abi.Keep()
        self.n_faces = int(n_faces)
        cfg = cfg if cfg is not None else snow_config()
        h = C.c_void_p()
        rc = L.pbsm3d_snow_create(C.byref(h), device, self.n_faces,
                                  self._keep.dptr(swe0) if swe0 is not None else None, C.byref(cfg))
        if rc != 0:
            raise Pbsm3dError(rc, L.pbsm3d_snow_last_error(None).decode())
        self._h = h

    def _ck(self, rc):
        if rc != 0:
            raise Pbsm3dError(rc, lib().pbsm3d_snow_last_error(self._h).decode())

    def drift(self, drift_mass):
        """Apply one step's drift_mass: numpy array (host) or float64 CUDA tensor (device)."""
        if hasattr(drift_mass, "data_ptr"):
            self._ck(lib().pbsm3d_snow_drift(self._h, C.c_void_p(drift_mass.data_ptr()), 1))
        else:
            d = np.ascontiguousarray(drift_mass, dtype=np.float64)
            self._ck(lib().pbsm3d_snow_drift(self._h, d.ctypes.data_as(C.c_void_p), 0))

    def outputs(self, device: bool = False):
        """(swe, snowdepthavg): numpy arrays, or float64 CUDA tensors with device=True."""
        if device:
            import torch

            swe = torch.empty(self.n_faces, dtype=torch.float64, device="cuda:0")
            sd = torch.empty_like(swe)
            self._ck(lib().pbsm3d_snow_outputs(self._h, C.c_void_p(swe.data_ptr()), C.c_void_p(sd.data_ptr()), 1))
            return swe, sd
        swe, sd = np.zeros(self.n_faces), np.zeros(self.n_faces)
        self._ck(lib().pbsm3d_snow_outputs(self._h, swe.ctypes.data_as(C.c_void_p), sd.ctypes.data_as(C.c_void_p),
                                           0))
        return swe, sd

    def get_state(self) -> np.ndarray:
        """[len(abi.SNOW_FIELDS) x F] snowcover state."""
        st = np.zeros((len(abi.SNOW_FIELDS), self.n_faces))
        self._ck(lib().pbsm3d_snow_get_state(self._h, st.ctypes.data_as(C.POINTER(C.c_double))))
        return st

    def set_state(self, state):
        st = np.ascontiguousarray(state, dtype=np.float64)
        assert st.shape == (len(abi.SNOW_FIELDS), self.n_faces)
        self._ck(lib().pbsm3d_snow_set_state(self._h, st.ctypes.data_as(C.POINTER(C.c_double))))

    def close(self):
        if getattr(self, "_h", None):
            lib().pbsm3d_snow_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
