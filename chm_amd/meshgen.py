"""Seeded synthetic meshes and forcing for the PBSM3D blowing-snow step.

Everything here is deterministic from an integer seed through a counter-based
64-bit PRNG (splitmix64 finaliser over (seed, stream, counter)), so the same
mesh and forcing can be regenerated bit-for-bit on any machine without
depending on numpy's generator implementations.

Mesh conventions follow the reference triangulation loader
(src/mesh/triangulation.cpp:206-356):
  * ``elem[i] = (v0, v1, v2)`` counter-clockwise (the signed area of
    src/mesh/triangulation.hpp:1830-1856 is positive),
  * ``neigh[i][j]`` is the face across the edge opposite vertex j, -1 on the
    domain boundary (CGAL / Triangle convention, triangulation.hpp:249-260),
  * faces are numbered 0..F-1 and that number is both the local and the global
    id (triangulation.cpp:1540-1563, single process).

Faces are renumbered along a Hilbert curve over their centroids, the same kind
of locality permutation the reference's mesher applies (docs/meshgen.rst), so
that lateral neighbours are close in memory.
"""
from __future__ import annotations

import dataclasses
import json
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_STREAM = np.uint64(0xD1B54A32D192ED03)

# stream ids (one independent sequence per purpose)
S_POINTS_X, S_POINTS_Y, S_ELEV, S_DIR, S_SPEED, S_TEMP, S_FIELD = range(1, 8)


def _splitmix(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(27)
        z *= np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return z


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n doubles in [0, 1) from counters offset..offset+n-1 of (seed, stream)."""
    ctr = np.arange(offset, offset + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = np.uint64(seed) * _GOLDEN + np.uint64(stream) * _STREAM
        z = _splitmix(ctr * _GOLDEN + key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Box-Muller standard normals from two interleaved uniform streams."""
    u = uniform(seed, stream, 2 * n)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclasses.dataclass
class Mesh:
    """Triangle mesh in the reference's .mesh layout (vertex xyz, elem, neigh)."""

    vertex: np.ndarray  # (nv, 3) float64
    elem: np.ndarray  # (F, 3) int32, CCW
    neigh: np.ndarray  # (F, 3) int32, -1 = boundary
    area: np.ndarray | None = None  # optional per-face "area" parameter override

    @property
    def n_faces(self) -> int:
        return int(self.elem.shape[0])

    def to_json(self) -> dict:
        """The reference .mesh JSON schema (triangulation.cpp:206-356)."""
        return {
            "mesh": {
                "version": "1.2.0",
                "vertex": self.vertex.tolist(),
                "elem": self.elem.tolist(),
                "neigh": self.neigh.tolist(),
                "nvertex": int(self.vertex.shape[0]),
                "nelem": int(self.elem.shape[0]),
                "is_geographic": 0,
                "proj4": "+proj=utm +zone=11 +datum=WGS84 +units=m +no_defs",
            }
        }

    @staticmethod
    def from_json(obj: dict, param: dict | None = None) -> "Mesh":
        m = obj["mesh"]
        nv, ne = int(m["nvertex"]), int(m["nelem"])
        vertex = np.asarray(m["vertex"], dtype=np.float64).reshape(nv, 3)
        elem = np.asarray(m["elem"], dtype=np.int32).reshape(ne, 3)
        neigh = np.asarray(m["neigh"], dtype=np.int32).reshape(ne, 3)
        if (neigh >= ne).any() or (neigh < -1).any():
            raise ValueError("neighbour index out of bounds")
        area = None
        if param is not None and "area" in param:
            area = np.asarray(param["area"], dtype=np.float64)
        return Mesh(vertex, elem, neigh, area)


def hilbert_index(x: np.ndarray, y: np.ndarray, order: int = 16) -> np.ndarray:
    """Hilbert-curve index of integer grid coordinates in [0, 2**order)."""
    x = x.astype(np.int64).copy()
    y = y.astype(np.int64).copy()
    d = np.zeros_like(x)
    s = 1 << (order - 1)
    while s > 0:
        rx = ((x & s) > 0).astype(np.int64)
        ry = ((y & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate
        flip = ry == 0
        swap_mask = flip & (rx == 1)
        x = np.where(swap_mask, s - 1 - x, x)
        y = np.where(swap_mask, s - 1 - y, y)
        xt = np.where(flip, y, x)
        yt = np.where(flip, x, y)
        x, y = xt, yt
        s >>= 1
    return d


def delaunay_mesh(points: np.ndarray, z: np.ndarray, reorder: bool = True) -> Mesh:
    """Delaunay triangulation of 2-D points (scipy/Qhull), CCW, reference neighbour convention."""
    from scipy.spatial import Delaunay

    tri = Delaunay(points)
    elem = tri.simplices.astype(np.int32).copy()
    neigh = tri.neighbors.astype(np.int32).copy()  # neighbors[i][j] is opposite vertex j
    p = points
    a = elem
    area2 = (p[a[:, 1], 0] - p[a[:, 0], 0]) * (p[a[:, 2], 1] - p[a[:, 0], 1]) - (
        p[a[:, 2], 0] - p[a[:, 0], 0]
    ) * (p[a[:, 1], 1] - p[a[:, 0], 1])
    cw = area2 < 0
    elem[cw, 1], elem[cw, 2] = elem[cw, 2].copy(), elem[cw, 1].copy()
    neigh[cw, 1], neigh[cw, 2] = neigh[cw, 2].copy(), neigh[cw, 1].copy()
    if reorder:
        c = p[elem].mean(axis=1)
        lo, hi = c.min(axis=0), c.max(axis=0)
        g = ((c - lo) / np.maximum(hi - lo, 1e-300) * 65535.0).astype(np.int64)
        perm = np.argsort(hilbert_index(g[:, 0], g[:, 1]), kind="stable")
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size)
        elem = elem[perm]
        nb = neigh[perm]
        neigh = np.where(nb >= 0, inv[np.maximum(nb, 0)], -1).astype(np.int32)
    vertex = np.column_stack([p[:, 0], p[:, 1], z]).astype(np.float64)
    return Mesh(vertex, elem.astype(np.int32), neigh.astype(np.int32))


def sinusoid_elevation(seed: int, x: np.ndarray, y: np.ndarray, zmin=200.0, zmax=2000.0) -> np.ndarray:
    """Sum of 8 random-phase plane sinusoids, rescaled to [zmin, zmax]."""
    u = uniform(seed, S_ELEV, 24)
    s = np.zeros_like(x)
    for k in range(8):
        theta = 2.0 * math.pi * u[3 * k]
        wavelength = 1000.0 * math.exp(math.log(8.0) * u[3 * k + 1])  # 1-8 km
        phase = 2.0 * math.pi * u[3 * k + 2]
        kk = 2.0 * math.pi / wavelength
        s += np.sin(kk * math.cos(theta) * x + kk * math.sin(theta) * y + phase)
    return zmin + (zmax - zmin) * (s - s.min()) / (s.max() - s.min())


def synthetic_mesh(n_points: int, seed: int = 0, extent: float = 1.0e4, reorder: bool = True) -> Mesh:
    """Uniform random points in [0, extent]^2 -> Delaunay -> sinusoid elevation."""
    x = uniform(seed, S_POINTS_X, n_points) * extent
    y = uniform(seed, S_POINTS_Y, n_points) * extent
    z = sinusoid_elevation(seed, x, y)
    return delaunay_mesh(np.column_stack([x, y]), z, reorder=reorder)


# ---------------------------------------------------------------------------
# forcing
# ---------------------------------------------------------------------------

Z_U_R = 50.0  # reference wind height, src/physics/Atmosphere.h:31
Z0_SNOW = 0.01  # src/physics/Snow.h:31

FORCING_FIELDS = ("U_R", "U_2m_above_srf", "vw_dir", "swe", "t", "rh", "snowdepthavg", "fetch", "p_snow_hours")


def u_ref_from_u2(u2: np.ndarray, snow_depth: np.ndarray) -> np.ndarray:
    """U_R whose log-profile gives u2 at 2 m above the snow surface (no canopy).

    Inverse of the log scaling in src/modules/scale_wind_vert.cpp:133-134 with
    Atmosphere::log_scale_wind (src/physics/Atmosphere.cpp:32-37).
    """
    z0 = Z0_SNOW
    return u2 * np.log((Z_U_R - (snow_depth + z0)) / z0) / np.log((2.0 - z0) / z0)


def forcing_c1(n_faces: int, seed: int = 0) -> dict:
    """Uniform 10 m/s at 2 m, direction N(270, 20) deg, -10 C, 70 % RH, 0.5 m snow, fetch
    1000 m and 24 h since snowfall (read only by use_PomLi_probability)."""
    f = n_faces
    sd = np.full(f, 0.5)
    u2 = np.full(f, 10.0)
    return {
        "U_R": u_ref_from_u2(u2, sd),
        "U_2m_above_srf": u2,
        "vw_dir": 270.0 + 20.0 * normal(seed, S_DIR, f),
        "swe": np.full(f, 50.0),
        "t": np.full(f, -10.0),
        "rh": np.full(f, 70.0),
        "snowdepthavg": sd,
        "fetch": np.full(f, 1000.0),
        "p_snow_hours": np.full(f, 24.0),
    }


def forcing_c2(n_faces: int, seed: int = 0) -> dict:
    """Per-face lognormal 2 m wind clipped to [4, 20] m/s, T ~ U(-25, -2) C."""
    f = n_faces
    out = forcing_c1(f, seed)
    u2 = np.clip(np.exp(math.log(9.0) + 0.35 * normal(seed, S_SPEED, f)), 4.0, 20.0)
    out["U_2m_above_srf"] = u2
    out["U_R"] = u_ref_from_u2(u2, out["snowdepthavg"])
    out["t"] = -25.0 + 23.0 * uniform(seed, S_TEMP, f)
    return out


class SmoothField:
    """Seeded random-Fourier-mode Gaussian field with a given correlation length.

    Used to advect wind speed and direction perturbations through time for
    multi-step runs: value(x, y, step) = field(x - a*step, y - b*step).
    """

    def __init__(self, seed: int, corr_len: float = 2000.0, n_modes: int = 64, stream: int = S_FIELD):
        u = uniform(seed, stream, 4 * n_modes)
        g = normal(seed, stream + 100, 2 * n_modes)
        self.kx = g[0::2] / corr_len
        self.ky = g[1::2] / corr_len
        self.phase = 2.0 * math.pi * u[0::4]
        self.n = n_modes

    def __call__(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        acc = np.zeros_like(x)
        for k in range(self.n):
            acc += np.cos(self.kx[k] * x + self.ky[k] * y + self.phase[k])
        return acc * math.sqrt(2.0 / self.n)


def forcing_advected(mesh: Mesh, step: int, seed: int = 0, advect=(150.0, 40.0)) -> dict:
    """C1 forcing with wind speed and direction perturbed by an advected smooth field."""
    f = mesh.n_faces
    c = mesh.vertex[mesh.elem].mean(axis=1)
    xs = c[:, 0] - advect[0] * step
    ys = c[:, 1] - advect[1] * step
    fs = SmoothField(seed, stream=S_FIELD)(xs, ys)
    fd = SmoothField(seed + 7919, stream=S_FIELD)(xs, ys)
    out = forcing_c1(f, seed)
    u2 = np.clip(10.0 * np.exp(0.25 * fs), 4.0, 20.0)
    out["U_2m_above_srf"] = u2
    out["U_R"] = u_ref_from_u2(u2, out["snowdepthavg"])
    out["vw_dir"] = 270.0 + 20.0 * fd
    return out


def save_mesh_json(mesh: Mesh, path: str) -> None:
    with open(path, "w") as fh:
        json.dump(mesh.to_json(), fh)


def load_mesh_json(path: str, param_path: str | None = None) -> Mesh:
    with open(path) as fh:
        obj = json.load(fh)
    param = None
    if param_path is not None:
        with open(param_path) as fh:
            param = json.load(fh)
    return Mesh.from_json(obj, param)
