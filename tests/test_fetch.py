"""Upwind fetch (CHM fetchr, src/modules/fetchr.cpp:28-120).

CPU: the oracle's restatement against closed forms and an independent numpy
nearest-centroid search.  GPU: the fetch kernel and its grid nearest-face
search against the oracle, bit for bit (the outputs are whole multiples of the
search step, the search returns face indices)."""
import math

import numpy as np
import pytest

from chm_amd import abi, meshgen


def _centroids(mesh):
    c = mesh.vertex[mesh.elem].mean(axis=1)
    return np.ascontiguousarray(c[:, 0]), np.ascontiguousarray(c[:, 1]), np.ascontiguousarray(c[:, 2])


def _numpy_nearest(cx, cy, qx, qy):
    out = np.empty(len(qx), dtype=np.int32)
    for k in range(len(qx)):
        dx = qx[k] - cx
        dy = qy[k] - cy
        out[k] = np.argmin(dx * dx + dy * dy)  # first minimum = lowest index
    return out


def _queries(cx, cy, n, seed):
    rng = np.random.default_rng(seed)
    x0, x1, y0, y1 = cx.min(), cx.max(), cy.min(), cy.max()
    qx = rng.uniform(x0 - 2000, x1 + 2000, n)
    qy = rng.uniform(y0 - 2000, y1 + 2000, n)
    qx[:50], qy[:50] = cx[:50], cy[:50]  # exactly on centroids
    return qx, qy


def _fetch_cfg(**kw):
    c = abi.FetchConfig()
    c.steps, c.incl_veg, c.max_distance, c.I = 10, 1, 1000.0, 0.06
    for k, v in kw.items():
        setattr(c, k, v)
    return c


# ---------------------------------------------------------------------------
# oracle (CPU)
# ---------------------------------------------------------------------------
def test_oracle_nearest_matches_numpy_and_breaks_ties_low(oracle):
    mesh = meshgen.synthetic_mesh(1500, seed=3)
    cx, cy, _ = _centroids(mesh)
    qx, qy = _queries(cx, cy, 600, 1)
    np.testing.assert_array_equal(oracle.fetch_nearest(cx, cy, qx, qy), _numpy_nearest(cx, cy, qx, qy))
    # duplicated centroids: the lower index wins
    cx2, cy2 = np.concatenate([cx, cx[:10]]), np.concatenate([cy, cy[:10]])
    idx = oracle.fetch_nearest(cx2, cy2, cx[:10], cy[:10])
    np.testing.assert_array_equal(idx, np.arange(10))


def _plane(n_side=40, spacing=25.0, slope=0.0):
    """Centroids on a square lattice; z rises towards the west by `slope` per metre."""
    g = np.arange(n_side) * spacing
    X, Y = np.meshgrid(g, g)
    cx, cy = X.ravel().astype(float), Y.ravel().astype(float)
    cz = 1000.0 + slope * (cx.max() - cx)
    return cx, cy, cz


def test_oracle_fetch_closed_forms(oracle):
    cfg = _fetch_cfg()
    # flat terrain: nothing upwind is higher than z + distance * I -> max_distance
    cx, cy, cz = _plane()
    west = np.full(len(cx), 270.0)
    f = oracle.fetch(cx, cy, cz, None, cfg, west)
    np.testing.assert_array_equal(f, 1000.0)
    # terrain rising upwind faster than I: the first step (100 m) already hits,
    # except where the upwind point leaves the lattice and the nearest face is the
    # face itself or lower
    cx, cy, cz = _plane(slope=0.1)
    f = oracle.fetch(cx, cy, cz, None, cfg, west)
    inner = cx >= 100.0 + 1e-9
    np.testing.assert_array_equal(f[inner], 100.0)
    # wind from the east on the same slope: upwind is lower -> max_distance
    f = oracle.fetch(cx, cy, cz, None, cfg, np.full(len(cx), 90.0))
    np.testing.assert_array_equal(f, 1000.0)
    # canopy taller than 1 m on the face itself -> 0 (fetchr.cpp:66-77)
    cx, cy, cz = _plane()
    can = np.zeros(len(cx))
    can[::7] = 1.5
    f = oracle.fetch(cx, cy, cz, can, cfg, west)
    np.testing.assert_array_equal(f[::7], 0.0)
    # a uniform 0.5 m canopy: x_sss = (((33.33333333*5 - 0.025) / (ln(60) * 0.001))**1.25) * 0.001
    # ~ 578 m > 100 m, so the first step resets the fetch; the same heights with
    # incl_veg = 0 are ignored
    x_sss = (((33.33333333 * 5 - 25.0 * 0.001) / (math.log(0.06 / 0.001) * 0.001)) ** (1 / 0.8)) * 0.001
    assert x_sss > 100.0
    can = np.full(len(cx), 0.5)
    np.testing.assert_array_equal(oracle.fetch(cx, cy, cz, can, cfg, west), 100.0)
    np.testing.assert_array_equal(oracle.fetch(cx, cy, cz, can, _fetch_cfg(incl_veg=0), west), 1000.0)


def test_oracle_fetch_steps_and_rise(oracle):
    """A ridge upwind (east) of a lattice: fetch is the first step whose nearest
    centroid lies on the ridge.  Lattice spacing 20 m, ridge x > 1000 m, steps of
    50 m; the query x + d has the ridge point 1020 nearest iff x + d > 1010 (at
    exactly 1010 the tie goes to the lower index, x = 1000)."""
    cx, cy, cz = _plane(n_side=60, spacing=20.0)
    cz = cz.copy()
    cz[cx > 1000.0] += 500.0
    cfg = _fetch_cfg(steps=20, max_distance=1000.0, I=1e-6)  # flat ground never counts, the ridge does
    f = oracle.fetch(cx, cy, cz, None, cfg, np.full(len(cx), 90.0))  # wind from the east
    sel = (cx <= 1000.0) & (cx >= 400.0)
    expect = (np.floor((1010.0 - cx[sel]) / 50.0) + 1.0) * 50.0
    np.testing.assert_array_equal(f[sel], expect)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def pb():
    from chm_amd import pbsm3d

    pbsm3d.lib()
    return pbsm3d


@pytest.mark.gpu
@pytest.mark.parametrize("points", [2500, 20000])
def test_gpu_nearest_matches_brute_force(pb, oracle, points):
    print(f"[case] fetch nearest points={points}", flush=True)
    mesh = meshgen.synthetic_mesh(points, seed=0)
    cx, cy, cz = _centroids(mesh)
    ctx = pb.Fetch(cx, cy, cz)
    qx, qy = _queries(cx, cy, 20000, 2)
    np.testing.assert_array_equal(ctx.nearest(qx, qy), oracle.fetch_nearest(cx, cy, qx, qy))
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("veg,kw", [(False, {}), (True, {}), (True, {"incl_veg": 0}),
                                    (False, {"steps": 7, "max_distance": 700.0, "I": 0.02})])
def test_gpu_fetch_matches_oracle(pb, oracle, mesh_c1, veg, kw):
    print(f"[case] fetch veg={veg} {kw}", flush=True)
    cx, cy, cz = _centroids(mesh_c1)
    F = len(cx)
    can = np.random.default_rng(5).uniform(0.0, 2.0, F) if veg else None
    cfg = _fetch_cfg(**kw)
    vw = meshgen.forcing_c1(F, seed=0)["vw_dir"]
    ctx = pb.Fetch(cx, cy, cz, can, cfg)
    got = ctx.run(vw)
    want = oracle.fetch(cx, cy, cz, can, cfg, vw)
    np.testing.assert_array_equal(got, want)
    assert len(np.unique(got)) > 2  # the search actually varies across the mesh
    ctx.close()


@pytest.mark.gpu
def test_gpu_fetch_device_arrays_feed_the_step(pb, oracle, mesh_c1):
    """fetch computed on the GPU from device-resident vw_dir, handed to the PBSM3D
    step as device forcing (tanh fetch), gives the oracle chain's fields."""
    import torch

    print("[case] fetch -> step", flush=True)
    cx, cy, cz = _centroids(mesh_c1)
    F = len(cx)
    fo = meshgen.forcing_c1(F, seed=4)
    fctx = pb.Fetch(cx, cy, cz)
    dfo = {k: torch.tensor(v, device="cuda:0") for k, v in fo.items()}
    dfo["fetch"] = fctx.run(dfo["vw_dir"])
    ofetch = oracle.fetch(cx, cy, cz, None, _fetch_cfg(), fo["vw_dir"])
    np.testing.assert_array_equal(dfo["fetch"].cpu().numpy(), ofetch)
    g = pb.face_geometry(mesh_c1)
    ctx = pb.Context(g, pb.default_config(nLayer=5, tol=1e-12, use_tanh_fetch=1))
    outs = {k: torch.zeros(F, dtype=torch.float64, device="cuda:0") for k in abi.OUT_FIELDS}
    ctx.step(dfo, outs=outs, on_device=True)
    orc = oracle.Oracle(g, oracle.oracle_config(nLayer=5, tol=1e-13, use_tanh_fetch=1))
    og, _ = orc.step(dict(fo, fetch=ofetch))
    for k in ("Qsalt", "Qsusp", "drift_mass"):
        got = outs[k].cpu().numpy()
        assert np.abs(got - og[k]).max() <= 1e-8 * np.abs(og[k]).max(), k
    fctx.close()
