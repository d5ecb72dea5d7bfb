"""Vertical wind scaling (CHM scale_wind_vert, src/modules/scale_wind_vert.cpp:28-242).

CPU: the oracle's E1 against scipy.special.exp1, its point scaling against the
closed-form log / exponential profiles, its neighbour spline smoothing against
an independent numpy restatement.  GPU: the HIP kernels against the oracle, and
the chain wind -> fetch -> PBSM3D step on device-resident arrays."""
import math

import numpy as np
import pytest
from scipy.special import exp1

from chm_amd import abi, meshgen


def _cfg(**kw):
    c = abi.WindConfig()
    c.ignore_canopy, c.smooth, c.z0_snow = 0, 1, 0.01
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _centroids(mesh):
    c = mesh.vertex[mesh.elem].mean(axis=1)
    return np.ascontiguousarray(c[:, 0]), np.ascontiguousarray(c[:, 1])


def _log_scale(u, z_in, z_out, sd, z0=0.01):
    return u * math.log((z_out - (sd + z0)) / z0) / math.log((z_in - (sd + z0)) / z0)


# ---------------------------------------------------------------------------
# oracle (CPU)
# ---------------------------------------------------------------------------
def test_oracle_e1_matches_scipy(oracle):
    xs = np.concatenate([np.logspace(-8, 1.7, 400), [0.5, 1.0, 1.0 + 1e-12, 2.0, 10.0]])
    got = np.array([oracle.expint_e1(x) for x in xs])
    np.testing.assert_allclose(got, exp1(xs), rtol=2e-14, atol=0)
    assert abs(oracle.expint_e1(1.0) - 0.21938393439552027368) < 1e-15  # Abramowitz & Stegun table 5.1


def test_oracle_point_scaling_closed_forms(oracle):
    F = 8
    nb = -np.ones((F, 3), dtype=np.int32)  # isolated faces: the smoothing keeps each own value
    cx, cy = np.arange(F) * 100.0, np.zeros(F)
    U_R = np.array([16.0, 16.0, 0.05, 16.0, 12.0, 12.0, 12.0, 9.0])
    sd = np.array([0.5, -9999.0, 0.5, 49.0, 0.3, 0.3, 5.0, np.nan])
    can = np.array([0.0, 0.0, 0.0, 0.0, 10.0, 3.0, 10.0, 0.0])
    lai = np.array([0.0, 0.0, 0.0, 0.0, 2.0, 1.5, 2.0, 0.0])
    got = oracle.scale_wind(nb, cx, cy, can, lai, _cfg(), U_R, sd)
    want = [
        _log_scale(16.0, 50.0, 2.5, 0.5),  # no canopy
        _log_scale(16.0, 50.0, 2.0, 0.0),  # -9999 snow depth counts as 0
        0.1,  # floor
        16.0,  # 2 m above the snow reaches the 50 m reference height
    ]
    # canopy 10 m, LAI 2, snow 0.3: log to the canopy top, exponential to mid-canopy (5 m),
    # log from there down to 2.3 m
    ut = _log_scale(12.0, 50.0, 10.0, 0.3)
    ub = ut * math.exp(2.0 * (5.0 / 10.0 - 1))
    want.append(_log_scale(ub, 5.0, 2.3, 0.3))
    # canopy 3 m: 2.3 m lies above the canopy bottom (1.5 m): exponential from the top
    ut3 = _log_scale(12.0, 50.0, 3.0, 0.3)
    want.append(ut3 * math.exp(1.5 * (2.3 / 3.0 - 1)))
    # 5 m of snow under a 10 m canopy: 2 m above the snow (7 m) lies above the canopy bottom
    ut7 = _log_scale(12.0, 50.0, 10.0, 5.0)
    want.append(ut7 * math.exp(2.0 * (7.0 / 10.0 - 1)))
    want.append(_log_scale(9.0, 50.0, 2.0, 0.0))  # NaN snow depth counts as 0
    np.testing.assert_allclose(got, want, rtol=1e-14)
    # ignore_canopy: every face takes the open-ground profile
    got2 = oracle.scale_wind(nb, cx, cy, can, lai, _cfg(ignore_canopy=1), U_R, sd)
    np.testing.assert_allclose(got2[4], _log_scale(12.0, 50.0, 2.3, 0.3), rtol=1e-14)


def _numpy_tps(samples, q):
    """thin_plate_spline::operator() restated with numpy.linalg.solve and scipy's E1."""
    n = len(samples)
    c, w = 0.577215, 0.01

    def R(p, r):
        d = math.hypot(p[0] - r[0], p[1] - r[1])
        d = (d * w / 2.0) ** 2
        return -(math.log(d) + c + exp1(d))

    A = np.zeros((n + 1, n + 1))
    for i in range(n):
        for j in range(i + 1, n):
            A[i, j + 1] = A[j, i + 1] = R(samples[i], samples[j])
    A[:, 0] = 1.0
    A[n, :] = 1.0
    A[n, 0] = 0.0
    b = np.array([s[2] for s in samples] + [0.0])
    x = np.linalg.solve(A, b)
    return x[0] + sum(x[i + 1] * R(samples[i], q) for i in range(n))


def test_oracle_smoothing_matches_numpy_restatement(oracle, mesh_small):
    F = mesh_small.n_faces
    cx, cy = _centroids(mesh_small)
    fo = meshgen.forcing_c2(F, seed=3)
    U_R, sd = fo["U_R"], fo["snowdepthavg"] + np.random.default_rng(0).uniform(0, 0.5, F)
    pt = oracle.scale_wind(mesh_small.neigh, cx, cy, None, None, _cfg(smooth=0), U_R, sd)
    got = oracle.scale_wind(mesh_small.neigh, cx, cy, None, None, _cfg(), U_R, sd)
    want = np.empty(F)
    for i in range(F):
        s = [(cx[n], cy[n], pt[n]) for n in mesh_small.neigh[i] if n >= 0]
        want[i] = max(0.1, _numpy_tps(s, (cx[i], cy[i]))) if s else max(0.1, pt[i])
    np.testing.assert_allclose(got, want, rtol=1e-11)
    # a constant field is reproduced exactly by the spline (constant term, weights summing to 0)
    flat = oracle.scale_wind(mesh_small.neigh, cx, cy, None, None, _cfg(), np.full(F, 16.0), np.full(F, 0.5))
    np.testing.assert_allclose(flat, _log_scale(16.0, 50.0, 2.5, 0.5), rtol=1e-12)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def pb():
    from chm_amd import pbsm3d

    pbsm3d.lib()
    return pbsm3d


@pytest.mark.gpu
@pytest.mark.parametrize("veg,smooth", [(False, 1), (True, 1), (True, 0)])
def test_gpu_wind_matches_oracle(pb, oracle, mesh_c1, veg, smooth):
    print(f"[case] wind veg={veg} smooth={smooth}", flush=True)
    F = mesh_c1.n_faces
    cx, cy = _centroids(mesh_c1)
    rng = np.random.default_rng(7)
    fo = meshgen.forcing_c2(F, seed=5)
    sd = rng.uniform(0.0, 3.0, F)
    sd[::97] = -9999.0
    can = rng.uniform(0.0, 12.0, F) if veg else None
    lai = rng.uniform(0.5, 3.0, F) if veg else None
    cfg = _cfg(smooth=smooth)
    w = pb.Wind(mesh_c1.neigh, cx, cy, can, lai, cfg)
    got = w.run(fo["U_R"], sd)
    want = oracle.scale_wind(mesh_c1.neigh, cx, cy, can, lai, cfg, fo["U_R"], sd)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    w.close()


@pytest.mark.gpu
def test_gpu_chain_wind_fetch_step(pb, oracle, mesh_c1):
    """U_R -> scale_wind_vert -> U_2m_above_srf and vw_dir -> fetchr -> fetch, both on
    device arrays, feeding the PBSM3D step: the same fields as the oracle chain."""
    import torch

    print("[case] wind -> fetch -> step", flush=True)
    F = mesh_c1.n_faces
    cx, cy = _centroids(mesh_c1)
    cz = np.ascontiguousarray(mesh_c1.vertex[mesh_c1.elem].mean(axis=1)[:, 2])
    fo = meshgen.forcing_c2(F, seed=11)
    d = {k: torch.tensor(v, device="cuda:0") for k, v in fo.items()}
    w = pb.Wind(mesh_c1.neigh, cx, cy)
    f = pb.Fetch(cx, cy, cz)
    d["U_2m_above_srf"] = w.run(d["U_R"], d["snowdepthavg"])
    d["fetch"] = f.run(d["vw_dir"])
    g = pb.face_geometry(mesh_c1)
    ctx = pb.Context(g, pb.default_config(nLayer=5, tol=1e-12))
    outs = {k: torch.zeros(F, dtype=torch.float64, device="cuda:0") for k in abi.OUT_FIELDS}
    ctx.step(d, outs=outs, on_device=True)
    ofo = dict(fo)
    ofo["U_2m_above_srf"] = oracle.scale_wind(mesh_c1.neigh, cx, cy, None, None, _cfg(), fo["U_R"],
                                              fo["snowdepthavg"])
    fcfg = abi.FetchConfig()
    fcfg.steps, fcfg.incl_veg, fcfg.max_distance, fcfg.I = 10, 1, 1000.0, 0.06
    ofo["fetch"] = oracle.fetch(cx, cy, cz, None, fcfg, fo["vw_dir"])
    np.testing.assert_allclose(d["U_2m_above_srf"].cpu().numpy(), ofo["U_2m_above_srf"], rtol=1e-12)
    np.testing.assert_array_equal(d["fetch"].cpu().numpy(), ofo["fetch"])
    orc = oracle.Oracle(g, oracle.oracle_config(nLayer=5, tol=1e-13))
    og, _ = orc.step(ofo)
    for k in ("Qsalt", "Qsusp", "drift_mass"):
        got = outs[k].cpu().numpy()
        assert np.abs(got - og[k]).max() <= 1e-8 * np.abs(og[k]).max(), k
    w.close()
    f.close()
