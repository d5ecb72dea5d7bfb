"""Pins the CPU oracle (oracle/pbsm3d_oracle.cpp) before it is used as the checker.

The reference has no PBSM3D tests or golden vectors (its googletest suite
covers other subsystems), so each oracle function is pinned by known-answer
values, invariants of the discretisation, an independent brute-force builder,
or an exact dense solve, chosen so that a dropped term, a flipped sign or a
wrong index fails at least one of them.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from chm_amd import abi, meshgen
from tests.conftest import rel_err

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _pattern_sha(rp, col):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rp, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(col, dtype="<i4").tobytes())
    return h.hexdigest()


def _one_face_mesh():
    """Right triangle (0,0), (1,0), (0,1) -- geometry known in closed form."""
    v = np.array([[0.0, 0.0, 10.0], [1.0, 0.0, 20.0], [0.0, 1.0, 30.0]])
    return meshgen.Mesh(v, np.array([[0, 1, 2]], dtype=np.int32), np.array([[-1, -1, -1]], dtype=np.int32))


# ---------------------------------------------------------------------------
# geometry (triangulation.hpp:1443-1498, 1577-1589, 1830-1856)
# ---------------------------------------------------------------------------
def test_geometry_right_triangle(oracle):
    g = oracle.face_geometry(_one_face_mesh())
    s = 1 / math.sqrt(2)
    # edge j is opposite vertex j; unit normals point out of the triangle
    np.testing.assert_allclose(g.normal_x, [s, -1.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(g.normal_y, [s, 0.0, -1.0], atol=1e-15)
    np.testing.assert_allclose(g.edge_length, [math.sqrt(2), 1.0, 1.0], rtol=1e-15)
    assert g.area[0] == 0.5
    assert g.center_z[0] == 20.0
    np.testing.assert_allclose([g.center_x[0], g.center_y[0]], [1 / 3, 1 / 3], rtol=1e-15)


def test_geometry_closure_and_outward(oracle, mesh_c1):
    g = oracle.face_geometry(mesh_c1)
    F = mesh_c1.n_faces
    nx, ny, E = (a.reshape(F, 3) for a in (g.normal_x, g.normal_y, g.edge_length))
    # sum_j E_j m_j = 0 (closed polygon)
    sx, sy = (E * nx).sum(1), (E * ny).sum(1)
    assert np.abs(sx).max() < 1e-9 * E.max() and np.abs(sy).max() < 1e-9 * E.max()
    # outward: normal . (edge midpoint - centroid) > 0
    v = mesh_c1.vertex[mesh_c1.elem]
    for j in range(3):
        mid = 0.5 * (v[:, (j + 1) % 3, :2] + v[:, (j + 2) % 3, :2])
        d = (mid[:, 0] - g.center_x) * nx[:, j] + (mid[:, 1] - g.center_y) * ny[:, j]
        assert (d > 0).all()
    assert (g.area > 0).all()  # CCW elements
    np.testing.assert_allclose(np.hypot(nx, ny), 1.0, rtol=1e-14)


def test_area_parameter_overrides(oracle):
    m = _one_face_mesh()
    m.area = np.array([123.5])
    assert oracle.face_geometry(m).area[0] == 123.5


# ---------------------------------------------------------------------------
# CSR pattern (LinearAlgebra.cpp:43-149)
# ---------------------------------------------------------------------------
def test_pattern_matches_brute_force_and_golden(oracle, mesh_c1):
    from tests.golden.make_fixtures import brute_force_pattern

    gold = json.load(open(os.path.join(GOLD, "patterns.json")))
    assert gold["c1_faces"] == mesh_c1.n_faces
    for L in (1, 5, 10):
        rp, col = oracle.csr_pattern(mesh_c1.n_faces, L, mesh_c1.neigh)
        assert _pattern_sha(rp, col) == gold[f"c1_L{L}"]["sha256"]
        deg = int((mesh_c1.neigh >= 0).sum())
        assert rp[-1] == L * (mesh_c1.n_faces + deg) + 2 * mesh_c1.n_faces * (L - 1)
    rp, col = oracle.csr_pattern(mesh_c1.n_faces, 3, mesh_c1.neigh)
    rb, cb = brute_force_pattern(mesh_c1.n_faces, 3, mesh_c1.neigh)
    assert np.array_equal(rp, rb) and np.array_equal(col, cb)


def test_pattern_reference_mesh(oracle):
    gold = json.load(open(os.path.join(GOLD, "patterns.json")))
    g = np.load(os.path.join(GOLD, "granger1m.npz"))
    F = g["elem"].shape[0]
    assert F == gold["granger_faces"] == 985
    for L in (1, 10):
        rp, col = oracle.csr_pattern(F, L, g["neigh"])
        assert _pattern_sha(rp, col) == gold[f"granger1m_L{L}"]["sha256"]


# ---------------------------------------------------------------------------
# per-face coefficients (PBSM3D.cpp:647-925) -- known answers
# ---------------------------------------------------------------------------
def _run(oracle, mesh, forcing, **cfg):
    g = oracle.face_geometry(mesh)
    o = oracle.Oracle(g, oracle.oracle_config(**cfg))
    outs, st = o.step(forcing)
    return o, outs, st, g


def test_saltation_known_answers(oracle, mesh_small):
    F = mesh_small.n_faces
    fo = meshgen.forcing_c1(F, seed=1)
    o, outs, st, g = _run(oracle, mesh_small, fo, nLayer=5)
    tr = o.face_trace()
    # u*_th(-10 C) = 0.35 - 10/150 + 100/8200 (Pomeroy & Li 2000 eqn 7)
    np.testing.assert_allclose(tr["uth"], 0.29552845528455285, rtol=1e-15)
    # u* = kappa u2 / ln(2 / 2e-4) at u2 = 10 m/s
    np.testing.assert_allclose(tr["ustar"], 0.43429448190325176, rtol=1e-15)
    assert (tr["saltation"] == 1).all()
    np.testing.assert_allclose(tr["hs"], 0.08436 * 0.43429448190325176 ** 1.27, rtol=1e-14)
    assert (tr["z0"] == 0.01).all()
    # c_salt = rho_f/(3.29 u*) (1 - u*_th^2/u*^2), rho_f the dry-air density of the standard
    # atmosphere at the centroid (pinned to the 1976 table in test_std_dry_air_density_us1976)
    rho = np.array([oracle.std_dry_air_density(z, 263.15) for z in g.center_z])
    cs = rho / (3.29 * 0.43429448190325176) * (1 - (0.29552845528455285 / 0.43429448190325176) ** 2)
    np.testing.assert_allclose(tr["c_salt"], cs, rtol=1e-13)
    # Qsalt = c_salt * 2.8 u*_th * hs
    np.testing.assert_allclose(outs["Qsalt"], cs * 2.8 * 0.29552845528455285 * tr["hs"], rtol=1e-13)


# U.S. Standard Atmosphere 1976, geometric altitude Z (m): pressure (Pa), temperature (K),
# density (kg m-3) as tabulated (five significant digits), and the 11 km geopotential
# tropopause (Z = 11,019.07 m): 22,632.06 Pa, 216.65 K.
_US1976 = [(0.0, 1.01325e5, 288.150, 1.2250), (1000.0, 8.9876e4, 281.651, 1.1117),
           (2000.0, 7.9501e4, 275.154, 1.0066), (3000.0, 7.0121e4, 268.659, 0.90925),
           (5000.0, 5.4048e4, 255.676, 0.73643)]
_TROPOPAUSE = (6356766.0 * 11000.0 / (6356766.0 - 11000.0), 22632.06, 216.65)


def test_std_dry_air_density_us1976(oracle):
    """MeteoIO's stdDryAirDensity (PBSM3D.cpp:784-786; MeteoIO is not in the reference
    tree) restated as p_std(z) / (R_dry T), checked against the published 1976 standard
    atmosphere: p = rho(z, T) * R_dry * T must reproduce the tabulated pressure (3e-5:
    the table's five digits plus MeteoIO's R_dry = 287.058 against the table's 287.0531)
    and rho(z, T_table) the tabulated density (7e-5).  A wrong lapse rate, exponent, sea
    level temperature, Earth radius (geometric vs geopotential height) or gas constant
    moves one of these by more (shown below on perturbed restatements)."""
    R = 287.058
    for z, p_tab, t_tab, rho_tab in _US1976:
        rho = oracle.std_dry_air_density(z, t_tab)
        assert abs(rho * R * t_tab - p_tab) / p_tab < 3e-5, (z, rho * R * t_tab, p_tab)
        assert abs(rho - rho_tab) / rho_tab < 7e-5, (z, rho, rho_tab)
        # the density is linear in 1/T at fixed altitude
        np.testing.assert_allclose(oracle.std_dry_air_density(z, 2 * t_tab), rho / 2, rtol=1e-15)
    zt, pt, tt = _TROPOPAUSE
    assert abs(oracle.std_dry_air_density(zt, tt) * R * tt - pt) / pt < 3e-5
    # the checks discriminate: plausible mistakes in the formula miss the table by more
    def p_variant(z, lapse=0.0065, t0=288.15, r0=6356766.0, rgas=287.058, g0=9.80665):
        h = z if r0 is None else r0 * z / (r0 + z)
        return 101325.0 * (1 - lapse * h / t0) ** (g0 / (lapse * rgas))
    for kw in (dict(r0=None), dict(lapse=0.0060), dict(t0=273.15), dict(rgas=287.0), dict(g0=9.81)):
        worst = max(abs(p_variant(z, **kw) - p_tab) / p_tab for z, p_tab, _, _ in _US1976 + [(zt, pt, 0, 0)])
        assert worst > 6e-5, (kw, worst)


def test_is_nan_band(oracle):
    """module_base::is_nan (module_base.hpp:471-479): a value counts as missing within
    1e-5 of -9999 (either side) or when NaN; just outside the band it is a value."""
    for v in (-9999.0, -9999.0 + 5e-6, -9999.0 - 5e-6, -9999.0 + 9.9e-6, float("nan")):
        assert oracle.is_nan(v), v
    for v in (-9999.0 + 2e-5, -9999.0 - 2e-5, -9998.0, 0.0, 9999.0, float("inf"), -float("inf")):
        assert not oracle.is_nan(v), v


def test_missing_swe_and_depth_act_as_zero(oracle, mesh_small):
    """swe and snowdepthavg within the -9999 band enter the step as 0 (PBSM3D.cpp:447, 469,
    1723): the fields equal those of a forcing with zeros in those faces."""
    F = mesh_small.n_faces
    fo = meshgen.forcing_c1(F, seed=4)
    rng = np.random.default_rng(4)
    pick = rng.choice(F, F // 5, replace=False)
    band = rng.choice([-9999.0, -9999.0 + 5e-6, -9999.0 - 5e-6, -9999.0 + 9e-6, np.nan], len(pick))
    zero = dict(fo, swe=fo["swe"].copy(), snowdepthavg=fo["snowdepthavg"].copy())
    zero["swe"][pick] = 0.0
    zero["snowdepthavg"][pick[: len(pick) // 2]] = 0.0
    miss = dict(fo, swe=fo["swe"].copy(), snowdepthavg=fo["snowdepthavg"].copy())
    miss["swe"][pick] = band
    miss["snowdepthavg"][pick[: len(pick) // 2]] = band[: len(pick) // 2]
    _, oz, sz, _ = _run(oracle, mesh_small, zero, nLayer=4, tol=1e-13)
    _, om, sm, _ = _run(oracle, mesh_small, miss, nLayer=4, tol=1e-13)
    assert sz["suspension_present"] == sm["suspension_present"] == 1
    for k in abi.OUT_FIELDS:
        np.testing.assert_array_equal(om[k], oz[k], err_msg=k)
    # eroding faces whose swe counts as 0 hit the more-than-available cap (PBSM3D.cpp:1724-1728)
    assert (om["pbsm_more_than_avail"][pick] == 1).any()


def test_tanh_fetch_factor_is_constant(oracle, mesh_small):
    F = mesh_small.n_faces
    fo = meshgen.forcing_c1(F, seed=1)
    _, o1, _, _ = _run(oracle, mesh_small, fo, nLayer=3)
    fo2 = dict(fo, fetch=np.full(F, 25.0))
    _, o2, _, _ = _run(oracle, mesh_small, fo2, nLayer=3)
    # PBSM3D.cpp:839-845 evaluates the tanh at fetch_ref=300: factor 0.98201379...
    np.testing.assert_allclose(o2["Qsalt"] / o1["Qsalt"], 0.5 * math.tanh(2.0 - 1e-10) + 0.5, rtol=1e-9)
    np.testing.assert_allclose(o2["Qsalt"] / o1["Qsalt"], 0.98201379, rtol=1e-8)


def test_no_saltation_branch(oracle, mesh_small):
    F = mesh_small.n_faces
    fo = meshgen.forcing_c1(F, seed=1)
    fo["U_2m_above_srf"] = np.full(F, 3.0)  # u* = 0.13 < u*_th
    o, outs, st, _ = _run(oracle, mesh_small, fo, nLayer=3)
    tr = o.face_trace()
    assert (tr["saltation"] == 0).all() and (tr["hs"] == 0).all()
    # u* from the reference-height wind with Z0_SNOW: kappa U_R / ln(50/0.01)
    np.testing.assert_allclose(tr["ustar"], np.maximum(0.01, 0.4 * fo["U_R"] / math.log(5000.0)), rtol=1e-15)
    assert (outs["Qsalt"] == 0).all()
    assert st["suspension_present"] == 0 and st["deposition_present"] == 0


def test_layer_wind_log_profile(oracle, mesh_small):
    F = mesh_small.n_faces
    L = 5
    fo = meshgen.forcing_c2(F, seed=2)
    o, _, _, _ = _run(oracle, mesh_small, fo, nLayer=L)
    tr, ct = o.face_trace(), o.cell_trace()
    dz = 5.0 / L
    sd = fo["snowdepthavg"]
    for z in range(L):
        cz = z * dz + tr["hs"] + dz / 2
        u = fo["U_R"] * np.log(cz / 0.01 - 1.0) / np.log((50.0 - sd - 0.01) / 0.01)
        np.testing.assert_allclose(ct["u_z"][z * F:(z + 1) * F], np.maximum(u, 0.01), rtol=1e-12)
    # wind increases with height in every column
    u = ct["u_z"].reshape(L, F)
    assert (np.diff(u, axis=0) > 0).all()


def test_ustar_coupling_root(oracle):
    """z0_ustar_coupling (PBSM3D.cpp:691-716): u* is the root of Li & Pomeroy 2000 eq. 5,
    u* = u2 k / ln(2 / (0.6131702345e-2 u*^2 + lambda / 2)), bracketed from the guess 1
    downward when g(1) < 0 and upward otherwise.  The equation has a second root near
    u* ~ 18 m/s (where the log's argument approaches 1); the search must return the one
    on the guess's side, as scipy's brentq on that side finds it.  With lambda / 2 > 2
    the log is negative for every u*, no bracket exists and the downward walk ends in
    Boost's escape bracket (0, a < DBL_MIN): u* ~ 0."""
    from scipy.optimize import brentq

    def g(u, u2, lam):
        return u2 * 0.4 / np.log(2.0 / (0.6131702345e-2 * u * u + 0.5 * lam)) - u

    for u2, lam, side in ((10.0, 0.0, (1e-6, 1.0)), (6.0, 0.1, (1e-6, 1.0)), (4.0, 0.24, (1e-6, 1.0)),
                          (20.0, 0.0, (1.0, 8.0)), (25.0, 0.3, (1.0, 8.0))):
        want = brentq(g, *side, args=(u2, lam), xtol=1e-15, rtol=1e-15)
        got, ok = oracle.ustar_coupled(u2, lam)
        assert ok
        assert abs(got - want) <= 4e-15 * want, (u2, lam, got, want)
        assert abs(g(got, u2, lam)) < 1e-13
        assert (g(1.0, u2, lam) < 0) == (got < 1.0)  # the root on the guess's side
    tiny, ok = oracle.ustar_coupled(10.0, 5.0)
    assert ok and 0 <= tiny < 1e-300


def test_z0_ustar_coupling_step(oracle, mesh_small):
    """Through the step: with vegetation (lambda = LAI hd / 2) saltating faces take the
    coupled u* and z0 = c u*^2 + lambda / 2 (PBSM3D.cpp:727-737, floored at Z0_SNOW);
    u* >= u*_th on exactly those faces, and the faces that do not saltate keep the
    reference-height u* with Z0_SNOW."""
    F = mesh_small.n_faces
    rng = np.random.default_rng(12)
    g = oracle.face_geometry(mesh_small)
    g.canopy_height = rng.uniform(0.4, 0.75, F)  # hd = canopy - 0.5 m snow in [0, 0.25] <= cutoff
    g.lai = rng.uniform(0.0, 3.0, F)
    fo = meshgen.forcing_c2(F, seed=12)
    o = oracle.Oracle(g, oracle.oracle_config(nLayer=4, z0_ustar_coupling=1))
    o.step(fo)
    tr = o.face_trace()
    hd = np.maximum(0.0, g.canopy_height - fo["snowdepthavg"])
    lam = 0.5 * g.lai * hd
    root = np.array([oracle.ustar_coupled(u, lm)[0] for u, lm in zip(fo["U_2m_above_srf"], lam)])
    # every face is a candidate (hd <= cutoff, depth >= min_sd_trans); those whose root
    # passes the threshold saltate with it (a face whose c_salt then comes out negative
    # drops its saltation flag but keeps u* and z0, PBSM3D.cpp:814-824)
    cand = root >= tr["uth"]
    assert cand.any() and (~cand).any()
    assert not (tr["saltation"][~cand] == 1).any()
    np.testing.assert_array_equal(tr["ustar"][cand], root[cand])
    np.testing.assert_allclose(tr["z0"][cand], np.maximum(0.01, 0.6131702345e-2 * root[cand] ** 2 + 0.5 * lam[cand]),
                               rtol=1e-15)
    assert (tr["z0"][cand] > 0.01).any()
    np.testing.assert_allclose(tr["ustar"][~cand], np.maximum(0.01, 0.4 * fo["U_R"][~cand] / math.log(5000.0)),
                               rtol=1e-15)
    assert (tr["z0"][~cand] == 0.01).all()


def test_pomli_probability(oracle, mesh_small):
    """use_PomLi_probability (PBSM3D.cpp:848-865): c_salt, hence Qsalt, is multiplied by
    the probability of blowing snow P = 1 / (1 + exp(sqrt(pi) (u_mean - u*) / delta)) of
    Essery, Li & Pomeroy 1999 (eqs 10-13: u_mean = 11.2 + 0.365 T + 0.00706 T^2 +
    0.9 ln A, delta = 0.145 T + 0.00196 T^2 + 4.3, A the hours since snowfall, u* the 10 m
    wind, log-scaled from U_R over snow, over sqrt(1 + 340 z0v), z0v = 0 without stalks).
    Pinned by the logistic's fixed points: A chosen so that u_mean equals the 10 m wind
    gives P = 1/2 exactly; u_mean = u10 -/+ delta / sqrt(pi) gives 1 / (1 + e^(-/+1)); and
    the two sides sum to one.  Faces that do not saltate keep the -9999 default."""
    F = mesh_small.n_faces
    fo = meshgen.forcing_c2(F, seed=13)
    T, sd = fo["t"], fo["snowdepthavg"]
    z10 = 10.0 + sd
    u10 = fo["U_R"] * np.log((z10 - (sd + 0.01)) / 0.01) / np.log((50.0 - (sd + 0.01)) / 0.01)
    base = 11.2 + 0.365 * T + 0.00706 * T * T
    delta = 0.145 * T + 0.00196 * T * T + 4.3
    _, off, _, _ = _run(oracle, mesh_small, fo, nLayer=3)
    salt = off["Qsalt"] > 0
    assert salt.sum() > F // 3
    probs = {}
    for name, shift, want in (("mid", 0.0, 0.5), ("lo", -1.0, 1 / (1 + 1 / math.e)), ("hi", 1.0, 1 / (1 + math.e))):
        u_mean = u10 + shift * delta / math.sqrt(math.pi)
        A = np.exp((u_mean - base) / 0.9)
        _, on, _, _ = _run(oracle, mesh_small, dict(fo, p_snow_hours=A), nLayer=3, use_PomLi_probability=1)
        p = on["blowingsnow_probability"]
        np.testing.assert_allclose(p[salt], want, rtol=1e-12, err_msg=name)
        assert (p[~salt] == -9999.0).all()
        np.testing.assert_allclose(on["Qsalt"][salt], want * off["Qsalt"][salt], rtol=1e-12, err_msg=name)
        probs[name] = p[salt]
    np.testing.assert_allclose(probs["lo"] + probs["hi"], 1.0, rtol=1e-12)
    # the key needs the p_snow_hours forcing
    g = oracle.face_geometry(mesh_small)
    with pytest.raises(RuntimeError):
        oracle.Oracle(g, oracle.oracle_config(nLayer=3, use_PomLi_probability=1)).step(
            {k: v for k, v in fo.items() if k != "p_snow_hours"})


def test_rouault_diffusion_coefficient(oracle, mesh_small):
    """rouault_diffusion_coef (PBSM3D.cpp:1164-1172): the vertical diffusivity uses
    1 / (1 + w^2 / (1.56 u*^2)) in place of snow_diffusion_const.  With a fixed settling
    velocity w the interior coupling to the cell above is area w + alpha3 with
    alpha3 = area K / dz (K = coef u* l), so alpha3 from the exported system, divided by
    alpha3 of a run with snow_diffusion_const = 1, is the coefficient itself; w = 0 makes
    it exactly 1."""
    F = mesh_small.n_faces
    L = 4
    fo = meshgen.forcing_c2(F, seed=14)
    g = oracle.face_geometry(mesh_small)
    a3 = {}
    for name, kw in (("rouault", dict(rouault_diffusion_coef=1)), ("one", dict(snow_diffusion_const=1.0))):
        o = oracle.Oracle(g, oracle.oracle_config(nLayer=L, do_fixed_settling=1, settling_velocity=0.4, **kw))
        o.step(fo)
        rp, col, val, _ = o.system(0)
        r = F + np.arange(F)  # layer 1: an interior layer for L = 4
        above = np.array([val[rp[k]:rp[k + 1]][col[rp[k]:rp[k + 1]] == k + F][0] for k in r])
        a3[name] = above - g.area * 0.4
        ustar = o.face_trace()["ustar"]
    np.testing.assert_allclose(a3["rouault"] / a3["one"], 1.0 / (1.0 + 0.4 ** 2 / (1.56 * ustar ** 2)), rtol=1e-9)
    for name, kw in (("rouault", dict(rouault_diffusion_coef=1)), ("one", dict(snow_diffusion_const=1.0))):
        o = oracle.Oracle(g, oracle.oracle_config(nLayer=L, do_fixed_settling=1, settling_velocity=0.0, **kw))
        o.step(fo)
        a3[name] = o.system(0)[2]
    np.testing.assert_array_equal(a3["rouault"], a3["one"])


# ---------------------------------------------------------------------------
# suspension / deposition assembly -- conservation invariants
# ---------------------------------------------------------------------------
def test_suspension_row_sums(oracle, mesh_c1):
    """With the sink off, interior middle-layer rows conserve mass (row sum 0);
    layer-0 rows sum to -alpha4' = rhs / c_salt; RHS lives in layer 0 only."""
    F, L = mesh_c1.n_faces, 5
    fo = meshgen.forcing_c1(F, seed=0)
    o, _, st, _ = _run(oracle, mesh_c1, fo, nLayer=L, do_sublimation=0)
    rp, col, val, rhs = o.system(0)
    rows = np.repeat(np.arange(F * L), np.diff(rp))
    rsum = np.bincount(rows, weights=val, minlength=F * L)
    scale = np.bincount(rows, weights=np.abs(val), minlength=F * L)
    interior = (mesh_c1.neigh >= 0).all(1)
    for z in range(1, L - 1):
        r = z * F + np.nonzero(interior)[0]
        assert np.abs(rsum[r]).max() < 1e-12 * scale[r].max()
    assert np.abs(rhs[F:]).max() == 0.0
    cs = o.face_trace()["c_salt"]
    sel = interior & (cs > 0)
    r0 = np.nonzero(sel)[0]
    np.testing.assert_allclose(rsum[r0], rhs[r0] / cs[sel], rtol=1e-9)
    assert (rhs[:F][cs > 0] < 0).all()
    # the -M structure of upwind FV: negative diagonal, non-negative couplings
    diag = val[col == rows]
    assert (diag < 0).all() and (val[col != rows] >= 0).all()


def test_sink_adds_area_dz_csubl(oracle, mesh_small):
    """V = area*dz/5 counted five times: switching sublimation on adds exactly
    area*dz*csubl to each interior row sum (PBSM3D.cpp:1222-1226)."""
    F, L = mesh_small.n_faces, 4
    fo = meshgen.forcing_c1(F, seed=4)
    o1, _, _, g = _run(oracle, mesh_small, fo, nLayer=L, do_sublimation=1)
    o0, _, _, _ = _run(oracle, mesh_small, fo, nLayer=L, do_sublimation=0)
    rp, col, v1, _ = o1.system(0)
    _, _, v0, _ = o0.system(0)
    rows = np.repeat(np.arange(F * L), np.diff(rp))
    d = np.bincount(rows, weights=v1 - v0, minlength=F * L)
    csubl = o1.cell_trace()["csubl"]
    expect = np.tile(g.area, L) * (5.0 / L) * csubl
    np.testing.assert_allclose(d, expect, rtol=1e-9, atol=1e-12 * np.abs(expect).max())
    assert (csubl < 0).all()  # undersaturated air (rh 70 %) sublimates


def test_deposition_matrix_structure_and_mass_balance(oracle, mesh_c1):
    F = mesh_c1.n_faces
    fo = meshgen.forcing_c1(F, seed=0)
    o, outs, st, g = _run(oracle, mesh_c1, fo, nLayer=5)
    assert st["deposition_present"] == 1
    rp, col, val, rhs = o.system(1)
    import scipy.sparse as sp

    A = sp.csr_matrix((val, col, rp), shape=(F, F))
    assert abs(A - A.T).max() == 0.0  # symmetric Laplacian + mass
    np.testing.assert_allclose(np.asarray(A.sum(1)).ravel(), g.area, rtol=1e-12)
    # 1^T A = area^T  =>  sum_i area_i qdep_i = sum_i rhs_i
    q = outs["qdep"]
    assert abs((g.area * q).sum() - rhs.sum()) < 1e-6 * np.abs(rhs).sum()


def test_smooth_coeff_zero_gives_diagonal_deposition(oracle, mesh_small):
    F = mesh_small.n_faces
    fo = meshgen.forcing_c1(F, seed=0)
    o, outs, st, g = _run(oracle, mesh_small, fo, nLayer=3, smooth_coeff=0.0)
    _, _, _, rhs = o.system(1)
    np.testing.assert_allclose(outs["qdep"], rhs / g.area, rtol=1e-9)


# ---------------------------------------------------------------------------
# solvers (LinearAlgebra.cpp:164-252)
# ---------------------------------------------------------------------------
def test_dense_and_krylov_agree_with_numpy(oracle, mesh_small):
    F, L = mesh_small.n_faces, 3
    fo = meshgen.forcing_c1(F, seed=0)
    o, _, _, _ = _run(oracle, mesh_small, fo, nLayer=L)
    rp, col, val, rhs = o.system(0)
    import scipy.sparse as sp

    A = sp.csr_matrix((val, col, rp), shape=(F * L, F * L)).toarray()
    x_np = np.linalg.solve(A, rhs)
    x_ge = oracle.dense_solve(rp, col, val, rhs)
    assert rel_err(x_ge, x_np) < 1e-11
    kappa = np.linalg.cond(A)
    assert 10 < kappa < 1e5
    for pc in (abi.PC_ILUT, abi.PC_ILU0, abi.PC_COLUMN, abi.PC_JACOBI, abi.PC_NONE):
        x, it, res, rc = oracle.solve_csr(rp, col, val, rhs, precond=pc, F=F, L=L, tol=1e-10)
        assert rc == 0 and res <= 1e-10
        assert np.linalg.norm(A @ x - rhs) / np.linalg.norm(rhs) <= 1.1e-10
        # forward error bounded by kappa * relative residual
        assert np.linalg.norm(x - x_np) / np.linalg.norm(x_np) <= kappa * 1.1e-10


def test_ilut_limits(oracle, mesh_small):
    """ILUT with tau=0 and unbounded fill is the exact LU; ILU(0) is ILUT
    restricted to the pattern with no dropping."""
    F, L = mesh_small.n_faces, 3
    fo = meshgen.forcing_c1(F, seed=0)
    o, _, _, _ = _run(oracle, mesh_small, fo, nLayer=L)
    rp, col, val, rhs = o.system(0)
    r = np.random.default_rng(0).standard_normal(F * L)
    import scipy.sparse as sp

    A = sp.csr_matrix((val, col, rp), shape=(F * L, F * L))
    lib = oracle.lib()
    try:
        lib.oracle_set_ilut(0.0, 1e6)
        z = oracle.precond_apply(rp, col, val, r, abi.PC_ILUT)
        assert np.linalg.norm(A @ z - r) / np.linalg.norm(r) < 1e-12
        lib.oracle_set_ilut(0.0, 1.0)  # fill 0 beyond the row's own counts
        z1 = oracle.precond_apply(rp, col, val, r, abi.PC_ILUT)
        z0 = oracle.precond_apply(rp, col, val, r, abi.PC_ILU0)
        # with no fill and no dropping, ILUT keeps at most the original counts: same quality class as ILU0
        e1 = np.linalg.norm(A @ z1 - r) / np.linalg.norm(r)
        e0 = np.linalg.norm(A @ z0 - r) / np.linalg.norm(r)
        assert e1 < 0.5 and e0 < 0.5
    finally:
        lib.oracle_set_ilut(1e-4, 3.0)
    z = oracle.precond_apply(rp, col, val, r, abi.PC_ILUT)
    assert np.linalg.norm(A @ z - r) / np.linalg.norm(r) < 1e-2


def test_column_preconditioner_is_exact_on_columns(oracle, mesh_small):
    """With lateral coupling removed the column (Thomas) preconditioner is A^-1."""
    F, L = mesh_small.n_faces, 6
    fo = meshgen.forcing_c1(F, seed=0)
    o, _, _, _ = _run(oracle, mesh_small, fo, nLayer=L)
    rp, col, val, rhs = o.system(0)
    rows = np.repeat(np.arange(F * L), np.diff(rp))
    keep = (col == rows) | (np.abs(col - rows) == F)
    import scipy.sparse as sp

    A = sp.csr_matrix((val * keep, col, rp), shape=(F * L, F * L))
    r = np.random.default_rng(1).standard_normal(F * L)
    z = oracle.precond_apply(rp, col, val * keep, r, abi.PC_COLUMN, F, L)
    assert np.linalg.norm(A @ z - r) / np.linalg.norm(r) < 1e-13


def _colour_key(oracle, neigh, F, L):
    """Row keys of the multicolour order (colour, layer, face) from the oracle's face colours."""
    colr = oracle.face_colors(neigh)[0].astype(np.int64)
    r = np.arange(F * L)
    return (colr[r % F] * L + r // F) * F + r % F, colr


def test_dilu_exact_on_triangular_parts(oracle, mesh_small):
    """D-ILU(0) in the multicolour order is the exact inverse when A has no
    strictly-upper (resp. strictly-lower) part in that order, and reduces to
    the column Thomas solve when lateral couplings are absent."""
    F, L = mesh_small.n_faces, 4
    fo = meshgen.forcing_c1(F, seed=0)
    o, _, _, _ = _run(oracle, mesh_small, fo, nLayer=L)
    rp, col, val, _ = o.system(0)
    key, colr = _colour_key(oracle, mesh_small.neigh, F, L)
    assert colr.max() <= 2 and all((colr[mesh_small.neigh[i][mesh_small.neigh[i] >= 0]] != colr[i]).all()
                                   for i in range(F))
    rows = np.repeat(np.arange(F * L), np.diff(rp))
    import scipy.sparse as sp

    r = np.random.default_rng(3).standard_normal(F * L)
    for keep in (key[col] <= key[rows], key[col] >= key[rows]):
        v = val * keep
        A = sp.csr_matrix((v, col, rp), shape=(F * L, F * L))
        z = oracle.precond_apply(rp, col, v, r, abi.PC_MCILU0, F, L)
        assert np.linalg.norm(A @ z - r) / np.linalg.norm(r) < 1e-13
    vert = (col == rows) | (np.abs(col - rows) == F)
    z1 = oracle.precond_apply(rp, col, val * vert, r, abi.PC_MCILU0, F, L)
    z2 = oracle.precond_apply(rp, col, val * vert, r, abi.PC_COLUMN, F, L)
    assert rel_err(z1, z2) < 1e-13


def test_tiny_golden_fields(oracle):
    """Dense-solve outputs committed in tests/golden/tiny_fields.npz."""
    d = np.load(os.path.join(GOLD, "tiny_fields.npz"))
    m = meshgen.Mesh(d["vertex"], d["elem"], d["neigh"])
    g = oracle.face_geometry(m)
    fo = {k[len("forcing_"):]: d[k] for k in d.files if k.startswith("forcing_")}
    o = oracle.Oracle(g, oracle.oracle_config(nLayer=3), direct=True)
    outs, st = o.step(fo)
    for k in ("Qsalt", "Qsusp", "Qsubl", "drift_mass", "conc", "qdep"):
        assert rel_err(outs[k], d[f"out_{k}"]) < 1e-13, k
    _, _, val, rhs = o.system(0)
    assert np.array_equal(val, d["susp_val"]) and np.array_equal(rhs, d["susp_rhs"])
    # Krylov at 1e-12 reproduces the dense-solve fields
    o2 = oracle.Oracle(g, oracle.oracle_config(nLayer=3, tol=1e-12))
    outs2, _ = o2.step(fo)
    for k in ("Qsusp", "conc", "qdep", "drift_mass"):
        assert rel_err(outs2[k], d[f"out_{k}"]) < 1e-9, k


# ---------------------------------------------------------------------------
# state semantics (PBSM3D.cpp:1661-1740, variablestorage.hpp:269-272)
# ---------------------------------------------------------------------------
def test_calm_step_keeps_state(oracle, mesh_small):
    F = mesh_small.n_faces
    fo = meshgen.forcing_c1(F)
    fo["U_2m_above_srf"] = np.full(F, 2.0)
    o, outs, st, _ = _run(oracle, mesh_small, fo, nLayer=3)
    assert st["suspension_present"] == 0
    assert (outs["drift_mass"] == -9999).all() and (outs["pbsm_more_than_avail"] == -9999).all()
    assert (outs["sum_drift"] == 0).all() and (outs["Qsusp"] == 0).all()


def test_multi_step_accumulates_and_caps(oracle, mesh_small):
    F = mesh_small.n_faces
    g = oracle.face_geometry(mesh_small)
    o = oracle.Oracle(g, oracle.oracle_config(nLayer=3))
    fo = meshgen.forcing_c1(F, seed=0)
    fo["swe"] = np.full(F, 0.05)  # little snow: erosion capped at -swe
    tot = np.zeros(F)
    subl = np.zeros(F)
    for _ in range(3):
        outs, st = o.step(fo)
        assert st["deposition_present"] == 1
        tot += outs["drift_mass"]
        subl += outs["Qsubl_mass"]
        assert (outs["drift_mass"] >= -0.05).all()
    np.testing.assert_allclose(outs["sum_drift"], tot, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(outs["sum_subl"], subl, rtol=1e-12, atol=1e-15)
    assert (outs["pbsm_more_than_avail"] == 1).any()
    sd, ss = o.get_state()
    np.testing.assert_array_equal(sd, outs["sum_drift"])


# ---------------------------------------------------------------------------
# multicolour D-ILU colouring
# ---------------------------------------------------------------------------
def _greedy_colors(nb):
    col = np.full(len(nb), -1)
    for i in range(len(nb)):
        used = {col[n] for n in nb[i] if n >= 0 and col[n] >= 0}
        c = 0
        while c in used:
            c += 1
        col[i] = c
    return col


@pytest.mark.parametrize("points", [2500, 20000])
def test_face_colouring_valid_three_colours_and_matches_library(oracle, points):
    """The library's host colouring and the oracle's restatement agree face by face;
    no two neighbours share a colour; the Kempe repair removes the fourth colour
    that plain greedy colouring leaves."""
    from chm_amd import pbsm3d as pb

    mesh = meshgen.synthetic_mesh(points, seed=0)
    nb = mesh.neigh
    col, nc = pb.face_colors(nb)
    ocol, onc = oracle.face_colors(nb)
    np.testing.assert_array_equal(col, ocol)
    assert nc == onc == 3
    for f in range(3):
        has = nb[:, f] >= 0
        assert (col[has] != col[nb[has, f]]).all()
    assert _greedy_colors(nb).max() == 3  # greedy alone needs a fourth colour on these meshes
