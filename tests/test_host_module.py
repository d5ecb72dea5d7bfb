"""The C++ PBSM3D module (chm_amd/host) driven like the reference's core drives
a module: constructed from config, init(mesh), run(mesh) per step, checkpoint.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from chm_amd import abi, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "chm_amd", "bin", "pbsm3d_host")


def _bin():
    if not os.path.exists(BIN):
        from chm_amd import build

        build.build_host()
    return BIN


def test_module_interface_and_face_api(tmp_path):
    """depends/provides (PBSM3D.cpp:103-219), config error on both fetch modes,
    -9999 defaults, and the module's face API geometry = pbsm3d_face_geometry."""
    m = meshgen.synthetic_mesh(300, seed=3)
    p = tmp_path / "m.json"
    meshgen.save_mesh_json(m, str(p))
    out = subprocess.run([_bin(), "check", str(p)], capture_output=True, text=True, check=True).stdout
    d = json.loads(out)
    assert d["faces"] == m.n_faces
    assert d["depends"] == ["U_2m_above_srf", "vw_dir", "swe", "t", "rh", "U_R", "fetch"]
    assert set(d["provides"]) == {"pbsm_more_than_avail", "global_cell_id", "blowingsnow_probability", "Qsubl",
                                  "Qsubl_mass", "sum_subl", "drift_mass", "Qsusp", "Qsalt", "sum_drift"}
    assert d["both_fetch_throws"] is True
    assert d["geometry_rc"] == 0 and d["geometry_max_abs_diff"] == 0.0
    assert d["default_nan"] is True


def test_reference_mesh_json_roundtrip(tmp_path):
    """The reference's granger1m test mesh, written in the .mesh JSON schema and
    read back by the module's loader, keeps its topology."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "granger1m.npz"))
    m = meshgen.Mesh(g["vertex"], g["elem"], g["neigh"])
    p = tmp_path / "granger.json"
    meshgen.save_mesh_json(m, str(p))
    back = meshgen.load_mesh_json(str(p))
    assert np.array_equal(back.elem, m.elem) and np.array_equal(back.neigh, m.neigh)
    d = json.loads(subprocess.run([_bin(), "check", str(p)], capture_output=True, text=True, check=True).stdout)
    assert d["faces"] == 985 and d["geometry_max_abs_diff"] == 0.0


_OPTION_KEYS = ("use_PomLi_probability", "rouault_diffusion_coef", "z0_ustar_coupling")


@pytest.mark.gpu
@pytest.mark.parametrize("keys", [(), ("mixed_precision=1", "solver_bicgstab_split=1", "solver_sweep_mode=1",
                                       "solver_spmv_variant=1", "checkpoint_sum_subl=1"),
                                  tuple(f"{k}=1" for k in _OPTION_KEYS)])
def test_module_run_matches_abi(tmp_path, keys):
    """Three steps of the C++ module equal three steps through the C ABI bit for bit; the
    split mixed solver with its launch shapes pinned on both sides (the module keys
    solver_sweep_mode / solver_spmv_variant), the reference's optional keys (PomLi
    probability from p_snow_hours, Rouault diffusivity, z0 / u* coupling), and the
    checkpoint restore continues identically -- with checkpoint_sum_subl also in sum_subl."""
    from chm_amd import pbsm3d as pb

    print(f"[case] host module run {keys}", flush=True)
    m = meshgen.synthetic_mesh(600, seed=9)
    F = m.n_faces
    p = tmp_path / "m.json"
    meshgen.save_mesh_json(m, str(p))
    fo = meshgen.forcing_c2(F, seed=9)
    fo["p_snow_hours"] = np.random.default_rng(9).uniform(1.0, 300.0, F)
    fbin = tmp_path / "f.bin"
    np.concatenate([fo[k] for k in abi.FORCING_FIELDS]).astype("<f8").tofile(fbin)
    obin = tmp_path / "o.bin"
    r = subprocess.run([_bin(), "run", str(p), str(fbin), "3", str(obin), "nLayer=6", *keys], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    st = json.loads(r.stdout)
    mixed = "mixed_precision=1" in keys
    assert st["restart_identical"] is True and st["susp_resid"] <= 1e-8
    assert st["restart_sum_subl_identical"] is mixed and st["checkpoint_keys"] == (2 if mixed else 1)
    got = np.fromfile(obin, dtype="<f8").reshape(len(abi.OUT_FIELDS), F)
    extra = dict(precision=abi.MIXED, krylov=abi.KRYLOV_BICGSTAB_SPLIT) if mixed else {}
    extra.update({k: 1 for k in _OPTION_KEYS if f"{k}=1" in keys})
    ctx = pb.Context(pb.face_geometry(m), pb.default_config(nLayer=6, **extra))
    if mixed:
        ctx.precond_select(1, 1)
        ctx.spmv_select(1, 1)
    for _ in range(3):
        ref, rst = ctx.step(fo)
    assert st["susp_iters"] == rst["susp_iters"] and st["susp_refinements"] == rst["susp_refinements"]
    if mixed:
        assert st["susp_refinements"] > 0
    if "use_PomLi_probability=1" in keys:
        assert ((ref["blowingsnow_probability"] > 0) & (ref["blowingsnow_probability"] < 1)).any()
    for k, name in enumerate(abi.OUT_FIELDS):
        np.testing.assert_array_equal(got[k], ref[name], err_msg=name)


@pytest.mark.gpu
def test_fetchr_module_feeds_pbsm3d(tmp_path, oracle):
    """The C++ fetchr module (reference keys steps / max_distance / I / incl_veg)
    provides `fetch` each step before PBSM3D runs, as the reference's core
    orders them; its output equals the oracle's fetch for the same directions."""
    print("[case] host fetchr module", flush=True)
    m = meshgen.synthetic_mesh(600, seed=9)
    F = m.n_faces
    p = tmp_path / "m.json"
    meshgen.save_mesh_json(m, str(p))
    fo = meshgen.forcing_c2(F, seed=9)
    fbin = tmp_path / "f.bin"
    np.concatenate([fo[k] for k in abi.FORCING_FIELDS]).astype("<f8").tofile(fbin)
    obin = tmp_path / "o.bin"
    r = subprocess.run([_bin(), "run", str(p), str(fbin), "2", str(obin), "nLayer=6", "use_fetchr=1", "steps=8",
                        "max_distance=800", "I=0.03"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(str(obin) + ".fetch", dtype="<f8")
    c = m.vertex[m.elem].mean(axis=1)
    cfg = abi.FetchConfig()
    cfg.steps, cfg.incl_veg, cfg.max_distance, cfg.I = 8, 1, 800.0, 0.03
    want = oracle.fetch(np.ascontiguousarray(c[:, 0]), np.ascontiguousarray(c[:, 1]), np.ascontiguousarray(c[:, 2]),
                        None, cfg, fo["vw_dir"])
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_scale_wind_vert_module_feeds_pbsm3d(tmp_path, oracle):
    """The C++ scale_wind_vert module (domain mode) provides U_2m_above_srf from U_R
    and snowdepthavg before PBSM3D runs; it equals the oracle's value."""
    print("[case] host scale_wind_vert module", flush=True)
    m = meshgen.synthetic_mesh(600, seed=9)
    F = m.n_faces
    p = tmp_path / "m.json"
    meshgen.save_mesh_json(m, str(p))
    fo = meshgen.forcing_c2(F, seed=9)
    fo["snowdepthavg"] = np.random.default_rng(2).uniform(0.1, 2.0, F)
    fbin = tmp_path / "f.bin"
    np.concatenate([fo[k] for k in abi.FORCING_FIELDS]).astype("<f8").tofile(fbin)
    obin = tmp_path / "o.bin"
    r = subprocess.run([_bin(), "run", str(p), str(fbin), "1", str(obin), "nLayer=6", "use_scale_wind_vert=1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(str(obin) + ".u2", dtype="<f8")
    c = m.vertex[m.elem].mean(axis=1)
    cfg = abi.WindConfig()
    cfg.ignore_canopy, cfg.smooth, cfg.z0_snow = 0, 1, 0.01
    want = oracle.scale_wind(m.neigh, np.ascontiguousarray(c[:, 0]), np.ascontiguousarray(c[:, 1]), None, None, cfg,
                             fo["U_R"], fo["snowdepthavg"])
    np.testing.assert_allclose(got, want, rtol=1e-12)


@pytest.mark.gpu
def test_snobal_module_closes_the_loop(tmp_path):
    """The C++ snobal module (drift part; keys drift_density / IC_rho) runs before PBSM3D
    each step, as CHM orders a provider of swe before its consumer: it takes the previous
    step's drift_mass and provides swe and snowdepthavg.  Three steps equal the same chain
    driven through the C-ABI (pbsm3d_snow_* then pbsm3d_step) bit for bit."""
    from chm_amd import pbsm3d as pb

    print("[case] host snobal module", flush=True)
    m = meshgen.synthetic_mesh(600, seed=9)
    F = m.n_faces
    p = tmp_path / "m.json"
    meshgen.save_mesh_json(m, str(p))
    fo = meshgen.forcing_c2(F, seed=9)
    fo["swe"] = np.random.default_rng(3).uniform(0.0, 60.0, F)
    fbin = tmp_path / "f.bin"
    np.concatenate([fo[k] for k in abi.FORCING_FIELDS]).astype("<f8").tofile(fbin)
    obin = tmp_path / "o.bin"
    r = subprocess.run([_bin(), "run", str(p), str(fbin), "3", str(obin), "nLayer=6", "use_snobal=1",
                        "IC_rho=250", "drift_density=280"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got_swe = np.fromfile(str(obin) + ".swe", dtype="<f8")
    got_sd = np.fromfile(str(obin) + ".sd", dtype="<f8")
    got = np.fromfile(obin, dtype="<f8").reshape(len(abi.OUT_FIELDS), F)
    sn = pb.Snow(F, fo["swe"], pb.snow_config(ic_rho=250.0, drift_density=280.0))
    ctx = pb.Context(pb.face_geometry(m), pb.default_config(nLayer=6))
    f2 = dict(fo)
    drift = np.full(F, -9999.0)
    for _ in range(3):
        sn.drift(drift)
        f2["swe"], f2["snowdepthavg"] = sn.outputs()
        out, _ = ctx.step(f2)
        drift = out["drift_mass"]
    np.testing.assert_array_equal(got_swe, f2["swe"])
    np.testing.assert_array_equal(got_sd, f2["snowdepthavg"])
    for k, name in enumerate(abi.OUT_FIELDS):
        np.testing.assert_array_equal(got[k], out[name], err_msg=name)
    assert np.abs(f2["swe"] - fo["swe"]).max() > 0  # drift changed the snowcover
    ctx.close()
    sn.close()
