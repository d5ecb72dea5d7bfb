"""The C-ABI library loads and exports what include/*.h declares (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from chm_amd import abi, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(pbsm3d_[a-z_0-9]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from chm_amd import pbsm3d

    L = pbsm3d.lib()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert L.pbsm3d_abi_version() == 7
    assert set(pbsm3d.EXPORTED) <= set(names)


def test_struct_layouts_match_header():
    # field counts of the ctypes mirror equal the C struct members
    txt = open(os.path.join(ROOT, "include", "pbsm3d.h")).read()
    for cname, py in (("pbsm3d_mesh", abi.Mesh), ("pbsm3d_config", abi.Config), ("pbsm3d_forcing", abi.Forcing),
                      ("pbsm3d_out", abi.Out), ("pbsm3d_stats", abi.Stats)):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (cname, cname), txt, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        members = re.findall(r"([A-Za-z_][A-Za-z_0-9]*)\s*;", body)
        assert [f[0] for f in py._fields_] == members, cname


def test_config_defaults_match_reference():
    from chm_amd import pbsm3d

    c = pbsm3d.default_config()
    # PBSM3D.cpp:123-145, 223-258 and LinearAlgebra.cpp:165-185
    assert (c.nLayer, c.settling_velocity, c.smooth_coeff, c.min_sd_trans, c.cutoff, c.snow_diffusion_const) == (
        10, 0.5, 820.0, 0.1, 0.3, 0.3)
    assert (c.do_fixed_settling, c.do_sublimation, c.do_lateral_diff, c.enable_veg) == (0, 1, 1, 1)
    assert (c.use_tanh_fetch, c.use_exp_fetch, c.use_R94_lambda) == (1, 0, 1)
    assert (c.tol, c.max_iters, c.restart) == (1e-8, 1000, 30)
    assert c.z0_snow == 0.01 and c.settling_scale == 1.0
    # PBSM3D.cpp:124-125, 258: off by default
    assert (c.use_PomLi_probability, c.z0_ustar_coupling, c.rouault_diffusion_coef) == (0, 0, 0)
    d = abi.default_config()
    for name, _ in abi.Config._fields_:
        assert getattr(c, name) == getattr(d, name), name


def test_host_face_geometry_matches_oracle(oracle, mesh_c1):
    from chm_amd import pbsm3d

    g1 = pbsm3d.face_geometry(mesh_c1)
    g0 = oracle.face_geometry(mesh_c1)
    for k in ("normal_x", "normal_y", "edge_length", "area", "center_x", "center_y", "center_z"):
        assert np.array_equal(getattr(g1, k), getattr(g0, k)), k


def test_create_rejects_bad_config_without_gpu():
    """Argument validation happens before any device call."""
    from chm_amd import pbsm3d

    m = meshgen.synthetic_mesh(30, seed=1)
    g = pbsm3d.face_geometry(m)
    for bad in (dict(nLayer=1), dict(settling_velocity=-1.0), dict(use_exp_fetch=1, use_tanh_fetch=1),
                dict(restart=0), dict(precond=abi.PC_ILUT), dict(krylov=7), dict(inner_rtol=1.5),
                dict(nLayer=17, krylov=abi.KRYLOV_BICGSTAB_SPLIT),
                dict(precond=abi.PC_JACOBI, krylov=abi.KRYLOV_BICGSTAB_SPLIT)):
        with pytest.raises(pbsm3d.Pbsm3dError) as e:
            pbsm3d.Context(g, pbsm3d.default_config(**bad))
        assert e.value.code == abi.PBSM3D_EINVAL


def test_context_calls_reject_null_context():
    """Calls on a context validate the handle before touching it (no GPU needed)."""
    from chm_amd import pbsm3d

    L = pbsm3d.lib()
    assert L.pbsm3d_profile_enable(None, 1) == abi.PBSM3D_EINVAL
    assert L.pbsm3d_profile_classes(None, 1) == abi.PBSM3D_EINVAL
    assert L.pbsm3d_precond_select(None, 0, 1) == abi.PBSM3D_EINVAL
    assert L.pbsm3d_spmv_select(None, 0, 1) == abi.PBSM3D_EINVAL
    assert L.pbsm3d_bicg_restart(None, 1e-5) == abi.PBSM3D_EINVAL
    assert L.pbsm3d_batch_training_pairs(None, 0, None, None) == abi.PBSM3D_EINVAL


def test_empty_inputs_rejected_without_gpu():
    """A mesh with no faces is refused by every entry point before any device call."""
    import ctypes as C

    from chm_amd import pbsm3d

    L = pbsm3d.lib()
    assert L.pbsm3d_face_geometry(0, None, None, None, *([None] * 7)) == abi.PBSM3D_EINVAL
    assert L.pbsm3d_face_colors(0, None, None) == abi.PBSM3D_EINVAL
    h = C.c_void_p()
    x = np.zeros(1)
    nb = np.full(3, -1, dtype=np.int32)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert L.pbsm3d_snow_create(C.byref(h), 0, 0, None, C.byref(abi.SnowConfig())) == abi.PBSM3D_EINVAL
    assert L.pbsm3d_fetch_create(C.byref(h), 0, 0, dp(x), dp(x), dp(x), None,
                                 C.byref(abi.FetchConfig())) == abi.PBSM3D_EINVAL
    assert L.pbsm3d_wind_create(C.byref(h), 0, 0, nb.ctypes.data_as(C.POINTER(C.c_int32)), dp(x), dp(x), None, None,
                                C.byref(abi.WindConfig())) == abi.PBSM3D_EINVAL
    assert h.value is None
    m = abi.Mesh()
    m.n_faces = 0
    assert L.pbsm3d_create(C.byref(h), C.byref(m), C.byref(pbsm3d.default_config()), 0) == abi.PBSM3D_EINVAL


def test_bad_neighbour_links_rejected_without_gpu():
    """The kernels read a neighbour's coupling back through the edge that points to the
    face, so pbsm3d_create and pbsm3d_face_colors refuse adjacency that is not a
    triangulation's: a face that is its own neighbour, a neighbour listed twice, a link
    not reciprocated (or reciprocated twice).  Refused before any device call."""
    import ctypes as C

    from chm_amd import pbsm3d

    m = meshgen.synthetic_mesh(60, seed=2)
    g0 = pbsm3d.face_geometry(m)
    nb0 = g0.neigh.reshape(-1, 3)
    i = int(np.flatnonzero((nb0 >= 0).all(axis=1))[0])  # an interior face
    n = int(nb0[i, 0])
    bad = []
    a = nb0.copy(); a[i, 1] = i; bad.append(("self", a))
    a = nb0.copy(); a[i, 1] = a[i, 0]; bad.append(("twice", a))
    a = nb0.copy(); a[n, list(a[n]).index(i)] = -1; bad.append(("one-way", a))
    a = nb0.copy(); a[i, 0] = -1; bad.append(("one-way back", a))
    L = pbsm3d.lib()
    col = np.zeros(m.n_faces, dtype=np.int32)
    for name, a in bad:
        g = pbsm3d.face_geometry(m)
        g.neigh = np.ascontiguousarray(a.reshape(-1).astype(np.int32))
        with pytest.raises(pbsm3d.Pbsm3dError) as e:
            pbsm3d.Context(g, pbsm3d.default_config())
        assert e.value.code == abi.PBSM3D_EINVAL, name
        assert L.pbsm3d_face_colors(m.n_faces, g.neigh.ctypes.data_as(C.POINTER(C.c_int32)),
                                    col.ctypes.data_as(C.POINTER(C.c_int32))) == abi.PBSM3D_EINVAL, name
    # the unmodified mesh colours fine
    assert L.pbsm3d_face_colors(m.n_faces, g0.neigh.ctypes.data_as(C.POINTER(C.c_int32)),
                                col.ctypes.data_as(C.POINTER(C.c_int32))) >= 3


def test_coarse_correction_config_rejected_without_gpu():
    """coarse_tiles is refused before any device call outside [0, 256], above the face
    count, without the learned start, or without the split BiCGStab; the batch has no
    coarse correction."""
    from chm_amd import pbsm3d

    m = meshgen.synthetic_mesh(60, seed=2)
    g = pbsm3d.face_geometry(m)
    split, learned = abi.KRYLOV_BICGSTAB_SPLIT, abi.WS_LEARNED
    for kw in ({"coarse_tiles": -1, "krylov": split, "warm_start": learned},
               {"coarse_tiles": 257, "krylov": split, "warm_start": learned},
               {"coarse_tiles": m.n_faces + 1, "krylov": split, "warm_start": learned},
               {"coarse_tiles": 4, "krylov": split},
               {"coarse_tiles": 4, "warm_start": learned}):
        with pytest.raises(pbsm3d.Pbsm3dError) as e:
            pbsm3d.Context(g, pbsm3d.default_config(nLayer=10, **kw))
        assert e.value.code == abi.PBSM3D_EINVAL, kw
    with pytest.raises(pbsm3d.Pbsm3dError) as e:
        pbsm3d.Batch([g], [pbsm3d.default_config(nLayer=10, krylov=split, warm_start=learned, coarse_tiles=4)])
    assert e.value.code == abi.PBSM3D_EINVAL
    assert pbsm3d.lib().pbsm3d_coarse_system(None, None, None, None) == abi.PBSM3D_EINVAL
