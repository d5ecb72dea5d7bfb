"""HDF5 mesh and parameter files (src/mesh/triangulation.cpp:724-1036 load_mesh_from_h5,
:1234-1330 load_hdf5_parameters), read by chm_amd/host/h5_file.cpp through the host
driver.

Pinned by the reference's own test data: functional_tests/mesh_versioning holds the
same mesh as .mesh / .param JSON (slope.mesh, slope.param) and as HDF5
(slope.metis_mesh.h5, slope.metis_param.h5), the latter with its faces reordered so
that face i is JSON face id[i] (id: a parameter of the .h5 parameter file).  So
the HDF5 reader must return the JSON vertices bit for bit, the JSON triangles and
areas in that order, and the JSON neighbours renumbered by the same permutation;
nothing is computed from the reader's own output.  Corrupted copies of the file
must raise a clear error, never crash or read out of bounds."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from chm_amd import abi, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
MESH_H5 = os.path.join(GOLD, "slope_mesh.h5")
PARAM_H5 = os.path.join(GOLD, "slope_param.h5")
BIN = os.path.join(ROOT, "chm_amd", "bin", "pbsm3d_host")


def _bin():
    if not os.path.exists(BIN):
        from chm_amd import build

        build.build_host()
    return BIN


def _dump(*paths):
    r = subprocess.run([_bin(), "mesh", *paths], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _reference_json_mesh():
    d = np.load(os.path.join(GOLD, "slope.npz"))
    return d["vertex"], d["elem"], d["neigh"], d["area"]


def _permutation(dump):
    ids = np.asarray(dump["parameters"]["id"])
    assert np.array_equal(ids, np.round(ids)) and sorted(ids.astype(int)) == list(range(len(ids)))
    ids = ids.astype(np.int64)
    inv = np.empty_like(ids)
    inv[ids] = np.arange(len(ids))
    return ids, inv


def test_h5_mesh_is_the_json_mesh_reordered():
    d = _dump(MESH_H5, PARAM_H5)
    V, E, N, A = _reference_json_mesh()
    F = len(E)
    assert d["faces"] == F == 2618
    ids, inv = _permutation(d)
    assert np.array_equal(np.array(d["vertex"]).reshape(-1, 3), V)
    assert np.array_equal(np.array(d["elem"]).reshape(-1, 3), E[ids])
    Np = N[ids]
    assert np.array_equal(np.array(d["neigh"]).reshape(-1, 3), np.where(Np >= 0, inv[np.maximum(Np, 0)], -1))
    assert np.array_equal(np.array(d["parameters"]["area"]), A[ids])
    assert sorted(d["parameters"]) == ["area", "id"]
    # metadata: a version 2.0.0, metis-ordered, unpartitioned file; owners all rank 0
    assert d["version"] == "2.0.0" and d["partition_method"] == "metis" and d["is_partition"] is False
    assert d["proj4"] == "+proj=utm +zone=8 +datum=NAD83 +units=m +no_defs"
    assert sum(d["local_sizes"]) == F and len(d["local_sizes"]) == 31
    assert d["owner"] == [0] * F and d["cell_global_id"] == list(range(F))


def test_h5_and_json_readers_agree(tmp_path):
    """The reordered JSON mesh written as .mesh JSON and read by the JSON loader
    equals what the HDF5 loader returns."""
    d = _dump(MESH_H5, PARAM_H5)
    V, E, N, A = _reference_json_mesh()
    ids, inv = _permutation(d)
    Np = N[ids]
    m = meshgen.Mesh(V, E[ids], np.where(Np >= 0, inv[np.maximum(Np, 0)], -1))
    p = tmp_path / "slope_reordered.json"
    meshgen.save_mesh_json(m, str(p))
    j = _dump(str(p))
    for k in ("vertex", "elem", "neigh"):
        assert j[k] == d[k], k


def test_h5_face_geometry():
    """The module's face API on the HDF5 mesh (area from the parameter file, as the
    reference's get_area prefers the "area" parameter) equals pbsm3d_face_geometry."""
    r = subprocess.run([_bin(), "check", MESH_H5, PARAM_H5], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    c = json.loads(r.stdout)
    assert c["faces"] == 2618 and c["geometry_rc"] == 0 and c["geometry_max_abs_diff"] == 0.0


def _corrupt(tmp_path, name, edit):
    b = bytearray(open(MESH_H5, "rb").read())
    b = edit(b)
    p = tmp_path / name
    p.write_bytes(bytes(b))
    return str(p)


def _fails(path, *extra):
    r = subprocess.run([_bin(), "mesh", path, *extra], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and r.returncode < 128, (r.returncode, r.stdout[:200])
    return r.stderr


def _patch_row(b, old, new):
    k = bytes(b).find(struct.pack("<3i", *old))
    assert k > 0
    b[k:k + 12] = struct.pack("<3i", *new)
    return b


@pytest.mark.parametrize("case", ["truncated", "not_hdf5", "superblock_v2", "bad_btree", "vertex_out_of_range",
                                  "self_neighbour", "missing_param_file"])
def test_h5_errors(tmp_path, case):
    if case == "truncated":
        p = _corrupt(tmp_path, "t.h5", lambda b: b[:60000])
        assert "truncated or corrupt" in _fails(p)
    elif case == "not_hdf5":
        p = tmp_path / "x.h5"
        p.write_text('{"mesh": {}}')
        assert "not an HDF5 file" in _fails(str(p))
    elif case == "superblock_v2":
        def v2(b):
            b[8] = 2
            return b
        assert "superblock version 2" in _fails(_corrupt(tmp_path, "v2.h5", v2))
    elif case == "bad_btree":
        def tree(b):
            k = bytes(b).find(b"TREE")
            b[k:k + 4] = b"XREE"
            return b
        assert "B-tree" in _fails(_corrupt(tmp_path, "bt.h5", tree))
    elif case == "vertex_out_of_range":
        # first triangle (315, 346, 353) -> vertex 1374 of 1374
        p = _corrupt(tmp_path, "e.h5", lambda b: _patch_row(b, (315, 346, 353), (1374, 346, 353)))
        assert "vertex out of range" in _fails(p)
    elif case == "self_neighbour":
        # face 0's neighbours (29, 31, 23) -> (0, 31, 23): the face itself
        p = _corrupt(tmp_path, "n.h5", lambda b: _patch_row(b, (29, 31, 23), (0, 31, 23)))
        assert "is face" in _fails(p)
    elif case == "missing_param_file":
        assert "cannot open" in _fails(MESH_H5, str(tmp_path / "none.h5"))


@pytest.mark.gpu
def test_module_step_on_h5_mesh(tmp_path):
    """Two PBSM3D steps through the C++ module on the HDF5 mesh with its parameter
    file equal the C-ABI's steps on the same arrays (area from the parameter)."""
    from chm_amd import pbsm3d as pb

    print("[case] host module on the HDF5 test mesh", flush=True)
    d = _dump(MESH_H5, PARAM_H5)
    F = d["faces"]
    m = meshgen.Mesh(np.array(d["vertex"]).reshape(-1, 3), np.array(d["elem"], dtype=np.int32).reshape(-1, 3),
                     np.array(d["neigh"], dtype=np.int32).reshape(-1, 3), np.array(d["parameters"]["area"]))
    fo = meshgen.forcing_c2(F, seed=4)
    fbin = tmp_path / "f.bin"
    np.concatenate([fo[k] for k in abi.FORCING_FIELDS]).astype("<f8").tofile(fbin)
    obin = tmp_path / "o.bin"
    r = subprocess.run([_bin(), "run", MESH_H5, str(fbin), "2", str(obin), "nLayer=5", f"param={PARAM_H5}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    st = json.loads(r.stdout)
    assert st["faces"] == F and st["restart_identical"] is True
    got = np.fromfile(obin, dtype="<f8").reshape(len(abi.OUT_FIELDS), F)
    ctx = pb.Context(pb.face_geometry(m), pb.default_config(nLayer=5))
    for _ in range(2):
        ref, _ = ctx.step(fo)
    ctx.close()
    for k, name in enumerate(abi.OUT_FIELDS):
        np.testing.assert_array_equal(got[k], ref[name], err_msg=name)
