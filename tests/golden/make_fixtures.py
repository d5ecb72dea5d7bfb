"""Writes the committed fixtures under tests/golden/.

* granger1m.npz -- vertex/elem/neigh/area arrays of the reference's own test
  mesh test_data/meshes/granger1m.{mesh,param} (read as JSON data; the
  reference's triangulation tests load the same files,
  src/tests/test_triangulation.cpp:42-62).
* patterns.json -- SHA-256 of (row_ptr, col) of the suspension / deposition
  CSR patterns built by an independent set-based numpy builder for the
  config-1 mesh (L = 5, 10) and the granger1m mesh (L = 10).
* tiny_fields.npz -- inputs and outputs of one step on a 12-point mesh solved
  by dense Gaussian elimination (no Krylov tolerance involved), written by the
  oracle itself: a regression fixture for the oracle and the GPU path, not an
  independent pin.
* slope_mesh.h5, slope_param.h5 -- the reference's HDF5 test mesh and its
  parameter file, copied as data (functional_tests/mesh_versioning/
  slope.metis_mesh.h5, slope.metis_param.h5).
* slope.npz -- vertex/elem/neigh of the same mesh in its .mesh JSON form and
  the area of its .param JSON (functional_tests/mesh_versioning/slope.mesh,
  slope.param), read as JSON data.  The .h5 holds these faces reordered: face
  i of the .h5 is face id[i] of the JSON mesh (id: the parameter file's "id").

Run from the repository root:  python tests/golden/make_fixtures.py
(or  python tests/golden/make_fixtures.py slope  for item 4 alone)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

REF_MESH = "/root/reference/test_data/meshes/granger1m.mesh"
REF_PARAM = "/root/reference/test_data/meshes/granger1m.param"
REF_SLOPE = "/root/reference/functional_tests/mesh_versioning"


def brute_force_pattern(n_faces: int, n_layer: int, neigh: np.ndarray):
    """Sorted-column CSR of the extruded nearest-neighbour graph, from Python sets."""
    rows = []
    for z in range(n_layer):
        for i in range(n_faces):
            cols = {z * n_faces + i}
            for f in range(3):
                if neigh[i, f] >= 0:
                    cols.add(z * n_faces + int(neigh[i, f]))
            if z > 0:
                cols.add((z - 1) * n_faces + i)
            if z < n_layer - 1:
                cols.add((z + 1) * n_faces + i)
            rows.append(sorted(cols))
    rp = np.zeros(len(rows) + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], dtype=np.int32)
    return rp, col


def pattern_sha(rp: np.ndarray, col: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rp, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(col, dtype="<i4").tobytes())
    return h.hexdigest()


def slope():
    import shutil

    from chm_amd import meshgen

    shutil.copyfile(os.path.join(REF_SLOPE, "slope.metis_mesh.h5"), os.path.join(HERE, "slope_mesh.h5"))
    shutil.copyfile(os.path.join(REF_SLOPE, "slope.metis_param.h5"), os.path.join(HERE, "slope_param.h5"))
    m = meshgen.load_mesh_json(os.path.join(REF_SLOPE, "slope.mesh"), os.path.join(REF_SLOPE, "slope.param"))
    np.savez_compressed(os.path.join(HERE, "slope.npz"), vertex=m.vertex, elem=m.elem, neigh=m.neigh, area=m.area)


def main():
    from chm_amd import meshgen

    # 1. the reference's granger1m test mesh
    mesh = meshgen.load_mesh_json(REF_MESH, REF_PARAM)
    np.savez_compressed(os.path.join(HERE, "granger1m.npz"), vertex=mesh.vertex, elem=mesh.elem,
                        neigh=mesh.neigh, area=mesh.area)

    # 2. pattern hashes
    c1 = meshgen.synthetic_mesh(2500, seed=0)
    pats = {"c1_faces": c1.n_faces, "granger_faces": mesh.n_faces}
    for name, m, layers in (("c1", c1, (1, 5, 10)), ("granger1m", mesh, (1, 10))):
        for L in layers:
            rp, col = brute_force_pattern(m.n_faces, L, m.neigh)
            pats[f"{name}_L{L}"] = {"sha256": pattern_sha(rp, col), "nnz": int(rp[-1])}
    with open(os.path.join(HERE, "patterns.json"), "w") as fh:
        json.dump(pats, fh, indent=1, sort_keys=True)

    # 3. tiny mesh, dense solve
    from oracle import oracle_py as O

    tiny = meshgen.synthetic_mesh(12, seed=5, extent=2000.0)
    g = O.face_geometry(tiny)
    fo = meshgen.forcing_c1(tiny.n_faces, seed=5)
    o = O.Oracle(g, O.oracle_config(nLayer=3), direct=True)
    outs, st = o.step(fo)
    rp, col, val, rhs = o.system(0)
    np.savez_compressed(
        os.path.join(HERE, "tiny_fields.npz"),
        vertex=tiny.vertex, elem=tiny.elem, neigh=tiny.neigh,
        **{f"forcing_{k}": v for k, v in fo.items()},
        **{f"out_{k}": v for k, v in outs.items()},
        susp_val=val, susp_rhs=rhs,
    )
    print(json.dumps(pats, indent=1))

    # 4. the HDF5 test mesh and its JSON form
    slope()


if __name__ == "__main__":
    slope() if sys.argv[1:] == ["slope"] else main()
