"""Writes tests/golden/c2_oracle_sample.npz: the production-size workload checked by
tests/test_c2_full.py without re-solving it on the GPU box.

Workload: meshgen.synthetic_mesh(500000, seed=0) (999,968 triangles), 10 layers,
meshgen.forcing_c2(F, seed=0), one 900 s step.  The fixture holds

* mesh_sha -- SHA-256 of the mesh's elem and neigh arrays (the test first checks that
  the generator rebuilt the same mesh);
* pattern_sha_L10 / pattern_sha_L1, nnz_L10 / nnz_L1 -- SHA-256 of (row_ptr, col) of the
  suspension and deposition CSR patterns built by the vectorised numpy builder below
  (columns of each row sorted, LinearAlgebra.cpp:67-149), cross-checked against the
  oracle's own pattern when the fixture is written;
* the oracle step (oracle/pbsm3d_oracle.cpp, the reference's GMRES(30) + ILUT, solved
  to a relative residual of 1e-12) at a seeded sample of faces: the per-face fields at
  `idx` (16,000 faces), the 10-layer concentration at `idx_conc` (4,000 faces), and the
  largest magnitude of every field over all faces (the test's error scale).

Run from the repository root:  python tests/golden/make_c2_fixture.py
(about a minute on 8 cores; the oracle runs with every available thread).
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

FIELDS = ("Qsalt", "Qsusp", "Qsubl", "qdep", "drift_mass", "sum_drift", "sum_subl")


def pattern_numpy(n_faces: int, n_layer: int, neigh: np.ndarray):
    """Sorted-column CSR of the extruded nearest-neighbour graph: each row r = z F + i
    holds {r, z F + neighbours of i, r - F if z > 0, r + F if z < L - 1}, built for all
    rows at once by sorting a padded candidate table."""
    F, L = n_faces, n_layer
    big = np.iinfo(np.int64).max
    z = np.repeat(np.arange(L, dtype=np.int64), F)
    i = np.tile(np.arange(F, dtype=np.int64), L)
    r = z * F + i
    nb = np.asarray(neigh, dtype=np.int64).reshape(F, 3)
    cand = np.full((F * L, 6), big, dtype=np.int64)
    cand[:, 0] = r
    for f in range(3):
        n = nb[i, f]
        cand[:, 1 + f] = np.where(n >= 0, z * F + n, big)
    cand[:, 4] = np.where(z > 0, r - F, big)
    cand[:, 5] = np.where(z < L - 1, r + F, big)
    cand.sort(axis=1)
    valid = cand != big
    rp = np.zeros(F * L + 1, dtype=np.int64)
    rp[1:] = np.cumsum(valid.sum(axis=1))
    return rp.astype(np.int32), cand[valid].astype(np.int32)


def pattern_sha(rp: np.ndarray, col: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(rp, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(col, dtype="<i4").tobytes())
    return h.hexdigest()


def mesh_sha(mesh) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(mesh.elem, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(mesh.neigh, dtype="<i4").tobytes())
    return h.hexdigest()


def main():
    from chm_amd import meshgen
    from oracle import oracle_py

    t0 = time.time()
    mesh = meshgen.synthetic_mesh(500000, seed=0)
    F, L = mesh.n_faces, 10
    out = {"faces": F, "layers": L, "mesh_sha": mesh_sha(mesh)}
    for layers in (L, 1):
        rp, col = pattern_numpy(F, layers, mesh.neigh)
        orp, ocol = oracle_py.csr_pattern(F, layers, mesh.neigh)
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol), "oracle pattern differs"
        out[f"pattern_sha_L{layers}"] = pattern_sha(rp, col)
        out[f"nnz_L{layers}"] = int(rp[-1])
    print(f"mesh + patterns {time.time() - t0:.1f} s", flush=True)
    g = oracle_py.face_geometry(mesh)
    fo = meshgen.forcing_c2(F, seed=0)
    orc = oracle_py.Oracle(g, oracle_py.oracle_config(nLayer=L, tol=1e-12),
                           threads=oracle_py.lib().oracle_max_threads())
    outs, st = orc.step(fo)
    print(f"oracle step {st['ms_total'] / 1e3:.1f} s: {st['susp_iters']} + {st['dep_iters']} iterations, "
          f"residuals {st['susp_resid']:.2e} / {st['dep_resid']:.2e}", flush=True)
    assert st["susp_resid"] <= 1e-12 and st["dep_resid"] <= 1e-12
    rng = np.random.default_rng(2024)
    idx = np.sort(rng.choice(F, 16000, replace=False)).astype(np.int32)
    idx_conc = np.sort(rng.choice(F, 4000, replace=False)).astype(np.int32)
    for k in FIELDS:
        out[k] = outs[k][idx]
        out[f"max_{k}"] = float(np.abs(outs[k]).max())
    conc = outs["conc"].reshape(L, F)
    out["conc"] = conc[:, idx_conc]
    out["max_conc"] = float(np.abs(conc).max())
    out["idx"] = idx
    out["idx_conc"] = idx_conc
    out["oracle_susp_iters"] = st["susp_iters"]
    out["oracle_dep_iters"] = st["dep_iters"]
    np.savez_compressed(os.path.join(HERE, "c2_oracle_sample.npz"), **out)
    print(f"wrote c2_oracle_sample.npz ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
