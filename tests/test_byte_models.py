"""The algorithmic-byte models behind bench.py's roofline, recounted element by element.

bench.split_apply_bytes is a closed form for one application of the split (Eisenstat)
D-ILU operator, averaged over BiCGStab's two applications per iteration.  Here the five
launches are walked face by face and layer by layer on small meshes, following what each
launch reads and writes at its own rows (the kernels in chm_amd/csrc/precond.hip): the
backward sweeps of colours top .. 1, the fused colour-0 launch, the forward sweeps of
colours 1 .. top.  Each launch counts every element it touches once; reads of other
colours' t and v at neighbour rows are reuse and are not counted (the SpMV convention of
perfect reuse of x).  The dot vector is a separate array in the first application (r^ . v)
and the input itself in the second (s . t).  A dropped or doubled term in the closed form
(a vector, the top colour's missing v store or lateral loads, the dot vector's second
count) fails here.  No GPU: the colouring is the library's host routine.
"""
import numpy as np
import pytest

import bench
from chm_amd import meshgen
from chm_amd import pbsm3d as pb

S = 4  # FP32 elements
IDX = 13  # neighbour positions (3 x int32) and the colour mask (uint8) of one face


def _launch_bytes(neigh, colors, L, dot_is_input):
    top = int(colors.max())
    faces_of = [np.flatnonzero(colors == c) for c in range(top + 1)]

    def lateral(i, upper):
        out = []
        for f in range(3):
            n = neigh[i, f]
            if n >= 0 and (colors[n] > colors[i]) == upper:
                out.append(f)
        return out

    total = 0
    # backward sweeps, colours top .. 1: w, 1/d~, the coupling above, the upper lateral
    # couplings in, t out; the top colour has no upper neighbours and reads no index data
    for c in range(top, 0, -1):
        touched = set()
        for i in faces_of[c]:
            for z in range(L):
                touched |= {("w", i, z), ("dinv", i, z), ("t", i, z)}
                if z < L - 1:
                    touched.add(("vert_up", i, z))
                for f in lateral(i, upper=True):
                    touched.add(("lat", f, i, z))
        total += len(touched) * S + (0 if c == top else IDX * len(faces_of[c]))
    # fused colour 0 (both sweeps, t in registers) and the forward sweeps of colours 1 .. top
    for c in range(0, top + 1):
        touched = set()
        for i in faces_of[c]:
            for z in range(L):
                touched |= {("w", i, z), ("dinv", i, z), ("E", i, z), ("o", i, z)}
                touched.add(("w" if dot_is_input else "dw", i, z))
                if c != top:
                    touched.add(("v", i, z))  # the top colour's v is gathered by no one
                if c > 0:
                    touched.add(("t", i, z))
                if z > 0:
                    touched.add(("vert_down", i, z))
                if c == 0 and z < L - 1:
                    touched.add(("vert_up", i, z))
                for f in lateral(i, upper=(c == 0)):
                    touched.add(("lat", f, i, z))
        total += len(touched) * S + IDX * len(faces_of[c])
    return total


@pytest.mark.parametrize("n_points,L", [(60, 2), (150, 5), (400, 10), (300, 7)])
def test_split_apply_bytes_recount(n_points, L):
    mesh = meshgen.synthetic_mesh(n_points, seed=3)
    neigh = np.asarray(mesh.neigh, dtype=np.int64).reshape(-1, 3)
    F = neigh.shape[0]
    colors, nc = pb.face_colors(mesh.neigh)
    assert nc >= 3
    deg = int((neigh >= 0).sum())
    N = F * L
    nnz = L * (F + deg) + 2 * F * (L - 1)  # diagonal, lateral per layer, vertical
    f0 = int((colors == 0).sum())
    f_top = int((colors == colors.max()).sum())
    first = _launch_bytes(neigh, colors, L, dot_is_input=False)
    second = _launch_bytes(neigh, colors, L, dot_is_input=True)
    assert (first + second) % 2 == 0
    assert bench.split_apply_bytes(nnz, N, F, f0, L, S, f_top) == (first + second) // 2
    # the two applications differ by exactly the dot vector
    assert first - second == N * S


# The vector kernels of one iteration as bicg_run (chm_amd/csrc/krylov.hip) passes them:
# arrays read and written; with the split operator (no preconditioner) p^ is p and s^ is s.
def _vector_arrays(preconditioned):
    ph, sh = ("ph", "sh") if preconditioned else ("p", "s")
    kernels = [
        ({"r", "v", "p"}, {"p"}),                       # k_bicg_p_dev
        ({"v", "r"}, {"s"}),                            # k_bicg_s_dev
        ({ph, sh, "s", "t", "rh", "x"}, {"x", "r"}),    # k_bicg_xr_dev
    ]
    return sum(len(rd) + len(wr) for rd, wr in kernels)


@pytest.mark.parametrize("preconditioned", [False, True])
def test_bicg_vectors_recount(preconditioned):
    assert bench.bicg_vectors_per_iteration(preconditioned) == _vector_arrays(preconditioned)
