"""Learned warm start: the network kernels against the FP64 reference MLP, the
training loop, and what the trained network gains over a cold start."""
import numpy as np
import pytest

from chm_amd import abi, meshgen
from oracle import mlp_oracle
from tests.conftest import rel_err


def _norm(rng):
    return (rng.normal(0, 1, 16), rng.uniform(0.5, 2, 16), rng.normal(0, 1, 11), rng.uniform(0.5, 2, 11))


def test_reference_mlp_gradient_by_finite_differences():
    """Pins the FP64 reference: its analytic gradient equals central differences."""
    rng = np.random.default_rng(0)
    p = (rng.normal(0, 0.2, mlp_oracle.param_count())).astype(np.float32)
    p[-16 + 11:] = 0
    X = rng.normal(0, 1, (8, 16))
    Y = rng.normal(0, 1, (8, 11))
    norm = _norm(rng)
    loss, g = mlp_oracle.loss_and_grad(p, X, Y, norm, bf16=False)
    p64 = p.astype(np.float64)
    for k in rng.choice(p.size - 16, 40, replace=False):
        e = 1e-4

        def f(v):
            q = p64.copy()
            q[k] = v
            return mlp_oracle.loss_and_grad(q, X, Y, norm, bf16=False)[0]

        fd = (f(p64[k] + e) - f(p64[k] - e)) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265], dtype=np.float32)
    r = mlp_oracle.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12


@pytest.mark.gpu
def test_network_gradients_match_reference():
    from chm_amd import pbsm3d as pb

    print("[case] mlp gradients", flush=True)
    rng = np.random.default_rng(1)
    net = pb.Mlp(seed=3)
    p = net.get_params()
    assert p.size == mlp_oracle.param_count() == pb.Mlp.param_count()
    assert np.all(p[-5:] == 0)
    norm = _norm(rng)
    net.set_normalization(*norm)
    n = 777  # not a tile multiple: exercises the edge handling
    X = rng.normal(0, 1.5, (n, 16)).astype(np.float32) * norm[1] + norm[0]
    Y = rng.normal(0, 1, (n, 11)).astype(np.float32)
    loss, g = net.gradients(X, Y)
    rloss, rg = mlp_oracle.loss_and_grad(p, X, Y, norm, bf16=True)
    assert abs(loss - rloss) <= 1e-3 * rloss
    o = 0
    for (r, c) in mlp_oracle.SHAPES:
        for size in (r * c, r):
            a, b = g[o:o + size], rg[o:o + size]
            assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7, (o, size)
            o += size
    # forward agrees with the reference
    import torch

    Xd = torch.tensor(X, dtype=torch.float32, device="cuda:0")
    P = torch.empty((n, 11), dtype=torch.float32, device="cuda:0")
    net.predict(Xd, P)
    ref = mlp_oracle.predict(p, X, norm, bf16=True)
    assert np.abs(P.cpu().numpy() - ref).max() <= 1e-2 * np.abs(ref).max()
    # the fused one-launch inference kernel and the layer-by-layer GEMMs agree bit for bit
    # (same MFMA operand layout, K order, epilogue and BF16 rounding), also on a ragged n
    P1 = torch.empty_like(P)
    P2 = torch.empty_like(P)
    net.predict(Xd, P1, path=1)
    net.predict(Xd, P2, path=2)
    assert torch.equal(P1, P2)
    assert torch.equal(P, P1)


@pytest.mark.gpu
def test_training_reduces_loss_on_a_known_function():
    """Fits y = smooth function of x; the loss must fall by an order of magnitude."""
    import torch

    from chm_amd import pbsm3d as pb

    print("[case] mlp training", flush=True)
    g = torch.Generator(device="cuda:0").manual_seed(0)
    X = torch.randn((65536, 16), device="cuda:0", generator=g)
    W = torch.randn((16, 11), device="cuda:0", generator=g) / 4
    Y = torch.tanh(X @ W) + 0.1 * (X[:, :11] ** 2)
    net = pb.Mlp(seed=1)
    net.set_normalization(np.zeros(16), np.ones(16), np.zeros(11), np.ones(11))
    l0, _ = net.gradients(X[:4096].cpu().numpy(), Y[:4096].cpu().numpy())
    loss = net.train(X, Y, epochs=8, batch=2048, lr=2e-3, seed=2)
    assert loss < 0.1 * l0, (l0, loss)
    # training is deterministic (split-K partial products summed in a fixed order): a second
    # network from the same seed on the same data ends with bitwise identical parameters
    net2 = pb.Mlp(seed=1)
    net2.set_normalization(np.zeros(16), np.ones(16), np.zeros(11), np.ones(11))
    loss2 = net2.train(X, Y, epochs=8, batch=8192, lr=2e-3, seed=2)
    net3 = pb.Mlp(seed=1)
    net3.set_normalization(np.zeros(16), np.ones(16), np.zeros(11), np.ones(11))
    loss3 = net3.train(X, Y, epochs=8, batch=8192, lr=2e-3, seed=2)
    assert loss2 == loss3
    np.testing.assert_array_equal(net2.get_params(), net3.get_params())


@pytest.fixture(scope="module")
def trained_small():
    """The 2,500-point mesh at 10 layers and a network trained on 1,024 FP64 solves of
    seeded forcing draws (seeds 10000 ..), collected 64 at a time through the batch."""
    from chm_amd import pbsm3d as pb
    from chm_amd import warmstart

    mesh = meshgen.synthetic_mesh(2500, seed=0)
    geom = pb.face_geometry(mesh)
    F = mesh.n_faces
    X, Y, info = warmstart.collect_pairs_batch(geom, lambda s: meshgen.forcing_c2(F, seed=s), range(10000, 11024))
    net, tinfo = warmstart.fit(X, Y, epochs=4, batch=16384, lr=2e-3)
    print("[warm] collect", info, "train", tinfo, flush=True)
    assert info["solves"] == 1024
    return mesh, geom, net


@pytest.mark.gpu
def test_warm_start_gain_over_held_out_draws(trained_small):
    """Cold against learned start on twelve held-out draws (seeds 7 .. 18, disjoint from the
    training draws) with the FP64 split BiCGStab at 1e-8: both reach the tolerance on
    every draw, the learned start begins below half the cold residual on every draw,
    never falls back to x0 = 0, and the total suspension iterations over the draws are
    fewer than the cold start's; the fields of the two starts agree within the solve
    tolerance's forward-error band."""
    from chm_amd import pbsm3d as pb
    from chm_amd import warmstart

    print("[case] warm start over held-out draws", flush=True)
    mesh, geom, net = trained_small
    F = mesh.n_faces
    held = [meshgen.forcing_c2(F, seed=s) for s in range(7, 19)]
    cfg = pb.default_config(nLayer=10, krylov=abi.KRYLOV_BICGSTAB_SPLIT, tol=1e-8)
    res = warmstart.compare_starts(geom, net, held, cfg, warmup_draw=meshgen.forcing_c2(F, seed=99))
    print("[warm] iters cold", res["cold"]["susp_iters"], "learned", res["learned"]["susp_iters"],
          "resid0", res["learned"]["resid0"], "ratio", res["iters_ratio_learned_to_cold"], flush=True)
    assert all(r < 0.5 for r in res["learned"]["resid0"]), res["learned"]["resid0"]
    assert not any(res["learned"]["cold_fallback"])
    assert sum(res["learned"]["susp_iters"]) < sum(res["cold"]["susp_iters"])
    cold = pb.Context(geom, cfg)
    wcfg = pb.default_config(nLayer=10, krylov=abi.KRYLOV_BICGSTAB_SPLIT, tol=1e-8, warm_start=abi.WS_LEARNED)
    warm = pb.Context(geom, wcfg)
    warm.set_warm_start(net)
    oc, _ = cold.step(held[0], want_fields=True)
    ow, sw = warm.step(held[0], want_fields=True)
    assert sw["susp_resid"] <= 1e-8 and sw["dep_resid"] <= 1e-8
    for k in ("Qsusp", "drift_mass"):
        d = np.abs(ow[k] - oc[k]).max() / np.abs(oc[k]).max()
        assert d < 1e-4, (k, d)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_learned_coarse_correction(trained_small, precision):
    """The learned coarse correction on 32 tiles: the tiles are the face-index ranges
    [a F / 32, (a + 1) F / 32), the basis is the learned start's x0 (non-negative, zero
    where no suspension is predicted), the Galerkin matrix W^T A W equals the one numpy
    builds from the exported system, the solve still reaches 1e-12 and its fields agree
    with those of the same solve without the correction (both at 1e-12)."""
    import scipy.sparse as sp

    from chm_amd import pbsm3d as pb

    print(f"[case] learned coarse correction precision={precision}", flush=True)
    mesh, geom, net = trained_small
    F, L, T = mesh.n_faces, 10, 32
    N = F * L
    fo = meshgen.forcing_c2(F, seed=7)
    cfg = pb.default_config(nLayer=L, krylov=abi.KRYLOV_BICGSTAB_SPLIT, tol=1e-12, precision=precision,
                            warm_start=abi.WS_LEARNED, coarse_tiles=T)
    ctx = pb.Context(geom, cfg)
    ctx.set_warm_start(net)
    out, st = ctx.step(fo, want_fields=True)
    print("[coarse] stats", {k: st[k] for k in ("susp_iters", "susp_resid0", "susp_resid_coarse", "susp_resid")},
          flush=True)
    assert st["susp_resid"] <= 1e-12 and st["dep_resid"] <= 1e-12
    assert 0 < st["susp_resid_coarse"] < 1
    tile, w, G = ctx.coarse_system()
    off = (np.arange(T + 1) * F) // T
    np.testing.assert_array_equal(tile, np.repeat(np.arange(T), np.diff(off)))
    assert (w >= 0).all() and (w > 0).any()
    rp, col, val, _ = ctx.system(0)
    A = sp.csr_matrix((val, col, rp), shape=(N, N))
    rows = np.arange(N)
    W = sp.csr_matrix((w, (rows, tile[rows % F])), shape=(N, T))
    Gn = (W.T @ A @ W).toarray()
    empty = np.flatnonzero(np.diag(Gn) == 0)
    Gn[empty, empty] = 1.0
    assert np.abs(G - Gn).max() <= 1e-12 * np.abs(Gn).max(), np.abs(G - Gn).max()
    ref = pb.Context(geom, pb.default_config(nLayer=L, krylov=abi.KRYLOV_BICGSTAB_SPLIT, tol=1e-12,
                                             precision=precision, warm_start=abi.WS_LEARNED))
    ref.set_warm_start(net)
    o2, _ = ref.step(fo, want_fields=True)
    for k in ("conc", "Qsusp", "qdep", "drift_mass"):
        assert rel_err(out[k], o2[k]) < 1e-8, (k, rel_err(out[k], o2[k]))


def _tile_degree(neigh, T):
    """Largest number of tiles (itself included) that one face-index tile touches."""
    F = neigh.shape[0]
    tof = np.repeat(np.arange(T), np.diff((np.arange(T + 1) * F) // T))
    i, j = np.nonzero(neigh >= 0)
    pairs = np.unique(np.stack([tof[i], tof[neigh[i, j]]], axis=1), axis=0)
    others = pairs[pairs[:, 0] != pairs[:, 1], 0]
    return int((np.bincount(others, minlength=T) + 1).max())


@pytest.mark.gpu
def test_coarse_tiles_checked_at_create():
    """The coarse tiles and their adjacency are built when the context is created, so a tile
    count the mesh cannot carry fails at pbsm3d_create and not at the first step: on the
    Hilbert-ordered 2,500-point mesh each of 32 face-index tiles touches at most 15 others
    and the context is made; with the faces renumbered by a random permutation the same 32
    tiles touch nearly all others and creation fails with PBSM3D_EINVAL."""
    from chm_amd import pbsm3d as pb

    print("[case] coarse tiles checked at create", flush=True)
    T = 32
    mesh = meshgen.synthetic_mesh(2500, seed=0)
    cfg = pb.default_config(nLayer=10, krylov=abi.KRYLOV_BICGSTAB_SPLIT, warm_start=abi.WS_LEARNED, coarse_tiles=T)
    assert _tile_degree(mesh.neigh, T) <= 16
    pb.Context(pb.face_geometry(mesh), cfg).close()
    F = mesh.n_faces
    perm = np.random.default_rng(3).permutation(F)
    inv = np.empty(F, dtype=np.int64)
    inv[perm] = np.arange(F)
    nb = mesh.neigh[perm]
    nb = np.where(nb >= 0, inv[np.maximum(nb, 0)], -1).astype(np.int32)
    shuffled = meshgen.Mesh(mesh.vertex, mesh.elem[perm], nb, None if mesh.area is None else mesh.area[perm])
    assert _tile_degree(shuffled.neigh, T) > 16
    with pytest.raises(pb.Pbsm3dError) as e:
        pb.Context(pb.face_geometry(shuffled), cfg)
    assert e.value.code == abi.PBSM3D_EINVAL and "coarse tile" in str(e.value)
