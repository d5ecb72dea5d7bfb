"""The batched parameter ensemble (include/pbsm3d_batch.h): members of uneven size, layer
count and parameters advanced by one pbsm3d_batch_step, against the same members run
one at a time through their own contexts and against the oracle."""
import numpy as np
import pytest

from chm_amd import abi, ensemble, meshgen
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _members(n, seed, lo, hi):
    return ensemble.draw_members(n, seed=seed, min_points=lo, max_points=hi)


@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_batch_matches_members_one_at_a_time(precision, oracle):
    """Six members (2k-9k points, four layer counts, perturbed z0 / settling / wind): three
    steps of the batch equal three steps of each member in its own context, both solved
    to 1e-12, within the solve tolerance; member 0 also against the oracle; every
    member's own residual meets the tolerance; the per-member state round-trips."""
    from chm_amd import pbsm3d as pb

    print(f"[case] batch vs members precision={precision}", flush=True)
    members = _members(6, 5, 2000, 9000)
    meshes = [meshgen.synthetic_mesh(m.n_points, seed=m.seed) for m in members]
    assert len({m.n_layer for m in members}) >= 3
    b = ensemble.build_batch(members, meshes, precision=precision, tol=1e-12)
    ctxs = [pb.Context(pb.face_geometry(mesh), ensemble.member_config(m, precision, tol=1e-12))
            for m, mesh in zip(members, meshes)]
    g0 = pb.face_geometry(meshes[0])
    orc = oracle.Oracle(g0, oracle.oracle_config(nLayer=members[0].n_layer, tol=1e-13, z0_snow=members[0].z0,
                                                 settling_scale=members[0].settling_scale))
    for step in range(3):
        fos = [ensemble.member_forcing(mesh, m, step) for m, mesh in zip(members, meshes)]
        got, st = b.step(fos, want_fields=True)
        for k, (ctx, fo) in enumerate(zip(ctxs, fos)):
            want, wst = ctx.step(fo, want_fields=True)
            assert st[k]["suspension_present"] == wst["suspension_present"]
            assert st[k]["deposition_present"] == wst["deposition_present"]
            assert st[k]["susp_resid"] <= 1e-12 and st[k]["dep_resid"] <= 1e-12, (k, st[k])
            np.testing.assert_allclose(got[k]["Qsalt"], want["Qsalt"], rtol=1e-13, atol=0)
            for f in ("conc", "Qsusp", "Qsubl", "qdep", "drift_mass", "sum_drift", "sum_subl"):
                assert rel_err(got[k][f], want[f]) < 1e-8, (step, k, f, rel_err(got[k][f], want[f]))
            np.testing.assert_array_equal(got[k]["pbsm_more_than_avail"], want["pbsm_more_than_avail"])
        og, _ = orc.step(fos[0])
        for f in ("conc", "Qsusp", "qdep", "drift_mass", "sum_drift"):
            assert rel_err(got[0][f], og[f]) < 1e-8, (step, f)
    info = b.info()
    assert info["unions"] == len({m.n_layer for m in members})
    print(f"[batch] {info}", flush=True)
    sd, ss = b.get_state(2)
    np.testing.assert_array_equal(sd, got[2]["sum_drift"])
    b.set_state(2, np.zeros_like(sd), np.zeros_like(ss))
    assert (b.get_state(2)[0] == 0).all()


@pytest.mark.parametrize("precision", [abi.FP64, abi.MIXED])
def test_batch_tiny_members_straddle_tiles(precision):
    """Members of 26, 63, 65 and 768 faces at one layer count form one union whose member
    runs start and end inside the 64-position tiles of the position order, so the
    element chunks of the per-member vector kernels are cut at member boundaries inside
    tiles: two steps of the batch equal the members' own contexts (1e-12 solves)."""
    from chm_amd import pbsm3d as pb

    print(f"[case] batch tiny members precision={precision}", flush=True)
    members = [ensemble.Member(seed=2000 + k, n_points=n, n_layer=10, z0=0.01, settling_scale=1.0, wind_scale=1.0)
               for k, n in enumerate((18, 38, 39, 392))]
    meshes = [meshgen.synthetic_mesh(n, seed=0) for n in (18, 38, 39, 392)]
    assert [m.n_faces for m in meshes] == [26, 63, 65, 768]
    b = ensemble.build_batch(members, meshes, precision=precision, tol=1e-12)
    ctxs = [pb.Context(pb.face_geometry(mesh), ensemble.member_config(m, precision, tol=1e-12))
            for m, mesh in zip(members, meshes)]
    assert b.info()["unions"] == 1
    for step in range(2):
        fos = [meshgen.forcing_c1(mesh.n_faces, seed=10 + step) for mesh in meshes]
        got, st = b.step(fos, want_fields=True)
        for k, (ctx, fo) in enumerate(zip(ctxs, fos)):
            want, wst = ctx.step(fo, want_fields=True)
            assert st[k]["suspension_present"] == wst["suspension_present"] == 1
            assert st[k]["susp_resid"] <= 1e-12 and st[k]["dep_resid"] <= 1e-12, (k, st[k])
            for f in ("conc", "Qsusp", "qdep", "drift_mass"):
                assert rel_err(got[k][f], want[f]) < 1e-8, (step, k, f, rel_err(got[k][f], want[f]))


def test_batch_member_without_suspension():
    """A calm member (no suspension, hence no deposition) beside windy ones keeps the
    -9999 defaults and x = 0, as a lone module would; the windy members still solve."""
    from chm_amd import pbsm3d as pb

    print("[case] batch calm member", flush=True)
    members = _members(3, 9, 2000, 4000)
    meshes = [meshgen.synthetic_mesh(m.n_points, seed=m.seed) for m in members]
    b = ensemble.build_batch(members, meshes, precision=abi.MIXED)
    fos = [ensemble.member_forcing(mesh, m, 0) for m, mesh in zip(members, meshes)]
    fos[1]["U_2m_above_srf"] = np.full(meshes[1].n_faces, 2.0)
    fos[1]["U_R"] = meshgen.u_ref_from_u2(fos[1]["U_2m_above_srf"], fos[1]["snowdepthavg"])
    got, st = b.step(fos, want_fields=True)
    assert st[1]["suspension_present"] == 0 and st[1]["deposition_present"] == 0
    assert (got[1]["conc"] == 0).all() and (got[1]["drift_mass"] == -9999).all()
    for k in (0, 2):
        assert st[k]["suspension_present"] == 1 and st[k]["susp_resid"] <= 1e-8 and st[k]["dep_resid"] <= 1e-8
        ctx = pb.Context(pb.face_geometry(meshes[k]), ensemble.member_config(members[k], abi.MIXED))
        want, _ = ctx.step(fos[k], want_fields=True)
        for f in ("Qsusp", "qdep", "drift_mass"):
            assert rel_err(got[k][f], want[f]) < 1e-4, (k, f)


def test_batch_c5_shape():
    """The ensemble workload at its full shape: 64 seeded members (25k-200k points,
    nLayer 5-10, z0 1e-4 - 1e-2, settling x 0.7-1.3, wind x 0.8-1.2), mixed precision
    with the learned start on the 10-layer members, two advected-forcing steps through
    pbsm3d_batch_step.  Every member meets ||b - A x|| <= 1e-8 ||b|| in both solves of both
    steps, and its fields agree with the same member run one at a time in its own
    context within the forward-error band of that tolerance."""
    import torch

    from chm_amd import pbsm3d as pb
    from chm_amd import warmstart

    print("[case] batch c5 shape: 64 members", flush=True)
    members = ensemble.draw_members(64, seed=0)
    meshes = ensemble.member_meshes(members, workers=16)
    print(f"[batch] {sum(m.n_faces for m in meshes)} faces in {len(meshes)} meshes", flush=True)
    k10 = next(k for k, m in enumerate(members) if m.n_layer == 10)
    g10 = pb.face_geometry(meshes[k10])
    X, Y = warmstart.collect_pairs(g10, [meshgen.forcing_advected(meshes[k10], s, seed=77) for s in (0, 5)])
    net, _ = warmstart.fit(X, Y, epochs=2, batch=8192)
    b = ensemble.build_batch(members, meshes, precision=abi.MIXED, warm_net=net)
    steps = 2
    r = ensemble.run_batch(b, members, meshes, steps=steps)
    print(f"[batch] {r['members']} members x {steps} steps: {r['member_steps_per_s']:.1f} member-steps/s, "
          f"{r['unions']} unions, max residual {r['max_resid']:.2e}, {b.info()}", flush=True)
    assert r["unions"] == len({m.n_layer for m in members})
    assert r["max_resid"] <= 1e-8
    worst = {f: 0.0 for f in ("Qsalt", "Qsusp", "Qsubl", "drift_mass", "sum_drift")}
    for k, (m, mesh) in enumerate(zip(members, meshes)):
        cfg = ensemble.member_config(m, abi.MIXED, warm=m.n_layer == 10)
        ctx = pb.Context(pb.face_geometry(mesh), cfg)
        if m.n_layer == 10:
            ctx.set_warm_start(net)
        for s, fo in enumerate(ensemble.member_forcing_device(mesh, m, steps, "cuda:0")):
            want, st = ctx.step(fo)
            assert st["susp_resid"] <= 1e-8 and st["dep_resid"] <= 1e-8
        for f in worst:
            got = r["outs"][k][f].cpu().numpy()
            worst[f] = max(worst[f], rel_err(got, want[f]))
        ctx.close()
        del ctx
    print("[batch] c5 batched vs one at a time, max over members of max|dx| / max|x|: "
          + ", ".join(f"{f} {v:.2e}" for f, v in worst.items()), flush=True)
    assert worst["Qsalt"] < 1e-13
    for f, v in worst.items():
        assert v < 5e-4, (f, v)
    torch.cuda.synchronize()


def test_batch_graph_replays_match_stream_launches():
    """The captured-graph replays of the BiCGStab iterations give the stream launches'
    results bit for bit, step after step (each step replays the two graphs many times,
    with the refinement loop's own launches and synchronisations between replays)."""
    print("[case] batch graphs vs stream launches", flush=True)
    members = _members(4, 11, 3000, 12000)
    meshes = [meshgen.synthetic_mesh(m.n_points, seed=m.seed) for m in members]
    runs = []
    for graphs in (True, False):
        b = ensemble.build_batch(members, meshes, precision=abi.MIXED)
        b.use_graphs(graphs)
        outs = []
        for step in range(3):
            fos = [ensemble.member_forcing(mesh, m, step) for m, mesh in zip(members, meshes)]
            got, st = b.step(fos, want_fields=True)
            outs.append((got, st))
        runs.append(outs)
        b.close()
    for (ga, sa), (gb, sb) in zip(*runs):
        for k in range(len(members)):
            assert sa[k]["susp_iters"] == sb[k]["susp_iters"] and sa[k]["dep_iters"] == sb[k]["dep_iters"]
            for f in ga[k]:
                np.testing.assert_array_equal(ga[k][f], gb[k][f], err_msg=f"member {k} {f}")


def test_batch_training_pairs_match_member_context():
    """pbsm3d_batch_training_pairs: three forcing draws of one 10-layer mesh solved as the
    members of one batch give each member the features and targets that
    pbsm3d_training_pairs gives for the same draw solved in a context of its own."""
    import torch

    from chm_amd import pbsm3d as pb

    print("[case] batch training pairs", flush=True)
    mesh = meshgen.synthetic_mesh(2500, seed=0)
    g = pb.face_geometry(mesh)
    F = mesh.n_faces
    cfg = pb.default_config(nLayer=10, krylov=abi.KRYLOV_BICGSTAB_SPLIT, tol=1e-12)
    fos = [meshgen.forcing_c2(F, seed=s) for s in (1, 2, 3)]
    b = pb.Batch([g] * 3, [cfg] * 3)
    _, st = b.step(fos)
    ctx = pb.Context(g, cfg)
    for m, fo in enumerate(fos):
        assert st[m]["suspension_present"] == 1
        Xb = torch.empty((F, pb.N_FEATURES), dtype=torch.float32, device="cuda:0")
        Yb = torch.empty((F, pb.N_TARGETS), dtype=torch.float32, device="cuda:0")
        b.training_pairs(m, Xb, Yb)
        ctx.step(fo)
        Xs, Ys = torch.empty_like(Xb), torch.empty_like(Yb)
        ctx.training_pairs(Xs, Ys)
        # features come from the per-face assembly (the same expressions in both); targets
        # from the two solves at 1e-12 (log concentrations, asinh of the deposition flux)
        assert torch.allclose(Xb, Xs, rtol=1e-6, atol=1e-6), (m, (Xb - Xs).abs().max().item())
        assert (Yb - Ys).abs().max().item() < 1e-4, (m, (Yb - Ys).abs().max().item())
    b.close()
    ctx.close()
