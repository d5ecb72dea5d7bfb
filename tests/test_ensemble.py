"""Ensemble members run concurrently give the same fields as run one by one."""
import numpy as np
import pytest

from chm_amd import abi, ensemble, meshgen


def test_member_draws_follow_the_configuration():
    ms = ensemble.draw_members(64, seed=0)
    assert len(ms) == 64
    assert all(25000 <= m.n_points <= 200000 for m in ms)
    assert all(5 <= m.n_layer <= 10 for m in ms)
    assert all(1e-4 <= m.z0 <= 1e-2 for m in ms)
    assert all(0.7 <= m.settling_scale <= 1.3 and 0.8 <= m.wind_scale <= 1.2 for m in ms)
    assert len({m.n_layer for m in ms}) >= 4  # uneven layer counts


@pytest.mark.gpu
def test_concurrent_members_match_sequential():
    import torch

    print("[case] ensemble concurrent vs sequential", flush=True)
    members = ensemble.draw_members(6, seed=3, min_points=2000, max_points=8000)
    a = ensemble.build(members, precision=abi.FP64, krylov=abi.KRYLOV_GMRES)
    b = ensemble.build(members, precision=abi.FP64, krylov=abi.KRYLOV_GMRES)
    ra = ensemble.run(a, steps=3, workers=6)
    rb = ensemble.run(b, steps=3, workers=1)
    for (m, mesh, ca), (_, _, cb) in zip(a, b):
        sa, _ = ca.get_state()
        sb, _ = cb.get_state()
        np.testing.assert_array_equal(sa, sb)
    for x, y in zip(ra["per_member"], rb["per_member"]):
        assert x["susp_iters"] == y["susp_iters"] and x["last_resid"] <= 1e-8
    print("[ensemble]", {k: v for k, v in ra.items() if k != "per_member"}, flush=True)


def test_device_forcing_matches_host_forcing():
    """The ensemble's device-built forcing (torch, FP64) equals member_forcing's numpy
    fields step by step (same seeded advected modes) to rounding."""
    import torch

    m = ensemble.draw_members(2, seed=3, min_points=2000, max_points=4000)[1]
    mesh = meshgen.synthetic_mesh(m.n_points, seed=m.seed)
    dev = ensemble.member_forcing_device(mesh, m, 4, "cpu")
    for s in range(4):
        host = ensemble.member_forcing(mesh, m, s)
        assert set(dev[s]) == set(host)
        for k, v in host.items():
            np.testing.assert_allclose(dev[s][k].numpy(), v, rtol=1e-14, atol=0, err_msg=f"{k} step {s}")
    assert not torch.equal(dev[0]["vw_dir"], dev[3]["vw_dir"])  # the fields move with the steps
