"""pbsm3d_wind_run_noalloc, the allocation-free variant of pbsm3d_wind_create +
pbsm3d_wind_run: through the C ABI on device arrays, equal bit for bit to the existing
entry point, on two calls in a row (the second with a smaller mesh, so the shared
scratch array holds the first call's larger contents), with and without vegetation,
and with the point-mode (no smoothing) path; sizes beyond the scratch and
out-of-range neighbour indices are refused."""
import numpy as np
import pytest

from chm_amd import abi, meshgen

pytestmark = pytest.mark.gpu


def _case(mesh, seed, veg):
    rng = np.random.default_rng(seed)
    F = mesh.n_faces
    c = mesh.vertex[mesh.elem].mean(axis=1)
    cx, cy = np.ascontiguousarray(c[:, 0]), np.ascontiguousarray(c[:, 1])
    can = rng.uniform(0.0, 6.0, F) if veg else None
    lai = rng.uniform(0.0, 3.0, F) if veg else None
    U_R = rng.uniform(4.0, 25.0, F)
    sd = rng.uniform(0.0, 2.5, F)
    return cx, cy, can, lai, U_R, sd


@pytest.mark.parametrize("veg,smooth", [(False, 1), (True, 1), (True, 0)])
def test_noalloc_matches_wind_run(veg, smooth):
    import torch

    from chm_amd import pbsm3d as pb

    print(f"[case] wind noalloc veg={veg} smooth={smooth}", flush=True)
    cfg = pb.wind_config(smooth=smooth)
    for n_points, seed in ((40000, 1), (3000, 2)):  # the second call is smaller
        mesh = meshgen.synthetic_mesh(n_points, seed=seed)
        cx, cy, can, lai, U_R, sd = _case(mesh, seed, veg)
        want = pb.Wind(mesh.neigh, cx, cy, can, lai, cfg).run(U_R, sd)

        def dev(a, dtype=torch.float64):
            return None if a is None else torch.tensor(a, dtype=dtype, device="cuda:0")

        out = torch.full((mesh.n_faces,), -1.0, dtype=torch.float64, device="cuda:0")
        pb.wind_run_noalloc(dev(mesh.neigh, torch.int32), dev(cx), dev(cy), dev(U_R), out, snowdepthavg=dev(sd),
                            canopy_height=dev(can), lai=dev(lai), cfg=cfg)
        np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_noalloc_refuses_bad_input():
    import torch

    from chm_amd import pbsm3d as pb

    print("[case] wind noalloc refusals", flush=True)
    mesh = meshgen.synthetic_mesh(2000, seed=3)
    cx, cy, _, _, U_R, sd = _case(mesh, 3, False)
    d = {k: torch.tensor(v, device="cuda:0") for k, v in (("cx", cx), ("cy", cy), ("u", U_R), ("sd", sd))}
    out = torch.zeros(mesh.n_faces, dtype=torch.float64, device="cuda:0")
    nb = mesh.neigh.copy()
    nb[5, 1] = mesh.n_faces + 7  # out of range: checked on the device, the read is skipped
    with pytest.raises(pb.Pbsm3dError) as e:
        pb.wind_run_noalloc(torch.tensor(nb, device="cuda:0"), d["cx"], d["cy"], d["u"], out, snowdepthavg=d["sd"])
    assert e.value.code == abi.PBSM3D_EINVAL
    # the same call with valid neighbours succeeds afterwards (the flag is cleared per call)
    pb.wind_run_noalloc(torch.tensor(mesh.neigh, device="cuda:0"), d["cx"], d["cy"], d["u"], out,
                        snowdepthavg=d["sd"])
    assert (out.cpu().numpy() >= 0.1).all()
    # more faces than the scratch holds: refused before any device work
    import ctypes as C

    lib = pb.lib()
    big = pb.PBSM3D_WIND_NOALLOC_MAX_FACES + 1
    nbd = torch.tensor(mesh.neigh, device="cuda:0")
    ptr = [C.c_void_p(t.data_ptr()) for t in (nbd, d["cx"], d["cy"])]
    rc = lib.pbsm3d_wind_run_noalloc(0, big, *ptr, None, None, C.byref(pb.wind_config()), C.c_void_p(d["u"].data_ptr()),
                                     None, C.c_void_p(out.data_ptr()))
    assert rc == abi.PBSM3D_EINVAL
    assert "scratch" in lib.pbsm3d_wind_last_error(None).decode()
