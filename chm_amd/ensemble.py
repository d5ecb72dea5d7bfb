"""Parameter ensembles of the PBSM3D step on one GPU.

Members are independent PBSM3D instances (own mesh, layer count, roughness
length, settling-velocity scale and wind scale), each a context with its own
HIP stream.  A pool of host threads drives them concurrently -- the ctypes call
into pbsm3d_step releases the GIL -- so the kernels of different members overlap
on the GPU and small members fill the CUs that one member alone leaves idle.

Member draws follow the ensemble configuration: meshes of 25k-200k seeded
points (about 50k-400k triangles), 5-10 layers, z0 log-uniform in
[1e-4, 1e-2] m, settling velocity x U(0.7, 1.3), wind x U(0.8, 1.2).
Forcing is the advected smooth random field of the multi-step configuration.
"""
from __future__ import annotations

import concurrent.futures
import dataclasses
import math
import time

import numpy as np

from . import abi, meshgen
from . import pbsm3d as pb


@dataclasses.dataclass
class Member:
    seed: int
    n_points: int
    n_layer: int
    z0: float
    settling_scale: float
    wind_scale: float


def draw_members(n: int, seed: int = 0, min_points: int = 25000, max_points: int = 200000) -> list[Member]:
    u = meshgen.uniform(seed, 41, 5 * n)
    out = []
    for k in range(n):
        a, b, c, d, e = u[5 * k:5 * k + 5]
        pts = int(min_points + (max_points - min_points) * a)
        out.append(Member(seed=1000 + k, n_points=pts, n_layer=5 + int(6 * b) if b < 1 else 10,
                          z0=math.exp(math.log(1e-4) + (math.log(1e-2) - math.log(1e-4)) * c),
                          settling_scale=0.7 + 0.6 * d, wind_scale=0.8 + 0.4 * e))
    return out


def member_forcing(mesh: meshgen.Mesh, m: Member, step: int) -> dict:
    fo = meshgen.forcing_advected(mesh, step, seed=m.seed)
    fo["U_2m_above_srf"] = fo["U_2m_above_srf"] * m.wind_scale
    fo["U_R"] = fo["U_R"] * m.wind_scale
    return fo


def member_forcing_device(mesh: meshgen.Mesh, m: Member, steps: int, device: str) -> list[dict]:
    """member_forcing for steps 0 .. steps-1 as device tensors, built on the GPU: the
    step's wind speed and direction come from the same seeded advected fields
    (meshgen.SmoothField modes, evaluated with torch in FP64), the other fields from
    the step's host draw, which does not depend on the step."""
    import torch

    base = {k: torch.tensor(v, device=device) for k, v in member_forcing(mesh, m, 0).items()}
    c = mesh.vertex[mesh.elem].mean(axis=1)
    cx = torch.tensor(c[:, 0], device=device)
    cy = torch.tensor(c[:, 1], device=device)
    fields = (meshgen.SmoothField(m.seed, stream=meshgen.S_FIELD),
              meshgen.SmoothField(m.seed + 7919, stream=meshgen.S_FIELD))

    def value(sf, xs, ys):
        acc = torch.zeros_like(xs)
        for k in range(sf.n):
            acc += torch.cos(float(sf.kx[k]) * xs + float(sf.ky[k]) * ys + float(sf.phase[k]))
        return acc * math.sqrt(2.0 / sf.n)

    z0 = meshgen.Z0_SNOW
    lift = torch.log((meshgen.Z_U_R - (base["snowdepthavg"] + z0)) / z0) / math.log((2.0 - z0) / z0)
    out = []
    for s in range(steps):
        xs, ys = cx - 150.0 * s, cy - 40.0 * s  # meshgen.forcing_advected's advection
        u2 = torch.clamp(10.0 * torch.exp(0.25 * value(fields[0], xs, ys)), 4.0, 20.0)
        f = dict(base)
        f["U_2m_above_srf"] = u2 * m.wind_scale
        f["U_R"] = u2 * lift * m.wind_scale
        f["vw_dir"] = 270.0 + 20.0 * value(fields[1], xs, ys)
        out.append(f)
    return out


def build(members: list[Member], precision=abi.MIXED, krylov=abi.KRYLOV_BICGSTAB_SPLIT, warm_net=None,
          device: int = 0):
    """Meshes, geometry and contexts of all members (set-up, not timed)."""
    built = []
    for m in members:
        mesh = meshgen.synthetic_mesh(m.n_points, seed=m.seed)
        geom = pb.face_geometry(mesh)
        cfg = pb.default_config(nLayer=m.n_layer, precision=precision, krylov=krylov, z0_snow=m.z0,
                                settling_scale=m.settling_scale)
        use_warm = warm_net is not None and m.n_layer == 10
        if use_warm:
            cfg.warm_start = abi.WS_LEARNED
        ctx = pb.Context(geom, cfg, device=device)
        if use_warm:
            ctx.set_warm_start(warm_net)
        built.append((m, mesh, ctx))
    return built


def _mesh_of(args):
    n_points, seed = args
    return meshgen.synthetic_mesh(n_points, seed=seed)


def member_meshes(members: list[Member], workers: int = 16) -> list[meshgen.Mesh]:
    """The members' seeded meshes, generated in a process pool (set-up, not timed)."""
    jobs = [(m.n_points, m.seed) for m in members]
    if workers <= 1 or len(jobs) <= 1:
        return [_mesh_of(j) for j in jobs]
    import multiprocessing as mp

    with concurrent.futures.ProcessPoolExecutor(max_workers=min(workers, len(jobs)),
                                                mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_mesh_of, jobs))


def member_config(m: Member, precision=abi.MIXED, krylov=abi.KRYLOV_BICGSTAB_SPLIT, warm: bool = False,
                  **overrides) -> abi.Config:
    cfg = pb.default_config(nLayer=m.n_layer, precision=precision, krylov=krylov, z0_snow=m.z0,
                            settling_scale=m.settling_scale, **overrides)
    if warm:
        cfg.warm_start = abi.WS_LEARNED
    return cfg


def build_batch(members: list[Member], meshes, precision=abi.MIXED, warm_net=None, device: int = 0, **overrides):
    """All members in one pbsm3d_batch (members of equal nLayer share one union system;
    the learned start applies to the 10-layer members).  Set-up, not timed."""
    geoms = [pb.face_geometry(mesh) for mesh in meshes]
    cfgs = [member_config(m, precision, abi.KRYLOV_BICGSTAB_SPLIT, warm_net is not None, **overrides)
            for m in members]
    b = pb.Batch(geoms, cfgs, device=device)
    if warm_net is not None:
        b.set_warm_start(warm_net)
    return b


def run_batch(batch, members: list[Member], meshes, steps: int, device: int = 0) -> dict:
    """`steps` timesteps of every member through one pbsm3d_batch_step call per step,
    forcing and outputs resident in HBM."""
    import torch

    dev = f"cuda:{device}"
    forcings = [member_forcing_device(mesh, m, steps, dev) for m, mesh in zip(members, meshes)]
    outs = [{k: torch.zeros(mesh.n_faces, dtype=torch.float64, device=dev) for k in abi.OUT_FIELDS}
            for mesh in meshes]
    torch.cuda.synchronize()
    iters = [0] * len(members)
    worst = 0.0
    t0 = time.perf_counter()
    for s in range(steps):
        _, st = batch.step([f[s] for f in forcings], outs=outs, on_device=True)
        for k, x in enumerate(st):
            iters[k] += x["susp_iters"]
            worst = max(worst, x["susp_resid"], x["dep_resid"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = sum(mesh.n_faces * m.n_layer for m, mesh in zip(members, meshes))
    return {"members": len(members), "steps": steps, "wall_s": wall, "member_steps_per_s": len(members) * steps / wall,
            "cell_steps_per_s": cells * steps / wall, "max_resid": worst, "susp_iters": iters,
            "unions": batch.info()["unions"], "outs": outs}


def run(built, steps: int, workers: int = 16, on_device: bool = True, device: int = 0) -> dict:
    """Run every member for `steps` timesteps, members in parallel on `workers` host threads."""
    import torch

    dev = f"cuda:{device}"
    forcings = []
    for m, mesh, ctx in built:
        if on_device:
            forcings.append(member_forcing_device(mesh, m, steps, dev))
        else:
            forcings.append([member_forcing(mesh, m, s) for s in range(steps)])
    outs = [{k: torch.zeros(ctx.F, dtype=torch.float64, device=dev) for k in abi.OUT_FIELDS} if on_device else None
            for _, _, ctx in built]
    torch.cuda.synchronize()

    def one(k):
        m, mesh, ctx = built[k]
        t0 = time.perf_counter()
        iters, last = 0, None
        for s in range(steps):
            _, st = ctx.step(forcings[k][s], outs=outs[k], on_device=on_device)
            iters += st["susp_iters"]
            last = st
        return {"member": k, "faces": ctx.F, "layers": ctx.L, "seconds": time.perf_counter() - t0,
                "susp_iters": iters, "last_resid": last["susp_resid"]}

    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        res = list(ex.map(one, range(len(built))))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = sum(r["faces"] * r["layers"] for r in res)
    return {"members": len(built), "steps": steps, "wall_s": wall, "member_steps_per_s": len(built) * steps / wall,
            "cell_steps_per_s": cells * steps / wall, "per_member": res}
