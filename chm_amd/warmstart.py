"""Training pipeline of the learned warm start.

(features, solution) pairs come from the FP64 GPU solver itself: each forcing
draw is one pbsm3d_step, after which pbsm3d_training_pairs writes the per-face
features and the solved 10-layer concentration profile + deposition flux.
The network is then trained on the GPU by the library's own BF16-MFMA forward
and backward kernels (pbsm3d_mlp_train).
"""
from __future__ import annotations

import time

import numpy as np

from . import abi
from . import pbsm3d as pb


def collect_pairs(geom: abi.Geometry, forcings, cfg: abi.Config | None = None, dt: int = 900, device: int = 0):
    """Run one FP64 step per forcing draw; return float32 device tensors X [n x 16], Y [n x 11]."""
    import torch

    cfg = cfg if cfg is not None else pb.default_config(nLayer=10)
    if cfg.nLayer != 10:
        raise ValueError("the warm-start network predicts 10 layers")
    ctx = pb.Context(geom, cfg, device=device)
    F = geom.n_faces
    xs, ys = [], []
    dev = f"cuda:{device}"
    for fo in forcings:
        _, st = ctx.step(fo)
        if not st["suspension_present"]:
            continue
        X = torch.empty((F, pb.N_FEATURES), dtype=torch.float32, device=dev)
        Y = torch.empty((F, pb.N_TARGETS), dtype=torch.float32, device=dev)
        ctx.training_pairs(X, Y)
        xs.append(X)
        ys.append(Y)
    ctx.close()
    return torch.cat(xs), torch.cat(ys)


def collect_pairs_batch(geom: abi.Geometry, draw, seeds, members: int = 64, cfg: abi.Config | None = None,
                        dt: int = 900, device: int = 0):
    """FP64 solves of many forcing draws of one mesh, `members` draws at a time as the
    members of one pbsm3d_batch (one union system; the split D-ILU BiCGStab, to the
    config's tolerance).  draw(seed) returns a forcing dict; draws without suspension
    are skipped.  Returns float32 device tensors X [n x 16], Y [n x 11] and a summary."""
    import torch

    cfg = cfg if cfg is not None else pb.default_config(nLayer=10, krylov=abi.KRYLOV_BICGSTAB_SPLIT)
    if cfg.nLayer != 10:
        raise ValueError("the warm-start network predicts 10 layers")
    seeds = list(seeds)
    F = geom.n_faces
    dev = f"cuda:{device}"
    nb = min(members, len(seeds))
    b = pb.Batch([geom] * nb, [cfg] * nb, device=device)
    X = torch.empty((len(seeds) * F, pb.N_FEATURES), dtype=torch.float32, device=dev)
    Y = torch.empty((len(seeds) * F, pb.N_TARGETS), dtype=torch.float32, device=dev)
    rows, solves, iters, t0 = 0, 0, 0, time.perf_counter()
    try:
        for k0 in range(0, len(seeds), nb):
            chunk = seeds[k0:k0 + nb]
            fos = [draw(s) for s in chunk]
            fos += [fos[-1]] * (nb - len(fos))  # a short last chunk repeats its last draw
            _, st = b.step(fos, dt=dt)
            for m in range(len(chunk)):
                if not st[m]["suspension_present"]:
                    continue
                b.training_pairs(m, X[rows:rows + F], Y[rows:rows + F])
                rows += F
                solves += 1
                iters += st[m]["susp_iters"]
    finally:
        b.close()
    torch.cuda.synchronize()
    return X[:rows], Y[:rows], {"solves": solves, "draws": len(seeds), "members_per_batch": nb,
                                "collect_s": round(time.perf_counter() - t0, 2),
                                "mean_susp_iters": round(iters / max(solves, 1), 2)}


def fit(X, Y, epochs: int = 6, batch: int = 4096, lr: float = 2e-3, seed: int = 0, device: int = 0,
        net: "pb.Mlp | None" = None) -> tuple["pb.Mlp", dict]:
    """Normalise with the data's mean / std and train; returns (network, info)."""
    import torch

    net = net if net is not None else pb.Mlp(device=device, seed=seed)
    xm, xsd = X.mean(0), X.std(0).clamp_min(1e-6)
    ym, ysd = Y.mean(0), Y.std(0).clamp_min(1e-6)
    net.set_normalization(xm.cpu().numpy(), xsd.cpu().numpy(), ym.cpu().numpy(), ysd.cpu().numpy())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = net.train(X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    secs = time.perf_counter() - t0
    flops = 6.0 * (16 * 128 + 128 * 128 * 2 + 128 * 16) * epochs * ((X.shape[0] + batch - 1) // batch) * batch
    return net, {"final_loss": loss, "train_s": secs, "samples": int(X.shape[0]), "epochs": epochs,
                 "train_tflops": flops / secs / 1e12}


def relative_error(net: "pb.Mlp", X, Y) -> np.ndarray:
    """Per-output RMS error of the prediction, in target units."""
    import torch

    P = torch.empty_like(Y)
    net.predict(X, P)
    return torch.sqrt(((P - Y) ** 2).mean(0)).cpu().numpy()


def compare_starts(geom: abi.Geometry, net: "pb.Mlp", draws, cfg: abi.Config, warmup_draw=None,
                   coarse_tiles: int = 0, device: int = 0) -> dict:
    """Cold start (x0 = 0) against the learned start on the same forcing draws with the
    solver of `cfg` (and, with coarse_tiles > 0, the learned start plus the learned coarse
    correction on that many tiles): suspension iterations, ||b - A x0|| / ||b||, whether
    the learned start fell back to x0 = 0, the residual after the first coarse correction
    and the step time with forcing and outputs resident on the device (no PCIe in the
    timing).  Each context first steps once on `warmup_draw` so its launch-shape tuning
    is not timed."""
    import torch

    dev = f"cuda:{device}"
    F = geom.n_faces
    dd = [{k: torch.as_tensor(v, device=dev) for k, v in d.items()} for d in draws]
    outs = {k: torch.zeros(F, dtype=torch.float64, device=dev) for k in abi.OUT_FIELDS}
    arms = [("cold", abi.WS_ZERO, 0), ("learned", abi.WS_LEARNED, 0)]
    if coarse_tiles > 0:
        arms.append(("learned_coarse", abi.WS_LEARNED, coarse_tiles))
    res = {}
    for label, ws, tiles in arms:
        c = abi.Config.from_buffer_copy(cfg)
        c.warm_start = ws
        c.coarse_tiles = tiles
        ctx = pb.Context(geom, c, device=device)
        if ws == abi.WS_LEARNED:
            ctx.set_warm_start(net)
        w0 = warmup_draw if warmup_draw is not None else draws[0]
        ctx.step({k: torch.as_tensor(v, device=dev) for k, v in w0.items()}, outs=outs, on_device=True)
        its, dits, r0, rc, fb, ms = [], [], [], [], [], []
        for d in dd:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, st = ctx.step(d, outs=outs, on_device=True)
            torch.cuda.synchronize()
            ms.append(round(1000.0 * (time.perf_counter() - t0), 3))
            if st["susp_resid"] > c.tol or st["dep_resid"] > c.tol:
                raise RuntimeError(f"{label} start missed the tolerance: {st}")
            its.append(int(st["susp_iters"]))
            dits.append(int(st["dep_iters"]))
            r0.append(round(float(st["susp_resid0"]), 4))
            rc.append(round(float(st["susp_resid_coarse"]), 4))
            fb.append(int(st["susp_cold_fallback"]))
        ctx.close()
        res[label] = {"susp_iters": its, "dep_iters": dits, "resid0": r0, "cold_fallback": fb, "ms_per_step": ms}
        if tiles:
            res[label]["resid_after_coarse"] = rc
            res[label]["coarse_tiles"] = tiles
    res["draws"] = len(draws)
    c_it, c_ms = sum(res["cold"]["susp_iters"]), sum(res["cold"]["ms_per_step"])
    for label, _, _ in arms[1:]:
        its = res[label]["susp_iters"]
        res[f"iters_ratio_{label}_to_cold"] = round(sum(its) / max(1, c_it), 4)
        res[f"time_ratio_{label}_to_cold"] = round(sum(res[label]["ms_per_step"]) / max(1e-9, c_ms), 4)
        res[f"draws_{label}_fewer_iters"] = sum(a < b for a, b in zip(its, res["cold"]["susp_iters"]))
        res[f"draws_{label}_more_iters"] = sum(a > b for a, b in zip(its, res["cold"]["susp_iters"]))
    return res


def learned_start_study(n_train: int = 20000, held_seeds=range(100, 116), epochs: int = 4, batch: int = 32768,
                        lr: float = 2e-3, points: int = 2500, coarse_tiles: int = 32,
                        device: int = 0) -> tuple["pb.Mlp", dict]:
    """The learned warm start at training scale: the 2,500-point mesh (about 5k triangles)
    at 10 layers, n_train FP64 solves of seeded forcing draws (per-face lognormal 2 m wind
    in [4, 20] m/s, direction N(270, 20) deg, T ~ U(-25, -2) C; seeds 10000 ..) collected
    64 at a time through the batch, the network trained on the GPU, then cold against
    learned start on held-out draws (seeds disjoint from the training ones) with FP64
    GMRES, the FP64 and the mixed split BiCGStab at tolerance 1e-8 (the split solvers also
    with the learned coarse correction on coarse_tiles tiles)."""
    from . import meshgen

    mesh = meshgen.synthetic_mesh(points, seed=0)
    geom = pb.face_geometry(mesh)
    F = mesh.n_faces
    seeds = range(10000, 10000 + n_train)
    X, Y, cinfo = collect_pairs_batch(geom, lambda s: meshgen.forcing_c2(F, seed=s), seeds, device=device)
    net, tinfo = fit(X, Y, epochs=epochs, batch=batch, lr=lr, device=device)
    del X, Y
    held = [meshgen.forcing_c2(F, seed=s) for s in held_seeds]
    warm = meshgen.forcing_c2(F, seed=99)
    out = {"mesh_faces": F, "layers": 10, "collect": cinfo, "train": tinfo, "held_out_seeds": list(held_seeds)}
    for label, prec, kry in (("fp64_gmres", abi.FP64, abi.KRYLOV_GMRES),
                             ("fp64_split", abi.FP64, abi.KRYLOV_BICGSTAB_SPLIT),
                             ("mixed_split", abi.MIXED, abi.KRYLOV_BICGSTAB_SPLIT)):
        cfg = pb.default_config(nLayer=10, precision=prec, krylov=kry, tol=1e-8)
        out[label] = compare_starts(geom, net, held, cfg, warmup_draw=warm, device=device,
                                    coarse_tiles=coarse_tiles if kry == abi.KRYLOV_BICGSTAB_SPLIT else 0)
    return net, out
