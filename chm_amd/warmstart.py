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
