#!/usr/bin/env python3
"""PBSM3D blowing-snow step throughput on one MI355X.

Workload: the production-size synthetic case -- 500,000 seeded uniform points
in a 10 x 10 km square -> Delaunay (999,968 triangles), 10 suspension layers
(9,999,680 unknowns), per-face lognormal 2 m wind in [4, 20] m/s, direction
N(270, 20) deg, T ~ U(-25, -2) C; one 900 s step solved to relative residual
1e-8.  Forcing and outputs stay resident in HBM; a step is one full
pbsm3d_step (coefficients, suspension assembly + solve, Qsusp/Qsubl,
deposition assembly + solve, drift mass), as the reference's run(mesh&).

Variants, timed back to back on the same mesh and forcing:
  plain          FP64 GMRES(30) + multicolour D-ILU(0), x0 = 0 (the reference's
                 algorithm on the GPU)
  mixed          FP32 SpMV / preconditioner / Krylov vectors inside FP64
                 refinement (FP64 residuals and dot accumulation)
  mixed_learned  mixed + the learned warm start; the network is trained during
                 set-up (not timed) on FP64 solves of other seeded forcing draws
`value` is mixed_learned's steps/s; `speedup_vs_plain` divides it by plain's.

Prints ONE JSON line.  --gpus must be 1 (one GPU per box).
"""
from __future__ import annotations

import argparse
import json
import statistics
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "PBSM3D steps/s @1M tri x 10 layers, rel-resid 1e-8: plain vs mixed+learned"


def spmv_bytes(nnz: int, n: int, f: int, s: int) -> int:
    """Algorithmic bytes of one suspension SpMV in the D-ILU position order: every
    nonzero value once (diag / lateral / vertical arrays), x read once, y written
    once, and the int32 neighbour positions of each face (3 per face)."""
    return nnz * s + 2 * n * s + 12 * f


def split_apply_bytes(nnz: int, n: int, f: int, f0: int, L: int, s: int, f_top: int = 0) -> int:
    """Algorithmic bytes of one application of the split (Eisenstat) D-ILU operator with
    the BiCGStab dots fused in: every off-diagonal coupling once (upper ones in the
    backward sweep, lower ones in the forward sweep); per row of colours >= 1 the
    backward sweep reads w, 1/d~ and writes t, the forward sweep reads w, 1/d~, D - d~,
    t and the dot vector and writes v and the result (10 vectors); colour 0 (f0 faces)
    does both sweeps in one pass with t in registers (6 vectors); the top colour (f_top
    faces) does not write v, which no higher colour reads (one vector less); the
    neighbour positions and colour mask once per sweep (13 bytes per face and sweep),
    except in the top colour's backward sweep, which has no lateral terms.  The dot vector is
    a separate array only in the first of BiCGStab's two applications (r^ . v); in the second
    (s . t) it is the input s itself, read once: the class average counts it for half of the
    rows."""
    n0 = f0 * L
    return ((nnz - n) * s + (10 * (n - n0) + 6 * n0 - f_top * L) * s + 26 * f - 13 * f0 - 13 * f_top
            - (n * s) // 2)


def bicg_vectors_per_iteration(preconditioned: bool) -> int:
    """n-vectors of HBM traffic of one BiCGStab iteration's vector kernels: p (r, v, p in;
    p out), s (r, v in; s out), x / r (p^, s^, s, t, r^, x in; x, r out).  With the split
    operator there is no separate preconditioner, p^ is p and s^ is s -- one array each, read
    once -- so the x / r kernel moves 7 vectors instead of 8."""
    return 4 + 3 + (8 if preconditioned else 7)


def spmv_dots_bytes(nnz: int, n: int, f: int, s: int) -> int:
    """The SpMV of the BiCGStab iteration, which also reads the vector w of its fused
    dot products (w.y, y.y)."""
    return spmv_bytes(nnz, n, f, s) + n * s


def measured_traffic(F: int, L: int) -> dict:
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes for this
    workload (profiles/hbm_traffic_c2.json), or {} for other workloads."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_c2.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except OSError:
        return {}
    return t if (t.get("faces"), t.get("layers")) == (F, L) else {}


def dilu_apply_bytes(nnz: int, n: int, f: int, s: int) -> int:
    """Algorithmic bytes of one multicolour D-ILU(0) apply in position order:
    forward sweep reads r, the strictly lower couplings, 1/d and writes y;
    backward sweep reads the strictly upper couplings, 1/d, y and writes z
    (off-diagonal values nnz - n once in total, 6 vectors), plus per face the
    neighbour positions and the colour mask in each sweep (2 x 13 bytes)."""
    return (nnz - n) * s + 6 * n * s + 26 * f


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def variant_config(pb, abi, name: str, L: int, krylov_mixed: int):
    if name == "plain":
        return pb.default_config(nLayer=L, precision=abi.FP64), False
    if name == "plain_bicgstab":
        return pb.default_config(nLayer=L, precision=abi.FP64, krylov=abi.KRYLOV_BICGSTAB), False
    if name == "plain_split":
        return pb.default_config(nLayer=L, precision=abi.FP64, krylov=abi.KRYLOV_BICGSTAB_SPLIT), False
    if name == "mixed_bicgstab":
        return pb.default_config(nLayer=L, precision=abi.MIXED, krylov=abi.KRYLOV_BICGSTAB), False
    if name == "mixed_split":
        return pb.default_config(nLayer=L, precision=abi.MIXED, krylov=abi.KRYLOV_BICGSTAB_SPLIT), False
    if name == "mixed":
        return pb.default_config(nLayer=L, precision=abi.MIXED, krylov=krylov_mixed), False
    if name == "mixed_learned":
        return pb.default_config(nLayer=L, precision=abi.MIXED, krylov=krylov_mixed, warm_start=abi.WS_LEARNED), True
    if name == "learned":
        return pb.default_config(nLayer=L, precision=abi.FP64, warm_start=abi.WS_LEARNED), True
    raise SystemExit(f"unknown variant {name}")


def _pin(pb, ctx, cfg, spmv_shape, dilu_mode):
    """Pinned launch shapes (profiling runs), else the tuner decides.  The non-temporal
    sweep modes (bit 1) exist for the split operator only; other solvers keep bit 0."""
    for which in (0, 1):
        if spmv_shape >= 0:
            ctx.spmv_select(which, spmv_shape)
        if dilu_mode >= 0:
            split = cfg.krylov == pb.abi.KRYLOV_BICGSTAB_SPLIT
            ctx.precond_select(which, dilu_mode if split else dilu_mode & 1)


def host_io_rate(pb, g, cfg, net, forcing, steps, spmv_shape=-1, dilu_mode=-1):
    """steps/s of the same step with host forcing and host outputs (the PCIe copies of
    the forcing in and the eight output fields out included)."""
    import numpy as np

    ctx = pb.Context(g, cfg)
    _pin(pb, ctx, cfg, spmv_shape, dilu_mode)
    if net is not None:
        ctx.set_warm_start(net)
    outs = {k: np.zeros(g.n_faces) for k in pb.abi.OUT_FIELDS}
    ctx.step(forcing, outs=outs)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.step(forcing, outs=outs)
    dt = time.perf_counter() - t0
    ctx.close()
    return steps / dt


def stream_probe(torch, mb: int = 1000, reps: int = 5) -> dict:
    """GB/s (read + written bytes, best of reps) of two plain streaming kernels on `mb` MB
    operands: torch's device-to-device copy (1 read : 1 write) and c = a + b (2 : 1)."""
    n = mb * 1000000 // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a).fill_(2.0)
    c = torch.empty_like(a)
    out = {}
    for name, fn, nbytes in (("copy", lambda: c.copy_(a), 2 * 4 * n),
                             ("add", lambda: torch.add(a, b, out=c), 3 * 4 * n)):
        fn()
        best = 0.0
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = max(best, nbytes / (e0.elapsed_time(e1) * 1e-3) / 1e9)
        out[name] = round(best, 1)
    del a, b, c
    return out


# timed steps: one launch in seven of a reported class is bracketed.  The period is odd on
# purpose: BiCGStab applies the operator twice per iteration (the second time with its input
# as the dot vector, one array less to read), so an even period would sample one of the two
# applications only; an odd one alternates between them.
PROFILE_EVERY = 7


def run_variant(pb, g, dev_forcing, dev_outs, cfg, net, steps, warmup, torch, spmv_shape=-1, dilu_mode=-1):
    ctx = pb.Context(g, cfg)
    _pin(pb, ctx, cfg, spmv_shape, dilu_mode)
    if net is not None:
        ctx.set_warm_start(net)
    for _ in range(warmup):
        ctx.step(dev_forcing, outs=dev_outs, on_device=True)
    torch.cuda.synchronize()
    # the timed steps bracket only the classes whose rooflines are reported, and only every
    # PROFILE_EVERY-th launch of them (each bracket's two event records leave the device idle
    # for ~6 us); two further steps bracket every launch of every class for the breakdown
    ctx.profile(True, classes=timed_classes(pb, cfg), every=PROFILE_EVERY)
    ctx.profile_read(reset=True)
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    marks = [t0]
    for _ in range(steps):
        _, st = ctx.step(dev_forcing, outs=dev_outs, on_device=True)  # returns with the step complete
        stats.append(st)
        marks.append(time.perf_counter())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    step_ms = [1000.0 * (b - a) for a, b in zip(marks, marks[1:])]
    prof = ctx.profile_read(reset=True)
    ctx.profile(True, every=1)
    for _ in range(BREAKDOWN_STEPS):
        ctx.step(dev_forcing, outs=dev_outs, on_device=True)
    torch.cuda.synchronize()
    breakdown = ctx.profile_read(reset=True)
    ctx.profile(False)
    nnz = pb.lib().pbsm3d_system_nnz(ctx._h, 0)
    tuning = {"spmv_fp64": ctx.spmv_tuning(0), "spmv_fp32": ctx.spmv_tuning(1),
              "dilu_fp64": ctx.precond_tuning(0), "dilu_fp32": ctx.precond_tuning(1)}
    ctx.close()
    return (t1 - t0), stats, prof, breakdown, nnz, tuning, step_ms


BREAKDOWN_STEPS = 2


def timed_classes(pb, cfg):
    """Kernel classes bracketed during the timed steps: the ones the rooflines report."""
    if cfg.krylov == pb.abi.KRYLOV_BICGSTAB_SPLIT:
        return ("precond",)
    return ("spmv64" if cfg.precision == pb.abi.FP64 else "spmv32", "precond")


def cpu_baseline(g_full, forcing, n_layer, faces, threads, steps):
    """Times the CPU oracle (GMRES(30) + ILUT(1e-4, 3), the reference's solver setup) on
    the same mesh and forcing: `steps` consecutive 900 s steps after init (the CSR
    patterns and the solver set-up of init are not timed, as in the reference's
    timestep loop), steps/s from the median step.  faces > 0 restricts it to a
    Hilbert-contiguous patch of that many faces and scales steps/s by the face ratio
    (a quicker, flattering sample); 0 = the whole mesh, no scaling."""
    import numpy as np

    from chm_amd import abi
    from oracle import oracle_py

    F = g_full.n_faces
    fs = min(faces, F) if faces > 0 else F
    if fs < F:
        nb = g_full.neigh.reshape(F, 3)[:fs].copy()
        nb[nb >= fs] = -1
        sl3 = slice(0, 3 * fs)
        gs = abi.Geometry(nb, g_full.normal_x[sl3], g_full.normal_y[sl3], g_full.edge_length[sl3], g_full.area[:fs],
                          g_full.center_x[:fs], g_full.center_y[:fs], g_full.center_z[:fs])
        fo = {k: np.ascontiguousarray(v[:fs]) for k, v in forcing.items()}
    else:
        gs, fo = g_full, forcing
    orc = oracle_py.Oracle(gs, oracle_py.oracle_config(nLayer=n_layer), threads=threads)
    times, st = [], None
    for k in range(steps):
        t0 = time.perf_counter()
        _, st = orc.step(fo, want_fields=False)
        times.append(time.perf_counter() - t0)
        log(f"[bench] CPU baseline step {k + 1}/{steps}: {times[-1]:.1f} s on {threads} threads")
    orc.close()
    dt = statistics.median(times)
    where = (f"the whole mesh ({F} faces) x {n_layer} layers" if fs == F else
             f"a {fs}-face Hilbert-contiguous patch of the mesh x {n_layer} layers, steps/s scaled by faces {fs}/{F}")
    return {
        "value": (1.0 / dt) * (fs / F),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{steps} consecutive 900 s steps of the CPU FP64 restatement (GMRES(30) + ILUT(1e-4, fill 3), "
                   f"tol 1e-8, OpenMP assembly / SpMV / dots, serial ILUT) on {where}, init not timed; median "
                   f"step {dt:.1f} s (steps {', '.join(f'{t:.1f}' for t in times)} s) on {threads} threads; "
                   f"susp/dep iterations {st['susp_iters']}/{st['dep_iters']}"),
    }


def train_network(pb, abi, meshgen, warmstart, g, F, draws, samples, epochs, torch):
    """FP64 GPU solves of seeded forcing draws (seeds 1..draws, disjoint from the
    timed forcing's seed 0) -> (features, solution) pairs -> GPU training."""
    t0 = time.perf_counter()
    forcings = [meshgen.forcing_c2(F, seed=1 + k) for k in range(draws)]
    X, Y = warmstart.collect_pairs(g, forcings, pb.default_config(nLayer=10))
    t_collect = time.perf_counter() - t0
    if samples and X.shape[0] > samples:
        gen = torch.Generator(device=X.device).manual_seed(1)
        sel = torch.randperm(X.shape[0], device=X.device, generator=gen)[:samples]
        X, Y = X[sel].contiguous(), Y[sel].contiguous()
    net, info = warmstart.fit(X, Y, epochs=epochs, batch=8192, lr=2e-3, seed=0)
    info.update({"draws": draws, "collect_s": round(t_collect, 2)})
    return net, info


def warm_start_gain(pb, abi, meshgen, warmstart, g, F, L, krylov_mixed, net, seeds):
    """One step per held-out forcing draw (seeds disjoint from the training draws and the
    timed forcing) with the headline solver, cold start vs learned start: suspension
    iterations, initial residual ||b - A x0|| / ||b|| and step time with the forcing
    resident in HBM, and the learned start with the learned coarse correction on 256
    tiles.  Each context first steps once on the timed forcing so its launch-shape
    tuning is not counted."""
    cfg = pb.default_config(nLayer=L, precision=abi.MIXED, krylov=krylov_mixed)
    out = warmstart.compare_starts(g, net, [meshgen.forcing_c2(F, seed=sd) for sd in seeds], cfg,
                                   warmup_draw=meshgen.forcing_c2(F, seed=0),
                                   coarse_tiles=256 if krylov_mixed == abi.KRYLOV_BICGSTAB_SPLIT else 0)
    out["seeds"] = seeds
    return out


def inner_target_sweep(pb, abi, meshgen, g, F, L, net, seeds, torch):
    """Total suspension-solve time and iterations of the mixed split BiCGStab over held-out
    draws for each candidate inner target, from x0 = 0 and from the learned start (the
    selection behind the library's defaults; seeds disjoint from the timed forcing, the
    training draws and the warm-start comparison draws)."""
    dev = "cuda:0"
    draws = [{k: torch.as_tensor(v, device=dev) for k, v in meshgen.forcing_c2(F, seed=sd).items()} for sd in seeds]
    warm0 = {k: torch.as_tensor(v, device=dev) for k, v in meshgen.forcing_c2(F, seed=0).items()}
    outs = {k: torch.zeros(F, dtype=torch.float64, device=dev) for k in abi.OUT_FIELDS}
    res = {"seeds": list(seeds)}
    for label, ws, targets in (("cold", abi.WS_ZERO, (1e-3, 1.5e-3, 2e-3, 3e-3, 5e-3)),
                               ("learned", abi.WS_LEARNED, (1e-4, 1.5e-4, 2e-4, 3e-4, 5e-4, 1e-3))):
        rows = []
        for tgt in targets:
            cfg = pb.default_config(nLayer=L, precision=abi.MIXED, krylov=abi.KRYLOV_BICGSTAB_SPLIT,
                                    warm_start=ws, inner_rtol=tgt)
            ctx = pb.Context(g, cfg)
            if ws == abi.WS_LEARNED:
                ctx.set_warm_start(net)
            ctx.step(warm0, outs=outs, on_device=True)
            its = ms = 0.0
            for d in draws:
                _, st = ctx.step(d, outs=outs, on_device=True)
                its += st["susp_iters"]
                ms += st["ms_susp_solve"]
            ctx.close()
            rows.append({"inner_rtol": tgt, "susp_iters": int(its), "ms_susp_solve": round(ms, 2)})
            log(f"[bench] inner sweep {label} {tgt:g}: {int(its)} iterations, {ms:.1f} ms")
        res[label] = rows
        res[f"best_{label}"] = min(rows, key=lambda r: r["ms_susp_solve"])["inner_rtol"]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=500000)
    ap.add_argument("--layers", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--variants", default="plain,mixed,mixed_learned")
    ap.add_argument("--krylov-mixed", default="split", choices=["gmres", "bicgstab", "split"])
    ap.add_argument("--train-draws", type=int, default=8)
    ap.add_argument("--train-samples", type=int, default=4000000)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--gain-seeds", default="100,101,102,103,104,105,106,107",
                    help="held-out forcing draws on which the learned start is compared with the cold start")
    ap.add_argument("--cpu-faces", type=int, default=0,
                    help="CPU baseline on a patch of this many faces, scaled (default 0: the whole mesh)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-one-core", action="store_true",
                    help="also time the CPU baseline on one thread (the reference's default build)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fetch", action="store_true", help="skip timing the fetch (fetchr) kernel")
    ap.add_argument("--spmv-shape", type=int, default=-1, help="pin the SpMV launch shape (default: tuned)")
    ap.add_argument("--ensemble", type=int, default=64, help="parameter-ensemble members to time (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=24)
    ap.add_argument("--host-io-steps", type=int, default=3, help="steps timed with host forcing / outputs (0: skip)")
    ap.add_argument("--dilu-mode", type=int, default=-1, choices=list(range(-1, 8)),
                    help="pin the D-ILU sweep mode: bit 0 XCD workgroup mapping; split operator only: bit 1 "
                         "non-temporal factor loads, bit 2 layer-parallel sweeps (default: tuned)")
    ap.add_argument("--inner-sweep-seeds", default="",
                    help="e.g. 200-211: time the mixed split solve's inner-target candidates over these "
                         "held-out draws, cold and learned (off by default)")
    ap.add_argument("--study-draws", type=int, default=20000,
                    help="learned-start study: FP64 solves of forcing draws on the 2,500-point mesh at 10 layers "
                         "that train its network, before cold vs learned on 16 held-out draws (0: skip)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one streaming kernel (y = -x, float32) reading and writing this many MB first: "
                         "the calibration dispatch for rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the step is built and measured for "
              f"one MI355X, and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from chm_amd import abi, meshgen, warmstart
    from chm_amd import pbsm3d as pb

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    pb.lib()
    t = time.perf_counter()
    mesh = meshgen.synthetic_mesh(args.points, seed=args.seed)
    F, L = mesh.n_faces, args.layers
    g = pb.face_geometry(mesh)
    forcing = meshgen.forcing_c2(F, seed=args.seed)
    log(f"[bench] mesh F={F} L={L} N={F * L} built in {time.perf_counter() - t:.1f}s")
    dev_forcing = {k: torch.tensor(v, device="cuda:0") for k, v in forcing.items()}
    dev_outs = {k: torch.zeros(F, dtype=torch.float64, device="cuda:0") for k in abi.OUT_FIELDS}
    krylov_mixed = {"bicgstab": abi.KRYLOV_BICGSTAB, "split": abi.KRYLOV_BICGSTAB_SPLIT,
                    "gmres": abi.KRYLOV_GMRES}[args.krylov_mixed]

    if args.copy_probe_mb:
        n = args.copy_probe_mb * 1000000 // 4
        xs = torch.ones(n, dtype=torch.float32, device="cuda:0")
        ys = torch.empty_like(xs)
        torch.cuda.synchronize()
        torch.neg(xs, out=ys)
        torch.cuda.synchronize()
        log(f"[bench] calibration kernel: read {4 * n} B, wrote {4 * n} B")
        del xs, ys

    names = args.variants.split(",")
    net, train_info = None, None
    if any(n in ("mixed_learned", "learned") for n in names):
        log(f"[bench] training the warm-start network ({args.train_draws} FP64 draws)")
        net, train_info = train_network(pb, abi, meshgen, warmstart, g, F, args.train_draws, args.train_samples,
                                        args.train_epochs, torch)
        log(f"[bench] trained: {train_info}")

    variants = {}
    plain_outs = None
    for name in names:
        cfg, learned = variant_config(pb, abi, name, L, krylov_mixed)
        log(f"[bench] variant {name}: warmup {args.warmup}, steps {args.steps}")
        secs, stats, prof, breakdown, nnz, tuning, step_ms = run_variant(
            pb, g, dev_forcing, dev_outs, cfg, net if learned else None, args.steps, args.warmup, torch,
            args.spmv_shape, args.dilu_mode)
        st = stats[-1]
        variants[name] = {
            "steps_per_s": args.steps / secs,
            "ms_per_step": 1000.0 * secs / args.steps,
            "ms_per_step_median": round(statistics.median(step_ms), 3),
            "krylov": {abi.KRYLOV_BICGSTAB: "bicgstab", abi.KRYLOV_BICGSTAB_SPLIT: "bicgstab_split"}.get(
                cfg.krylov, f"gmres({cfg.restart})"),
            "precision": "fp64" if cfg.precision == abi.FP64 else "fp32-inner/fp64-refinement",
            "warm_start": bool(learned),
            "susp_iters": st["susp_iters"],
            "dep_iters": st["dep_iters"],
            "susp_resid": st["susp_resid"],
            "dep_resid": st["dep_resid"],
            "refinements": st["susp_refinements"],
            "susp_resid0": st["susp_resid0"],
            "cold_fallback": bool(st["susp_cold_fallback"]),
            "launch_tuning": tuning,
            "ms_phase": {k: round(st[k], 3) for k in ("ms_assembly", "ms_susp_solve", "ms_post", "ms_dep_solve")},
            "kernels": {k: {"launches": n, "ms_avg": (ms / n if n else 0.0), "ms_total": ms}
                        for k, (n, ms) in prof.items() if n},
            f"kernel_breakdown_{BREAKDOWN_STEPS}_steps": {k: {"launches": n, "ms_total": round(ms, 3)}
                                                          for k, (n, ms) in breakdown.items() if n},
            "nnz": nnz,
        }
        # per-field difference of this variant's last-step outputs from the plain FP64 variant's
        # (both solved to ||b - A x|| <= 1e-8 ||b||): max|x - x_plain| / max|x_plain|
        outs_now = {k: v.clone() for k, v in dev_outs.items()}
        if name == "plain":
            plain_outs = outs_now
        elif plain_outs is not None:
            variants[name]["field_rel_diff_vs_plain"] = {
                k: float((outs_now[k] - plain_outs[k]).abs().max() / plain_outs[k].abs().max().clamp_min(1e-300))
                for k in ("Qsalt", "Qsusp", "Qsubl", "drift_mass", "sum_drift")}
        log(f"[bench] {name}: {variants[name]['steps_per_s']:.2f} steps/s, iters {st['susp_iters']}/{st['dep_iters']}")

    # upstream producer of the fetch input (fetchr): the GPU upwind search on the same mesh
    fetch_info = None
    if not args.no_fetch:
        cxyz = mesh.vertex[mesh.elem].mean(axis=1)
        fctx = pb.Fetch(np.ascontiguousarray(cxyz[:, 0]), np.ascontiguousarray(cxyz[:, 1]),
                        np.ascontiguousarray(cxyz[:, 2]))
        fetch_out = fctx.run(dev_forcing["vw_dir"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            fetch_out = fctx.run(dev_forcing["vw_dir"])
        torch.cuda.synchronize()
        fetch_info = {"ms_per_call": round(1000.0 * (time.perf_counter() - t0) / 5, 3), "faces": F,
                      "queries_per_face": 10, "mean_fetch_m": round(float(fetch_out.mean()), 2)}
        fctx.close()
        log(f"[bench] fetch: {fetch_info}")

    # parameter ensemble: every member advanced by one pbsm3d_batch_step per step (members of
    # equal nLayer in one union, per-member BiCGStab scalars), against the members run one at
    # a time and concurrently (one context and stream each, host threads)
    ens = None
    if args.ensemble > 0:
        from chm_amd import ensemble

        log(f"[bench] ensemble: {args.ensemble} members x {args.ensemble_steps} steps")
        members = ensemble.draw_members(args.ensemble, seed=0)
        t_set = time.perf_counter()
        meshes = ensemble.member_meshes(members, workers=16)
        batch = ensemble.build_batch(members, meshes, precision=abi.MIXED, warm_net=net)
        t_set = time.perf_counter() - t_set
        ensemble.run_batch(batch, members, meshes, 1)  # first step: launch shapes of the transforms warm up
        bat = ensemble.run_batch(batch, members, meshes, args.ensemble_steps)
        info = batch.info()
        batch.close()
        del batch
        ref_steps = min(args.ensemble_steps, 2)
        built = ensemble.build(members, precision=abi.MIXED, krylov=krylov_mixed, warm_net=net)
        ensemble.run(built, 1)  # first steps include the launch-shape tuning
        seq = ensemble.run(built, ref_steps, workers=1)
        # at most 16 host threads: the process has 4 hardware queues, and more streams in
        # flight only contend (64 threads for 64 members ran at 0.7x the rate of 16)
        par = ensemble.run(built, ref_steps, workers=min(len(built), 16))
        ens = {"members": len(members), "steps": args.ensemble_steps,
               "faces": [mesh.n_faces for mesh in meshes], "layers": [m.n_layer for m in members],
               "member_steps_per_s_batched": round(bat["member_steps_per_s"], 3),
               "cell_steps_per_s_batched": round(bat["cell_steps_per_s"], 1),
               "member_steps_per_s_concurrent": round(par["member_steps_per_s"], 3),
               "member_steps_per_s_one_at_a_time": round(seq["member_steps_per_s"], 3),
               "speedup_batched_vs_one_at_a_time": round(bat["member_steps_per_s"] / seq["member_steps_per_s"], 3),
               "speedup_batched_vs_concurrent": round(bat["member_steps_per_s"] / par["member_steps_per_s"], 3),
               "max_resid_batched": bat["max_resid"],
               "max_resid_concurrent": max(r["last_resid"] for r in par["per_member"]),
               "unions": bat["unions"], "batched_launches_last_step": info["launches"],
               "batched_lockstep_iterations_last_step": info["iterations"],
               "reference_steps": ref_steps, "setup_s": round(t_set, 1),
               "note": "mixed precision, learned start on the 10-layer members, forcing built on the GPU; batched: "
                       "one pbsm3d_batch_step per step (members of equal nLayer in one union system, one launch "
                       "per colour sweep and per BiCGStab kernel for all of them); concurrent: one context and "
                       "HIP stream per member on 16 host threads; one at a time: the same contexts in turn"}
        for _, _, ctx in built:
            ctx.close()
        log(f"[bench] ensemble: {ens}")

    pcie = None
    if args.host_io_steps > 0 and names:
        hn = "mixed_learned" if "mixed_learned" in names else names[-1]
        hcfg, hlearned = variant_config(pb, abi, hn, L, krylov_mixed)
        pcie = {"variant": hn, "steps_per_s": round(host_io_rate(pb, g, hcfg, net if hlearned else None, forcing,
                                                                 args.host_io_steps, args.spmv_shape,
                                                                 args.dilu_mode), 4),
                "note": "host (numpy) forcing and outputs: H2D of 8 forcing fields and D2H of 8 output fields "
                        "per step included"}
        log(f"[bench] host-array (PCIe-inclusive) rate: {pcie}")

    gain = None
    if net is not None and "mixed_learned" in names and args.gain_seeds:
        log("[bench] learned vs cold start on held-out forcing draws")
        gain = warm_start_gain(pb, abi, meshgen, warmstart, g, F, L, krylov_mixed, net,
                               [int(x) for x in args.gain_seeds.split(",")])
        log(f"[bench] gain: {gain}")

    sweep = None
    if net is not None and args.inner_sweep_seeds:
        a, b = (int(x) for x in args.inner_sweep_seeds.split("-"))
        sweep = inner_target_sweep(pb, abi, meshgen, g, F, L, net, range(a, b + 1), torch)
        log(f"[bench] inner sweep: {json.dumps(sweep)}")

    study = None
    if args.study_draws > 0:
        log(f"[bench] learned-start study: {args.study_draws} FP64 solves on the 2,500-point mesh at 10 layers")
        snet, study = warmstart.learned_start_study(n_train=args.study_draws)
        snet.close()
        log(f"[bench] study: {json.dumps(study)}")

    N = F * L
    head_name = "mixed_learned" if "mixed_learned" in variants else names[-1]
    head = variants[head_name]
    # rooflines of the suspension solve's HBM-bound kernels in the headline variant; the
    # reported one is the kernel class with the largest device time
    # element width of the solve's operator and preconditioner: FP32 in the mixed variants
    s_bytes = 8 if head["precision"] == "fp64" else 4
    kspmv = "spmv32" if s_bytes == 4 else "spmv64"
    nnz = head["nnz"]
    bicg = head["krylov"] == "bicgstab"
    if head["krylov"] == "bicgstab_split":
        colors, _ = pb.face_colors(mesh.neigh)
        f0 = int((colors == 0).sum())
        f_top = int((colors == colors.max()).sum())
        algos = {"precond": (split_apply_bytes(nnz, N, F, f0, L, s_bytes, f_top),
                             "split D-ILU operator (Eisenstat): backward + forward sweep, fused BiCGStab dots",
                             "split_apply")}
    else:
        algos = {kspmv: ((spmv_dots_bytes if bicg else spmv_bytes)(nnz, N, F, s_bytes),
                         "suspension SpMV, position order" + (", fused BiCGStab dots" if bicg else ""),
                         "spmv_dots" if bicg else "spmv"),
                 "precond": (dilu_apply_bytes(nnz, N, F, s_bytes), "multicolour D-ILU(0) apply, position order",
                             "dilu_apply")}
    traffic = measured_traffic(F, L)
    rooflines = {}
    for kname, (algo, what, kid) in algos.items():
        k = head["kernels"].get(kname)
        if not k or k["ms_avg"] <= 0:
            continue
        achieved = algo / (k["ms_avg"] * 1e-3) / 1e9
        tr = traffic.get(f"{kid}_fp{8 * s_bytes}", {}).get("bytes_per_launch")
        rooflines[kname] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr, "kernel": f"{kname} ({what})",
                            "algorithmic_bytes": algo, "ms_avg": round(k["ms_avg"], 5), "launches": k["launches"],
                            "ms_total": round(k["ms_total"], 3), "sampled_every": PROFILE_EVERY}
        if tr:
            rooflines[kname]["traffic_source"] = traffic["source"]
    dom = max(rooflines, key=lambda n: rooflines[n]["ms_total"]) if rooflines else None
    roofline = rooflines.get(dom)
    if roofline is not None:
        # context for frac: what plain streaming kernels reach on this device (1 GB operands)
        stream = stream_probe(torch)
        roofline["measured_stream_gbs"] = stream
        roofline["frac_of_measured_stream"] = round(roofline["achieved"] / max(stream.values()), 4)
    # BiCGStab vector work (p; alpha + s; omega + x, r; the rho / stop stage), from the untimed
    # steps with every class bracketed: bicg_vectors_per_iteration n-vectors of traffic per
    # iteration, suspension (n = N, the inner precision) and deposition (n = F, FP64)
    bd = head.get(f"kernel_breakdown_{BREAKDOWN_STEPS}_steps", {}).get("ortho")
    if head["krylov"] == "bicgstab_split" and bd and bd["ms_total"] > 0:
        nvs = bicg_vectors_per_iteration(False)
        algo = BREAKDOWN_STEPS * nvs * (head["susp_iters"] * N * s_bytes + head["dep_iters"] * F * 8)
        achieved = algo / (bd["ms_total"] * 1e-3) / 1e9
        rooflines["ortho"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                              "kernel": "ortho (BiCGStab vector kernels k_bicg_p_dev / k_bicg_s_dev / k_bicg_xr_dev "
                                        "and k_bicg_stage, suspension and deposition)",
                              "algorithmic_bytes": algo, "ms_total": bd["ms_total"], "brackets": bd["launches"],
                              "source": f"kernel_breakdown_{BREAKDOWN_STEPS}_steps (HIP events, every class bracketed)"}
        tv = traffic.get(f"bicg_vectors_fp{8 * s_bytes}", {}).get("bytes_per_iteration")
        if tv:  # measured HBM bytes of one suspension iteration's vector kernels, against nvs N s
            rooflines["ortho"]["traffic_per_suspension_iteration"] = {"measured": tv, "algorithmic": nvs * N * s_bytes,
                                                                      "source": traffic["source"]}

    cpu = cpu1 = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"[bench] CPU baseline: oracle, {args.cpu_faces or F} faces, {threads} threads, {args.cpu_steps} steps")
        cpu = cpu_baseline(g, forcing, L, args.cpu_faces, threads, args.cpu_steps)
        if args.cpu_one_core and threads != 1:
            # one core, as the reference's default build runs (no OpenMP, one MPI rank)
            log(f"[bench] CPU baseline: oracle, {args.cpu_faces or F} faces, 1 thread")
            cpu1 = cpu_baseline(g, forcing, L, args.cpu_faces, 1, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": round(head["steps_per_s"], 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head["precision"] == "fp64" else "f32+f64",
        "data": "synthetic (seeded Delaunay mesh and forcing)",
        "config": {"workload": f"PBSM3D C2: {args.points} points -> {F} triangles x {L} layers, 900 s step, tol 1e-8",
                   "faces": F, "layers": L, "unknowns": N, "headline_variant": head_name},
        "variants": variants,
        "speedup_vs_plain": (round(head["steps_per_s"] / variants["plain"]["steps_per_s"], 3)
                             if "plain" in variants else None),
        "warm_start_training": train_info,
        "warm_start_gain": gain,
        "learned_start_study": study,
        "inner_target_sweep": sweep,
        "fetch": fetch_info,
        "ensemble": ens,
        "pcie_inclusive": pcie,
        "roofline": roofline,
        "kernel_rooflines": rooflines,
        "cpu_baseline": cpu,
        "cpu_baseline_one_core": cpu1,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
