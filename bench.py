#!/usr/bin/env python3
"""fwd+bwd ms/iter of the equirectangular Gaussian rasterizer on one MI355X.

Workload: the 1.5M-Gaussian room scene at 1920x960 (FP32, SH degree 3, seeded;
omnigs_amd.scenes.room_scene).  One step = forward (preprocess, binning,
render) + the L1 upstream gradient sign(img - target)/(3HW) from the fused
loss kernel (omnigs_photometric_loss with lambda 0) + backward (render
backward, preprocess backward), all through the C ABI with inputs resident in
HBM.  Prints ONE JSON line.

The roofline object prices the dominant stage (by time, from the warm-up
steps' per-stage events) against HBM: achieved = its algorithmic bytes per
launch / its average launch time, from HIP events the library records around
that stage on its launch stream during the timed steps.  The CPU baseline runs the FP32 oracle over full fwd+bwd
iterations of the same workload on the host's cores (and one thread beside).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

HBM_PEAK_GBS = 8000.0
VALU_SIMDS = 1024          # 256 CUs x 4 SIMDs
VALU_CLOCK_GHZ = 2.4       # peak engine clock


def algorithmic_bytes(P, Pv, R, tiles, N, depth_passes, tile_passes):
    """Compulsory HBM bytes per launch of each stage (every dataflow tensor
    read or written once; gathers counted once per instance)."""
    return {
        # means 12 + scales 12 + rotations 16 + opacity 4 + SH 192 read;
        # radii/tiles/key 12 for all, xy 8 + conic 16 + rgb 12 + clamped 1 for visible
        "preprocess": 236 * P + 12 * P + 37 * Pv,
        # per pass: key read twice (count + scatter), value read, key+value written
        "depth_sort": depth_passes * 20 * P,
        # tiles_touched gathered through the order twice, offsets written
        "scan": 20 * P,
        # order, offsets, xy, radii read; tile key + id written per instance
        "duplicate": 20 * P + 8 * R,
        "tile_sort": tile_passes * 20 * R,
        "ranges": 4 * R + 8 * tiles,
        # per instance: id 4 + xy 8 + conic 16 + rgb 12; per pixel 3 colours + T + n_contrib written
        "render": 40 * R + 8 * tiles + 20 * N,
        # per instance gather 40; per pixel T, n_contrib, dL/dpix read; 9 gradient floats per visible Gaussian
        "render_backward": 40 * R + 8 * tiles + 20 * N + 36 * Pv,
        # the 2D gradients summed per Gaussian (rec 32 + tiles_touched 4 read; dmean2D 12 + dconic 16 +
        # dopacity 4 + dcolor 12 written; the per-instance records between the two kernels are this
        # design's own traffic, not counted), then the 3D terms: means, scales, rot, SH, radii, clamped
        # read (237 B), dmeans, dcov3D, dSH, dscale, drot written (256 B); one fused kernel
        "preprocess_backward": 573 * P,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--gaussians", type=int, default=1_500_000)
    ap.add_argument("--no-decoder", action="store_true", help="skip the learned-colour leg")
    ap.add_argument("--batch", action="store_true",
                    help="only the batched multi-view workload: 1.5M Gaussians, 8/16/32 views of uneven size, "
                         "autotuned tile shape and views per launch (views/sec)")
    ap.add_argument("--batch-views", default="8,16,32",
                    help="with --batch: the batch sizes V to time (comma-separated)")
    ap.add_argument("--batch-vpl", default="",
                    help="with --batch: only these views per launch (comma-separated; default all), tile 16x16 only "
                         "when given")
    ap.add_argument("--memory-compare", action="store_true",
                    help="only the checkpointed-vs-stored backward comparison (3M Gaussians at 4096x2048)")
    ap.add_argument("--memory-modes", default="stored,checkpointed,checkpointed_one_band",
                    help="with --memory-compare: the modes to run (comma-separated)")
    ap.add_argument("--distill-steps", type=int, default=2000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (known HBM bytes) to scale rocprofv3 FETCH/WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "only one GPU is measured: the rasterizer runs on one GPU per process, "
                                   "and multi-GPU sharding is a design, not a measured path", "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch

    from omnigs_amd import capi, scenes

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        return copy_calibration(dev, args.steps)
    if args.memory_compare:
        return memory_compare(dev, args)
    if args.batch:
        return batch_workload(dev, args)
    sc = scenes.room_scene(P=args.gaussians)
    P, W, H = sc["means3D"].shape[0], sc["width"], sc["height"]
    N = W * H
    tiles = ((W + 15) // 16) * ((H + 15) // 16)
    ds = {k: (torch.from_numpy(v).to(dev) if isinstance(v, np.ndarray) else v) for k, v in sc.items()}
    ds["opacities"] = ds["opacities"].reshape(-1).contiguous()
    target = ds["target"]
    bufs = (capi.TorchBuffer(dev), capi.TorchBuffer(dev), capi.TorchBuffer(dev))
    l1 = capi.PhotometricLoss(lambda_dssim=0.0)
    dL = torch.empty((3, H, W), dtype=torch.float32, device=dev)

    def step():
        R, img, radii, _ = capi.forward_scene(ds, buffers=bufs)
        l1(img, target, dL_dimage=dL)
        g = capi.backward_scene(ds, R, radii, bufs, dL)
        return R, radii, g

    torch.cuda.reset_peak_memory_stats(dev)
    median_ms, ms_per_step, stages, dominant, dom_ms, (R, radii, g) = time_steps(step, args.warmup, args.steps,
                                                                                 dev)
    peak_hbm = torch.cuda.max_memory_allocated(dev)
    Pv = int((radii > 0).sum())

    depth_passes = 4
    tile_passes = (int(tiles).bit_length() + 7) // 8
    alg = algorithmic_bytes(P, Pv, R, tiles, N, depth_passes, tile_passes)
    stage_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items() if v[1]}
    achieved = alg[dominant] / (dom_ms * 1e-3) / 1e9
    whole_bytes = 848 * P + 124 * R + 40 * N
    traffic = None
    valu_issue = None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj.get(dominant)
        # the render kernels are VALU-issue-bound: wave64 VALU instructions per launch (rocprofv3
        # SQ_INSTS_VALU, same workload) at 4 cycles each on 1024 SIMDs, against the launch time
        valu = tj.get("_valu_instructions", {}).get(dominant)
        if valu:
            cap = VALU_SIMDS * VALU_CLOCK_GHZ * 1e9 * dom_ms * 1e-3 / 4.0
            valu_issue = {"instructions": valu, "frac": round(valu / cap, 4), "simds": VALU_SIMDS,
                          "clock_ghz": VALU_CLOCK_GHZ, "cycles_per_instruction": 4}

    out = {
        "metric": "fwd+bwd ms/iter, 1.5M Gaussians @1920x960 lonlat (plain SH path)",
        "value": round(median_ms, 4),
        "unit": "ms/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded room scene; random-init Gaussians)",
        "config": {"workload": "plain SH fwd+bwd, 1.5M Gaussians, 1920x960 equirectangular, tile 16x16",
                   "gaussians": P, "width": W, "height": H, "num_rendered": R, "visible": Pv,
                   "sh_degree": 3, "peak_hbm_bytes": peak_hbm},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": alg[dominant], "avg_launch_ms": round(dom_ms, 5),
                     "valu_issue": valu_issue},
        "whole_path": {"algorithmic_bytes": whole_bytes,
                       "achieved_GBps": round(whole_bytes / (median_ms * 1e-3) / 1e9, 2),
                       "frac_of_hbm_roofline": round(whole_bytes / (median_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "stages_ms": {k: round(v, 5) for k, v in stage_ms.items()},
    }

    out["pose_cycle"] = pose_cycle_leg(ds, bufs, target, l1, dL, dev)
    if not args.no_decoder:
        out["mlp_colour"], dec = decoder_leg(sc, ds, bufs, target, l1, dL, args, dev)
        out["batched_mlp_colour"] = batched_decoder_leg(sc, ds, dec, args, dev)
    if not args.no_cpu:
        target = sc["target"]
        out["cpu_baseline"] = cpu_baseline(sc)
    print(json.dumps(out))
    return 0


def pose_cycle_leg(ds, bufs, target, l1, dL, dev, n_poses: int = 8, steps: int = 40) -> dict:
    """The same fwd + L1 + bwd over a cycle of seeded poses (centres uniform in
    the 1 m cube, as a training loop moves from view to view), so that R
    changes from call to call: ms per step, and how many forwards kept their
    speculative binning (a layout of the previous call's capacity with
    headroom) rather than binning again for the new R."""
    import statistics

    import numpy as np
    import torch

    from omnigs_amd import capi, scenes
    rng = np.random.default_rng(99)
    poses = []
    for _ in range(n_poses):
        vm, cc = scenes.camera_at(rng.uniform(-0.5, 0.5, size=3))
        poses.append((torch.from_numpy(vm).to(dev), torch.from_numpy(cc).to(dev)))
    lib = capi.load()
    kept = []

    def step(k):
        d = dict(ds)
        d["viewmatrix"], d["campos"] = poses[k % n_poses]
        R, img, radii, _ = capi.forward_scene(d, buffers=bufs)
        kept.append(bufs[1].requested != lib.omnigs_lonlat_binning_bytes(R))
        l1(img, target, dL_dimage=dL)
        capi.backward_scene(d, R, radii, bufs, dL)
        return R

    for k in range(n_poses):
        step(k)
    kept.clear()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    Rs = []
    torch.cuda.synchronize()
    for k in range(steps):
        evs[k][0].record()
        Rs.append(step(k))
        evs[k][1].record()
    torch.cuda.synchronize()
    ts = [a.elapsed_time(b) for a, b in evs]
    # restore the headline pose's buffers for the legs that follow
    capi.forward_scene(ds, buffers=bufs)
    return {"poses": n_poses, "steps": steps, "ms_per_step_median": round(statistics.median(ts), 4),
            "ms_per_step_mean": round(statistics.mean(ts), 4), "num_rendered_range": [min(Rs), max(Rs)],
            "forwards_on_larger_capacity": int(sum(kept))}


def time_steps(step, warmup: int, steps: int, dev):
    """W untimed steps with every stage timed by the library's HIP events (the
    per-stage breakdown, and which stage dominates), then K timed steps
    bracketed by synchronisation, each step also bracketed by events on the
    launch stream, with the library timing only the dominant stage (two events
    per step; timing every stage would cost the stream a few microseconds per
    stage boundary).  Returns (median step ms, wall ms per step, breakdown from
    the warm-up, dominant stage, its mean ms over the timed steps, last step's
    result)."""
    import torch

    from omnigs_amd import capi
    res = step()  # the first step grows the buffers: not in the breakdown
    torch.cuda.synchronize()
    capi.profile_read()  # drop anything recorded so far
    capi.profile_enable(True)
    for _ in range(max(1, warmup - 1)):
        res = step()
    torch.cuda.synchronize()
    capi.profile_enable(False)
    stages = capi.profile_read()
    stage_ms = {k: v[0] / v[1] for k, v in stages.items() if v[1]}
    dominant = max(stage_ms, key=lambda k: stage_ms[k])
    capi.profile_enable({dominant})
    stream = torch.cuda.current_stream(dev)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        evs[i][0].record(stream)
        res = step()
        evs[i][1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    capi.profile_enable(False)
    dom = capi.profile_read()[dominant]
    return (statistics.median(a.elapsed_time(b) for a, b in evs), (t1 - t0) * 1000.0 / steps, stages, dominant,
            dom[0] / dom[1], res)


# Camera-centre box of the single-view workload (lo xyz, hi xyz): the room
# scene's camera and the decoder's held-out poses are drawn from it.
VIEW_BOX = (-0.5, -0.5, -0.5, 0.5, 0.5, 0.5)

# PSNR of the reference on 360Roam (supp.pdf p.5, Table 3): the quality level
# at which the decoder's extra error is priced as a PSNR loss.
REFERENCE_PSNR_DB = 25.464


def random_rotation(rng):
    import numpy as np
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def decoder_leg(sc, ds, bufs, target, l1, dL, args, dev) -> dict:
    """The learned-colour form: distils the decoder from the scene's own SH
    colours (2000 passes, 64 seeded directions per Gaussian from seeded camera
    centres in VIEW_BOX, Adam with network lr 1e-3 and latent lr 1e-2), judges
    it against the SH render on 16 held-out random poses (other centres in the
    same box, random rotations), then times fwd+bwd with decoder colours and
    the decoder's backward (latent and weight gradients) on the timed
    workload."""
    import math

    import numpy as np
    import torch

    from omnigs_amd import capi, scenes
    from omnigs_amd.decoder import ColourDecoder

    P, W, H = sc["means3D"].shape[0], sc["width"], sc["height"]
    shs = ds["shs"].contiguous()
    dec = ColourDecoder(P, dev, seed=0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    # the cameras of this workload stand in the 1 m cube around the origin (the
    # scene's own camera and the held-out poses below): the 64 directions per
    # Gaussian come from seeded centres in that cube
    losses = dec.distill(shs, steps=args.distill_steps, lr=1e-3, num_dirs=64, seed=1, means3D=ds["means3D"],
                         view_box=VIEW_BOX)
    e1.record()
    torch.cuda.synchronize()
    distill_ms = e0.elapsed_time(e1)
    L = losses.cpu().numpy()

    rng = np.random.default_rng(2024)
    mses, maxerr = [], 0.0
    for _ in range(16):
        R_ = random_rotation(rng)
        c = rng.uniform(-0.5, 0.5, size=3)
        vm = torch.from_numpy(scenes.viewmatrix_from_pose(R_, -R_ @ c).astype(np.float32)).to(dev)
        cp = torch.from_numpy(c.astype(np.float32)).to(dev)
        common = dict(background=ds["background"], means3D=ds["means3D"], opacities=ds["opacities"],
                      viewmatrix=vm, campos=cp, width=W, height=H, scales=ds["scales"], rotations=ds["rotations"])
        _, ish, _, _ = capi.forward(shs=shs, **common)
        _, imlp, _, _ = capi.forward(mlp=dec.mlp_args(), **common)
        d = imlp - ish
        mses.append(float((d * d).mean()))
        maxerr = max(maxerr, float(d.abs().max()))
    mse = float(np.mean(mses))
    psnr_vs_sh = 10 * math.log10(1.0 / mse)
    mse_ref = 10 ** (-REFERENCE_PSNR_DB / 10)
    loss_db = 10 * math.log10(1.0 + mse / mse_ref)

    mlp = dec.mlp_args()
    ws = capi.mlp_workspace(P, dev)

    def step():
        R, img, radii, _ = capi.forward_scene(ds, buffers=bufs, mlp=mlp)
        l1(img, target, dL_dimage=dL)
        g = capi.backward_scene(ds, R, radii, bufs, dL, mlp=mlp, mlp_workspace=ws)
        return R, radii, g

    median_ms, ms_per_step, stages, _, _, (R_dec, _, g_dec) = time_steps(step, args.warmup, args.steps, dev)
    # the training pass runs on the Gaussians with a colour gradient only
    trained = int((g_dec["dL_dcolors"] != 0).any(dim=1).sum())
    stage_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items() if v[1]}
    # compulsory traffic with the decoder: the 192 B of SH per Gaussian become 32 B of BF16
    # latents in the forward read, the backward read and the backward gradient write
    alg_dec = 368 * P + 124 * int(R_dec) + 40 * W * H
    # training pass FLOPs per trained Gaussian: forward, backward chain and weight gradients, 5888 MACs each
    flop_per_pass = 3 * 2 * 5888 * trained
    dec_ms = stage_ms.get("decoder_backward", 0.0)
    return {
        "ms_per_iter": round(median_ms, 4),
        "roofline": {"bound": "hbm", "algorithmic_bytes": alg_dec,
                     "achieved_GBps": round(alg_dec / (median_ms * 1e-3) / 1e9, 1),
                     "frac": round(alg_dec / (median_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "ms_per_step_wall": round(ms_per_step, 4),
        "stages_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "distill": {"passes": args.distill_steps, "ms_total": round(distill_ms, 1),
                    "ms_per_pass": round(distill_ms / max(1, args.distill_steps), 4),
                    "loss_first": float(L[:10].mean()), "loss_last": float(L[-50:].mean())},
        "heldout_poses": 16,
        "psnr_vs_sh_db": round(psnr_vs_sh, 3),
        "max_channel_err": round(maxerr, 4),
        "psnr_loss_db_at_reference_quality": round(loss_db, 4),
        "gate_db": 0.1,
        "gate_passed": bool(loss_db <= 0.1),
        "colour_path_used": "decoder" if loss_db <= 0.1 else "SH (fallback: gate failed)",
        "decoder_backward_mfma": {"gaussians_trained": trained, "flop": flop_per_pass, "ms": round(dec_ms, 5),
                                  "TFLOPs": round(flop_per_pass / (dec_ms * 1e-3) / 1e12, 2) if dec_ms else None,
                                  "peak_TFLOPs_bf16": 2500.0},
    }, dec


def random_views(rng, V, sizes_choice, cube: float):
    """V seeded equirectangular views: sizes drawn from sizes_choice in equal
    shares (each size V / len(sizes_choice) times, rounded, in a seeded
    shuffled order -- so that batches of different V hold the same mix and
    their views/s compare), random rotation, centre uniform in a cube of side
    `cube` m around the origin."""
    import numpy as np

    from omnigs_amd import scenes
    sizes, vms, cps = [], [], []
    draw = [sizes_choice[i % len(sizes_choice)] for i in range(V)]
    order = rng.permutation(V)
    for i in range(V):
        sizes.append(tuple(draw[order[i]]))
        R = random_rotation(rng)
        c = rng.uniform(-cube / 2, cube / 2, size=3)
        vms.append(scenes.viewmatrix_from_pose(R, -R @ c))
        cps.append(c.astype(np.float32))
    return sizes, np.stack(vms).astype(np.float32), np.stack(cps).astype(np.float32)


def batched_decoder_leg(sc, ds, dec, args, dev) -> dict:
    """The batched + learned-colour form on the timed workload: 8 views of
    1920x960 of the 1.5M-Gaussian scene per launch (poses in a 1 m cube), one
    captured graph per step (forward, fused L1 gradient, backward with the
    decoder backward), ms per view."""
    import numpy as np
    import torch

    from omnigs_amd.batch import build_launches, capture, make_step, time_graph

    P = sc["means3D"].shape[0]
    V = 8
    rng = np.random.default_rng(77)
    sizes, vms, cps = random_views(rng, V, [(1920, 960)], 1.0)
    inputs = dict(means3D=ds["means3D"], opacities=ds["opacities"], background=ds["background"], shs=ds["shs"],
                  scales=ds["scales"], rotations=ds["rotations"])
    target = torch.rand((3 * sum(w * h for w, h in sizes),), generator=torch.Generator(device=dev).manual_seed(3),
                        device=dev)
    out = {}
    for name, mlp in (("sh", None), ("decoder", dec.mlp_args())):
        rs, ts, lfs, dls = build_launches(P, sizes, (16, 16), V, inputs, torch.from_numpy(vms).to(dev),
                                          torch.from_numpy(cps).to(dev), target, dev, mlp)
        g = capture(make_step(rs, ts, lfs, dls), warmup=args.warmup)
        ms = time_graph(g, args.steps)
        assert not any(r.overflowed() for r in rs)
        inst = sum(r.needed()[0] for r in rs)
        # the compulsory-traffic model of one view (848 P + 124 R + 40 N bytes with SH,
        # 368 P + ... with the decoder's 16 BF16 latents), charged per view as if each view
        # were its own call, against HBM peak
        per_g = 848 if mlp is None else 368
        alg = sum(per_g * P + 40 * w * h for w, h in sizes) + 124 * inst
        out[name] = {"ms_per_step": round(ms, 3), "ms_per_view": round(ms / V, 4),
                     "views_per_sec": round(1000.0 * V / ms, 1), "instances": inst,
                     "roofline_per_view_model": {"bound": "hbm", "algorithmic_bytes": alg,
                                                 "achieved_GBps": round(alg / (ms * 1e-3) / 1e9, 1),
                                                 "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
        del g, rs, ts, lfs, dls
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    return {"workload": f"{V} views of 1920x960 per launch, 1.5M Gaussians, tile 16x16, hipGraph replay",
            **out, "colour_path_used": "decoder only if its quality gate passed (see mlp_colour)"}


def batch_workload(dev, args) -> int:
    """Batched multi-view fwd+bwd: 1.5M Gaussians with the small-scene
    distributions, V = 8, 16, 32 views with sizes drawn from {960x480,
    1920x960, 2048x1024} and poses in a 2 m cube; for each V the autotuner
    times every tile shape and views-per-launch as a captured graph and keeps
    the fastest.  Prints one JSON line."""
    import numpy as np
    import torch

    from omnigs_amd import scenes
    from omnigs_amd.batch import autotune

    P = args.gaussians
    sc = scenes.small_scene(P=P, width=64, height=32, seed=0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    inputs = dict(means3D=t(sc["means3D"]), opacities=t(sc["opacities"].reshape(-1)),
                  background=t(sc["background"]), shs=t(sc["shs"]), scales=t(sc["scales"]),
                  rotations=t(sc["rotations"]))
    choice = [(960, 480), (1920, 960), (2048, 1024)]
    rows = {}
    vpl_only = tuple(int(x) for x in args.batch_vpl.split(",") if x)
    for V in (int(x) for x in args.batch_views.split(",") if x):
        rng = np.random.default_rng(100 + V)
        sizes, vms, cps = random_views(rng, V, choice, 2.0)
        target = torch.rand((3 * sum(w * h for w, h in sizes),), generator=torch.Generator(device=dev).manual_seed(V),
                            device=dev)
        extra = {"shapes": ((16, 16),)} if vpl_only else {}
        res = autotune(P, sizes, inputs, t(vms), t(cps), target, device=dev, reps=max(3, args.steps // 4),
                       views_per_launch=tuple(v for v in (vpl_only or (1, 2, 4, 8, 16, 32)) if v <= V),
                       log=lambda r: print(json.dumps(r), file=sys.stderr, flush=True), **extra)
        pix = sum(w * h for w, h in sizes)
        rows[str(V)] = {"best_tile": f"{res.tile[0]}x{res.tile[1]}", "best_views_per_launch": res.views_per_launch,
                        "views_per_sec": round(res.views_per_sec, 1), "ms_per_view": round(res.ms_per_view, 4),
                        "pixels": pix, "gigapixels_per_sec": round(pix / (res.ms_per_view * V * 1e-3) / 1e9, 3),
                        "table": res.table}
    print(json.dumps({"metric": "batched fwd+bwd views/sec (autotuned)", "unit": "views/s",
                      "config": {"workload": "1.5M Gaussians, uneven equirectangular views, L1 step per launch",
                                 "gaussians": P, "sizes_from": ["960x480", "1920x960", "2048x1024"]},
                      "results": rows}))
    return 0


def memory_compare(dev, args) -> int:
    """Checkpointed vs stored backward through the LibTorch autograd Function
    (caller contract renderLonlat with activations), 3M Gaussians at
    4096x2048: ms per fwd+bwd iteration, peak HBM over the iteration and the
    bytes the render holds between its forward and its backward."""
    import numpy as np
    import torch

    from omnigs_amd import rasterizer, scenes

    P = args.gaussians if args.gaussians != 1_500_000 else 3_000_000
    W, H = 4096, 2048
    sc = scenes.room_scene(P=P, width=W, height=H)
    rng = np.random.default_rng(5)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)  # noqa: E731
    params = {
        "xyz": t(sc["means3D"]).requires_grad_(),
        "features_dc": t(sc["shs"][:, :1, :]).requires_grad_(),
        "features_rest": t(sc["shs"][:, 1:, :]).requires_grad_(),
        "scaling": t(np.log(sc["scales"])).requires_grad_(),
        "rotation": t(sc["rotations"] * rng.uniform(0.5, 2.0, size=(P, 1))).requires_grad_(),
        "opacity": t(np.log(sc["opacities"] / (1.0 - sc["opacities"]))).reshape(P, 1).requires_grad_(),
    }
    vm, cp, bg, target = t(sc["viewmatrix"]), t(sc["campos"]), t(sc["background"]), t(sc["target"])
    res = {}
    # checkpointed: the backward bins in bands of tile rows (the library's
    # default budget, R / 4 instances here); one_band: the same with a budget
    # above R, i.e. the whole binning at once
    modes = (("stored", False, 0), ("checkpointed", True, 0), ("checkpointed_one_band", True, 2**31 - 1),
             ("checkpointed_budget_8M", True, 8_000_000), ("checkpointed_budget_13M", True, 13_000_000))
    wanted = set(args.memory_modes.split(","))
    for mode, ck, budget in (m for m in modes if m[0] in wanted):

        def it():
            for p in params.values():
                p.grad = None
            base = torch.cuda.memory_allocated(dev)
            img, _, _, _ = rasterizer.render_lonlat(params, vm, cp, H, W, bg, checkpoint=ck, band_instances=budget)
            held = torch.cuda.memory_allocated(dev) - base
            (img - target).abs().mean().backward()
            return held

        for _ in range(2):
            it()
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        e = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
        held = 0
        for i in range(args.steps):
            e[i][0].record()
            held = it()
            e[i][1].record()
        torch.cuda.synchronize()
        res[mode] = {"ms_per_iter": round(statistics.median(a.elapsed_time(b) for a, b in e), 3),
                     "peak_hbm_bytes_over_params": torch.cuda.max_memory_allocated(dev) - base,
                     "held_fwd_to_bwd_bytes": held}
    print(json.dumps({"metric": "checkpointed vs stored backward, fwd+bwd ms/iter and HBM",
                      "config": {"workload": "autograd renderLonlat (activations included) + L1",
                                 "gaussians": P, "width": W, "height": H},
                      "steps": args.steps, **res}))
    return 0


def copy_calibration(dev, steps: int) -> int:
    """A plain device-to-device copy of 1 GiB: 1 GiB read + 1 GiB written per
    launch, the yardstick that scales rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = 1 << 28  # float32 elements = 1 GiB
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(max(1, steps)):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                      "launches": max(1, steps)}))
    return 0


def host_cpu() -> dict:
    """The host's CPU: model name, logical CPUs, and those this process may
    run on (its affinity mask) and is allotted (OMP_NUM_THREADS)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": omp or None}


def cpu_baseline(sc, repeats: int = 3) -> dict:
    """The FP32 oracle timed on the host over one full fwd+bwd iteration of the
    same workload: the OpenMP tile-parallel variant on every core this process
    is allotted (its affinity mask, capped by OMP_NUM_THREADS where the host
    sets one -- on a shared GPU host that is the CPU share of one GPU), median
    of ``repeats`` iterations, as the reported baseline; and the single-thread
    line-for-line port beside it (one iteration; about 20 s in all)."""
    import numpy as np

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import Oracle

    target = sc["target"]
    sc = {k: v for k, v in sc.items() if k != "target"}
    cpu = host_cpu()
    want = cpu["affinity"] if not cpu["omp_num_threads"] else min(cpu["affinity"], cpu["omp_num_threads"])

    def one(threads):
        o = Oracle(sc, 32)
        used = o.set_threads(threads)
        t = time.perf_counter()
        o.forward()
        img = o.get("out_color").reshape(3, sc["height"], sc["width"])
        dL = (np.sign(img - target) / (3.0 * sc["height"] * sc["width"])).astype(np.float32)
        o.backward(dL)
        return time.perf_counter() - t, used

    runs = [one(want) for _ in range(max(1, repeats))]
    used = runs[0][1]
    dt_par = statistics.median(r[0] for r in runs)
    dt_1, _ = one(1)
    return {"value": round(dt_par * 1000.0, 1), "unit": "ms/iter", "cores": used, "kind": "port",
            "sample": f"full fwd+bwd iterations (1.5M Gaussians, 1920x960) of the FP32 oracle, OpenMP "
                      f"tile-parallel with deterministic reductions on {used} host threads, median of "
                      f"{len(runs)}",
            "host": cpu,
            "single_thread": {"value": round(dt_1 * 1000.0, 1), "unit": "ms/iter", "cores": 1}}


if __name__ == "__main__":
    sys.exit(main())
