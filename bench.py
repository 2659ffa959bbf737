"""Forward+backward rays/s of the plain renderer on one MI355X.

Workload (the plain full-frame configuration): 128^3 quadric grid of a
torus (R=1, r=0.35 world units, + N(0, 0.005) noise), SH degree 2
(32 appearance channels, coefficients N(0, 0.1)), 100 camera poses on a
radius-4 sphere, FOV 40 deg, 800x800, 192 uniform samples per ray, logistic
coefficient 5.  One step = one pose: ray generation + forward + MSE loss
against a random target + backward into zeroed gradient grids, over all
640,000 pixel rays.  Synthetic data, random-init grids.

Prints ONE JSON line (rank 0).  Extra fields: the op-surface path's rays/s,
the proposal-guided figure when its network is available, the dominant
kernel's roofline and the CPU oracle baseline.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("fwd+bwd rays/s at 128³ grid 800×800, plain vs proposal at "
          "≤0.1 dB PSNR loss")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--op-steps", type=int, default=2, help="steps of the op-surface path (0: skip)")
    ap.add_argument("--cpu-rays", type=int, default=65536,
                    help="rays in the CPU oracle sample (0: skip the CPU baseline)")
    ap.add_argument("--cpu-literal-rays", type=int, default=200,
                    help="rays of the 64^3 / 100x100 workload timed through the literal per-ray "
                         "compositing loop (0: skip)")
    ap.add_argument("--scatter", type=int, default=0, help="fused scatter mode (0 tile, 1 atomic)")
    ap.add_argument("--proposal-s", type=str, default="40,5",
                    help="logistic coefficients at which the proposal path is measured "
                         "(comma list; empty: skip)")
    ap.add_argument("--proposal-train-steps", type=int, default=4000)
    ap.add_argument("--train-steps", type=int, default=10,
                    help="timed steps of the 256^3 random-ray training-step form per scatter "
                         "mode (0: skip)")
    ap.add_argument("--large-steps", type=int, default=5,
                    help="timed steps of the 384^3 / SH 3 / 1920x1080 plain-vs-proposal "
                         "comparison (0: skip)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run a plain device copy of this many bytes first (a counter "
                         "calibration kernel of known traffic for rocprofv3 --pmc runs)")
    return ap.parse_args()


def load_profile_record(name):
    """A JSON record of rocprofv3 counter passes kept in profiles/ (None if
    absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def measured_traffic():
    """HBM bytes per launch of the fused backward on this workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counter passes recorded next to the
    kernel-trace summaries (None when the record is absent)."""
    path = os.path.join(ROOT, "profiles", "fused_backward_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def measured_utilization():
    """Issue-rate figures of the fused kernels (VALU busy fraction, MFMA
    rate) from the rocprofv3 SQ counter passes recorded next to the
    kernel-trace summaries (None when the record is absent)."""
    path = os.path.join(ROOT, "profiles", "fused_utilization.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def algorithmic_bytes(R, C, P):
    """Compulsory HBM bytes of one fused backward launch over P rays: read the
    per-ray colour, depth and their gradients (32 B), read both grids once,
    read-modify-write both gradient grids once."""
    G = 28 * R ** 3 + 4 * C * (R + 1) ** 3
    return 32 * P + 3 * G


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "only one GPU is measured here; --gpus must be 1 (the path "
                                   "would shard by pose with replicated grids and one gradient "
                                   "all-reduce per step)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from qsdvr_amd import _lib, fused, proposal, scenes
    from qsdvr_amd import render as hr
    from qsdvr_amd import qsdvr
    from qsdvr_amd.camera import Camera

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = "cuda"
    w = scenes.WORKLOADS["plain"]
    t0 = time.time()
    grid, data, gen = scenes.build_grids(w, seed=0, device=dev)
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    cam_buf = fused.camera_buffer(cams, dev)
    P = w.width * w.height
    targets = torch.rand((w.n_poses, P, 3), generator=torch.Generator().manual_seed(1)).to(dev)
    log(f"setup {time.time() - t0:.1f}s")
    if args.copy_bytes > 0:
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        ce = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ce[0].record()
        for i in range(3):
            dst.copy_(src)
            ce[i + 1].record()
        torch.cuda.synchronize()
        copy_ms = min(ce[i].elapsed_time(ce[i + 1]) for i in range(3))
        log(f"calibration copy: {args.copy_bytes} bytes read + written, 3 launches; fastest "
            f"{copy_ms:.3f} ms = {2 * args.copy_bytes / copy_ms / 1e6:.0f} GB/s (read + write)")
        del src, dst
    G = grid.requires_grad_(True)
    D = data.requires_grad_(True)
    CAM_BYTES = 80

    def fused_step(i):
        p = i % w.n_poses
        G.grad = None
        D.grad = None
        cb = cam_buf[p * CAM_BYTES:(p + 1) * CAM_BYTES]
        c, _ = fused.FusedRender.apply(G, D, cb, 1, w.width, w.height, None, w.interval, w.s,
                                       args.scatter)
        loss = ((c - targets[p]) ** 2).mean()
        loss.backward()
        return loss

    for i in range(args.warmup):
        fused_step(i)
    torch.cuda.synchronize()
    fused.TIMERS = []
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for i in range(args.steps):
        fused_step(args.warmup + i)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    timers, fused.TIMERS = fused.TIMERS, None
    ms_step = ev0.elapsed_time(ev1) / args.steps
    rays_s = P / (ms_step / 1e3)
    kern = {}
    for name, a, b in timers:
        kern.setdefault(name, []).append(a.elapsed_time(b))
    kern_ms = {k: statistics.mean(v) for k, v in kern.items()}
    log(f"fused: {ms_step:.3f} ms/step, {rays_s:.4g} rays/s (wall {wall:.2f}s); kernels {kern_ms}")

    # ------------------------------------------------ op-surface path ----
    op_rays = None
    if args.op_steps > 0:
        G2 = grid.detach().clone().requires_grad_(True)
        D2 = data.detach().clone().requires_grad_(True)
        cam_objs = []
        for p in range(args.op_steps + 1):
            a = qsdvr.CameraArgs.from_camera(cams[p])
            c = Camera(a, w.interval, w.R)
            c.GenerateRayPoints(None)
            cam_objs.append(c)

        def op_step(i):
            G2.grad = None
            D2.grad = None
            inp = cam_objs[i].GetRayPoints()
            out = hr.render_lists(G2, D2, inp, w.s)
            ((out - targets[i]) ** 2).mean().backward()

        # untimed warm-up over the timed poses themselves: the caching
        # allocator then holds blocks for each pose's sample count (a pose
        # with more samples than any before it otherwise pays device
        # allocations inside the timed step)
        for i in range(args.op_steps + 1):
            op_step(i)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.op_steps + 1)]
        ev[0].record()
        for i in range(args.op_steps):
            op_step(1 + i)
            ev[i + 1].record()
        torch.cuda.synchronize()
        op_each = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.op_steps)]
        op_ms = ev[0].elapsed_time(ev[-1]) / args.op_steps
        op_rays = P / (op_ms / 1e3)
        op_renders = sum(int(cam_objs[1 + i].args.renderCount) for i in range(args.op_steps))
        log(f"op-surface: {op_ms:.3f} ms/step, {op_rays:.4g} rays/s; steps "
            f"{[round(t, 2) for t in op_each]}, render samples "
            f"{[int(cam_objs[1 + i].args.renderCount) for i in range(args.op_steps)]}")
        del G2, D2, cam_objs
        torch.cuda.empty_cache()

    # ---------------------------------------------------- proposal path ----
    prop_results = []
    for s_txt in [x for x in args.proposal_s.split(",") if x.strip()]:
        prop_results.append(measure_proposal(w, grid.detach(), data.detach(), cams, cam_buf,
                                             targets, float(s_txt), args))
        log(f"proposal: {prop_results[-1]}")

    # ------------------------------------------ training-step form ----
    train = None
    if args.train_steps > 0:
        train = measure_training_step(args)
        log(f"training step: {train}")

    large = None
    if args.large_steps > 0:
        large = measure_large(args)
        log(f"large: {large}")

    # ------------------------------------------------------- roofline ----
    # The dominant kernel against its two rooflines, with the per-launch counts
    # from the rocprofv3 counter passes recorded in profiles/: HBM (measured
    # FETCH/WRITE bytes) and float32 compute (executed VALU and MFMA float32
    # operations).  `bound` is the roofline it runs closer to.  For "hbm",
    # `achieved` is its algorithmic bytes per launch over this run's launch
    # time; for "mfma" (float32 compute: matrix cores and VALU), its executed
    # float32 FLOPs per launch over that time.
    dom = max(kern_ms, key=kern_ms.get)
    t_launch = kern_ms[dom] / 1e3
    byts = algorithmic_bytes(w.R, 32, P) if dom == "fused_backward" else (
        28 * w.R ** 3 + 4 * 32 * (w.R + 1) ** 3 + 16 * P)
    traffic = measured_traffic()
    util = measured_utilization() or {}
    flops = (util.get(dom) or {}).get("fp32_flops_per_launch")
    traffic_frac = None
    if dom == "fused_backward" and traffic is not None:
        traffic_frac = traffic["bytes_per_launch"] / t_launch / 1e9 / HBM_PEAK_GBS
    compute_frac = flops / t_launch / 1e12 / FP32_PEAK_TFLOPS if flops else None
    bound = "mfma" if (compute_frac is not None and traffic_frac is not None
                       and compute_frac > traffic_frac) else "hbm"
    if bound == "hbm":
        achieved, peak, unit = byts / t_launch / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        achieved, peak, unit = flops / t_launch / 1e12, FP32_PEAK_TFLOPS, "TFLOP/s"
    roofline = {"kernel": dom, "bound": bound, "achieved": round(achieved, 2), "peak": peak,
                "unit": unit, "frac": round(achieved / peak, 5),
                "traffic": traffic["bytes_per_launch"] if traffic_frac is not None else None,
                "traffic_frac": round(traffic_frac, 5) if traffic_frac is not None else None,
                "algorithmic_bytes_per_launch": byts,
                "algorithmic_hbm_frac": round(byts / t_launch / 1e9 / HBM_PEAK_GBS, 5),
                "fp32_flops_per_launch": flops,
                "fp32_frac": round(compute_frac, 5) if compute_frac is not None else None,
                "avg_launch_ms": round(kern_ms[dom], 4)}
    if traffic_frac is not None:
        roofline["traffic_source"] = traffic["source"]
    if flops:
        roofline["flops_source"] = util.get("source")

    # ------------------------------------------ op-surface HBM fraction ----
    # The op-surface path (the reference's 14 ops, each reading its inputs
    # and writing its outputs once) moves B_op = (344 + 20 C) s_r + 150 + 4 G/P
    # bytes per ray, s_r the render samples per ray of the timed poses (from
    # RayCut's totals), G the bytes of both grids; its HBM rate is its own
    # rays/s times B_op.
    op_surface = None
    if op_rays:
        s_r = op_renders / (args.op_steps * P)
        G_bytes = 28 * w.R ** 3 + 4 * 32 * (w.R + 1) ** 3
        b_op = (344 + 20 * 32) * s_r + 150 + 4 * G_bytes / P
        op_surface = {"rays_per_s": round(op_rays, 1), "ms_per_step": round(op_ms, 3),
                      "render_samples_per_ray": round(s_r, 2),
                      "algorithmic_bytes_per_ray": round(b_op, 1),
                      "achieved_GBs": round(b_op * op_rays / 1e9, 1),
                      "frac": round(b_op * op_rays / 1e9 / HBM_PEAK_GBS, 4)}
        # measured HBM bytes of one op-surface step (rocprofv3 FETCH/WRITE
        # passes of all its kernels) over this run's step time
        rec = load_profile_record("op_surface_traffic.json")
        if rec is not None:
            op_surface["traffic_bytes_per_step"] = rec["bytes_per_step"]
            op_surface["traffic_GBs"] = round(rec["bytes_per_step"] / (op_ms / 1e3) / 1e9, 1)
            op_surface["traffic_frac"] = round(op_surface["traffic_GBs"] / HBM_PEAK_GBS, 4)
            op_surface["traffic_source"] = rec["source"]

    # --------------------------------------------------- CPU baseline ----
    cpu = None
    if args.cpu_rays > 0:
        cpu = cpu_baseline(w, cams[0], grid.detach().cpu(), data.detach().cpu(),
                           targets[0].cpu(), args.cpu_rays, args.cpu_literal_rays)

    out = {"metric": METRIC, "value": round(rays_s, 1), "unit": "rays/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (torus SDF quadric grid, random appearance, random targets)",
           "config": {"workload": w.name, "grid": w.R, "sh_degree": w.sh_degree,
                      "width": w.width, "height": w.height, "samples": w.n_samples,
                      "poses": w.n_poses, "logistic_coef": w.s, "path": "fused plain",
                      "scatter_mode": args.scatter},
           "library": _lib.loaded_path(),
           "op_surface_rays_per_s": round(op_rays, 1) if op_rays else None,
           "op_surface": op_surface,
           "proposal": prop_results,
           "kernel_ms": {k: round(v, 4) for k, v in kern_ms.items()},
           "training_step": train, "large": large,
           "roofline": roofline, "utilization": measured_utilization(),
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def measure_proposal(w, grid, data, cams, cam_buf, targets, s, args, K=32, steps=None,
                     extra_gammas=()):
    """Proposal-guided vs plain at logistic coefficient s on the same scene:
    train the proposal on the first 90% of the poses, then time plain and
    proposal forward+backward steps on the rest and compare PSNR on the last
    5 poses against a dense plain render with 8x the samples (GT)."""
    import torch

    from qsdvr_amd import fused, proposal

    P = w.width * w.height
    n = len(cams)
    n_train = (9 * n) // 10
    held = n - n_train
    steps = steps or args.steps
    prop = proposal.Proposal(seed=0)
    t0 = time.perf_counter()
    proposal.train(prop, grid, data, cams[:n_train], w.interval, s,
                   steps=args.proposal_train_steps, batch=8192, lr=1e-2)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    G = grid.clone().requires_grad_(True)
    D = data.clone().requires_grad_(True)
    CB = 80

    def plain(p):
        G.grad = None
        D.grad = None
        c, _ = fused.FusedRender.apply(G, D, cam_buf[p * CB:(p + 1) * CB], 1, w.width, w.height,
                                       None, w.interval, s, args.scatter)
        ((c - targets[p]) ** 2).mean().backward()

    def guided(p, gamma=proposal.SAMPLE_GAMMA):
        G.grad = None
        D.grad = None
        c, _ = proposal.render_proposal(G, D, cam_buf[p * CB:(p + 1) * CB], 1, w.width, w.height,
                                        w.interval, s, prop, K=K, gamma=gamma)
        ((c - targets[p]) ** 2).mean().backward()

    out = {"logistic_coef": s, "fine_samples": K, "coarse_bins": 64,
           "sampler_eps": proposal.SAMPLE_EPS, "sampler_gamma": proposal.SAMPLE_GAMMA,
           "train_steps": args.proposal_train_steps, "train_seconds": round(train_s, 2)}
    runs = [("plain", plain), ("proposal", guided)]
    for g in extra_gammas:  # sharper samplers: same kernels, timed as well
        runs.append((f"proposal_gamma{g:g}", lambda p, g=g: guided(p, g)))
    for name, fn in runs:
        for i in range(2):
            fn(n_train + i)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for i in range(steps):
            fn(n_train + (i % held))
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out[name + "_ms_per_step"] = round(ms, 3)
        out[name + "_rays_per_s"] = round(P / (ms / 1e3), 1)
    out["speedup"] = round(out["plain_ms_per_step"] / out["proposal_ms_per_step"], 3)
    for g in extra_gammas:
        out[f"speedup_gamma{g:g}"] = round(out["plain_ms_per_step"] /
                                           out[f"proposal_gamma{g:g}_ms_per_step"], 3)
    with torch.no_grad():
        test = cams[n - 5:]
        tb = fused.camera_buffer(test)
        gt, _ = fused.render(grid, data, test, w.interval / 8, s)
        pl, _ = fused.render(grid, data, test, w.interval, s)
        pr, _ = proposal.render_proposal(grid, data, tb, len(test), w.width, w.height, w.interval,
                                         s, prop, K=K)
        out["psnr_plain_vs_gt_db"] = round(proposal.psnr(pl, gt), 3)
        out["psnr_proposal_vs_gt_db"] = round(proposal.psnr(pr, gt), 3)
        out["psnr_loss_db"] = round(out["psnr_plain_vs_gt_db"] - out["psnr_proposal_vs_gt_db"], 3)
        out["psnr_proposal_vs_plain_db"] = round(proposal.psnr(pr, pl), 3)
        for g in extra_gammas:
            pg, _ = proposal.render_proposal(grid, data, tb, len(test), w.width, w.height,
                                             w.interval, s, prop, K=K, gamma=g)
            out[f"psnr_loss_db_gamma{g:g}"] = round(out["psnr_plain_vs_gt_db"] -
                                                   proposal.psnr(pg, gt), 3)
        out["gt"] = (f"plain renderer at 8x samples ({8 * w.n_samples} per diagonal), "
                     f"poses {n - 5}-{n - 1}")
    del G, D
    torch.cuda.empty_cache()
    return out


def measure_large(args, s=40.0, K=48):
    """The large-scale form: 384^3 noisy gyroid shell, SH degree 3 (53
    channels), 50 poses at 1920x1080; plain renderer with 384 samples per
    diagonal against the proposal with 48 fine samples, forward+backward,
    trained and timed as the 128^3 proposal comparison."""
    import torch

    from qsdvr_amd import fused, scenes

    w = scenes.WORKLOADS["large"]
    t0 = time.time()
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    cam_buf = fused.camera_buffer(cams)
    P = w.width * w.height
    tg = torch.Generator(device="cuda").manual_seed(9)
    targets = {p: torch.rand((P, 3), generator=tg, device="cuda")
               for p in range((9 * w.n_poses) // 10, w.n_poses)}
    log(f"large setup {time.time() - t0:.1f}s")
    out = {"workload": w.name, "grid": w.R, "sh_degree": w.sh_degree, "channels": 53,
           "width": w.width, "height": w.height, "plain_samples": w.n_samples}
    out.update(measure_proposal(w, grid, data, cams, cam_buf, targets, s, args, K=K,
                                steps=args.large_steps, extra_gammas=(2.0,)))
    del grid, data, targets
    torch.cuda.empty_cache()
    return out


def measure_training_step(args):
    """The training-step form: 256^3 union of 8 random spheres, SH degree 2,
    each step a random batch of 65,536 rays over 100 poses at 800x800 with
    N=256; step = draw the batch + fused forward + MSE + adjoint into zeroed
    gradient grids.  Timed per scatter mode: deterministic (sorted-segment
    sums) and atomic (one atomic per sample and channel); the tile mode is
    added for reference (its wave-level reduction needs coherent rays, which
    a random batch does not have)."""
    import torch

    from qsdvr_amd import fused, scenes

    w = scenes.WORKLOADS["train"]
    t0 = time.time()
    grid, data, gen = scenes.build_grids(w, seed=0, device="cuda")
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    cam_buf = fused.camera_buffer(cams)
    log(f"training-step setup {time.time() - t0:.1f}s")
    B = 65536
    n_all = w.n_poses * w.width * w.height
    G = grid.requires_grad_(True)
    D = data.requires_grad_(True)
    out = {"workload": w.name, "grid": w.R, "sh_degree": w.sh_degree, "rays_per_step": B,
           "poses": w.n_poses, "samples": w.n_samples, "logistic_coef": w.s}
    names = {fused.SCATTER_DET: "deterministic", fused.SCATTER_ATOMIC: "atomic",
             fused.SCATTER_TILE: "tile"}
    for mode in (fused.SCATTER_DET, fused.SCATTER_ATOMIC, fused.SCATTER_TILE):
        g = torch.Generator(device="cuda").manual_seed(7)
        tg = torch.Generator(device="cuda").manual_seed(8)

        def step():
            G.grad = None
            D.grad = None
            pick = torch.randint(0, n_all, (B,), generator=g, device="cuda").to(torch.int32)
            target = torch.rand((B, 3), generator=tg, device="cuda")
            c, _ = fused.FusedRender.apply(G, D, cam_buf, w.n_poses, w.width, w.height, pick,
                                           w.interval, w.s, mode)
            ((c - target) ** 2).mean().backward()

        for _ in range(2):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.train_steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.train_steps
        out[names[mode]] = {"ms_per_step": round(ms, 3), "rays_per_s": round(B / (ms / 1e3), 1)}
    # the whole GPU-resident step: the deterministic adjoint, then Adam (HIP)
    # on the quadric and appearance grids
    from qsdvr_amd.train import FusedTrainer

    tr = FusedTrainer(D.detach(), cam_buf, w.n_poses, w.width, w.height, w.interval, w.s,
                      sdf_grid=G.detach(), lr=1e-3, scatter_mode=fused.SCATTER_DET, seed=7)
    tg = torch.Generator(device="cuda").manual_seed(8)

    def full_step():
        rays = tr.draw(B)
        tr.step(rays, torch.rand((B, 3), generator=tg, device="cuda"))

    for _ in range(2):
        full_step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.train_steps):
        full_step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.train_steps
    out["deterministic_with_adam"] = {"ms_per_step": round(ms, 3),
                                      "rays_per_s": round(B / (ms / 1e3), 1)}
    del tr, G, D, grid, data
    torch.cuda.empty_cache()
    return out


def host_cores():
    """(threads to use, description): the CPUs this process may run on
    (``os.sched_getaffinity``), capped by OMP_NUM_THREADS when it is set,
    and lscpu's model / socket / core summary."""
    import subprocess

    avail = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    use = min(avail, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else avail
    desc = {"affinity_cpus": avail, "omp_num_threads": cap}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket",
                             "Thread(s) per core", "CPU(s)"):
                desc[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        desc["lscpu"] = "unavailable"
    return use, desc


def cpu_baseline(w, cam, grid, data, target, n_rays, n_literal):
    """The CPU oracle (render_torch restated; torch CPU, float32,
    vectorised) timed on a fixed subset of one pose's rays: ray generation
    (RayCut + GenerateRayPoints restated) + forward + MSE + backward.  Also
    the 64^3 / 100x100 / N=128 workload in full, and the literal per-ray
    compositing loop of render_torch.py:155-167 on a sample of its rays."""
    import numpy as np
    import torch

    from oracle import raygen
    from oracle import render as orc
    from qsdvr_amd import scenes

    threads, cores = host_cores()
    torch.set_num_threads(threads)
    interval = np.float32(w.interval)
    P = w.width * w.height
    sel = np.linspace(0, P - 1, n_rays).astype(np.int64)
    tgt = target[torch.as_tensor(sel)]
    G = grid.clone().requires_grad_(True)
    D = data.clone().requires_grad_(True)
    t0 = time.perf_counter()
    smp = raygen.ray_samples(cam.as_dict(), interval, w.R, pixels=sel)
    col = orc.render(G, D, smp, w.R, w.s)
    orc.mse_loss(col, tgt).backward()
    dt = time.perf_counter() - t0
    out = {"value": round(n_rays / dt, 1), "unit": "rays/s", "cores": threads, "kind": "port",
           "sample": f"{n_rays} rays (every {P // n_rays}th pixel of pose 0 of "
                     f"{w.name}), {smp['renderCount']} render samples; ray generation + "
                     f"fwd + MSE + bwd, float32, torch CPU",
           "seconds": round(dt, 2), "host": cores}
    # the 64^3 sphere / SH 1 / 100x100 / N=128 workload, vectorised, in full
    ws = scenes.WORKLOADS["small"]
    g1, d1, gen = scenes.build_grids(ws, seed=0)
    cam1 = scenes.orbit_cameras(1, ws.width, ws.height)[0]
    t1 = torch.rand((ws.width * ws.height, 3), generator=gen)
    G1 = g1.clone().requires_grad_(True)
    D1 = d1.clone().requires_grad_(True)
    t0 = time.perf_counter()
    s1 = raygen.ray_samples(cam1.as_dict(), np.float32(ws.interval), ws.R)
    c1 = orc.render(G1, D1, s1, ws.R, ws.s)
    orc.mse_loss(c1, t1).backward()
    dt1 = time.perf_counter() - t0
    out["small_vectorised"] = {"rays_per_s": round(ws.width * ws.height / dt1, 1),
                               "seconds": round(dt1, 3), "workload": ws.name}
    if n_literal > 0:
        # literal per-ray, per-sample loop (render_torch.py:155-167 with each
        # ray's colour in its own row), forward + backward, on evenly spaced rays
        pick = np.linspace(0, ws.width * ws.height - 1, n_literal).astype(np.int64)
        G1 = g1.clone().requires_grad_(True)
        D1 = d1.clone().requires_grad_(True)
        t0 = time.perf_counter()
        sp = raygen.ray_samples(cam1.as_dict(), np.float32(ws.interval), ws.R, pixels=pick)
        sdf = orc.sample_sdf(G1, torch.as_tensor(sp["sdfPointList"]),
                             torch.as_tensor(sp["sdfIndexList"]), ws.R)
        rp = torch.as_tensor(sp["renderPointList"])
        ri = torch.as_tensor(sp["renderIndexList"])
        rgb, _ = orc.shade(orc.sample_normal(G1, rp, ri), torch.as_tensor(sp["viewDirList"]),
                           orc.grid_interpolation(D1, rp, ri, ws.R))
        cl = orc.composite_literal(rgb, sdf, sp["rayList"], ws.s)
        orc.mse_loss(cl, t1[torch.as_tensor(pick)]).backward()
        dtl = time.perf_counter() - t0
        out["small_literal_loop"] = {"rays_per_s": round(n_literal / dtl, 2),
                                     "seconds": round(dtl, 2), "rays": n_literal,
                                     "workload": ws.name}
    return out


if __name__ == "__main__":
    main()
