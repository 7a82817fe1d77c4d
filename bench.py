"""Benchmark: rays/sec of one 800x800 view of the 100k-voxel value-noise scene
(128^3 grid, SVO intersection with max_hits 60, inverse-CDF sampling at step
voxel/8, 64-sample hierarchical pass, field 32 -> 3x128 -> 4) on one MI355X,
with p50/p99 view latency, the plain FP32 GPU path and the CPU oracle timed
beside it, and the roofline of the dominant kernel.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--plain-views V]
                  [--cpu-rays C] [--no-plain] [--no-cpu]

Prints one JSON line.  A step is one view: rays already on the device to the
RGB, depth and alpha image.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/sec @800x800 (intersect+sample+render) vs plain FP32; p50/p99 view latency"
PEAK_BF16_TFLOPS = 2500.0     # MI355X dense BF16 MFMA
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--res", type=int, default=800)
    ap.add_argument("--voxels", type=int, default=100_000)
    ap.add_argument("--plain-views", type=int, default=2)
    ap.add_argument("--cpu-rays", type=int, default=24000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--camera-seed", type=int, default=11)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first copy a buffer of this many MiB (a known-size kernel for "
                         "calibrating PMC byte counters)")
    return ap.parse_args()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the PMC runs (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated), recorded in profiles/traffic.json; None
    when that measurement is absent."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from nsvf_amd import scene as S
    from nsvf_amd.render import DeviceScene, Renderer

    dev = torch.device("cuda:0")
    if args.calib_copy_mib > 0:
        src = torch.ones(args.calib_copy_mib * (1 << 20) // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t0 = time.time()
    sc = S.noise_scene(res=128, num_voxels=args.voxels)
    cam = S.sphere_camera(2.5, args.camera_seed)
    rs_np, rd_np = S.camera_rays(args.res, args.res, cam)
    R = rs_np.shape[0]
    dsc = DeviceScene(sc, dev)
    rs = torch.from_numpy(rs_np).to(dev)
    rd = torch.from_numpy(rd_np).to(dev)
    log(f"scene {sc.num_voxels} voxels, {sc.num_keys} keys, {dsc.num_nodes} octree nodes "
        f"({time.time() - t0:.1f}s setup)")

    # ---- fused BF16 path
    rnd = Renderer(dev, precision="bf16", hierarchical=True, fine_steps=64.0)
    for _ in range(args.warmup):
        rnd.render(dsc, rs, rd)
    torch.cuda.synchronize()
    kev = rnd.record_kernels(True)
    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        rnd.render(dsc, rs, rd)
        e1.record()
        lat.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    rnd.record_kernels(False)
    lat_ms = np.array([a.elapsed_time(b) for a, b in lat])
    ms_per_step = wall * 1000.0 / args.steps
    value = R / (wall / args.steps)
    st = rnd.last
    # exact sample counts of the view (identical for every repetition)
    rnd.render(dsc, rs, rd, keep=True)
    n_coarse, n_fine = rnd.count_samples()
    rnd.kept = None
    log(f"fused bf16: {ms_per_step:.2f} ms/view, {value / 1e6:.2f} Mrays/s; hit rays {st.hit_rays}, "
        f"coarse {n_coarse} samples ({st.coarse_slots} slots), fine {n_fine} samples "
        f"({st.fine_slots} slots allocated)")

    # roofline of the dominant kernel: the BF16 MLP kernel (MFMA-bound), with
    # algorithmic FLOPs of the samples it evaluates / its HIP-event time; the
    # gather kernel feeding it reported beside it (HBM/L2-bound)
    flops_per_sample = sc.mlp.flops_per_sample()
    n_views = args.steps
    mlp_ev = [(m, a, b) for k, m, a, b in kev if k == "mlp"]
    emb_ev = [(m, a, b) for k, m, a, b in kev if k == "embed"]
    mlp_ms = sum(a.elapsed_time(b) for _, a, b in mlp_ev)
    emb_ms = sum(a.elapsed_time(b) for _, a, b in emb_ev)
    samples = (n_coarse + n_fine) * n_views
    achieved = flops_per_sample * samples / (mlp_ms * 1e-3) / 1e12
    D = dsc.D
    emb_bytes = samples * (16 + 2 * D) + n_views * 2 * (
        sc.num_voxels * 44 + sc.num_keys * 4 * D)
    roofline = {"bound": "mfma", "kernel": "mlp_bf16_kernel<32,128,3>",
                "achieved": round(achieved, 2), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_BF16_TFLOPS, 4),
                "traffic": measured_traffic("mlp_bf16_kernel"),
                "avg_launch_ms": round(mlp_ms / max(len(mlp_ev), 1), 4),
                "flops_per_sample": flops_per_sample,
                "samples_per_launch": round(samples / max(len(mlp_ev), 1))}
    kernels = {"embed_bf16_kernel": {"avg_launch_ms": round(emb_ms / max(len(emb_ev), 1), 4),
                                     "achieved_GBs": round(emb_bytes / (emb_ms * 1e-3) / 1e9, 1),
                                     "bound": "hbm/l2"}}

    # per-stage breakdown of one traced view (outside the timed region)
    rnd.trace = []
    rnd.render(dsc, rs, rd)
    torch.cuda.synchronize()
    stages = {}
    for (n0, e0), (n1, e1) in zip(rnd.trace[:-1], rnd.trace[1:]):
        stages[n1] = round(e0.elapsed_time(e1), 3)
    rnd.trace = None

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "rays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"noise{128}: 128^3 grid, {sc.num_voxels} voxels, SVO max_hits 60, "
                               f"inverse-CDF step voxel/8 + 64 hierarchical samples, "
                               f"{args.res}x{args.res} view, field 32->3x128->4 BF16 MFMA",
                   "rays_per_view": R, "voxels": sc.num_voxels, "octree_nodes": dsc.num_nodes,
                   "parallelism": "single GPU"},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "samples_per_view": {"hit_rays": st.hit_rays, "coarse": n_coarse, "fine": n_fine,
                             "coarse_slots": st.coarse_slots},
        "stages_ms": stages,
        "roofline": roofline,
        "kernels": kernels,
    }

    # ---- plain FP32 path (the comparator)
    if not args.no_plain and args.plain_views > 0:
        from nsvf_amd.plain import PlainNSVF
        pl = PlainNSVF(sc, dev)
        prgb, _, _ = pl.render(rs, rd)   # warm-up (also builds the octree cache)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.plain_views):
            prgb, _, _ = pl.render(rs, rd)
        torch.cuda.synchronize()
        pms = (time.perf_counter() - t1) * 1000.0 / args.plain_views
        frgb, _, _ = rnd.render(dsc, rs, rd)
        mse = float(torch.mean((frgb.double() - prgb.double()) ** 2))
        out["plain_fp32"] = {"rays_per_sec": round(R / (pms / 1000.0), 1), "ms_per_view": round(pms, 2),
                             "views": args.plain_views,
                             "psnr_fused_vs_plain_db": round(10 * math.log10(1.0 / mse), 2) if mse > 0 else None}
        out["speedup_vs_plain_fp32"] = round(value / (R / (pms / 1000.0)), 2)
        log(f"plain fp32: {pms:.1f} ms/view; speedup {out['speedup_vs_plain_fp32']}x")
        del pl
        torch.cuda.empty_cache()

    # ---- CPU baseline: the oracle on a bounded ray sample of the same view
    if not args.no_cpu and args.cpu_rays > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import nsvf_oracle as O
        cores = len(os.sched_getaffinity(0))
        rng = np.random.default_rng(0)
        sel = np.sort(rng.choice(R, size=min(args.cpu_rays, R), replace=False))
        t2 = time.perf_counter()
        O.render_view(sc, rs_np[sel], rd_np[sel], dtype=np.float32)
        cpu_s = time.perf_counter() - t2
        out["cpu_baseline"] = {"value": round(len(sel) / cpu_s, 1), "unit": "rays/s", "cores": cores,
                               "kind": "port",
                               "sample": f"{len(sel)} seeded random rays of the {args.res}x{args.res} "
                                         f"view, octree build + intersection + sampling (C++ "
                                         f"restatement, 1 thread) + gather/MLP/compositing (NumPy "
                                         f"FP32), {cpu_s:.1f} s"}
        log(f"cpu oracle: {len(sel) / cpu_s:.0f} rays/s on {cores} cores")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
