"""Benchmark: rays/sec of one 800x800 view of the 100k-voxel value-noise scene
(128^3 grid, SVO intersection with max_hits 60, inverse-CDF sampling at step
voxel/8, 64-sample hierarchical pass, field 32 -> 3x128 -> 4) on one MI355X,
with p50/p99 view latency, the plain FP32 GPU path and the CPU oracle timed
beside it, and the roofline of the dominant kernel.

The headline runs the field at the fastest precision whose per-pixel RGB
error against the FP64 oracle stays within 1e-3 on this view: split BF16
("bf16x3": three BF16 MFMAs per product on hi / lo operand parts, FP32
accumulation).  Plain BF16 MFMA and FP32 MFMA are timed and checked beside it.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--plain-views V]
                  [--cpu-rays C] [--check-rays A] [--no-plain] [--no-cpu]

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
PEAK_FP32_TFLOPS = 157.3      # MI355X dense FP32 MFMA
PEAK_HBM_GBS = 8000.0
HEADLINE = "bf16x3"
RGB_BAR = 1e-3                # per-pixel RGB error target against FP64
GRAZE_EPS = 1e-6              # a hit shorter than this grazes a voxel face


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--res", type=int, default=800)
    ap.add_argument("--voxels", type=int, default=100_000)
    ap.add_argument("--plain-views", type=int, default=2)
    ap.add_argument("--cpu-rays", type=int, default=6000,
                    help="rays of the view in each timed CPU-baseline run (median of --cpu-reps)")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--check-rays", type=int, default=3000,
                    help="random rays of the view checked against the FP64 oracle (plus every "
                         "ray whose hit row is full)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-headline", action="store_true",
                    help="time the headline precision only (for profiling runs)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the from-camera and voxel-statistics views (so that a kernel "
                         "trace holds the headline's identical launches only)")
    ap.add_argument("--camera-seed", type=int, default=11)
    ap.add_argument("--field", choices=["auto", "split", "fused"], default="auto",
                    help="BF16 field as a gather kernel + an MLP kernel, one fused kernel, or "
                         "fused where the fused kernel supports the field's shape")
    ap.add_argument("--serve-requests", type=int, default=40,
                    help="least requests per arrival rate in the serving measurement (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="arrivals per rate span at least this long (more requests at high rates)")
    ap.add_argument("--serve-rates", default="10,20,40,80,120,160,inf",
                    help="Poisson arrival rates (requests/s) of the serving measurement; inf: "
                         "every request arrives at once (the saturation throughput)")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the backward / training / pruning / splitting measurements")
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


def time_views(rnd, dsc, rs, rd, width, steps, warmup):
    """Mean wall ms per view over `steps` views after `warmup`, the per-view
    HIP-event latencies, and the field-kernel events of the timed views."""
    import torch
    for _ in range(warmup):
        rnd.render(dsc, rs, rd, width=width)
    torch.cuda.synchronize()
    kev = rnd.record_kernels(True)
    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        rnd.render(dsc, rs, rd, width=width)
        e1.record()
        lat.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    rnd.record_kernels(False)
    lat_ms = np.array([a.elapsed_time(b) for a, b in lat])
    return wall * 1000.0 / steps, lat_ms, kev


def kernel_ms(kev, name):
    ev = [(a, b) for k, _, a, b in kev if k == name]
    return sum(a.elapsed_time(b) for a, b in ev), len(ev)


def stage_rooflines(stages, R, n_hit, nodes, P, n_coarse_slots, n_fine_slots, n_fine_eval):
    """Algorithmic HBM bytes of the bandwidth-bound stages of a view (padded
    contract of the reference ops: R rays, P = max_hits, T octree nodes, M
    samples of 16-byte records) over their traced times."""
    rows = {
        # rays in, node records once, sorted hit rows out (lo, hi, idx)
        "intersect": R * 24 + nodes * 48 + R * P * 12,
        # hit rows in, sample records out
        "coarse_sample": n_hit * (16 * P + 4) + n_coarse_slots * 16,
        # coarse samples and field in (bins + weights), fine records and marks out
        "fine_sample": n_coarse_slots * (16 + 16) + n_hit * 4 + n_fine_slots * (16 + 2),
        # samples and field in, one rgb / depth / missed per ray out
        "fine_composite": n_fine_slots * (16 + 16 + 2) + R * 20,
    }
    out = {}
    for k, b in rows.items():
        ms = stages.get(k)
        if ms:
            gbs = b / (ms * 1e-3) / 1e9
            out[k] = {"ms": ms, "algorithmic_bytes": int(b), "achieved_GBs": round(gbs, 1),
                      "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)}
    return out


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
    split_arg = {"auto": None, "split": True, "fused": False}[args.field]
    flops_per_sample = sc.mlp.flops_per_sample()

    # ---- the view at each field precision (headline first)
    precs = [HEADLINE] if args.only_headline else [HEADLINE, "bf16", "fp32"]
    res = {}
    for prec in precs:
        rnd = Renderer(dev, precision=prec, hierarchical=True, fine_steps=64.0,
                       split_field=split_arg if prec == "bf16" else None)
        steps = args.steps if prec != "fp32" else min(args.steps, 5)
        warm = args.warmup if prec != "fp32" else min(args.warmup, 2)
        ms, lat_ms, kev = time_views(rnd, dsc, rs, rd, args.res, steps, warm)
        res[prec] = {"rnd": rnd, "ms": ms, "lat": lat_ms, "kev": kev, "steps": steps}
        log(f"{prec}: {ms:.2f} ms/view, {R / ms / 1e3:.2f} Mrays/s")

    h = res[HEADLINE]
    rnd = h["rnd"]
    ms_per_step = h["ms"]
    value = R / (ms_per_step / 1000.0)
    st = rnd.last
    # exact sample counts of the view (identical for every repetition)
    rgb_h, depth_h, missed_h = rnd.render(dsc, rs, rd, keep=True, width=args.res)
    n_coarse, n_fine = rnd.count_samples()
    n_fine_eval = rnd.kept["f_evaluated"]
    counts = rnd.kept["count"].cpu().numpy()
    gpu_rows = rnd.kept["hits"].cpu().numpy()
    headline_probs = (rnd.kept["c_cnt"].cpu().numpy(), rnd.kept["c_off"].cpu().numpy(),
                      rnd.kept["c_prob"].cpu().numpy())
    rnd.kept = None
    samples_per_view = n_coarse + n_fine_eval

    # roofline of the dominant kernel: the headline's MLP kernel (MFMA-bound),
    # MFMA work = 3 BF16 products per product of the model's FLOPs
    mlp_ms, n_mlp = kernel_ms(h["kev"], "mlp")
    emb_ms, n_emb = kernel_ms(h["kev"], "embed")
    x3_flops = 3 * flops_per_sample
    samples = samples_per_view * h["steps"]
    achieved = x3_flops * samples / (mlp_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "kernel": "mlp_bf16x3_kernel<32,128,3>",
                "achieved": round(achieved, 2), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_BF16_TFLOPS, 4),
                "traffic": measured_traffic("mlp_bf16x3_kernel"),
                "avg_launch_ms": round(mlp_ms / max(n_mlp, 1), 4),
                "flops_per_sample": x3_flops,
                "model_flops_per_sample": flops_per_sample,
                "model_TFLOPs": round(achieved / 3, 2),
                "samples_per_launch": round(samples / max(n_mlp, 1))}
    D = dsc.D
    emb_bytes = samples * (16 + 4 * D) + n_emb * (sc.num_voxels * 44 + sc.num_keys * 4 * D)
    kernels = {"embed_bf16x3_kernel": {
        "avg_launch_ms": round(emb_ms / max(n_emb, 1), 4),
        "algorithmic_bytes_per_launch": round(emb_bytes / max(n_emb, 1)),
        "achieved_GBs": round(emb_bytes / (emb_ms * 1e-3) / 1e9, 1) if emb_ms else None,
        "bound": "hbm", "traffic": measured_traffic("embed_bf16x3_kernel")}}

    variants = {}
    if not args.no_variants:
        # the same view from its camera: rays generated inside the traversal
        # kernel (outside the timed region; reported beside the metric)
        from nsvf_amd.rays import ViewCamera
        vcam = ViewCamera.looking_at(cam, args.res, args.res)
        rnd.render_camera(dsc, vcam)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            rnd.render_camera(dsc, vcam)
        torch.cuda.synchronize()
        cam_ms = (time.perf_counter() - t1) * 1000.0 / args.steps

        # the same view with train-set voxel statistics on (both passes fold
        # their sample probabilities into a per-voxel running max)
        rnd.track = torch.zeros(sc.num_voxels, dtype=torch.float32, device=dev)
        rnd.render(dsc, rs, rd, width=args.res)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            rnd.render(dsc, rs, rd, width=args.res)
        torch.cuda.synchronize()
        track_ms = (time.perf_counter() - t1) * 1000.0 / args.steps
        tracked_voxels = int((rnd.track > 0).sum())
        rnd.track = None
        # early ray termination as the reference's README renders with it
        # (raymarching_tolerance 0.01, chunk_size 512: chunks of 512 x 1024
        # rows / hits); off by default
        ref_img = rnd.render(dsc, rs, rd, width=args.res)[0].clone()
        et = Renderer(dev, precision=HEADLINE, hierarchical=True, compat=True, tolerance=0.01,
                      chunk_size=512 * 1024)
        img = et.render(dsc, rs, rd, width=args.res, keep=True)[0]
        torch.cuda.synchronize()
        ev = (int(et.kept["c_eval"].sum()), int(et.kept["f_eval"].sum()))
        stopped = int((et.kept["f_active"] == 0).sum())
        t1 = time.perf_counter()
        for _ in range(max(1, args.steps // 4)):
            et.render(dsc, rs, rd, width=args.res)
        torch.cuda.synchronize()
        et_ms = (time.perf_counter() - t1) * 1000.0 / max(1, args.steps // 4)
        et_diff = float((img - ref_img).abs().max())
        del et, img, ref_img
        torch.cuda.empty_cache()
        variants = {
            "early_termination": {"ms_per_view": round(et_ms, 3), "tolerance": 0.01,
                                  "chunk": 512 * 1024, "coarse_evaluated": ev[0],
                                  "fine_evaluated": ev[1], "rays_stopped_fine": stopped,
                                  "max_rgb_change": et_diff,
                                  "note": "the reference's chunked marching (renderer.py:"
                                          "136-174); off by default"},
            "from_camera": {"ms_per_view": round(cam_ms, 3),
                            "note": "rays generated inside the traversal kernel from the camera"},
            "with_voxel_stats": {"ms_per_view": round(track_ms, 3),
                                 "voxels_seen": tracked_voxels,
                                 "note": "track_voxel_probs after both passes"}}

    # per-stage breakdown of one traced view (outside the timed region)
    rnd.trace = []
    rnd.render(dsc, rs, rd, width=args.res)
    torch.cuda.synchronize()
    stages = {}
    for (n0, e0), (n1, e1) in zip(rnd.trace[:-1], rnd.trace[1:]):
        stages[n1] = round(e0.elapsed_time(e1), 3)
    rnd.trace = None

    precisions = {}
    for prec, r in res.items():
        entry = {"ms_per_view": round(r["ms"], 3), "rays_per_sec": round(R / r["ms"] * 1e3, 1),
                 "p50_ms": round(float(np.percentile(r["lat"], 50)), 3),
                 "p99_ms": round(float(np.percentile(r["lat"], 99)), 3), "views": r["steps"]}
        fk = "mlp" if prec == HEADLINE else "field"
        fms, nf = kernel_ms(r["kev"], fk)
        if prec == "bf16" and r["rnd"].uses_split_field(dsc):
            fms, nf = kernel_ms(r["kev"], "mlp")
        if nf:
            mult = 3 if prec == HEADLINE else 1
            peak = PEAK_FP32_TFLOPS if prec == "fp32" else PEAK_BF16_TFLOPS
            tf = mult * flops_per_sample * samples_per_view * r["steps"] / (fms * 1e-3) / 1e12
            entry["field_kernel"] = {"avg_launch_ms": round(fms / nf, 4),
                                     "achieved_TFLOPs": round(tf, 1), "peak_TFLOPs": peak,
                                     "frac": round(tf / peak, 4)}
        precisions[prec] = entry

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "rays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
        "data": "synthetic",
        "config": {"workload": f"noise{128}: 128^3 grid, {sc.num_voxels} voxels, SVO max_hits 60, "
                               f"inverse-CDF step voxel/8 + 64 hierarchical samples, "
                               f"{args.res}x{args.res} view, field 32->3x128->4 split-BF16 MFMA "
                               f"(bf16x3, FP32 accumulation)",
                   "rays_per_view": R, "voxels": sc.num_voxels, "octree_nodes": dsc.num_nodes,
                   "sampler": "reference mode (compat quirks reproduced)",
                   "parallelism": "single GPU"},
        "p50_ms": round(float(np.percentile(h["lat"], 50)), 3),
        "p99_ms": round(float(np.percentile(h["lat"], 99)), 3),
        "samples_per_view": {"hit_rays": st.hit_rays, "coarse": n_coarse, "fine": n_fine,
                             "fine_evaluated": n_fine_eval,
                             "coarse_slots": st.coarse_slots},
        "stages_ms": stages,
        "roofline": roofline,
        "kernels": kernels,
        "stage_rooflines": stage_rooflines(stages, R, st.hit_rays, dsc.num_nodes, sc.max_hits,
                                           st.coarse_slots, st.fine_slots, n_fine_eval),
        "precisions": precisions,
        **variants,
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
        plain = {"rays_per_sec": round(R / (pms / 1000.0), 1), "ms_per_view": round(pms, 2),
                 "views": args.plain_views}
        for prec, r in res.items():
            frgb, _, _ = r["rnd"].render(dsc, rs, rd, width=args.res)
            d = frgb.double() - prgb.double()
            mse = float(torch.mean(d * d))
            plain[f"vs_{prec}"] = {"max_abs_rgb": float(d.abs().max()),
                                   "psnr_db": round(10 * math.log10(1.0 / mse), 2) if mse > 0 else None}
        out["plain_fp32"] = plain
        out["speedup_vs_plain_fp32"] = round(value / (R / (pms / 1000.0)), 2)
        log(f"plain fp32: {pms:.1f} ms/view; speedup {out['speedup_vs_plain_fp32']}x")
        del pl
        torch.cuda.empty_cache()

    # ---- CPU leg: the oracle on the host, timed as the CPU baseline and run
    # in FP64 as the checker of the GPU outputs on a subset of the view
    if not args.no_cpu:
        views = {}
        for p, r in res.items():
            img = [t.cpu().numpy() for t in r["rnd"].render(dsc, rs, rd, keep=True, width=args.res)]
            k = r["rnd"].kept
            fine = (k["f_cnt"].cpu().numpy(), k["f_off"].cpu().numpy(),
                    k["f_samples"][0].cpu().numpy(), k["flag"].cpu().numpy().astype(bool))
            r["rnd"].kept = None
            views[p] = (*img, fine)
        cpu_base, accuracy = cpu_leg(args, sc, dsc, rs_np, rd_np, views, counts, gpu_rows,
                                     headline_probs)
        if cpu_base:
            out["cpu_baseline"] = cpu_base
        out["accuracy_vs_fp64"] = accuracy

    if not args.no_rows:
        out["rows"] = rows(args, dev, sc, dsc)

    if args.serve_requests > 0:
        out["serving"] = serving(args, dev)

    print(json.dumps(out), flush=True)


def csr_to_padded(cnt, off, arr, width, fill):
    """Rows off[n] .. off[n] + cnt[n] of arr as a [len(cnt), width] array."""
    cnt, off = np.asarray(cnt, np.int64), np.asarray(off, np.int64)
    out = np.full((len(cnt), width), fill, dtype=arr.dtype)
    col = np.arange(width)[None]
    m = col < cnt[:, None]
    out[m] = arr[(off[:, None] + col)[m]]
    return out


def fine_flips(fine, ref, sel):
    """Rays of `sel` whose GPU fine samples (fine = counts, offsets and voxel
    ids of the stream, hit flags) do not carry the oracle's voxel ids."""
    f_cnt, f_off, f_idx, flag = fine
    ev = ref["eval_rows"]
    ri = ref["f_idx"][ev]
    cnt = f_cnt[ev]
    W = max(ri.shape[1], int(cnt.max()) if len(cnt) else 0)
    ri = np.pad(ri, ((0, 0), (0, W - ri.shape[1])), constant_values=-1)
    gi = csr_to_padded(cnt, f_off[ev], f_idx, W, -1)
    out = np.zeros(len(sel), bool)
    out[np.nonzero(flag[sel])[0][(gi != ri).any(1)]] = True
    return out


def cpu_leg(args, sc, dsc, rs_np, rd_np, views, counts, gpu_rows, headline_probs):
    """The oracle on the host cores.  (1) CPU baseline: the reference path
    restated on the CPU (octree build, intersection and sampling in C++, one
    thread; gather, MLP and compositing in NumPy float32 on BLAS threads) on
    a seeded random sample of the view's rays, the median of --cpu-reps
    runs.  (2) Checker: the same path in float64 on a seeded subset of the
    view (plus every ray with a full hit row) against each precision's GPU
    view -- max abs error per output, end to end and, for the headline,
    with the same samples (the oracle's fine pass fed the GPU's coarse
    probabilities) -- and the hit rows against the FP32 restatement
    (bit-exact) and an FP64 slab test (grazing rays).  End to end, a ray
    whose fine samples do not carry the oracle's voxel ids (count or
    sequence) is a sampling flip: the rounding of the coarse probabilities
    changed a sampler decision (in the reference's mode, for instance, the
    fine pass's trailing loop stops on ray 0's slot of the batch row,
    sample_gpu.cu:194); flips are counted per precision, and each
    precision's error is also given on the other rays."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import nsvf_oracle as O
    from threadpoolctl import threadpool_limits

    cores = len(os.sched_getaffinity(0))
    R = rs_np.shape[0]
    base = None
    if args.cpu_rays > 0 and args.cpu_reps > 0:
        rng = np.random.default_rng(0)
        sel = np.sort(rng.choice(R, size=min(args.cpu_rays, R), replace=False))
        times = []
        with threadpool_limits(limits=cores):
            t0 = time.perf_counter()
            O.build_easy_octree(sc.points, np.float32(sc.voxel_size) * np.float32(0.5))
            build_s = time.perf_counter() - t0
            for _ in range(args.cpu_reps):
                t1 = time.perf_counter()
                O.render_view(sc, rs_np[sel], rd_np[sel], dtype=np.float32)
                times.append(time.perf_counter() - t1)
        med = float(np.median(times))
        # render_view builds the octree once per call: the rate excludes it
        per_view = max(med - build_s, 1e-9)
        base = {"value": round(len(sel) / per_view, 1), "unit": "rays/s", "cores": cores,
                "kind": "port",
                "sample": f"{len(sel)} seeded random rays of the {args.res}x{args.res} view, median "
                          f"of {len(times)} runs ({', '.join(f'{t:.2f}' for t in times)} s, each "
                          f"with a {build_s:.2f} s octree build, excluded from the rate); "
                          f"intersection + sampling in the C++ restatement (1 thread), gather / "
                          f"MLP / compositing in NumPy float32 ({cores} BLAS threads)"}
        log(f"cpu oracle: {base['value']:.0f} rays/s (median of {len(times)})")

    # checker subset: random rays plus the rays whose hit row is full or has
    # equal entry depths (where traversal order decides the row)
    rng = np.random.default_rng(1)
    P = sc.max_hits
    full = np.nonzero(counts >= P)[0]
    valid = np.arange(P)[None] < counts[:, None]
    lo_bits = gpu_rows[:, :, 0]
    tied = np.nonzero(((lo_bits[:, 1:] == lo_bits[:, :-1]) & valid[:, 1:]).any(1))[0]
    sel = np.union1d(rng.choice(R, size=min(args.check_rays, R), replace=False),
                     np.union1d(full, tied))
    t0 = time.perf_counter()
    with threadpool_limits(limits=cores):
        # geometry and sampling over the whole view (the reference sampler's
        # quirks read neighbouring rays), the FP64 field on the subset
        ref = O.render_view(sc, rs_np, rd_np, compat=True, dtype=np.float64, subset=sel)
        W = ref["c_idx"].shape[1]
        gp = csr_to_padded(*headline_probs, W, np.float32(0))
        ref_s = O.render_view(sc, rs_np, rd_np, compat=True, dtype=np.float64, subset=sel,
                              coarse_probs=gp)
    acc = {"rays": int(len(sel)), "full_rows": int(len(full)), "tied_rows": int(len(tied)),
           "bar_rgb": RGB_BAR,
           "oracle_s": round(time.perf_counter() - t0, 1)}
    for prec, (rgb, depth, missed, fine) in views.items():
        e = np.abs(rgb[sel] - ref["rgb"]).max(1)
        flips = fine_flips(fine, ref, sel)
        mse = float(np.mean((rgb[sel].astype(np.float64) - ref["rgb"]) ** 2))
        rest = float(e[~flips].max()) if (~flips).any() else 0.0
        acc[prec] = {"end_to_end": {
            "max_abs_rgb": float(e.max()),
            "max_abs_depth": float(np.abs(depth[sel] - ref["depth"]).max()),
            "max_abs_alpha": float(np.abs(missed[sel] - ref["missed"]).max()),
            "psnr_rgb_db": round(10 * math.log10(1.0 / mse), 2) if mse > 0 else None,
            "sampling_flips": int(flips.sum()),
            "max_abs_rgb_on_flips": float(e[flips].max()) if flips.any() else 0.0,
            "max_abs_rgb_without_flips": rest},
            "meets_rgb_bar": rest <= RGB_BAR}
    rgb, depth, missed, _ = views[HEADLINE]
    acc[HEADLINE]["same_samples"] = {
        "max_abs_rgb": float(np.abs(rgb[sel] - ref_s["rgb"]).max()),
        "max_abs_depth": float(np.abs(depth[sel] - ref_s["depth"]).max()),
        "max_abs_alpha": float(np.abs(missed[sel] - ref_s["missed"]).max())}
    acc[HEADLINE]["meets_rgb_bar"] = (acc[HEADLINE]["meets_rgb_bar"] and
                                      acc[HEADLINE]["same_samples"]["max_abs_rgb"] <= RGB_BAR)
    # hit rows: bit-exact against the FP32 restatement, and the rays whose
    # hit set an FP64 slab test decides differently (grazing)
    idx32, lo32, hi32 = ref["sorted_idx"][sel], ref["sorted_lo"][sel], ref["sorted_hi"][sel]
    g = gpu_rows[sel]
    cnt = counts[sel]
    gi = np.where(np.arange(P)[None] < cnt[:, None], g[:, :, 2], -1)
    glo = np.where(gi != -1, g[:, :, 0].view(np.float32), np.float32(O.MAX_DEPTH))
    ghi = np.where(gi != -1, g[:, :, 1].view(np.float32), np.float32(O.MAX_DEPTH))
    acc["hit_rows_bit_exact_vs_fp32_oracle"] = bool(
        np.array_equal(gi, idx32) and np.array_equal(glo, lo32) and np.array_equal(ghi, hi32))
    leaves = O.build_easy_octree(sc.points, np.float32(sc.voxel_size) * np.float32(0.5))[0][:sc.num_voxels]
    acc["grazing"] = O.grazing_report(gi, glo, ghi, rs_np[sel], rd_np[sel], leaves,
                                      sc.voxel_size, P, GRAZE_EPS)
    log("accuracy vs fp64: " + ", ".join(
        f"{p} rgb {acc[p]['end_to_end']['max_abs_rgb_without_flips']:.2e} "
        f"({acc[p]['end_to_end']['sampling_flips']} flips)" for p in views)
        + f"; grazing {acc['grazing']}")
    return base, acc


def _time_ms(fn, reps):
    """Mean device time of fn() over reps calls (HIP events on the current
    stream, after one untimed call)."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def rows(args, dev, sc, dsc):
    """The other rows of the path on the same 100k-voxel scene: backward
    kernels and one training step on a 256x256 batch of rays, the pruning
    pass (BF16 and FP32 MFMA) and the split."""
    import ctypes

    import torch

    from nsvf_amd import scene as S
    from nsvf_amd._lib import lib
    from nsvf_amd.render import DeviceScene, Renderer
    from nsvf_amd.train import TrainableField, TrainRenderer, mse_loss, trilinear_gather_backward
    from nsvf_amd.voxels import prune_scores, split_scene

    res = {}
    rs_np, rd_np = S.camera_rays(256, 256, S.sphere_camera(2.5, args.camera_seed))
    rs, rd = torch.from_numpy(rs_np).to(dev), torch.from_numpy(rd_np).to(dev)
    R = rs.shape[0]
    D = dsc.D
    # the coarse stream of the batch (what training differentiates through)
    g = Renderer(dev, precision="fp32", hierarchical=False).coarse_stream(dsc, rs, rd).owned()
    M = g.slots
    rec = g.samples
    s_idx = rec[:, 0].contiguous()
    ray = rec[:, 3].long()
    xyz = (rs.index_select(0, ray) + rd.index_select(0, ray) * rec[:, 1].view(torch.float32)[:, None]).contiguous()
    grad_emb = torch.randn((M, D), device=dev)
    K, N = dsc.values.shape[0], dsc.centers.shape[0]
    from nsvf_amd.train import gather_backward_plan
    plan = gather_backward_plan(dsc.feats, K)
    ms_np = _time_ms(lambda: trilinear_gather_backward(s_idx, xyz, dsc.centers, dsc.feats, K,
                                                       grad_emb, dsc.voxel_size), 10)
    ms = _time_ms(lambda: trilinear_gather_backward(s_idx, xyz, dsc.centers, dsc.feats, K, grad_emb,
                                                    dsc.voxel_size, plan), 10)
    nbytes = M * 4 * D + M * (4 + 12) + N * 44 + K * 4 * D
    res["gather_backward"] = {"ms": round(ms, 3), "samples": M, "algorithmic_bytes": nbytes,
                              "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                              "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                              "ms_sorting_incidences_per_call": round(ms_np, 3),
                              "note": "the scene's corner-incidence order computed once"}
    field = torch.rand((M, 4), device=dev)
    field[:, 0] = field[:, 0] * 20 - 5
    gr, gd, gm = torch.randn((R, 3), device=dev), torch.randn(R, device=dev), torch.randn(R, device=dev)
    gf = torch.zeros_like(field)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    ms = _time_ms(lambda: lib().nsvf_render_composite_backward(
        p(g.hit_list), p(g.scal[0:1]), g.n_hit, p(g.off), p(g.cnt), p(g.samples), p(field), p(gr),
        p(gd), p(gm), p(gf), st), 20)
    nbytes = 2 * M * 32 + M * 16 + R * 20
    res["composite_backward"] = {"ms": round(ms, 3), "samples": M, "algorithmic_bytes": nbytes,
                                 "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                                 "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    # one training step: forward, MSE against U(0,1) targets, backward --
    # with the HIP training MLP, and with torch's Linear layers for reference
    target = torch.rand((R, 3), device=dev)
    mlp_names = {"hip": "HIP FP32 MFMA",
                 "hip_x3": "HIP split BF16 (the backward's head in FP32)",
                 "torch": "torch FP32 nn.Linear"}
    for key, impl in (("train_step", "hip"), ("train_step_split_bf16", "hip_x3"),
                      ("train_step_torch_mlp", "torch")):
        tf = TrainableField(sc, dev, mlp=impl)
        tr = TrainRenderer(dev)

        def step():
            tf.zero_grad(set_to_none=True)
            rgb, _, _ = tr.forward(dsc, tf, rs, rd)
            mse_loss(rgb, target).backward()

        torch.cuda.synchronize()
        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            step()
        torch.cuda.synchronize()
        res[key] = {"ms": round((time.perf_counter() - t0) * 1e3 / 5, 2), "rays": R,
                    "samples": M,
                    "mlp": mlp_names[impl],
                    "kernels": "HIP gather / composite forward and backward"}
        del tf, tr
        torch.cuda.empty_cache()
    # the training configuration as specified: the shell scene, 4 views x
    # 2048 random pixels (64x64 views), with and without the traversal's
    # coherent grouping of the batch's rays, and with the split-BF16 MLP
    shell = S.shell_scene()
    dshell = DeviceScene(shell, dev)
    rng = np.random.default_rng(7)
    parts = []
    for v in range(4):
        a_, b_ = S.camera_rays(64, 64, S.sphere_camera(2.5, 100 + v))
        pick = rng.choice(a_.shape[0], 2048, replace=False)
        parts.append((a_[pick], b_[pick]))
    mix = rng.permutation(4 * 2048)
    crs = torch.from_numpy(np.concatenate([q[0] for q in parts])[mix]).to(dev)
    crd = torch.from_numpy(np.concatenate([q[1] for q in parts])[mix]).to(dev)
    ctarget = torch.rand((crs.shape[0], 3), device=dev)
    c3 = {"rays": int(crs.shape[0]), "scene": f"shell, {shell.num_voxels} voxels"}
    for key, coherent, impl in (("ms", True, "hip"), ("ms_rays_in_batch_order", False, "hip"),
                                ("ms_split_bf16", True, "hip_x3")):
        tf = TrainableField(shell, dev, mlp=impl)
        tr = TrainRenderer(dev, coherent=coherent)

        def step():
            tf.zero_grad(set_to_none=True)
            rgb, _, _ = tr.forward(dshell, tf, crs, crd)
            mse_loss(rgb, ctarget).backward()

        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            step()
        torch.cuda.synchronize()
        c3[key] = round((time.perf_counter() - t0) * 1e3 / 20, 3)
        del tf, tr
    res["train_step_4x2048_random_rays"] = c3
    # the training MLP kernels alone on the batch's samples
    from nsvf_amd.train import FieldMLP, FieldMLPx3
    tf = TrainableField(sc, dev)
    prm = []
    for w_, b_ in zip(tf.weights, tf.biases):
        prm += [w_.detach().requires_grad_(True), b_.detach().requires_grad_(True)]
    emb = torch.randn((M, D), device=dev).mul_(0.1).requires_grad_(True)
    dout = torch.randn((M, 4), device=dev)
    fl = sc.mlp.flops_per_sample()
    ms_f = _time_ms(lambda: FieldMLP.apply(emb, *prm), 5)
    ms_f3 = _time_ms(lambda: FieldMLPx3.apply(emb, *prm), 5)
    ms_fb = _time_ms(lambda: FieldMLP.apply(emb, *prm).backward(dout), 5)
    ms_fb3 = _time_ms(lambda: FieldMLPx3.apply(emb, *prm).backward(dout), 5)
    # backward: recomputed forward + input gradients + weight gradients, plus
    # the first layer once more (about 3 forward FLOPs and one layer)
    ms_b = ms_fb - ms_f
    bw_flops = M * (3 * fl + 2 * dsc.D * dsc.H)
    res["mlp_train"] = {"forward_ms": round(ms_f, 3), "backward_ms": round(ms_b, 3),
                        "forward_split_bf16_ms": round(ms_f3, 3),
                        "backward_split_bf16_ms": round(ms_fb3 - ms_f3, 3), "samples": M,
                        "forward_TFLOPs": round(M * fl / (ms_f * 1e-3) / 1e12, 1),
                        "backward_TFLOPs": round(bw_flops / (ms_b * 1e-3) / 1e12, 1),
                        "peak_TFLOPs": 157.3}
    del tf, prm, emb, dout
    torch.cuda.empty_cache()
    # pruning: 16^3 lattice per voxel through the field MLP, max relu(sigma)
    flops = sc.num_voxels * 4096 * sc.mlp.flops_per_sample()
    for prec, peak in (("bf16", PEAK_BF16_TFLOPS), ("bf16x3", PEAK_BF16_TFLOPS),
                       ("fp32", 157.3)):
        ms = _time_ms(lambda: prune_scores(dsc, prec), 2)
        tf_s = flops / (ms * 1e-3) / 1e12
        # split BF16 runs three BF16 MFMAs per product
        mfma_tf = tf_s * (3 if prec == "bf16x3" else 1)
        res[f"prune_{prec}"] = {"ms": round(ms, 2), "lattice_points": sc.num_voxels * 4096,
                                "achieved_TFLOPs": round(tf_s, 1), "peak_TFLOPs": peak,
                                "frac": round(mfma_tf / peak, 4)}
    # splitting (device key rebuild + interpolation; includes the uploads and
    # the readback of the new scene)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    new = split_scene(sc, dev)
    torch.cuda.synchronize()
    res["split"] = {"ms_including_transfers": round((time.perf_counter() - t0) * 1e3, 1),
                    "voxels": sc.num_voxels, "new_voxels": new.num_voxels, "new_keys": new.num_keys}

    # scene files: the encoder's checkpoint buffers written and read back,
    # then made resident (octree build, node records, upload)
    import tempfile

    from nsvf_amd import scene_io as IO
    from nsvf_amd.render import DeviceScene
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "scene.pt")
        t0 = time.perf_counter()
        IO.save_checkpoint(sc, path)
        t1 = time.perf_counter()
        loaded = IO.load_checkpoint(path, mlp=sc.mlp)
        t2 = time.perf_counter()
        DeviceScene(loaded, dev)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
    res["scene_files"] = {"ms": round((t3 - t1) * 1e3, 1), "save_ms": round((t1 - t0) * 1e3, 1),
                          "load_ms": round((t2 - t1) * 1e3, 1),
                          "resident_ms": round((t3 - t2) * 1e3, 1), "voxels": loaded.num_voxels}
    def row_ms(v):
        if "forward_ms" in v:
            return f"{v['forward_ms']} + {v['backward_ms']}"
        return v.get("ms", v.get("ms_including_transfers"))

    log("rows: " + ", ".join(f"{k} {row_ms(v)} ms" for k, v in res.items()))
    return res


SERVE_SCENES = [  # (grid resolution, embedding seed, MLP seed): shells of ~20k-400k voxels
    (46, 11, 21), (79, 12, 22), (107, 13, 23), (125, 14, 24)]
SERVE_RES = (256, 400, 512, 800)


def serving(args, dev):
    """Four resident shell scenes (spherical-shell generator at grids of 46,
    79, 107 and 125 cells: about 20k, 100k, 250k and 400k voxels); requests
    with a uniform scene, a resolution uniform over 256^2 / 400^2 / 512^2 /
    800^2 and a camera at radius U(2, 3), arriving as a Poisson process at
    each rate; rays are on the device before the clock starts.  Reports
    rays/s, p50/p99 latency (arrival to completion), and the error of a
    256^2 view of each scene against the plain FP32 path (beside the fused
    FP32 path's against the same)."""
    import torch

    from nsvf_amd import scene as S
    from nsvf_amd.plain import PlainNSVF
    from nsvf_amd.render import DeviceScene, Renderer
    from nsvf_amd.serve import RenderServer, latency_summary

    t0 = time.time()
    scenes = {f"shell{r}": S.shell_scene(res=r, emb_seed=e, mlp_seed=m) for r, e, m in SERVE_SCENES}
    srv = RenderServer(dev, slots=4, precision=HEADLINE, hierarchical=True)
    for name, sc in scenes.items():
        srv.load_scene(name, sc)
    log(f"serving: {len(scenes)} scenes resident "
        f"({', '.join(f'{n}: {sc.num_voxels}' for n, sc in scenes.items())} voxels; "
        f"{time.time() - t0:.1f}s setup)")
    names = list(scenes)
    rng = np.random.default_rng(5)

    def make_requests(n):
        reqs = []
        for _ in range(n):
            name = names[int(rng.integers(len(names)))]
            res = int(rng.choice(SERVE_RES))
            cam = S.sphere_camera(float(rng.uniform(2.0, 3.0)), int(rng.integers(1 << 30)))
            rs, rd = S.camera_rays(res, res, cam)
            reqs.append((name, torch.from_numpy(rs).to(dev), torch.from_numpy(rd).to(dev), res))
        return reqs

    # warm-up: every slot sees the largest job (so the measured runs
    # allocate nothing) and every scene renders once
    big = names[-1]
    for _ in range(2 * len(srv.scheduler.slots)):
        rs, rd = S.camera_rays(800, 800, S.sphere_camera(2.5, int(rng.integers(1 << 30))))
        srv.submit(big, torch.from_numpy(rs).to(dev), torch.from_numpy(rd).to(dev), width=800)
    srv.drain()
    for n in names:
        srv.submit(n, *make_requests(1)[0][1:])
    srv.drain()
    torch.cuda.synchronize()
    # each (scene, resolution) alone: submitted to the idle server, arrival
    # to completion (the second of two)
    solo = {}
    for n in names:
        for res in SERVE_RES:
            rs, rd = S.camera_rays(res, res, S.sphere_camera(2.5, 99))
            rs, rd = torch.from_numpy(rs).to(dev), torch.from_numpy(rd).to(dev)
            for _ in range(2):
                req = srv.submit(n, rs, rd, width=res)
                srv.drain()
            solo[(n, res * res)] = req.latency
    log("serving: alone (ms) " + ", ".join(f"{n}@{int(math.sqrt(k))}: {v * 1e3:.1f}"
                                           for (n, k), v in solo.items()))
    allocs0 = torch.cuda.memory_stats(dev).get("num_alloc_retries", 0), \
        torch.cuda.memory_stats(dev).get("segment.all.allocated", 0)
    runs = []
    for rate_s in [x for x in args.serve_rates.split(",") if x]:
        rate = float(rate_s)
        n = args.serve_requests if math.isinf(rate) else \
            max(args.serve_requests, int(rate * args.serve_seconds))
        items = make_requests(n)
        if math.isinf(rate):
            arrivals = [(0.0, *it) for it in items]
        else:
            gaps = rng.exponential(1.0 / rate, len(items))
            arrivals = [(float(t), *it) for t, it in zip(np.cumsum(gaps) - gaps[0], items)]
        torch.cuda.synchronize()
        reqs = srv.run_arrivals(arrivals)
        torch.cuda.synchronize()
        summ = latency_summary(reqs)
        summ["rate_req_per_s"] = rate_s
        done = [r for r in reqs if r.state == "done"]
        slow = np.array([r.latency / solo[(r.scene, r.rays)] for r in done])
        span = max(r.finished for r in done) - min(r.arrival for r in done)
        summ["slowdown_vs_alone"] = {"p50": round(float(np.percentile(slow, 50)), 2),
                                     "p99": round(float(np.percentile(slow, 99)), 2)}
        # the requests' time alone, summed, over the run's span: above 1 when
        # overlapping them beats running them one after another
        summ["alone_time_per_span"] = round(sum(solo[(r.scene, r.rays)] for r in done) / span, 3)
        runs.append(summ)
        ph = summ["phases_ms"]
        log(f"serving @ {rate_s} req/s ({n} requests): {summ['rays_per_sec'] / 1e6:.1f} Mrays/s, "
            f"p50 {summ['p50_ms']:.1f} ms, p99 {summ['p99_ms']:.1f} ms; queued p50/p99 "
            f"{ph['queued']['p50']}/{ph['queued']['p99']}, geometry {ph['geometry']['p50']}/"
            f"{ph['geometry']['p99']}, shading {ph['shading']['p50']}/{ph['shading']['p99']} ms; "
            f"mean batch {summ['mean_batch']}; slowdown vs alone p50/p99 "
            f"{summ['slowdown_vs_alone']['p50']}/{summ['slowdown_vs_alone']['p99']}, "
            f"alone time / span {summ['alone_time_per_span']}")
    allocs1 = torch.cuda.memory_stats(dev).get("num_alloc_retries", 0), \
        torch.cuda.memory_stats(dev).get("segment.all.allocated", 0)
    # error against the plain FP32 path, one 256^2 view per scene.  The fine
    # samples' depths are continuous in the coarse field values, so two
    # renders whose fields differ in the last bits place most fine samples a
    # few ulps apart, and a ray whose sample crosses a gap between voxels
    # changes by up to a few 1e-3; the fused FP32 path against the same plain
    # path shows that spread between two FP32 implementations
    errs = {}
    for name in names:
        rs_np, rd_np = S.camera_rays(256, 256, S.sphere_camera(2.5, 3))
        rs, rd = torch.from_numpy(rs_np).to(dev), torch.from_numpy(rd_np).to(dev)
        req = srv.submit(name, rs, rd, width=256)
        srv.drain()
        pl = PlainNSVF(scenes[name], dev, compat=srv.compat)
        prgb = pl.render(rs, rd)[0].double()
        frgb = Renderer(dev, precision="fp32", compat=srv.compat).render(
            DeviceScene(scenes[name], dev), rs, rd, width=256)[0].double()
        d = (req.result[0].double() - prgb).abs().amax(1)
        df = (frgb - prgb).abs().amax(1)
        mse = float(torch.mean((req.result[0].double() - prgb) ** 2))
        errs[name] = {"max_abs_rgb": float(d.max()),
                      "pixels_over_1e-3": int((d > 1e-3).sum()),
                      "psnr_db": round(10 * math.log10(1.0 / mse), 2) if mse > 0 else None,
                      "fp32_fused_max_abs_rgb": float(df.max()),
                      "fp32_fused_pixels_over_1e-3": int((df > 1e-3).sum())}
        del pl
        torch.cuda.empty_cache()
    sat = max(runs, key=lambda r: r["rays_per_sec"] or 0)
    return {"scenes": {n: sc.num_voxels for n, sc in scenes.items()}, "slots": 4,
            "precision": HEADLINE, "sampler_compat": bool(srv.compat),
            "max_batch": srv.scheduler.max_batch, "max_batch_rays": srv.scheduler.max_batch_rays,
            "resolutions": list(SERVE_RES), "runs": runs,
            "saturation": {"rays_per_sec": sat["rays_per_sec"], "at_rate": sat["rate_req_per_s"],
                           "p50_ms": sat["p50_ms"], "p99_ms": sat["p99_ms"]},
            "alone_ms": {f"{n}@{int(math.sqrt(k))}": round(v * 1e3, 2) for (n, k), v in solo.items()},
            "device_segments_allocated_during_runs": int(allocs1[1] - allocs0[1]),
            "error_vs_plain_fp32_256": errs, "budget_peak_bytes": srv.budget.peak}


if __name__ == "__main__":
    main()
