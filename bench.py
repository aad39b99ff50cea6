"""Benchmark of the per-frame view-synthesis path on one MI355X.

Workload (default): four equirectangular 4096x2048 10-bit sources with 16-bit
inverse depth (0.5-25 m) on a 0.3 m square -> one 1920x1080 perspective view
(90 degree horizontal FOV) at the square's centre; synthetic seeded scene.
One step = one synthesized view, sources resident in HBM: reproject + bin +
rasterise + blend all four sources, inpaint, encode YUV420 10-bit.

Prints one JSON line.  `value` = views/s of the timed steps; p50/p99 are the
per-view latencies (host call to completion, synchronous like the reference).
roofline: algorithmic bytes of one view (each source sample once at its stored
precision + each output sample once) / the view's device time, against 8 TB/s.
cpu_baseline: the CPU oracle (single-threaded restatement of the reference CPU
path) on the same view, on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rvs-cuda_amd", "python"))

HBM_PEAK_GBS = 8000.0


def b_alg(rg):
    """Algorithmic bytes of one view: sum_s W*H*(1.5*b_color + b_depth) + W*H*1.5*b_out."""
    b = lambda bd: 1 if bd <= 8 else 2  # noqa: E731
    tot = 0
    for s in rg["sources"]:
        c = s["cam"]
        tot += c.w * c.h * (1.5 * b(s["bd_color"]) + b(s["bd_depth"]))
    t = rg["target"]
    tot += t["cam"].w * t["cam"].h * 1.5 * b(t["bitdepth"])
    return int(tot)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="erp4_1080p")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--batch", type=int, default=64, help="views per batch of the batched measurement (0: skip)")
    ap.add_argument("--batch-reps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import rvs_amd as R
    import workloads as W

    rg = W.rig(args.workload)
    t0 = time.time()
    planes = W.render_sources(rg)
    print("# scene rendered in %.1fs" % (time.time() - t0), file=sys.stderr, flush=True)
    srcs = W.device_sources(rg, planes)
    tgt = W.device_target(rg)
    syn = R.Synthesizer(srcs, tgt, want_maps=False)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        syn.run()
    torch.cuda.synchronize()

    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        syn.run()            # returns after the view is complete (synchronous plain path)
        lat.append(time.perf_counter() - a)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    views_s = args.steps / t_total

    # device time per stage (HIP events on the synthesis stream), separate pass
    R.profile_enable(True)
    stage = {k: [] for k in R.PROF_STAGES}
    for _ in range(args.profile_steps):
        syn.run()
        for k, v in R.profile_read().items():
            stage[k].append(v)
    counters = R.profile_counters()
    R.profile_enable(False)
    stage_ms = {k: float(np.mean(v)) for k, v in stage.items()}
    dev_ms = sum(stage_ms.values())
    balg = b_alg(rg)
    achieved = balg / (dev_ms * 1e-3) / 1e9

    out = {
        "metric": "served views/s + p50/p99 ms vs plain per-view; 4x4096x2048 ERP->1080p",
        "value": round(views_s, 3),
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_total / args.steps, 3),
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%s: 4x ERP 4096x2048 10-bit + 16-bit depth -> 1920x1080 perspective, blending 1.0"
                        % args.workload,
            "mode": "plain per-view path (one view per launch sequence, synchronous)",
            "sources_resident": True,
        },
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "counters": counters,
        "roofline": {
            "kernel": "one view's launch set (bin, raster+blend, inpaint, encode)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "alg_bytes": balg,
            "device_ms": round(dev_ms, 4),
            "traffic": None,
        },
    }

    if args.batch > 0:
        # batched: `batch` 1080p views at random-walk poses, one launch set per batch
        tgts = W.batch_targets(rg, args.batch, seed=3, sizes=((rg["target"]["cam"].w, rg["target"]["cam"].h),))
        bs = R.BatchSynthesizer(srcs, tgts)
        bs.run()
        torch.cuda.synchronize()
        times = []
        for _ in range(args.batch_reps):
            a = time.perf_counter()
            bs.run()
            times.append(time.perf_counter() - a)
        R.profile_enable(True)
        bs.run()
        bstage = R.profile_read()
        bcount = R.profile_counters()
        R.profile_enable(False)
        bt = float(np.median(times))
        out["batched"] = {
            "views": args.batch,
            "views_s": round(args.batch / bt, 2),
            "batch_ms": round(1e3 * bt, 3),
            "stage_ms": {k: round(v, 4) for k, v in bstage.items()},
            "counters": bcount,
            "vs_plain": round(args.batch / bt / views_s, 2),
        }
        del bs

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_bind as ob
        t = rg["target"]
        a = time.perf_counter()
        ref = ob.synthesize(W.oracle_sources(rg, planes, ob), W.oracle_cam(t["cam"], ob), t["pos"], t["rot"],
                            t["bitdepth"], t["blending_exp"])
        cpu_s = time.perf_counter() - a
        outp = syn.run()
        torch.cuda.synchronize()
        npdt = np.uint16 if t["bitdepth"] > 8 else np.uint8
        same = all(np.array_equal(outp[k].cpu().numpy().view(npdt), ref[k]) for k in ("y", "u", "v"))
        out["cpu_baseline"] = {
            "value": round(1.0 / cpu_s, 5),
            "unit": "views/s",
            "cores": 1,
            "kind": "port",
            "sample": "one full view of the same workload (single-threaded oracle, %.1f s)" % cpu_s,
        }
        out["parity_full_size"] = {"yuv_bit_exact_vs_oracle": bool(same)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
