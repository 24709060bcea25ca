"""Benchmark: ray samples/s and frames/s of the HIP raycaster on the medium
workload (512^3 uint16, 256^3 bricks + 2-voxel apron, 3 LODs, 4096-entry
Gaussian-bump TF, 1920x1080 perspective 60 degree FoV, 36-view orbit,
sampling rate 1.5, min/max empty-space skipping), rendered at LOD 0.

A step is one frame: one raycast of the whole 1080p image for the next
view of the orbit, with the volume already bricked into the device pool.
`value` is nominal samples per second (sum over pixels of
floor(chord * rate) through the LOD-0 box, identical for any renderer);
frames/s and evaluated samples/s are reported beside it.  The roofline
line prices the raycast kernel at its algorithmic gather traffic:
evaluated samples x 8 corners x 2 bytes + 4 bytes per pixel, over the
kernel's HIP-event time, against 8 TB/s.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gsamples/s + FPS @1080p 512³u16; learned-LOD FPS at ≤1/255 RMSE; p99 ms"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "raycast_medium_traffic.json")


def cpu_baseline(vol, w, cams, threads: int) -> dict:
    """The oracle's FP32 multi-threaded CPU raycaster on a bounded sample of
    the same workload (the first `len(cams)` orbit views at full size)."""
    from oracle import oracle as O
    from imagevis3d_amd import workloads
    nominal, secs = 0, 0.0
    for cam in cams:
        p = O.make_params(vol, w.tf, **workloads.oracle_kwargs(w, cam))
        nominal += O.nominal_samples(p)
        t0 = time.perf_counter()
        O.baseline_render(p, threads)
        secs += time.perf_counter() - t0
    return {"value": nominal / secs / 1e9, "unit": "Gsamples/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(cams)} of the 36 orbit views at 1920x1080 on the 512^3 volume, "
                      f"LOD 0, {secs:.1f} s of CPU time (FP32 OpenMP raycaster, same model)",
            "fps": len(cams) / secs}


def calibrate_copy(nbytes: int = 1 << 30) -> int:
    """A plain device copy of known size (nbytes read + nbytes written), the
    reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings."""
    import torch
    n = nbytes // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "dispatches": 3}))
    return 0


def main() -> int:
    if "--calibrate-copy" in sys.argv:
        return calibrate_copy()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=36)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8, help="samples per march step")
    ap.add_argument("--tile-group", type=int, default=4, help="ray tiles per XCD turn")
    ap.add_argument("--xcd-split", type=int, default=0, help="0 or frame regions per row")
    ap.add_argument("--tile-width", type=int, default=8, help="ray tile width (4, 8, 16)")
    ap.add_argument("--cell", type=int, default=16, help="empty-space cell edge in voxels")
    ap.add_argument("--no-skip", action="store_true", help="march whole bricks (no cell skipping)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the given launch settings instead of autotuning them")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(every GPU box here has one GPU; the sharding design is in DESIGN.md)",
              file=sys.stderr)
        return 2

    from imagevis3d_amd import tuvok, workloads

    w = workloads.medium()
    print(f"[bench] generating {w.name} volume", file=sys.stderr, flush=True)
    vol = w.volume()
    r = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, False)
    assert r.initialize(tuvok.createContext(w.width, w.height, device=0)), r.lastError()
    r.setRendererTarget(tuvok.renderer.types.RT_Headless)
    r.setEmptySpaceCell(args.cell)
    r.setRayTileShape(args.tile_width)
    t0 = time.perf_counter()
    r.setVolume(vol, w.brick, w.apron, w.n_lod)
    brick_s = time.perf_counter() - t0
    r.set1DTrans(w.tf)
    r.setLODLimits((0, 0))
    r.setKernelTuning(args.batch, args.tile_group, args.xcd_split)
    workloads.apply(r, w, w.cameras[0])
    if args.no_skip:
        r.setEmptySpaceSkipping(False)
    tuning = None
    if not args.no_autotune:
        # empty-space cell and launch shape for this workload (the brick
        # size stays the configured 256^3), timed on 4 orbit views at full
        # size (untimed setup)
        tv = [dict(view_pos=c.view_pos, view_dir=c.view_dir, up_dir=c.up_dir, fov=c.fov,
                   width=w.width, height=w.height) for c in w.cameras[::9]]
        t0 = time.perf_counter()
        t = r.autotune(tv, flags=r.TUNE_CELL | r.TUNE_LAUNCH)
        tuning = {"brick": t.brick_size, "cell": t.cell, "samples_per_batch": t.samples_per_batch,
                  "tile_width": t.tile_width, "tile_group": t.tile_group,
                  "tuned_ms": round(t.kernel_ms, 4), "start_ms": round(t.baseline_ms, 4),
                  "settings_timed": t.candidates, "seconds": round(time.perf_counter() - t0, 2)}
        print(f"[bench] autotuned {tuning}", file=sys.stderr, flush=True)
    print(f"[bench] bricked into the device pool in {brick_s:.2f} s; warmup {args.warmup}",
          file=sys.stderr, flush=True)

    cams = w.cameras
    for i in range(args.warmup):
        workloads.set_camera(r, cams[i % len(cams)])
        assert r.paint(), r.lastError()
        r.frameStats()
    r.synchronize()

    kernel_ms, evaluated, nominal = [], 0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        workloads.set_camera(r, cams[i % len(cams)])
        if not r.paint():
            raise RuntimeError(r.lastError())
        st = r.frameStats()  # waits for this frame's events and counters
        kernel_ms.append(st.kernel_ms)
        evaluated += st.evaluated_samples
        nominal += st.nominal_samples
    r.synchronize()
    wall = time.perf_counter() - t0
    print(f"[bench] {args.steps} frames in {wall:.3f} s", file=sys.stderr, flush=True)

    ksum = sum(kernel_ms) / 1e3
    gather_bytes = evaluated * 8 * 2 + args.steps * w.width * w.height * 4
    achieved = gather_bytes / ksum / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("workload") == w.name:
            traffic = t.get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": nominal / wall / 1e9,
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "volume": "512x512x512 u16", "image": "1920x1080",
                   "views": len(cams), "rate": w.rate, "bricks": "256^3 + 2-voxel apron",
                   "lod": 0, "tf": "4096 x RGBA f32, 3 Gaussian bumps", "fov_deg": 60},
        "launch": tuning or {"samples_per_batch": args.batch, "tile_group": args.tile_group,
                             "xcd_split": args.xcd_split, "tile_width": args.tile_width,
                             "brick": w.brick, "cell": 0 if args.no_skip else args.cell},
        "fps": args.steps / wall,
        "evaluated_gsamples_per_s": evaluated / wall / 1e9,
        "nominal_samples_per_frame": nominal // args.steps,
        "evaluated_samples_per_frame": evaluated // args.steps,
        "kernel_ms_per_frame": ksum / args.steps * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": None,
    }
    r.cleanup()
    if not args.no_cpu_baseline:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            threads = min(16, os.cpu_count() or 1)
        print(f"[bench] CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(vol, w, cams[:args.cpu_views], threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
