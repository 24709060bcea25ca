"""Benchmark: ray samples/s and frames/s of the HIP raycaster on the medium
workload (512^3 uint16, 256^3 bricks + 2-voxel apron, 3 LODs, 4096-entry
Gaussian-bump TF, 1920x1080 perspective 60 degree FoV, 36-view orbit,
sampling rate 1.5, min/max empty-space skipping), rendered at LOD 0.

A step is one orbit: all 36 views of the 1080p orbit raycast in one
persistent-kernel launch (tvk_render_batch), with the volume already
bricked into the device pool and the images left in HBM; each step waits
for its per-view sample counters.  `value` is nominal samples per second (sum over pixels of
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "raycast_medium_orbit_batch_traffic.json")


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
    isa = "AVX-512" if O.lib().baseline_uses_avx512() else "AVX2"
    return {"value": nominal / secs / 1e9, "unit": "Gsamples/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(cams)} of the 36 orbit views at 1920x1080 on the 512^3 volume, "
                      f"LOD 0, {secs:.1f} s wall on {threads} threads = every host core "
                      f"(nproc {threads}; FP32 OpenMP raycaster built for {isa}, same model, "
                      f"empty-space cells skipped span by span)",
            "fps": len(cams) / secs, "isa": isa}


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


def run_interactive(r, w, views, tune=None, warmup: int = 4) -> dict:
    """Sequential interactive frames on the medium workload: the 36 orbit
    cameras painted one per frame (paint + frame stats), after retuning the
    launch shape for single-view launches."""
    from imagevis3d_amd import workloads
    tuning = tune(views[4:5], "single-view paints") if tune else None
    cams = w.cameras
    for i in range(warmup):
        workloads.set_camera(r, cams[i % len(cams)])
        assert r.paint(), r.lastError()
        r.frameStats()
    r.synchronize()
    kernel_ms, nominal = [], 0
    t0 = time.perf_counter()
    for cam in cams:
        workloads.set_camera(r, cam)
        if not r.paint():
            raise RuntimeError(r.lastError())
        st = r.frameStats()  # waits for this frame's events and counters
        kernel_ms.append(st.kernel_ms)
        nominal += st.nominal_samples
    r.synchronize()
    wall = time.perf_counter() - t0
    return {"frames": len(cams), "fps": len(cams) / wall, "ms_per_frame": wall / len(cams) * 1e3,
            "kernel_ms_per_frame": sum(kernel_ms) / len(cams), "gsamples_per_s": nominal / wall / 1e9,
            "launch": tuning}


def run_serving(duration_s: float, rate_hz: float, seed: int = 5, lanes: int = 4,
                tune_size=(512, 512), part_cost: int = 0) -> dict:
    """The serving workload: seeded Poisson arrivals at `rate_hz` for
    `duration_s` over a pool of 7 device-generated volumes (64^3, 256^3 and
    512^3 in uint8 and uint16, one 2048^3 uint16), each request a random
    camera, a random 1D or 2D TF (two 1D slots and one 2D slot per volume)
    and an image of 256^2, 512^2, 1920x1080 or 3840x2160; each volume's
    launch shape autotuned once at setup.  Returns latency percentiles
    (submit -> image in host memory) and aggregate samples/s."""
    import threading

    import numpy as np

    from imagevis3d_amd import workloads
    from imagevis3d_amd.renderer import RenderMode
    from imagevis3d_amd.serving import RenderServer

    t0 = time.perf_counter()
    rs = workloads.serving_pool()
    setup_s = time.perf_counter() - t0
    # each volume's launch shape tuned once on one oblique view (setup, not
    # timed) of tune_size
    t0 = time.perf_counter()
    tuned = []
    for r in rs:
        t = r.autotune([dict(view_pos=(0.9, 0.6, 1.4), view_dir=(-0.9, -0.6, -1.4), fov=50.0,
                             width=tune_size[0], height=tune_size[1], tf_slot=0)], flags=r.TUNE_LAUNCH)
        tuned.append({"samples_per_batch": t.samples_per_batch, "tile_width": t.tile_width,
                      "work_group_waves": t.wg_waves, "brick_staging": t.brick_staging})
    tune_s = time.perf_counter() - t0
    srv = RenderServer(arena_bytes=4 << 30, lanes=lanes)
    srv.setPartCost(part_cost)  # heavy requests render in parts (0: the default cost)
    ids = [srv.addVolume(r) for r in rs]
    # warm every kernel variant and graph once (not timed)
    for v in range(len(ids)):
        for mode, slot in ((RenderMode.RM_1DTRANS, 0), (RenderMode.RM_2DTRANS, 2)):
            srv.wait(srv.submit(ids[v], dict(view_pos=(0, 0, 2), view_dir=(0, 0, -1), fov=50.0,
                                             width=64, height=64, tf_slot=slot), mode))

    def drive(rate, duration, rng):
        """Poisson arrivals at `rate` for `duration` s: (requests, results, wall s)."""
        n = int(duration * rate)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
        rq = []
        for _ in range(n):
            v, view, mode = workloads.serving_request(rng, len(ids))
            rq.append((ids[v], view, mode))
        tickets = [None] * n
        res = [None] * n

        def waiter():
            for i in range(n):
                while tickets[i] is None:
                    time.sleep(0.0002)
                res[i] = srv.wait(tickets[i])[1]

        th_wait = threading.Thread(target=waiter)
        th_wait.start()
        start = time.perf_counter()
        for i, (vid, view, mode) in enumerate(rq):
            delay = start + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(vid, view, mode, readback=False)
        th_wait.join()
        return rq, res, time.perf_counter() - start

    rng = np.random.default_rng(seed)
    reqs, results, wall = drive(rate_hz, duration_s, rng)
    n_req = len(reqs)
    st = srv.stats()
    lat = np.array([r.latency_ms for r in results])
    nominal = sum(r.nominal_samples for r in results)
    # the same mix at half the rate, below this GPU's capacity: latency
    # without the queueing of the saturated run
    h_reqs, h_res, h_wall = drive(rate_hz / 2, duration_s / 2, np.random.default_rng(seed + 1))
    h_lat = np.array([r.latency_ms for r in h_res])
    half = {"offered_rate_hz": rate_hz / 2, "completed_rate_hz": len(h_reqs) / h_wall,
            "requests": len(h_reqs), "p50_ms": float(np.percentile(h_lat, 50)),
            "p99_ms": float(np.percentile(h_lat, 99)),
            "mean_queue_ms": float(np.mean([r.queue_ms for r in h_res]))}

    def by(key):
        # latency percentiles per request class
        groups = {}
        for (vid, view, mode), x in zip(reqs, lat):
            groups.setdefault(key(vid, view, mode), []).append(x)
        return {k: {"n": len(v), "p50_ms": round(float(np.percentile(v, 50)), 2),
                    "p99_ms": round(float(np.percentile(v, 99)), 2)}
                for k, v in sorted(groups.items())}

    vol_name = {vid: f"{d[0]}^3-u{bits}" for vid, (d, bits, _) in zip(ids, workloads.SERVING_POOL)}
    out = {
        "workload": f"serving-poisson{rate_hz:g}hz-{duration_s:g}s-7vols",
        "requests": n_req, "offered_rate_hz": rate_hz, "completed_rate_hz": n_req / wall,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "mean_queue_ms": float(np.mean([r.queue_ms for r in results])),
        "gsamples_per_s": nominal / wall / 1e9, "launches": int(st.launches),
        "max_jobs_per_launch": int(st.max_jobs_per_launch), "max_lanes_busy": int(st.max_lanes_busy),
        "split_requests": int(st.split_requests), "parts_launched": int(st.parts_launched),
        "volume_setup_s": round(setup_s, 2), "launch_tuning_s": round(tune_s, 2),
        "launch_tuning": tuned,
        # the tail by class: the heaviest class (4K on the 2048^3 volume)
        # sets p99 by its own render time
        "latency_by_image": by(lambda vid, view, mode: f"{view['width']}x{view['height']}"),
        "latency_by_volume_and_tf": by(
            lambda vid, view, mode: vol_name[vid] + ("-2d" if mode == RenderMode.RM_2DTRANS else "-1d")),
        "half_rate": half,
        "note": "image in pinned host memory counts as done; images 256^2..3840x2160",
    }
    srv.close()
    for r in rs:
        r.cleanup()
    return out


def run_batched(batches: int = 3, seed: int = 3, tune: bool = True, groups: bool = True,
                batch: int = 0) -> dict:
    """The out-of-core batched workload: a 2048x2048x1024 uint16 volume
    (device-generated fBm + layered slabs, copied to host memory) streamed
    into a 64 GB device brick pool (128^3 bricks + 2-voxel apron, 5 LODs);
    4 seeded 2D TFs (256x256 Gaussian lobes over value x gradient
    magnitude) with Phong shading; 8 orbit cameras x 4 TFs = 32 views at
    1920x1080 per persistent-kernel launch; brick size (32^3 ... 256^3,
    the paged volume set up again from the host copy for each), empty-space
    cell and launch shape autotuned on the 32 views.  With `groups` the 4
    views of each camera walk their rays once (TF groups; the images equal
    the ungrouped launch's bit for bit)."""
    from imagevis3d_amd import tuvok, workloads
    from imagevis3d_amd.renderer import VOL_FBM_PLANES

    dims, bits = workloads.BATCHED_DIMS, 16
    t0 = time.perf_counter()
    # exact gradient range from an in-core pass, then the host copy
    g = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, False)
    assert g.initialize(tuvok.createContext(64, 64, device=0)), g.lastError()
    g.setGeneratedVolume(VOL_FBM_PLANES, seed, dims, bits, 256, 2, 1)
    _, max_grad = g.histogram2D(256, 256)
    vol = g.readLOD(0)
    g.cleanup()
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = workloads.batched_renderer(vol, max_grad, 64 * 10**9)
    r.setTFGrouping(groups)
    if batch:  # samples per gather batch of the untuned launch (0: the default)
        r.setKernelTuning(batch, 4, 0)
    plan_s = time.perf_counter() - t0
    views = workloads.batched_views(dims)
    t0 = time.perf_counter()
    _, st0, _ = r.renderBatch(views, readback=False)  # pages the view set's bricks in
    page_s = time.perf_counter() - t0
    paged_bytes, paging_host_ms = r.pagingStats()
    t0 = time.perf_counter()
    # tune=False (counter runs): the configured 128^3 bricks, launch shape
    # timed only (flags 0 times the current settings)
    t = r.autotune(views, flags=(r.TUNE_BRICK | r.TUNE_CELL | r.TUNE_LAUNCH) if tune else 0)
    tune_s = time.perf_counter() - t0
    kms, nominal, evaluated, gathered = [], 0, 0, 0
    t0 = time.perf_counter()
    for _ in range(batches):
        _, st, k = r.renderBatch(views, readback=False)
        kms.append(k)
        nominal += sum(x.nominal_samples for x in st)
        evaluated += sum(x.evaluated_samples for x in st)
        gathered += sum(x.gathered_samples for x in st)
    wall = time.perf_counter() - t0
    ms = r.memStats()
    ksum = sum(kms) / 1e3
    # gather bytes: value + 6 gradient neighbours, 8 corners x 2 B each, per
    # sample whose voxels a march gathered (a TF group gathers each sample
    # of its walk once for its 4 views), + 4 B per output pixel
    achieved = (gathered * 7 * 16 + batches * len(views) * 1920 * 1080 * 4) / ksum / 1e9
    out = {
        "workload": "batched-2048x2048x1024-u16-paged64GB-2dtf-phong-8cams-x-4tfs-1080p",
        "views_per_launch": len(views), "launches": batches, "tf_groups": groups,
        "gsamples_per_s": nominal / wall / 1e9, "views_per_s": batches * len(views) / wall,
        "ms_per_launch": wall / batches * 1e3, "kernel_ms_per_launch": ksum / batches * 1e3,
        "evaluated_gsamples_per_s": evaluated / wall / 1e9,
        "gathered_gsamples_per_s": gathered / wall / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS},
        "tuned": {"brick": t.brick_size, "cell": t.cell, "samples_per_batch": t.samples_per_batch,
                  "tile_width": t.tile_width, "tile_group": t.tile_group,
                  "work_group_waves": t.wg_waves, "tuned_ms": round(t.kernel_ms, 2),
                  "start_ms": round(t.baseline_ms, 2), "settings_timed": t.candidates,
                  "seconds": round(tune_s, 1)},
        "pool": {"slots": int(ms.pool_slots), "resident": int(ms.resident_bricks),
                 "uploaded": int(ms.bricks_uploaded)},
        "setup_s": {"generate": round(gen_s, 2), "host_pyramid": round(plan_s, 2),
                    "first_batch_paging": round(page_s, 2)},
        # the cold first launch: every brick the 32 views sample streamed from
        # host memory (brick sources on worker threads, copies on the copy
        # stream) -- bytes of bricks over the launch's wall time
        "cold_paging": {"bytes": int(paged_bytes), "wall_s": round(page_s, 3),
                        "gb_per_s": paged_bytes / page_s / 1e9,
                        "host_paging_ms": round(paging_host_ms, 1)},
    }
    r.cleanup()
    return out


def run_learned_lod(n: int = 2048, edge: int = 128, n_lod: int = 6, eval_views: int = 24,
                    eval_size=(3840, 2160), train_size=(1920, 1080), train_views: int = 16,
                    cal_views_n: int = 8, cal_target: float = 0.9) -> dict:
    """The learned-LOD workload: an n^3 uint8 volume of multi-scale noise
    with 64 thin tubes (device-generated, in HBM), 128^3 bricks, 6 LODs,
    1D TF (tubes_tf: opaque tubes over a faint noise bump), rate 2.  The
    bf16 MLP is trained on the renderer's own fine-vs-coarse renders
    (`train_views` views x 4 TFs at `train_size`), its selection budget
    calibrated on `cal_views_n` views (mean RMSE `cal_target`/255), then 24
    new random views at 3840x2160 with a new TF are rendered at full
    resolution and with the learned per-region LOD at two operating points:
    the error target alone ("target") and under the 16 GB brick budget
    ("budgeted"), and at 1/2, 1/4 and 1/8 of the calibrated error budget
    ("target_curve": how far the budget must drop for every frame, not only
    their mean, to stay within 1/255).  "checked" holds every frame to
    1/255 at run time: the per-frame check estimates each learned frame's
    RMSE from a sample of level-0 tiles and renders the frame at level 0
    above the bound; its selection (the error target at 1, 1/2 or 1/4 of
    its factor, or the 16 GB brick budget) is the fastest whose
    calibration frames all end within 1/255; its frame rate is that of a
    frame loop in which the host selects the next frame while the device
    renders the current one.
    "out_of_core" replays the budgeted frames from a 16 GB paged pool over a
    host-memory copy of the volume.  Frame time of the learned path
    includes features, MLP and selection."""
    import math

    import numpy as np

    from imagevis3d_amd import learned_lod as LL
    from imagevis3d_amd import transfer, tuvok
    from imagevis3d_amd.renderer import VOL_NOISE_TUBES, brick_slot_bytes

    def views(k, seed, w, h):
        rng = np.random.default_rng(seed)
        out = []
        for _ in range(k):
            th, ph = rng.uniform(0, 2 * math.pi), rng.uniform(-1.0, 1.0)
            d = rng.uniform(1.2, 1.9)
            eye = (d * math.cos(ph) * math.sin(th), d * math.sin(ph), d * math.cos(ph) * math.cos(th))
            out.append(dict(view_pos=eye, view_dir=tuple(-e for e in eye), up_dir=(0.0, 1.0, 0.0),
                            fov=50.0, width=w, height=h))
        return out

    def say(msg):
        print(f"[bench] learned LOD: {msg}", file=sys.stderr, flush=True)

    dims = (n, n, n)
    r = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, False)
    assert r.initialize(tuvok.createContext(480, 270, device=0)), r.lastError()
    r.setRendererTarget(tuvok.renderer.types.RT_Capture)
    say(f"generating {n}^3")
    t0 = time.perf_counter()
    r.setGeneratedVolume(VOL_NOISE_TUBES, 4, dims, 8, edge, 2, n_lod)
    setup_s = time.perf_counter() - t0
    r.setSampleRateModifier(2.0)
    hist = LL.region_histograms(r)
    detail = LL.region_detail(r)
    say("training data")
    t0 = time.perf_counter()
    X, Y = LL.training_rows(r, dims, edge, n_lod, views(train_views, 11, *train_size),
                            [transfer.tubes_tf(256, seed=s) for s in range(4)], hist, detail)
    data_s = time.perf_counter() - t0
    say(f"training on {len(X)} rows")
    t0 = time.perf_counter()
    wbits, b, loss = LL.train(X, Y, steps=3000)
    train_s = time.perf_counter() - t0
    pred = LL.Predictor(wbits, b)
    tf = transfer.tubes_tf(256, seed=9)
    r.set1DTrans(tf)
    mom = LL.tf_moments(hist, tf)
    slot_bytes = brick_slot_bytes(edge, 2, 8)
    # the brick budget of the configuration: 16 GB of device slots (padded
    # with their apron and blocked layout) for the chosen levels' bricks
    mem_budget = 16 * 10**9
    cal_views = views(cal_views_n, 12, *eval_size)
    # two operating points: under the brick budget, and with the error
    # target alone (no brick cap); the budget factor is calibrated so that
    # the mean RMSE over the calibration views stays at cal_target/255
    points = {"budgeted": mem_budget, "target": 0}
    say("calibrating the budget factors")
    cal = {k: LL.calibrate_safety(r, pred, cal_views, dims, edge, n_lod, hist, mom, cal_target,
                                  slot_bytes, mb, detail, reduce="mean")
           for k, mb in points.items()}
    # The per-frame check: each learned frame also renders one ray tile in
    # check_k at level 0, estimates its own RMSE from them and, above
    # check_bound/255, is rendered again at level 0.  The budget-factor
    # scale it runs at is picked on the calibration views (the fastest
    # whose frames all end within 1/255), then measured on new views.
    check_k, check_bound = 32, 0.9
    cal_refs = []
    for v in cal_views:
        r.resize((v["width"], v["height"]))
        LL.set_view(r, v)
        cal_refs.append(LL.full_resolution(r, dims, edge))

    def checked_accuracy(vs, refs, sel):
        """Per-frame RMSE vs level 0 and fallbacks, one capture per view;
        sel = (operating point, budget-factor scale)."""
        point, scale = sel
        r.setRegionLODCheck(check_k, check_bound)
        errs, est, fell = [], [], 0
        for v, ref in zip(vs, refs):
            r.resize((v["width"], v["height"]))
            LL.set_view(r, v)
            lev, _, _ = LL.choose_levels(pred, v, dims, edge, n_lod, hist, mom, cal[point][0] * scale,
                                         slot_bytes, points[point], detail)
            r.setRegionLODs(lev)
            img = r.capture()
            c = r.regionLODCheck()
            fell += c.fell_back
            est.append(c.rmse_estimate)
            errs.append(LL.rmse255(img, ref))
            r.setRegionLODs(None)
        r.setRegionLODCheck(0, 0.0)
        return errs, est, fell

    def pipelined_ms(vs, sel=None):
        """Wall time per frame of a frame loop in which the host selects the
        next frame's levels while the device renders the current one
        (sel None: every frame at full resolution, no selection)."""
        r.resize((vs[0]["width"], vs[0]["height"]))
        if sel is not None:
            r.setRegionLODCheck(check_k, check_bound)
        r.synchronize()
        t0 = time.perf_counter()
        for v in vs:
            LL.set_view(r, v)
            if sel is not None:
                lev, _, _ = LL.choose_levels(pred, v, dims, edge, n_lod, hist, mom,
                                             cal[sel[0]][0] * sel[1], slot_bytes, points[sel[0]], detail)
                r.setRegionLODs(lev)
            else:
                r.setRegionLODs(zero_regions)
            assert r.paint(), r.lastError()
        r.synchronize()
        ms = (time.perf_counter() - t0) / len(vs) * 1e3
        r.setRegionLODCheck(0, 0.0)
        r.setRegionLODs(None)
        return ms

    zero_regions = np.zeros((n // edge) ** 3, np.uint8)
    # candidate selections of the checked loop: the error target alone at
    # fractions of its factor, and the 16 GB brick budget at its factor
    check_sels = (("target", 1.0), ("target", 0.5), ("target", 0.25), ("budgeted", 1.0))
    check_cal = {}
    say("calibrating the checked loop")
    for sel in check_sels:
        errs, _, fell = checked_accuracy(cal_views, cal_refs, sel)
        check_cal[sel] = {"rmse_max": float(np.max(errs)), "fallbacks": fell,
                          "ms": pipelined_ms(cal_views, sel)}
    ok = [sel for sel in check_sels if check_cal[sel]["rmse_max"] <= 1.0]
    check_sel = min(ok, key=lambda sel: check_cal[sel]["ms"]) if ok else ("target", 0.25)

    # full resolution at its best: the uniform LOD-0 path with its
    # empty-space cells and launch shape tuned for this volume (the default
    # 16^3 cells cost a march step each), and per view the faster of it and
    # the per-region path with every region at level 0
    say("tuning the full-resolution path")
    r.resize(eval_size)
    t_full = r.autotune(cal_views[:2], flags=r.TUNE_CELL | r.TUNE_LAUNCH)
    full_tuning = {"cell": t_full.cell, "samples_per_batch": t_full.samples_per_batch,
                   "tile_width": t_full.tile_width, "tile_group": t_full.tile_group,
                   "work_group_waves": t_full.wg_waves, "brick_staging": t_full.brick_staging,
                   "tuned_ms": round(t_full.kernel_ms, 2), "start_ms": round(t_full.baseline_ms, 2)}
    full_ms, full_uniform_ms, full_regions_ms, s_full = [], [], [], 0
    acc = {k: {"ms": [], "kernel_ms": [], "sel_ms": [], "rmse": [], "rmse_uniform": [], "mem": [],
               "samples": 0} for k in points}
    chk = {"rmse": [], "estimate": [], "fallbacks": 0}
    ooc_in = []  # (view, budgeted levels, in-core image, level-0 frame, selection ms)
    # the error/speed trade-off below the calibrated point: the same views
    # at fractions of the calibrated budget factor (no brick cap)
    curve_scales = (0.5, 0.25, 0.125)
    curve = {s: {"ms": [], "rmse": []} for s in curve_scales}
    for i, v in enumerate(views(eval_views, 13, *eval_size)):
        say(f"view {i + 1}/{eval_views}")
        r.resize(eval_size)
        LL.set_view(r, v)
        r.setRegionLODs(None)
        assert r.paint(), r.lastError()
        st = r.frameStats()
        s_full += st.evaluated_samples
        uni = r.capture()
        ref = LL.full_resolution(r, dims, edge)
        reg0_ms = r.frameStats().kernel_ms  # the level-0 region frame just captured
        full_uniform_ms.append(st.kernel_ms)
        full_regions_ms.append(reg0_ms)
        full_ms.append(min(st.kernel_ms, reg0_ms))
        for k, mb in points.items():
            a = acc[k]
            t0 = time.perf_counter()
            lev, _, m = LL.choose_levels(pred, v, dims, edge, n_lod, hist, mom, cal[k][0],
                                         slot_bytes, mb, detail)
            r.setRegionLODs(lev)
            sel = (time.perf_counter() - t0) * 1e3
            assert r.paint(), r.lastError()
            st = r.frameStats()
            a["samples"] += st.evaluated_samples
            img = r.capture()
            a["sel_ms"].append(sel)
            a["kernel_ms"].append(st.kernel_ms)
            a["ms"].append(st.kernel_ms + sel)
            a["rmse"].append(LL.rmse255(img, ref))
            a["rmse_uniform"].append(LL.rmse255(img, uni))
            a["mem"].append(m)
            if k == "budgeted":  # replayed out of core below
                ooc_in.append((v, lev, img, ref, sel))
            r.setRegionLODs(None)
        errs, est, fell = checked_accuracy([v], [ref], check_sel)
        chk["rmse"] += errs
        chk["estimate"] += est
        chk["fallbacks"] += fell
        for s, c in curve.items():
            t0 = time.perf_counter()
            lev, _, _ = LL.choose_levels(pred, v, dims, edge, n_lod, hist, mom, cal["target"][0] * s,
                                         slot_bytes, 0, detail)
            r.setRegionLODs(lev)
            sel = (time.perf_counter() - t0) * 1e3
            assert r.paint(), r.lastError()
            c["ms"].append(r.frameStats().kernel_ms + sel)
            c["rmse"].append(LL.rmse255(r.capture(), ref))
            r.setRegionLODs(None)

    def summary(k):
        a = acc[k]
        return {"fps_learned": 1e3 / float(np.mean(a["ms"])),
                "speedup": float(np.mean(full_ms) / np.mean(a["ms"])),
                "kernel_speedup": float(np.mean(full_ms) / np.mean(a["kernel_ms"])),
                "rmse_mean": float(np.mean(a["rmse"])), "rmse_max": float(np.max(a["rmse"])),
                "rmse_vs_global_sampling_mean": float(np.mean(a["rmse_uniform"])),
                "evaluated_sample_ratio": a["samples"] / max(s_full, 1),
                "selection_ms_mean": float(np.mean(a["sel_ms"])),
                "brick_bytes_mean": float(np.mean(a["mem"])),
                "brick_budget": points[k] or None,
                "budget_factor": cal[k][0], "calibration_worst_rmse": cal[k][1]}

    # frame rate of the checked loop and of full resolution, both pipelined
    # over the same held-out views
    say("pipelined frame loops")
    eval_vs = views(eval_views, 13, *eval_size)
    full_pipe = pipelined_ms(eval_vs)
    checked_pipe = pipelined_ms(eval_vs, check_sel)
    checked = {
        "tile_sample": check_k, "rmse_bound": check_bound, "selection": check_sel[0],
        "factor_scale": check_sel[1],
        "fps_learned": 1e3 / checked_pipe, "fps_full": 1e3 / full_pipe,
        "speedup": full_pipe / checked_pipe,
        # against the per-view best full-resolution kernel time
        "speedup_vs_full_kernel": float(np.mean(full_ms)) / checked_pipe,
        "rmse_mean": float(np.mean(chk["rmse"])), "rmse_max": float(np.max(chk["rmse"])),
        "frames_within_1": float(np.mean(np.asarray(chk["rmse"]) <= 1.0)),
        "estimate_mean": float(np.mean(chk["estimate"])),
        "fallback_frames": chk["fallbacks"], "frames": len(chk["rmse"]),
        "calibration": {f"{p}x{s}": c for (p, s), c in check_cal.items()},
    }

    # the predictor alone on one frame's inputs (every region x coarser level)
    xs, _, _ = LL.plan_features(views(1, 14, *eval_size)[0], dims, edge, n_lod, hist, mom, detail)
    mlp_ms = []
    for _ in range(20):
        pred(xs)
        mlp_ms.append(pred.kernel_ms)
    mlp_t = float(np.median(mlp_ms)) * 1e-3
    mlp_flops = 28800.0 * len(xs)  # 2*(32*64 + 3*64*64 + 64) per row
    pred.close()
    out_of_core = run_learned_lod_out_of_core(r, ooc_in, tf, edge, n_lod, mem_budget,
                                              float(np.mean(full_ms)))
    r = None  # released by the out-of-core leg
    out = {
        "workload": f"learned-lod-{n}^3-u8-noise-tubes-{edge}^3-bricks-{n_lod}lods-"
                    f"{eval_size[0]}x{eval_size[1]}-{eval_views}views",
        # full resolution per view: the faster of the tuned uniform LOD-0
        # path and the per-region path at level 0 (speed-ups are against it)
        "fps_full": 1e3 / float(np.mean(full_ms)),
        "fps_full_uniform": 1e3 / float(np.mean(full_uniform_ms)),
        "fps_full_regions_level0": 1e3 / float(np.mean(full_regions_ms)),
        "full_tuning": full_tuning,
        "rmse_target": 1.0,
        "calibration": {"views": cal_views_n, "mean_rmse_target": cal_target},
        "training": {"views": train_views, "tfs": 4, "size": list(train_size)},
        "target": summary("target"),
        "budgeted": summary("budgeted"),
        # every frame held to 1/255 by the per-frame check (frame loop with
        # the host's selection overlapped with the device's render)
        "checked": checked,
        # every point: speedup, mean / max RMSE and the share of frames
        # within 1/255 (the calibrated point first)
        "target_curve": [{"factor_scale": 1.0, "speedup": summary("target")["speedup"],
                          "rmse_mean": float(np.mean(acc["target"]["rmse"])),
                          "rmse_max": float(np.max(acc["target"]["rmse"])),
                          "frames_within_1": float(np.mean(np.asarray(acc["target"]["rmse"]) <= 1.0))}] +
                        [{"factor_scale": s, "speedup": float(np.mean(full_ms) / np.mean(c["ms"])),
                          "rmse_mean": float(np.mean(c["rmse"])), "rmse_max": float(np.max(c["rmse"])),
                          "frames_within_1": float(np.mean(np.asarray(c["rmse"]) <= 1.0))}
                         for s, c in curve.items()],
        "mlp": {"rows": int(len(xs)), "kernel_ms": mlp_t * 1e3, "tflops": mlp_flops / mlp_t / 1e12,
                "mfma_peak_frac": mlp_flops / mlp_t / 2.5e15, "dtype": "bf16 x bf16 -> f32"},
        "train_rows": int(len(X)), "train_loss": loss,
        "seconds": {"volume": round(setup_s, 2), "training_data": round(data_s, 2),
                    "training": round(train_s, 2)},
        "out_of_core": out_of_core,
    }
    return out


def run_learned_lod_out_of_core(r, frames, tf, edge: int, n_lod: int, pool_bytes: int,
                                full_ms_in_core: float) -> dict:
    """The learned-LOD volume out of core: LOD 0 copied to host memory, the
    in-core renderer `r` released (it is cleaned up here), and the bricks of
    each budgeted selection paged from host memory into a pool of
    `pool_bytes` (the configuration's 16 GB).  `frames` holds (view, levels,
    in-core image, level-0 frame, selection ms) of the budgeted point: each
    frame is rendered cold (its missing bricks streamed in, multi-pass) and
    again warm (every brick resident), its image compared with the in-core
    one and its RMSE with level 0."""
    import numpy as np

    from imagevis3d_amd import learned_lod as LL
    from imagevis3d_amd import tuvok

    def say(msg):
        print(f"[bench] learned LOD out of core: {msg}", file=sys.stderr, flush=True)

    say("copying LOD 0 to host memory")
    t0 = time.perf_counter()
    host_vol = r.readLOD(0)
    copy_s = time.perf_counter() - t0
    r.cleanup()
    rp = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, False)
    assert rp.initialize(tuvok.createContext(480, 270, device=0)), rp.lastError()
    rp.setRendererTarget(tuvok.renderer.types.RT_Capture)
    say("host pyramid and brick ranges")
    t0 = time.perf_counter()
    rp.setPagedFlatVolume(host_vol, edge, pool_bytes, apron=2, n_lod=n_lod)
    setup_s = time.perf_counter() - t0
    rp.setSampleRateModifier(2.0)
    rp.set1DTrans(tf)
    warm_ms, cold_ms, rmse, paged, equal = [], [], [], [], 0
    for i, (v, lev, img_ic, ref, sel) in enumerate(frames):
        say(f"frame {i + 1}/{len(frames)}")
        rp.resize((v["width"], v["height"]))
        LL.set_view(rp, v)
        rp.setRegionLODs(lev)
        b0, _ = rp.pagingStats()
        t0 = time.perf_counter()
        img = rp.capture()
        cold_ms.append((time.perf_counter() - t0) * 1e3 + sel)
        paged.append(rp.pagingStats()[0] - b0)
        assert rp.paint(), rp.lastError()  # warm: the selection's bricks are resident
        warm_ms.append(rp.frameStats().kernel_ms + sel)
        equal += int(np.array_equal(img, img_ic))
        rmse.append(LL.rmse255(img, ref))
    ms = rp.memStats()
    rp.cleanup()
    del host_vol
    return {
        "pool_bytes": pool_bytes, "pool_slots": int(ms.pool_slots),
        "host_copy_s": round(copy_s, 2), "paged_setup_s": round(setup_s, 2),
        "frames": len(frames), "frames_equal_in_core": equal,
        "fps_learned_warm": 1e3 / float(np.mean(warm_ms)),
        # against level 0 in core (which does not fit the 16 GB pool)
        "speedup_warm_vs_full_in_core": full_ms_in_core / float(np.mean(warm_ms)),
        "rmse_mean": float(np.mean(rmse)), "rmse_max": float(np.max(rmse)),
        "cold_frame_ms_mean": float(np.mean(cold_ms)),
        "paged_gb_per_frame_mean": float(np.mean(paged)) / 1e9,
        "paging_gb_per_s": float(np.sum(paged)) / (float(np.sum(cold_ms)) / 1e3) / 1e9,
        "bricks_evicted": int(ms.bricks_evicted),
    }


def run_mlp_only(rows: int = 163840, repeats: int = 20) -> dict:
    """The learned-LOD predictor alone (no torch in the process): random
    bf16 weights of the 32 -> 4 x 64 -> 1 MLP over `rows` feature rows (one
    4096^3 frame's: 32^3 regions x 5 coarser levels), timed with the
    library's HIP events; for kernel traces of lod_mlp_kernel."""
    import numpy as np

    from imagevis3d_amd import learned_lod as LL

    rng = np.random.default_rng(0)
    n_w = LL.N_IN * LL.HIDDEN + 3 * LL.HIDDEN * LL.HIDDEN + LL.HIDDEN
    w = LL.to_bf16_bits(rng.normal(0.0, 0.15, n_w).astype(np.float32))
    b = rng.normal(0.0, 0.05, 4 * LL.HIDDEN + 1).astype(np.float32)
    x = rng.normal(0.0, 1.0, (rows, LL.N_IN)).astype(np.float32)
    pred = LL.Predictor(w, b)
    ms = []
    for _ in range(repeats + 1):
        pred(x)
        ms.append(pred.kernel_ms)
    pred.close()
    t = float(np.median(ms[1:])) * 1e-3
    flops = 2.0 * (LL.N_IN * LL.HIDDEN + 3 * LL.HIDDEN * LL.HIDDEN + LL.HIDDEN) * rows
    return {"workload": f"lod-mlp-{rows}-rows", "kernel_ms": t * 1e3, "tflops": flops / t / 1e12,
            "mfma_peak_frac": flops / t / 2.5e15, "dtype": "bf16 x bf16 -> f32"}


def main() -> int:
    if "--calibrate-copy" in sys.argv:
        return calibrate_copy()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps; a step renders the 36-view orbit in one batched launch")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-interactive", action="store_true",
                    help="skip the sequential one-view-per-paint leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=36)
    ap.add_argument("--batch", type=int, default=8, help="samples per march step")
    ap.add_argument("--tile-group", type=int, default=4, help="ray tiles per XCD turn")
    ap.add_argument("--xcd-split", type=int, default=0, help="0 or frame regions per row")
    ap.add_argument("--tile-width", type=int, default=8, help="ray tile width (4, 8, 16)")
    ap.add_argument("--cell", type=int, default=16, help="empty-space cell edge in voxels")
    ap.add_argument("--no-skip", action="store_true", help="march whole bricks (no cell skipping)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the given launch settings instead of autotuning them")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--serving-seconds", type=float, default=20.0,
                    help="seconds of arrivals at the full rate (then half as long at half the rate)")
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--serving-only", action="store_true")
    ap.add_argument("--serving-lanes", type=int, default=4, help="streams of the render server")
    ap.add_argument("--serving-tune-size", type=int, nargs=2, default=(512, 512),
                    help="image size of the view each serving volume's launch shape is tuned on")
    ap.add_argument("--serving-part-cost", type=float, default=0,
                    help="estimated samples above which a request renders in parts (0: the server's "
                         "default; 1.8e19: never split)")
    ap.add_argument("--no-batched", action="store_true", help="skip the out-of-core batched workload")
    ap.add_argument("--batched-only", action="store_true")
    ap.add_argument("--no-learned-lod", action="store_true", help="skip the learned-LOD workload")
    ap.add_argument("--learned-lod-only", action="store_true")
    ap.add_argument("--learned-lod-size", type=int, default=4096)
    ap.add_argument("--staging", type=int, default=0,
                    help="with --no-autotune: 1 marches the 1D TF over LDS-staged cells")
    ap.add_argument("--batched-batch", type=int, default=0,
                    help="batched workload: samples per gather batch to start from (0: default)")
    ap.add_argument("--no-tf-groups", action="store_true",
                    help="batched workload: render each of a camera's TF views on its own")
    ap.add_argument("--mlp-only", action="store_true",
                    help="time the learned-LOD MLP alone (random weights, no torch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the renderer runs on one GPU; sharding across GPUs is not built)",
              file=sys.stderr)
        return 2

    from imagevis3d_amd import tuvok, workloads

    if args.serving_only:
        print(json.dumps(run_serving(args.serving_seconds, args.serving_rate, lanes=args.serving_lanes,
                                     tune_size=tuple(args.serving_tune_size),
                                     part_cost=min(int(args.serving_part_cost), 2**64 - 1))))
        return 0
    if args.batched_only:
        print(json.dumps(run_batched(tune=not args.no_autotune, groups=not args.no_tf_groups,
                                     batch=args.batched_batch)))
        return 0
    if args.learned_lod_only:
        print(json.dumps(run_learned_lod(args.learned_lod_size)))
        return 0
    if args.mlp_only:
        print(json.dumps(run_mlp_only()))
        return 0

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
    r.setBrickStaging(bool(args.staging))
    workloads.apply(r, w, w.cameras[0])
    if args.no_skip:
        r.setEmptySpaceSkipping(False)
    cams = w.cameras
    views = [dict(view_pos=c.view_pos, view_dir=c.view_dir, up_dir=c.up_dir, fov=c.fov,
                  width=w.width, height=w.height) for c in cams]

    def tune(tv, what):
        # empty-space cell and launch shape (the brick size stays the
        # configured 256^3), timed on the given views (untimed setup)
        t0 = time.perf_counter()
        t = r.autotune(tv, flags=r.TUNE_CELL | r.TUNE_LAUNCH)
        out = {"brick": t.brick_size, "cell": t.cell, "samples_per_batch": t.samples_per_batch,
               "tile_width": t.tile_width, "tile_group": t.tile_group,
               "work_group_waves": t.wg_waves, "brick_staging": t.brick_staging,
               "tuned_ms": round(t.kernel_ms, 4), "start_ms": round(t.baseline_ms, 4),
               "settings_timed": t.candidates, "seconds": round(time.perf_counter() - t0, 2)}
        print(f"[bench] autotuned for {what}: {out}", file=sys.stderr, flush=True)
        return out

    tuning = None if args.no_autotune else tune(views, "the 36-view orbit launch")
    print(f"[bench] bricked into the device pool in {brick_s:.2f} s; warmup {args.warmup}",
          file=sys.stderr, flush=True)

    # A step renders the whole orbit: 36 views of 1920x1080 in one persistent
    # launch (BatchRenderer-style capture of an orbit), images left in HBM.
    for _ in range(args.warmup):
        r.renderBatch(views, readback=False)
    r.synchronize()
    kernel_ms, evaluated, nominal = [], 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, st, kms = r.renderBatch(views, readback=False)  # waits for the launch's counters
        kernel_ms.append(kms)
        evaluated += sum(v.evaluated_samples for v in st)
        nominal += sum(v.nominal_samples for v in st)
    r.synchronize()
    wall = time.perf_counter() - t0
    frames = args.steps * len(views)
    print(f"[bench] {args.steps} orbit launches ({frames} frames) in {wall:.3f} s",
          file=sys.stderr, flush=True)

    ksum = sum(kernel_ms) / 1e3
    gather_bytes = evaluated * 8 * 2 + frames * w.width * w.height * 4
    achieved = gather_bytes / ksum / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        if t.get("workload") == w.name and t.get("launch") == "orbit-batch-36":
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
                   "views": len(cams), "views_per_launch": len(views), "rate": w.rate,
                   "bricks": "256^3 + 2-voxel apron", "lod": 0,
                   "tf": "4096 x RGBA f32, 3 Gaussian bumps", "fov_deg": 60},
        "launch": tuning or {"samples_per_batch": args.batch, "tile_group": args.tile_group,
                             "xcd_split": args.xcd_split, "tile_width": args.tile_width,
                             "brick": w.brick, "cell": 0 if args.no_skip else args.cell},
        "fps": frames / wall,
        "evaluated_gsamples_per_s": evaluated / wall / 1e9,
        "nominal_samples_per_frame": nominal // frames,
        "evaluated_samples_per_frame": evaluated // frames,
        "kernel_ms_per_launch": ksum / args.steps * 1e3,
        "kernel_ms_per_frame": ksum / frames * 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE per launch read from "
                                        + os.path.relpath(TRAFFIC_FILE, ROOT)
                                        + " (a committed profile, not measured in this run)")
                     if traffic is not None else None},
        "cpu_baseline": None,
    }
    if not args.no_interactive:
        # interactive frames: one view per paint, cameras in orbit order,
        # launch shape tuned for a single view
        out["interactive"] = run_interactive(r, w, views, None if args.no_autotune else tune)
        print(f"[bench] interactive {out['interactive']}", file=sys.stderr, flush=True)
    r.cleanup()
    if not args.no_learned_lod:
        print("[bench] learned-LOD workload", file=sys.stderr, flush=True)
        ll = run_learned_lod(args.learned_lod_size)
        out["learned_lod"] = ll
        # FPS of the frame loop whose every frame is held to 1/255 RMSE
        # (the per-frame check with its level-0 fallback)
        out["learned_lod_fps"] = ll["checked"]["fps_learned"]
        print(f"[bench] learned LOD {ll}", file=sys.stderr, flush=True)
    if not args.no_batched:
        print("[bench] out-of-core batched workload", file=sys.stderr, flush=True)
        out["batched"] = run_batched(groups=not args.no_tf_groups)
        print(f"[bench] batched {out['batched']}", file=sys.stderr, flush=True)
    if not args.no_serving:
        print(f"[bench] serving workload ({args.serving_seconds:g} s at {args.serving_rate:g} req/s)",
              file=sys.stderr, flush=True)
        sv = run_serving(args.serving_seconds, args.serving_rate, lanes=args.serving_lanes,
                         tune_size=tuple(args.serving_tune_size),
                         part_cost=min(int(args.serving_part_cost), 2**64 - 1))
        out["serving"] = sv
        out["p99_ms"] = sv["p99_ms"]
        print(f"[bench] serving {sv}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        try:
            threads = len(os.sched_getaffinity(0))  # every host core this process may use
        except AttributeError:
            threads = os.cpu_count() or 1
        print(f"[bench] CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(vol, w, cams[:args.cpu_views], threads)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
