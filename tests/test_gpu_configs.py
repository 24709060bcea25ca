"""The benchmark configurations at their own shapes on the GPU:

C3 -- the out-of-core 2D-TF + Phong workload: a crop of its fBm + slabs
      volume paged through a pool smaller than the volume (bricks evicted
      between frames, LOD 0 and LOD 1) against the FP64 oracle; and the
      full 2048x2048x1024 volume in a 64 GB pool, where the 8 cameras x 4
      TFs launch equals per-view captures bit for bit.
C4 -- the learned-LOD volume (multi-scale noise + thin tubes): a crop with
      mixed per-region levels against the oracle's per-region model; and
      2048^3 with every region at level 0 against the uniform render.
"""
import numpy as np
import pytest

from conftest import assert_parity, gpu_renderer, pixel_diff
from imagevis3d_amd import RenderMode, brick_slot_bytes, transfer, workloads
from imagevis3d_amd.renderer import VOL_FBM_PLANES, VOL_NOISE_TUBES, generate_volume_gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LIGHT = dict(ambient=(0.12, 0.12, 0.12), diffuse=(0.8, 0.8, 0.8), specular=(0.3, 0.3, 0.3),
             shininess=16.0, light_dir=(0.3, 0.6, 0.75))


def test_c3_crop_paged_2dtf_phong_matches_oracle():
    dims, B = (256, 256, 128), 64
    print("case C3 crop 256x256x128, paged 2D TF + Phong vs oracle", flush=True)
    vol = generate_volume_gpu(VOL_FBM_PLANES, 3, dims, 16)
    _, max_grad = O.histogram_2d_exact(vol, 256, 256)
    levels = O.lod_pyramid(vol, 3)
    sb = brick_slot_bytes(B, 2, 16)
    # LOD 0 has 32 bricks, LOD 1 has 4: a share of 33 slots cannot hold both
    r = workloads.batched_renderer(vol, max_grad, 33 * sb, brick=B, n_lod=3)
    r.resize((200, 120))
    r.setRendererTarget(1)
    r.setLightColors((*LIGHT["ambient"], 1.0), (*LIGHT["diffuse"], 1.0),
                     (*LIGHT["specular"], LIGHT["shininess"]), LIGHT["light_dir"])
    views = workloads.batched_views(dims, 200, 120, n_cams=4)
    w = workloads.Workload("c3-crop", dims, 16, 0, 3, B, 2, 3, None, 200, 120, 1.0, [])
    scale = float(np.float32(255.0 / float(vol.max())))
    for i, lod in enumerate([0, 1, 0, 1]):
        v = views[5 * i % len(views)]
        cam = workloads.Camera(v["view_pos"], v["view_dir"], v["up_dir"], v["fov"])
        r.setLODLimits((lod, lod))
        r.set2DTrans(transfer.gaussian_lobes_2d(256, 256, seed=40 + v["tf_slot"]), max_grad)
        workloads.apply(r, w, cam)
        img = r.capture()
        assert r.frameStats().lod == lod
        lv = levels[lod]
        kw = workloads.oracle_kwargs(w, cam, lod=lod)
        kw.update(brick_size=16, apron=1)
        ref, _, fl = O.render(O.make_params(
            lv, None, mode=1, tf2d=transfer.gaussian_lobes_2d(256, 256, seed=40 + v["tf_slot"]),
            max_grad=max_grad, tf_scale=scale, lighting=True, **LIGHT, **kw))
        assert (O.to_rgba8(ref)[..., :3] > 0).mean() > 0.05
        assert_parity(img, O.to_rgba8(ref), fl, f"C3 crop view {i} LOD {lod}")
    assert r.memStats().bricks_evicted > 0
    r.cleanup()


def test_c3_full_size_batch_equals_per_view_captures():
    print("case C3 full size: 2048x2048x1024 u16, 64 GB pool, 8 cams x 4 TFs", flush=True)
    g = gpu_renderer(width=64, height=64)
    g.setGeneratedVolume(VOL_FBM_PLANES, 3, workloads.BATCHED_DIMS, 16, 256, 2, 1)
    _, max_grad = g.histogram2D(256, 256)
    vol = g.readLOD(0)
    g.cleanup()
    r = workloads.batched_renderer(vol, max_grad, 64 * 10**9)
    views = workloads.batched_views(width=960, height=540)
    imgs, stats, _ = r.renderBatch(views)
    assert len(imgs) == 32 and all(s.lod == 0 for s in stats)
    r.setRendererTarget(1)
    r.resize((960, 540))
    for i in (0, 13, 31):
        v = views[i]
        r.set2DTrans(transfer.gaussian_lobes_2d(256, 256, seed=40 + v["tf_slot"]), max_grad)
        workloads.set_camera(r, workloads.Camera(v["view_pos"], v["view_dir"], v["up_dir"], v["fov"]))
        assert np.array_equal(r.capture(), imgs[i]), f"view {i}"
        assert r.frameStats().evaluated_samples == stats[i].evaluated_samples
    b, ms = r.pagingStats()
    print(f"paged {b / 1e9:.2f} GB in {ms:.0f} ms of host time", flush=True)
    assert b > 0
    r.cleanup()


def tubes(n, edge):
    vol = generate_volume_gpu(VOL_NOISE_TUBES, 4, (n, n, n), 8)
    tf = transfer.tubes_tf(256, seed=4)
    w = workloads.Workload("c4", (n, n, n), 8, 0, 4, edge, 2, 3, tf, 240, 135, 2.0,
                           workloads.orbit_cameras(12, fov=50.0, elevation_deg=15.0))
    return vol, tf, w


def test_c4_crop_mixed_region_levels_match_oracle():
    n, edge = 256, 64
    print("case C4 crop 256^3 noise + tubes, mixed region levels vs oracle", flush=True)
    vol, tf, w = tubes(n, edge)
    r = gpu_renderer(w)
    r.setVolume(vol, edge, 2, 3)
    r.set1DTrans(tf)
    g = n // edge
    levels = O.lod_pyramid(vol, 3)
    for view in (2, 7):
        rl = np.random.default_rng(view).integers(0, 3, (g, g, g)).astype(np.uint8)
        r.setRegionLODs(rl)
        cam = w.cameras[view]
        workloads.apply(r, w, cam)
        img = r.capture()
        kw = workloads.oracle_kwargs(w, cam)
        kw.pop("brick_size")
        kw["apron"] = 2
        ref, ns, fl = O.render_regions(O.make_params(vol, tf, **kw), levels, rl, edge)
        assert (O.to_rgba8(ref)[..., :3] > 0).mean() > 0.02
        assert_parity(img, O.to_rgba8(ref), fl, f"C4 crop regions view {view}")
        st = r.frameStats()
        assert abs(int(st.evaluated_samples) - int(ns.sum())) <= 2e-3 * ns.sum() + 64
    r.cleanup()


def test_c4_2048_level_zero_regions_equal_uniform_render_in_rmse():
    n, edge = 2048, 128
    print("case C4 2048^3 noise + tubes, all regions level 0 vs uniform", flush=True)
    w = workloads.Workload("c4", (n, n, n), 8, VOL_NOISE_TUBES, 4, edge, 2, 6,
                           transfer.tubes_tf(256, seed=4), 1920, 1080, 2.0,
                           workloads.orbit_cameras(4, fov=50.0, elevation_deg=15.0))
    r = gpu_renderer(w)
    r.setGeneratedVolume(VOL_NOISE_TUBES, 4, (n, n, n), 8, edge, 2, 6)
    r.set1DTrans(w.tf)
    g = n // edge
    for cam in w.cameras[:2]:
        workloads.apply(r, w, cam)
        r.setRegionLODs(None)
        r.setLODLimits((0, 0))
        uni = r.capture()
        r.setRegionLODs(np.zeros((g, g, g), np.uint8))
        reg = r.capture()
        n0 = r.frameStats().evaluated_samples
        d, rmse = pixel_diff(reg, uni)
        print(f"level-0 regions vs uniform: max {d.max()} rmse {rmse:.4f}", flush=True)
        # the same model sampled differently (segments restart at each of the
        # 16^3 region entries along a ray through thin tubes): close in RMSE
        assert rmse < 1.0
        # every region three levels coarser: an eighth of the samples or less
        r.setRegionLODs(np.full((g, g, g), 3, np.uint8))
        r.capture()
        assert r.frameStats().evaluated_samples < n0 / 6
    r.cleanup()
