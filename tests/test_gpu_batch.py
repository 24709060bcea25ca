"""Batched launches (many views x TF slots in one persistent-kernel launch,
replayed from a captured HIP graph): every view equals the single-view
capture of the same camera and TF bit for bit, with graphs on and off,
across graph replays, launch shapes and empty-space cell sizes."""
import numpy as np
import pytest

from conftest import assert_parity, gpu_renderer
from imagevis3d_amd import brick_slot_bytes, transfer, workloads
from imagevis3d_amd.workloads import Camera
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def views_for(cams, sizes, slots):
    out = []
    for cam, (w, h), s in zip(cams, sizes, slots):
        out.append(dict(view_pos=cam.view_pos, view_dir=cam.view_dir, up_dir=cam.up_dir,
                        fov=cam.fov, ortho=cam.ortho, ortho_half_height=cam.ortho_half_height,
                        width=w, height=h, tf_slot=s))
    return out


def single(r, w, cam, size, tf):
    r.resize(size)
    r.set1DTrans(tf)
    workloads.apply(r, w, cam)
    img = r.capture()
    return img, r.frameStats()


def medium_small():
    w = workloads.medium(dims=(128, 128, 128), width=160, height=96)
    w.brick = 32
    return w, w.volume()


def test_batch_equals_single_views_1d():
    w, vol = medium_small()
    print("case batch 1D: 6 cameras x 2 TFs, mixed sizes", flush=True)
    tf2 = transfer.gaussian_bumps_tf(4096, seed=11)
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, 1)
    r.set1DTransSlot(0, w.tf)
    r.set1DTransSlot(1, tf2)
    workloads.apply(r, w, w.cameras[0])
    cams = [w.cameras[i] for i in (0, 5, 11, 17, 23, 29)] * 2
    sizes = [(160, 96), (96, 160), (64, 64), (200, 72), (8, 8), (33, 47)] * 2
    slots = [0] * 6 + [1] * 6
    views = views_for(cams, sizes, slots)
    imgs, stats, kms = r.renderBatch(views)
    assert kms > 0
    # one TF shared by every view (LDS-staged TF)
    imgs0, stats0, _ = r.renderBatch(views[:6])
    ref = gpu_renderer(w)
    ref.setVolume(vol, w.brick, w.apron, 1)
    for i, v in enumerate(views):
        a, st = single(ref, w, cams[i], sizes[i], w.tf if slots[i] == 0 else tf2)
        assert np.array_equal(imgs[i], a), f"view {i}"
        assert stats[i].evaluated_samples == st.evaluated_samples
        assert stats[i].nominal_samples == st.nominal_samples
        if i < 6:
            assert np.array_equal(imgs0[i], a), f"shared-TF view {i}"
    r.cleanup()
    ref.cleanup()


def test_graph_replays_match_eager():
    w, vol = medium_small()
    print("case graph replay vs eager", flush=True)
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, 1)
    r.set1DTrans(w.tf)
    workloads.apply(r, w, w.cameras[0])
    views = views_for([w.cameras[i] for i in range(0, 36, 4)], [(128, 80)] * 9, [0] * 9)
    r.setLaunchGraphs(False)
    eager, est, _ = r.renderBatch(views)
    r.setLaunchGraphs(True)
    for rep in range(4):
        got, st, _ = r.renderBatch(views)
        for i in range(len(views)):
            assert np.array_equal(got[i], eager[i]), f"replay {rep} view {i}"
            assert st[i].evaluated_samples == est[i].evaluated_samples
        # eager work and a synchronisation between the replays
        r.setLaunchGraphs(False)
        again, _, _ = r.renderBatch(views[:2])
        assert np.array_equal(again[1], eager[1])
        r.synchronize()
        r.setLaunchGraphs(True)
    # a different job count replays the same graph
    got, _, _ = r.renderBatch(views[3:5])
    assert np.array_equal(got[0], eager[3]) and np.array_equal(got[1], eager[4])
    r.cleanup()


def test_batch_2d_tf_phong_and_iso():
    n = 96
    vol = workloads.medium(dims=(n, n, n)).volume()
    print("case batch 2D TF + Phong, iso", flush=True)
    _, mg = O.histogram_2d_exact(vol, 256, 256)
    lobes = [transfer.gaussian_lobes_2d(256, 256, seed=s) for s in (1, 2)]
    cams = [Camera((1.2, 0.5, 1.1), (-1.2, -0.5, -1.1), fov=50.0),
            Camera((-0.4, 1.6, 0.7), (0.4, -1.6, -0.7), (0.0, 0.0, 1.0), fov=40.0),
            Camera((0.0, 0.0, 1.9), (0.0, 0.0, -1.0), ortho=True, ortho_half_height=0.6)]
    r = gpu_renderer(width=120, height=90)
    r.setVolume(vol, 32, 2, 1)
    r.setRenderMode(tuvok_mode("2d"))
    r.setLightingEnabled(True)
    for s, t in enumerate(lobes):
        r.set2DTransSlot(s, t, mg)
    r.setSampleRateModifier(1.25)
    views = views_for(cams * 2, [(120, 90)] * 6, [0, 0, 0, 1, 1, 1])
    imgs, stats, _ = r.renderBatch(views)
    for i, v in enumerate(views):
        r.resize((120, 90))
        r.set2DTrans(lobes[v["tf_slot"]], mg)
        workloads.set_camera(r, cams[i % 3])
        a = r.capture()
        assert np.array_equal(imgs[i], a), f"2D view {i}"
    # isosurface through the batch path equals the capture too
    r.setRenderMode(tuvok_mode("iso"))
    r.setIsoValue(2600.0)
    imgs, _, _ = r.renderBatch(views[:3])
    for i in range(3):
        workloads.set_camera(r, cams[i])
        assert np.array_equal(imgs[i], r.capture()), f"iso view {i}"
    r.cleanup()


def tuvok_mode(name):
    from imagevis3d_amd.renderer import RenderMode
    return {"1d": RenderMode.RM_1DTRANS, "2d": RenderMode.RM_2DTRANS,
            "iso": RenderMode.RM_ISOSURFACE}[name]


@pytest.mark.parametrize("tile_width", [4, 16])
def test_tile_shapes_and_cells_do_not_change_the_image(tile_width):
    w, vol = medium_small()
    print(f"case tile width {tile_width}", flush=True)
    r = gpu_renderer(w)
    r.set1DTrans(w.tf)
    r.setVolume(vol, w.brick, w.apron, 1)
    workloads.apply(r, w, w.cameras[7])
    base = r.capture()
    st0 = r.frameStats()
    r.setRayTileShape(tile_width)
    assert np.array_equal(r.capture(), base)
    for cell in (8, 32):
        r.setEmptySpaceCell(cell)
        r.setVolume(vol, w.brick, w.apron, 1)
        img = r.capture()
        assert np.array_equal(img, base), f"cell {cell}"
        st = r.frameStats()
        assert st.nominal_samples == st0.nominal_samples
    r.cleanup()


def test_batch_on_paged_volume_matches_in_core():
    w, vol = medium_small()
    print("case batch on a paged pool", flush=True)
    slot_bytes = brick_slot_bytes(w.brick, 2, 16)
    pg = gpu_renderer(w)
    pg.setPagedFlatVolume(vol, w.brick, 70 * slot_bytes, n_lod=1)
    pg.set1DTrans(w.tf)
    workloads.apply(pg, w, w.cameras[0])
    ic = gpu_renderer(w)
    ic.setVolume(vol, w.brick, w.apron, 1)
    ic.set1DTrans(w.tf)
    workloads.apply(ic, w, w.cameras[0])
    views = views_for([w.cameras[i] for i in (2, 14, 26)], [(160, 96)] * 3, [0] * 3)
    a, sa, _ = pg.renderBatch(views)
    b, sb, _ = ic.renderBatch(views)
    for i in range(3):
        assert np.array_equal(a[i], b[i]), f"view {i}"
        assert sa[i].evaluated_samples == sb[i].evaluated_samples
    assert pg.memStats().bricks_uploaded > 0
    pg.cleanup()
    ic.cleanup()


def test_batch_views_against_oracle():
    w = workloads.small()
    print("case small workload batch vs oracle", flush=True)
    vol = w.volume()
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, w.n_lod)
    r.set1DTrans(w.tf)
    workloads.apply(r, w, w.cameras[0])
    cams = [w.cameras[0], Camera((0.9, 0.4, 1.3), (-0.9, -0.4, -1.3), fov=45.0)]
    views = views_for(cams, [(128, 128), (96, 64)], [0, 0])
    imgs, stats, _ = r.renderBatch(views)
    for i, cam in enumerate(cams):
        w.width, w.height = views[i]["width"], views[i]["height"]
        p = O.make_params(vol, w.tf, **workloads.oracle_kwargs(w, cam))
        ref, ns, fl = O.render(p)
        assert_parity(imgs[i], O.to_rgba8(ref), fl, f"batch view {i}")
        assert abs(int(stats[i].evaluated_samples) - int(ns.sum())) <= 2e-3 * ns.sum() + 64
    r.cleanup()
