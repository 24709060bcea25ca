"""HIP raycaster through the C ABI against the FP64 oracle (parity bar:
max |d| <= 2/255 on unflagged pixels, RMSE <= 0.5/255) and the
size-independent invariants at the benchmark's full sizes."""
import math

import numpy as np
import pytest

from conftest import assert_parity, gpu_renderer, pixel_diff
from imagevis3d_amd import transfer, workloads
from imagevis3d_amd.workloads import Camera
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def render_gpu(r, w, cam, lod_limits=None):
    if lod_limits is not None:
        r.setLODLimits(lod_limits)
    workloads.apply(r, w, cam)
    assert r.paint(), r.lastError()
    img = r.capture()
    return img, r.frameStats()


def oracle_ref(vol, w, cam, lod=0, level=None, **over):
    kw = workloads.oracle_kwargs(w, cam, lod=lod)
    kw.update(over)
    p = O.make_params(level if level is not None else vol, w.tf, **kw)
    img, ns, fl = O.render(p)
    return O.to_rgba8(img), ns, fl, p


def test_small_workload_full_size_parity():
    w = workloads.small()
    print("case small 64^3 u8 512x512 ortho", flush=True)
    vol = w.volume()
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, w.n_lod)
    r.set1DTrans(w.tf)
    img, st = render_gpu(r, w, w.cameras[0])
    ref8, ns, fl, p = oracle_ref(vol, w, w.cameras[0])
    assert_parity(img, ref8, fl, "small")
    assert st.nominal_samples == 512 * 512 * 64
    assert abs(int(st.evaluated_samples) - int(ns.sum())) <= 1e-3 * ns.sum()
    r.cleanup()


@pytest.mark.parametrize("view", [0, 7, 22])
def test_medium_workload_reduced_parity(view):
    """The medium workload's generator, TF and orbit at 128^3 / 32^3 bricks,
    320x180, every LOD level."""
    w = workloads.medium(dims=(128, 128, 128), width=320, height=180)
    w.brick = 32
    print(f"case medium-reduced view={view}", flush=True)
    vol = w.volume()
    levels = O.lod_pyramid(vol, w.n_lod)
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, w.n_lod)
    r.set1DTrans(w.tf)
    cam = w.cameras[view]
    for lod in range(w.n_lod):
        img, st = render_gpu(r, w, cam, (lod, lod))
        assert st.lod == lod
        ref8, ns, fl, _ = oracle_ref(vol, w, cam, lod=lod, level=levels[lod])
        assert_parity(img, ref8, fl, f"medium-reduced view {view} lod {lod}")
        assert abs(int(st.evaluated_samples) - int(ns.sum())) <= 2e-3 * ns.sum() + 64
    r.cleanup()


def test_medium_full_volume_reduced_image_parity():
    """The full 512^3 uint16 medium volume (256^3 bricks), one orbit view at
    240x135 against the oracle."""
    w = workloads.medium(width=240, height=135)
    print("case medium 512^3 at 240x135", flush=True)
    vol = w.volume()
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, w.n_lod)
    r.set1DTrans(w.tf)
    cam = w.cameras[3]
    img, st = render_gpu(r, w, cam, (0, 0))
    ref8, ns, fl, p = oracle_ref(vol, w, cam)
    assert_parity(img, ref8, fl, "medium full volume")
    assert st.nominal_samples == O.nominal_samples(p)
    r.cleanup()


def test_medium_full_size_invariants():
    """1920x1080 on the 512^3 volume: bricking invariance (256^3 vs 128^3
    bricks) and empty-space skipping on/off give the same image."""
    w = workloads.medium()
    print("case medium full size invariants", flush=True)
    vol = w.volume()
    cam = w.cameras[5]
    r = gpu_renderer(w)
    r.set1DTrans(w.tf)
    r.setVolume(vol, 256, 2, 1)
    a, sa = render_gpu(r, w, cam, (0, 0))
    r.setEmptySpaceSkipping(False)
    assert r.paint()
    b = r.capture()
    r.setEmptySpaceSkipping(True)
    r.setVolume(vol, 128, 2, 1)
    c, sc = render_gpu(r, w, cam, (0, 0))
    assert np.array_equal(a, b)
    d, rmse = pixel_diff(a, c)
    print(f"bricking invariance 256 vs 128: max {d.max()} rmse {rmse:.5f}")
    assert d.max() <= 1 and rmse < 0.05
    assert sa.nominal_samples == sc.nominal_samples > 0
    r.cleanup()


def test_empty_space_skipping_exact_and_counts():
    n = 96
    vol = np.zeros((n, n, n), np.uint8)
    vol[30:60, 20:80, 10:50] = 180
    vol[5:12, 70:90, 60:90] = 90
    tf = np.zeros((256, 4), np.float32)
    tf[150:] = (1.0, 0.5, 0.2, 0.05)
    w = workloads.Workload("skip", (n, n, n), 8, 0, 0, 16, 2, 1, tf, 200, 150, 1.3,
                           [Camera((1.1, 0.3, 1.4), (-1.1, -0.3, -1.4), fov=45.0)])
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, 1)
    r.set1DTrans(tf)
    a, sa = render_gpu(r, w, w.cameras[0])
    w.skip = False
    b, sb = render_gpu(r, w, w.cameras[0])
    assert np.array_equal(a, b)
    assert sa.evaluated_samples < 0.5 * sb.evaluated_samples
    assert sa.bricks_skipped > 0 and sb.bricks_skipped == 0
    w.skip = True  # the oracle counts the samples of the skipping render
    ref8, ns, fl, _ = oracle_ref(vol, w, w.cameras[0])
    assert_parity(a, ref8, fl, "skip")
    assert abs(int(sa.evaluated_samples) - int(ns.sum())) <= 2e-3 * ns.sum() + 64
    r.cleanup()


def test_constant_volume_closed_form():
    n = 40
    vol = np.full((n, n, n), 77, np.uint16)
    tf = np.tile(np.float32([0.8, 0.3, 0.1, 0.04]), (100, 1))
    w = workloads.Workload("const", (n, n, n), 16, 0, 0, 16, 2, 1, tf, 64, 64, 1.7,
                           [Camera((0.0, 0.0, 1.5), (0.0, 0.0, -1.0), ortho=True,
                                   ortho_half_height=0.5)], term_alpha=2.0)
    r = gpu_renderer(w)
    r.setVolume(vol, 16, 2, 1)
    r.set1DTrans(tf)
    img, _ = render_gpu(r, w, w.cameras[0])
    A = 1 - (1 - float(np.float32(0.04))) ** n
    exp = np.floor(np.float32([0.8, 0.3, 0.1]).astype(np.float64) * A * 255 + 0.5)
    assert np.abs(img[..., :3].astype(int) - exp.astype(int)).max() <= 1
    assert np.all(img[..., 3] == 255)
    r.cleanup()


def test_c60_dataset_views(c60_path, c60, c60_tf):
    vol = c60["blocks"][0]["volume"]
    views = [Camera((0.0, 0.0, 2.0), (0.0, 0.0, -1.0), fov=50.0),
             Camera((1.4, 0.6, 1.2), (-1.4, -0.6, -1.2), fov=50.0),
             Camera((-0.3, 1.9, 0.4), (0.3, -1.9, -0.4), (0.0, 0.0, 1.0), fov=50.0)]
    w = workloads.Workload("c60", (128, 128, 128), 8, 0, 0, 128, 2, 1, c60_tf, 160, 120, 1.0,
                           views, bg_top=(0.1, 0.2, 0.3), bg_bottom=(0.0, 0.0, 0.0))
    r = gpu_renderer(w)
    assert r.loadDataset(c60_path), r.lastError()
    r.set1DTrans(c60_tf)
    for i, cam in enumerate(views):
        img, _ = render_gpu(r, w, cam)
        ref8, _, fl, _ = oracle_ref(vol, w, cam)
        assert_parity(img, ref8, fl, f"c60 view {i}")
    r.cleanup()


def test_nearest_interpolant():
    w = workloads.small()
    w.width, w.height = 200, 120
    vol = w.volume()
    cam = Camera((0.9, 0.4, 1.3), (-0.9, -0.4, -1.3), fov=45.0)
    r = gpu_renderer(w)
    r.setVolume(vol, 32, 2, 1)
    r.set1DTrans(w.tf)
    r.setInterpolant(0)
    img, _ = render_gpu(r, w, cam)
    ref8, _, fl, _ = oracle_ref(vol, w, cam, interp=0)
    assert_parity(img, ref8, fl, "nearest")
    r.cleanup()


def test_preserve_transparency_and_background():
    w = workloads.small()
    w.width, w.height = 96, 64
    w.bg_top, w.bg_bottom = (0.9, 0.8, 0.1), (0.1, 0.3, 0.6)
    vol = w.volume()
    cam = Camera((0.2, 0.3, 1.8), (-0.2, -0.3, -1.8), fov=40.0)
    r = gpu_renderer(w)
    r.setVolume(vol, 64, 2, 1)
    r.set1DTrans(w.tf)
    img, _ = render_gpu(r, w, cam)
    ref8, _, fl, _ = oracle_ref(vol, w, cam)
    assert_parity(img, ref8, fl, "bg gradient")
    t = r.capture(preserve_transparency=True)
    ref8t, _, flt, _ = oracle_ref(vol, w, cam, bg_top=(0, 0, 0), bg_bottom=(0, 0, 0), bg_alpha=0.0)
    assert_parity(t, ref8t, flt, "transparent")
    assert t[..., 3].min() == 0  # background pixels keep zero alpha
    # the background comes back after the capture
    assert np.array_equal(r.capture(), img)
    r.cleanup()


def test_region_rotation_matches_oracle():
    w = workloads.small()
    w.width, w.height = 128, 96
    vol = w.volume()
    c, s = math.cos(0.9), math.sin(0.9)
    R = np.array([[1, 0, 0, 0], [0, c, s, 0], [0, -s, c, 0], [0, 0, 0, 1]], np.float64)
    T = np.eye(4)
    T[3, :3] = (0.1, -0.05, 0.0)
    cam = Camera((0.0, 0.2, 1.9), (0.0, -0.2, -1.9), fov=45.0)
    r = gpu_renderer(w)
    r.setVolume(vol, 32, 2, 1)
    r.set1DTrans(w.tf)
    r.setRotation4x4(R)
    r.setTranslation4x4(T)
    img, _ = render_gpu(r, w, cam)
    ref8, _, fl, _ = oracle_ref(vol, w, cam, model=R @ T)
    assert_parity(img, ref8, fl, "region rotation+translation")
    r.cleanup()


def test_ragged_sizes_and_eye_inside_volume():
    """Image sizes that are not multiples of the 8x8 ray tile, a volume that
    is not a multiple of the brick size, a 1-pixel image, and a camera
    inside the volume."""
    rng = np.random.default_rng(9)
    vol = rng.integers(0, 256, (23, 41, 37)).astype(np.uint8)
    tf = transfer.ramp_tf(256, 60, 250)
    tf[:, 3] *= 0.3
    cams = [Camera((0.05, 0.02, 0.1), (0.3, -0.2, -1.0), fov=70.0),
            Camera((0.8, 0.9, 1.1), (-0.8, -0.9, -1.1), fov=30.0)]
    for (W, H) in [(333, 77), (1, 1), (9, 130)]:
        w = workloads.Workload("ragged", (37, 41, 23), 8, 0, 0, 16, 2, 1, tf, W, H, 1.25, cams)
        r = gpu_renderer(w)
        r.setVolume(vol, 16, 2, 1)
        r.set1DTrans(tf)
        for i, cam in enumerate(cams):
            print(f"case ragged {W}x{H} cam {i}", flush=True)
            img, _ = render_gpu(r, w, cam)
            ref8, _, fl, _ = oracle_ref(vol, w, cam)
            assert_parity(img, ref8, fl, f"ragged {W}x{H} cam {i}")
        r.cleanup()


def test_tile_rect_culling_edges():
    """Only the tiles inside the volume box's projected rectangle are
    raycast; a fill writes the rest.  Views with the box cut by the image
    border (perspective, rolled, orthographic), a box in front of the camera
    but outside the frustum (no tile left), a box behind the camera and a
    camera inside the box (full frame) match the oracle, and one batched
    launch of all of them (graph replay, a job with no tiles) gives the
    paints' images byte for byte."""
    rng = np.random.default_rng(21)
    vol = rng.integers(0, 256, (23, 41, 37)).astype(np.uint8)
    tf = transfer.ramp_tf(256, 40, 250)
    tf[:, 3] *= 0.4
    cams = [Camera((0.0, 0.0, 2.0), (0.8, 0.1, -2.0), fov=40.0),
            Camera((0.3, -0.2, 1.6), (-0.5, 0.55, -1.6), up_dir=(0.4, 1.0, 0.0), fov=35.0),
            Camera((0.4, 0.3, 2.0), (0.0, 0.0, -1.0), ortho=True, ortho_half_height=0.3),
            Camera((0.0, 0.0, 2.0), (1.0, 0.0, -0.35), fov=40.0),
            Camera((0.0, 0.0, 2.0), (0.0, 0.0, 1.0), fov=60.0),
            Camera((0.1, 0.05, 0.0), (0.2, 0.1, -1.0), fov=80.0)]
    W, H = 203, 117
    w = workloads.Workload("culling", (37, 41, 23), 8, 0, 0, 16, 2, 1, tf, W, H, 1.5, cams)
    w.bg_top, w.bg_bottom = (0.9, 0.7, 0.2), (0.1, 0.2, 0.5)
    r = gpu_renderer(w)
    r.setVolume(vol, 16, 2, 1)
    r.set1DTrans(tf)
    paints = []
    for i, cam in enumerate(cams):
        print(f"case culling cam {i}", flush=True)
        img, st = render_gpu(r, w, cam)
        ref8, ns, fl, _ = oracle_ref(vol, w, cam)
        assert_parity(img, ref8, fl, f"culling cam {i}")
        if i in (3, 4):  # nothing of the box in view: the row gradient alone
            assert st.nominal_samples == 0 and ns.sum() == 0
            assert np.array_equal(img, np.broadcast_to(img[:, :1], img.shape))
            assert np.abs(img.astype(int) - ref8.astype(int)).max() <= 1
        paints.append(img)
    views = [dict(view_pos=c.view_pos, view_dir=c.view_dir, up_dir=c.up_dir, fov=c.fov, ortho=c.ortho,
                  ortho_half_height=c.ortho_half_height, width=W, height=H) for c in cams]
    for _ in range(2):  # capture, then replay
        got, st, _ = r.renderBatch(views)
        for i in range(len(cams)):
            assert np.array_equal(got[i], paints[i]), f"batched cam {i}"
    r.cleanup()


def test_empty_volume_and_thin_volume():
    tf = transfer.ramp_tf(256, 10, 200)
    for shape in [(16, 16, 16), (1, 40, 30)]:
        vol = np.zeros(shape, np.uint8)
        if shape[0] == 1:
            vol[0, 10:30, 5:25] = 150
        w = workloads.Workload("thin", shape[::-1], 8, 0, 0, 16, 2, 1, tf, 48, 40, 1.0,
                               [Camera((0.4, 0.5, 1.6), (-0.4, -0.5, -1.6), fov=50.0)],
                               bg_top=(0.2, 0.2, 0.2), bg_bottom=(0.6, 0.1, 0.1))
        r = gpu_renderer(w)
        r.setVolume(vol, 16, 2, 1)
        r.set1DTrans(tf)
        img, st = render_gpu(r, w, w.cameras[0])
        ref8, _, fl, _ = oracle_ref(vol, w, w.cameras[0])
        assert_parity(img, ref8, fl, f"volume {shape}")
        r.cleanup()


def test_progressive_lod_subframes():
    w = workloads.medium(dims=(96, 96, 96), width=128, height=72)
    w.brick = 32
    vol = w.volume()
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, 3)
    r.set1DTrans(w.tf)
    from imagevis3d_amd import tuvok
    r.setRendererTarget(tuvok.renderer.types.RT_Interactive)
    workloads.apply(r, w, w.cameras[1])
    assert r.getCurrentSubFrameCount() == 3
    lods = []
    while True:
        assert r.paint()
        lods.append(r.frameStats().lod)
        if not r.checkForRedraw():
            break
    assert lods == [2, 1, 0]
    interactive = r.capture()
    r.setRendererTarget(tuvok.renderer.types.RT_Capture)
    r.scheduleCompleteRedraw()
    assert np.array_equal(r.capture(), interactive)
    r.cleanup()


def test_memory_budget_is_enforced():
    """The budget is the device's (every renderer there shares it): a volume
    larger than it is refused with TVK_ERR_OOM; the budget is restored."""
    from imagevis3d_amd import TvkError
    from imagevis3d_amd.renderer import memman_stats
    r = gpu_renderer()
    before = memman_stats(0).gpu_budget
    r.setMaxUsableGPUMem(1 << 20)
    try:
        with pytest.raises(TvkError) as e:
            r.setVolume(np.zeros((128, 128, 128), np.uint8), 64, 2, 1)
        assert e.value.code == -5
        ms = r.memStats()
        assert ms.gpu_mem > (1 << 30) and ms.gpu_budget == 1 << 20
        assert memman_stats(0).gpu_budget == 1 << 20
    finally:
        r.setMaxUsableGPUMem(before)
    r.cleanup()


def test_cell_ranges_beyond_one_work_group_per_cell():
    """A 1024^3 level cut into 8^3 empty-space cells (2.1 M cells, more than
    the cell-range kernel's work-groups, so each loops over several cells):
    skipping on and off give the same image, so no cell was left with a
    wrong range."""
    from imagevis3d_amd.renderer import VOL_NOISE_TUBES
    n = 1024
    print("case cell ranges 1024^3 u8, 8^3 cells", flush=True)
    r = gpu_renderer(width=384, height=256)
    r.setEmptySpaceCell(8)
    r.setGeneratedVolume(VOL_NOISE_TUBES, 4, (n, n, n), 8, 128, 2, 1)
    r.set1DTrans(transfer.tubes_tf(256, seed=4))
    workloads.set_camera(r, Camera((1.2, 0.5, 1.4), (-1.2, -0.5, -1.4), fov=45.0))
    on = r.capture()
    s_on = r.frameStats().evaluated_samples
    r.setEmptySpaceSkipping(False)
    off = r.capture()
    s_off = r.frameStats().evaluated_samples
    print(f"evaluated samples: skipping {s_on}, without {s_off}", flush=True)
    assert (on[..., :3] > 0).mean() > 0.05
    assert np.array_equal(on, off)
    assert s_on <= s_off
    r.cleanup()
