"""TF groups: in a gradient-mode launch (2D TF, or 1D TF with lighting),
views that differ only in their TF slot walk each ray once and composite
every member's TF.  Every view's image and sample counts must equal those
of the same launch with grouping off bit for bit: members with different
empty-space cells, early termination at different samples, a transparent
and an opaque TF, groups larger than four, views left out of a group by
their image size, blend precision, the clip plane, nearest sampling,
graph replays and a paged pool that needs several passes."""
import numpy as np
import pytest

from conftest import gpu_renderer
from imagevis3d_amd import brick_slot_bytes, transfer, workloads
from imagevis3d_amd.renderer import BlendPrecision, RenderMode
from imagevis3d_amd.workloads import Camera
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CAMS = [Camera((1.2, 0.5, 1.1), (-1.2, -0.5, -1.1), fov=50.0),
        Camera((0.0, 0.0, 1.9), (0.0, 0.0, -1.0), ortho=True, ortho_half_height=0.6)]


def view(cam, size, slot):
    return dict(view_pos=cam.view_pos, view_dir=cam.view_dir, up_dir=cam.up_dir, fov=cam.fov,
                ortho=cam.ortho, ortho_half_height=cam.ortho_half_height,
                width=size[0], height=size[1], tf_slot=slot)


def grouped_vs_alone(r, views, label):
    r.setTFGrouping(False)
    ref, rst, _ = r.renderBatch(views)
    r.setTFGrouping(True)
    got, gst, _ = r.renderBatch(views)
    for i in range(len(views)):
        assert np.array_equal(got[i], ref[i]), f"{label}: view {i}"
        assert gst[i].evaluated_samples == rst[i].evaluated_samples, f"{label}: view {i} samples"
        assert gst[i].nominal_samples == rst[i].nominal_samples, f"{label}: view {i} nominal"
        assert rst[i].gathered_samples == rst[i].evaluated_samples
    return got, gst


def volume(n=96):
    return workloads.medium(dims=(n, n, n)).volume()


def test_groups_2d_phong_equal_separate_views():
    vol = volume()
    print("case TF groups: 2D TF + Phong, 5 TFs x 2 cameras + off-size view", flush=True)
    _, mg = O.histogram_2d_exact(vol, 256, 256)
    lobes = [transfer.gaussian_lobes_2d(256, 256, seed=s) for s in (1, 2, 3, 4)]
    clear = np.zeros((256, 256, 4), np.float32)
    opaque = np.ones((256, 256, 4), np.float32)
    r = gpu_renderer(width=120, height=90)
    r.setVolume(vol, 32, 2, 1)
    r.setRenderMode(RenderMode.RM_2DTRANS)
    r.setLightingEnabled(True)
    for s, t in enumerate(lobes + [clear, opaque]):
        r.set2DTransSlot(s, t, mg)
    # camera 0: slots 0..5 (a group of four, then a group of two), the
    # same camera at another size (alone); camera 1: slots 0, 4, 2
    views = [view(CAMS[0], (120, 90), s) for s in range(6)]
    views += [view(CAMS[0], (90, 120), 1)]
    views += [view(CAMS[1], (120, 90), s) for s in (0, 4, 2)]
    r.setKernelTuning(2, 4, 0)  # 2 samples per gather batch
    got, st = grouped_vs_alone(r, views, "2D Phong")
    assert st[4].evaluated_samples == 0  # transparent TF: every cell skipped
    # the walk of a group gathers the union of its members' samples, counted
    # on its leader; a view marched alone gathers what it evaluates
    ev = [int(x.evaluated_samples) for x in st]
    ga = [int(x.gathered_samples) for x in st]
    for lead, members in ((0, (1, 2, 3)), (4, (5,)), (7, (8, 9))):
        grp = (lead,) + members
        assert max(ev[i] for i in grp) <= ga[lead] <= sum(ev[i] for i in grp), (lead, ga, ev)
        assert all(ga[i] == 0 for i in members)
    assert ga[6] == ev[6] and ga[0] < sum(ev[:4])
    assert np.any(got[0] != got[1])
    # one sample per gather batch (grouped and alone) gives the same images
    r.setKernelTuning(1, 4, 0)
    got1, st1 = grouped_vs_alone(r, views, "2D Phong, batch 1")
    for i in range(len(views)):
        assert np.array_equal(got1[i], got[i]), f"batch 1: view {i}"
        assert st1[i].evaluated_samples == st[i].evaluated_samples
    # runs of 4 samples per ray gathered across lanes (grouped), against
    # the march alone
    r.setKernelTuning(4, 4, 0)
    got4, st4 = grouped_vs_alone(r, views, "2D Phong, runs of 4")
    for i in range(len(views)):
        assert np.array_equal(got4[i], got[i]), f"runs of 4: view {i}"
        assert st4[i].evaluated_samples == st[i].evaluated_samples
        assert st4[i].gathered_samples == st[i].gathered_samples
    r.setKernelTuning(8, 4, 0)
    # the group's images equal single captures too
    r.resize((120, 90))
    r.set2DTrans(lobes[2], mg)
    workloads.set_camera(r, CAMS[0])
    assert np.array_equal(r.capture(), got[2])
    for label, setup in (("blend 8 bit", lambda: r.setBlendPrecision(BlendPrecision.BP_8BIT)),
                         ("clip plane", lambda: (r.setBlendPrecision(BlendPrecision.BP_32BIT),
                                                 r.setClipPlane((0.3, -0.2, 0.9, 0.05)),
                                                 r.enableClipPlane(True)))):
        print(f"case TF groups: 2D TF + Phong, {label}", flush=True)
        setup()
        grouped_vs_alone(r, views, label)
    r.cleanup()


@pytest.mark.parametrize("interp,batch", [("linear", 1), ("nearest", 1), ("linear", 4), ("nearest", 4)])
def test_groups_1d_lighting_equal_separate_views(interp, batch):
    vol = volume(80)
    print(f"case TF groups: 1D TF + lighting, {interp}, batch {batch}", flush=True)
    tfs = [transfer.gaussian_bumps_tf(4096, seed=s) for s in (3, 5)]
    tfs.append(transfer.ramp_tf(4096, 500.0, 9000.0))  # dense: rays end early
    r = gpu_renderer(width=96, height=72)
    r.setVolume(vol, 16, 2, 1)
    r.setLightingEnabled(True)
    if interp == "nearest":
        from imagevis3d_amd.renderer import Interpolant
        r.setInterpolant(Interpolant.NearestNeighbor)
    for s, t in enumerate(tfs):
        r.set1DTransSlot(s, t)
    r.setKernelTuning(batch, 4, 0)
    views = [view(c, (96, 72), s) for c in CAMS for s in (0, 1, 2)]
    grouped_vs_alone(r, views, f"1D lighting {interp} batch {batch}")
    r.cleanup()


def test_groups_graph_replays():
    vol = volume(64)
    print("case TF groups: graph replays, grouped and ungrouped tables", flush=True)
    _, mg = O.histogram_2d_exact(vol, 128, 128)
    lobes = [transfer.gaussian_lobes_2d(128, 128, seed=s) for s in (5, 6)]
    r = gpu_renderer(width=64, height=48)
    r.setVolume(vol, 32, 2, 1)
    r.setRenderMode(RenderMode.RM_2DTRANS)
    for s, t in enumerate(lobes):
        r.set2DTransSlot(s, t, mg)
    grouped = [view(CAMS[0], (64, 48), 0), view(CAMS[0], (64, 48), 1)]
    alone = [view(CAMS[0], (64, 48), 0), view(CAMS[1], (64, 48), 1)]
    r.setLaunchGraphs(False)
    eg, _, _ = r.renderBatch(grouped)
    ea, _, _ = r.renderBatch(alone)
    r.setLaunchGraphs(True)
    for rep in range(3):
        for vs, ref in ((grouped, eg), (alone, ea)):
            got, _, _ = r.renderBatch(vs)
            for i in range(2):
                assert np.array_equal(got[i], ref[i]), f"replay {rep} view {i}"
        r.synchronize()
    r.cleanup()


def test_groups_on_paged_pool():
    w = workloads.medium(dims=(128, 128, 128), width=112, height=80)
    vol = w.volume()
    print("case TF groups: paged pool, bricks paged in over passes", flush=True)
    _, mg = O.histogram_2d_exact(vol, 256, 256)
    lobes = [transfer.gaussian_lobes_2d(256, 256, seed=s) for s in (7, 8, 9)]
    sb = brick_slot_bytes(32, 2, 16)
    pg = gpu_renderer(w)
    pg.setPagedFlatVolume(vol, 32, 70 * sb, n_lod=1)
    ic = gpu_renderer(w)
    ic.setVolume(vol, 32, 2, 1)
    for r in (pg, ic):
        r.setRenderMode(RenderMode.RM_2DTRANS)
        r.setLightingEnabled(True)
        for s, t in enumerate(lobes):
            r.set2DTransSlot(s, t, mg)
    views = [view(c, (112, 80), s) for c in (CAMS[0], w.cameras[9]) for s in range(3)]
    pg.setKernelTuning(4, 4, 0)  # grouped: runs of 4 samples per ray across lanes
    a, sa = grouped_vs_alone(pg, views, "paged")
    b, sb_, _ = ic.renderBatch(views)
    for i in range(len(views)):
        assert np.array_equal(a[i], b[i]), f"paged vs in-core view {i}"
        assert sa[i].evaluated_samples == sb_[i].evaluated_samples
    assert pg.memStats().bricks_uploaded > 0
    pg.cleanup()
    ic.cleanup()


def test_groups_allocation_free_launch():
    from test_zz_no_alloc import render_noalloc
    vol = volume(64)
    print("case TF groups: allocation-free launch vs ungrouped batch", flush=True)
    _, mg = O.histogram_2d_exact(vol, 128, 128)
    lobes = [transfer.gaussian_lobes_2d(128, 128, seed=s) for s in (11, 12, 13)]
    r = gpu_renderer(width=80, height=60)
    r.setVolume(vol, 32, 2, 1)
    r.setRenderMode(RenderMode.RM_2DTRANS)
    r.setLightingEnabled(True)
    for s, t in enumerate(lobes):
        r.set2DTransSlot(s, t, mg)
    views = [view(c, (80, 60), s) for c in CAMS for s in range(3)]
    r.setTFGrouping(False)
    ref, rst, _ = r.renderBatch(views)  # also puts every TF slot on the device
    r.setTFGrouping(True)
    got, gst = render_noalloc(r, views)
    for i in range(len(views)):
        assert np.array_equal(got[i], ref[i]), f"view {i}"
        assert gst[i].evaluated_samples == rst[i].evaluated_samples
    assert gst[0].gathered_samples > 0 and gst[1].gathered_samples == 0  # view 0 leads a group
    r.cleanup()
