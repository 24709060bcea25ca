"""The LDS-staged 1D march (tvk_set_brick_staging): the same frames and
sample counts as the gathering march, bit for bit -- u16 and u8 volumes,
perspective and orthographic views, ray tile shapes, the clip plane,
8-bit blend precision, a paged pool with evictions, a batched launch --
and against the FP64 oracle."""
import numpy as np
import pytest

from conftest import assert_parity, gpu_renderer
from imagevis3d_amd import BlendPrecision, brick_slot_bytes, transfer, workloads
from imagevis3d_amd.renderer import VOL_NOISE_TUBES, generate_volume_gpu
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def frames(r, w, cams):
    out = []
    for cam in cams:
        workloads.apply(r, w, cam)
        img = r.capture()
        out.append((img, r.frameStats().evaluated_samples))
    return out


def both(r, w, cams):
    r.setBrickStaging(False)
    a = frames(r, w, cams)
    r.setBrickStaging(True)
    b = frames(r, w, cams)
    r.setBrickStaging(False)
    for i, ((ia, sa), (ib, sb)) in enumerate(zip(a, b)):
        assert np.array_equal(ia, ib), f"view {i}: max |d| {np.abs(ia.astype(int) - ib.astype(int)).max()}"
        assert sa == sb, f"view {i}: samples {sa} vs {sb}"
    return b


@pytest.mark.parametrize("brick,tile_width", [(32, 8), (64, 4), (128, 16)])
def test_staged_equals_gathers_u16(brick, tile_width):
    print(f"case staging u16 brick {brick} tile {tile_width}", flush=True)
    w = workloads.medium(dims=(128, 128, 128), width=200, height=120)
    vol = w.volume()
    r = gpu_renderer(w)
    r.setVolume(vol, brick, 2, 2)
    r.set1DTrans(w.tf)
    r.setRayTileShape(tile_width)
    cams = [w.cameras[i] for i in (0, 7, 19, 30)]
    cams.append(workloads.Camera((0.2, 0.1, 1.5), (-0.1, -0.05, -1.0), (0.0, 1.0, 0.0), 60.0, True, 0.55))
    both(r, w, cams)
    r.cleanup()


def test_staged_equals_gathers_u8_clip_blend():
    print("case staging u8, clip plane, 8-bit blend", flush=True)
    n = 160
    vol = generate_volume_gpu(VOL_NOISE_TUBES, 4, (n, n, n), 8)
    tf = transfer.tubes_tf(256, seed=4)
    w = workloads.Workload("st", (n, n, n), 8, 0, 2, 32, 2, 2, tf, 192, 128, 2.0,
                           workloads.orbit_cameras(6, fov=50.0, elevation_deg=15.0))
    r = gpu_renderer(w)
    r.setVolume(vol, 32, 2, 2)
    r.set1DTrans(tf)
    both(r, w, w.cameras)
    r.setClipPlane((0.3, -0.5, 0.8, 0.07))
    r.enableClipPlane(True)
    both(r, w, w.cameras[:3])
    r.enableClipPlane(False)
    r.setBlendPrecision(BlendPrecision.BP_8BIT)
    both(r, w, w.cameras[:3])
    r.cleanup()


def test_staged_paged_pool_and_batch():
    print("case staging paged pool + batch", flush=True)
    w = workloads.medium(dims=(128, 128, 128), width=200, height=120)
    vol = w.volume()
    ref = gpu_renderer(w)
    ref.setVolume(vol, 32, 2, 2)
    ref.set1DTrans(w.tf)
    ref.setLODLimits((1, 1))
    # 40 slots cannot hold a view's 64 LOD-0 bricks: the paged frames fall
    # back to LOD 1, paging bricks in and out between views
    pg = gpu_renderer(w)
    pg.setPagedFlatVolume(vol, 32, 40 * brick_slot_bytes(32, 2, 16), n_lod=2)
    pg.set1DTrans(w.tf)
    pg.setLODLimits((0, 1))
    pg.setBrickStaging(True)
    for v in (2, 11, 25):
        a = frames(ref, w, [w.cameras[v]])[0]
        b = frames(pg, w, [w.cameras[v]])[0]
        assert pg.frameStats().lod == 1
        assert np.array_equal(a[0], b[0]) and a[1] == b[1], f"paged view {v}"
    assert pg.memStats().bricks_evicted > 0
    ref.setVolume(vol, 32, 2, 1)
    ref.set1DTrans(w.tf)
    views = [workloads.view_of(w.cameras[3 * i], 200, 120, 0) for i in range(8)]
    ref.setBrickStaging(False)
    ia, sa, _ = ref.renderBatch(views)
    ref.setBrickStaging(True)
    ib, sb, _ = ref.renderBatch(views)
    for i in range(len(views)):
        assert np.array_equal(ia[i], ib[i]), f"batch view {i}"
        assert sa[i].evaluated_samples == sb[i].evaluated_samples
    ref.cleanup()
    pg.cleanup()


def test_staged_against_oracle():
    print("case staging vs oracle", flush=True)
    w = workloads.medium(dims=(96, 96, 96), width=160, height=96)
    w.brick = 32
    vol = w.volume()
    r = gpu_renderer(w)
    r.setVolume(vol, 32, 2, 1)
    r.set1DTrans(w.tf)
    r.setBrickStaging(True)
    for view in (4, 21):
        cam = w.cameras[view]
        workloads.apply(r, w, cam)
        img = r.capture()
        ref, _, fl = O.render(O.make_params(vol, w.tf, **workloads.oracle_kwargs(w, cam)))
        assert_parity(img, O.to_rgba8(ref), fl, f"staged view {view}")
    r.cleanup()
