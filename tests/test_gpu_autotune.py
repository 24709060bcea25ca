"""Autotuner and re-bricking: the tuned settings are valid, leave the
images bit-identical (cell edge, launch shape) or within 1/255 (brick
size), and are never slower than the starting point on the tuning views."""
import numpy as np
import pytest

from conftest import gpu_renderer, pixel_diff
from imagevis3d_amd import workloads

pytestmark = pytest.mark.gpu


def setup():
    w = workloads.medium(dims=(128, 128, 128), width=160, height=96)
    w.brick = 32
    vol = w.volume()
    r = gpu_renderer(w)
    r.setVolume(vol, w.brick, w.apron, 1)
    r.set1DTrans(w.tf)
    workloads.apply(r, w, w.cameras[0])
    views = [dict(view_pos=c.view_pos, view_dir=c.view_dir, up_dir=c.up_dir, fov=c.fov,
                  width=w.width, height=w.height) for c in w.cameras[::9]]
    return w, vol, r, views


def test_cell_and_launch_tuning_keep_images():
    w, vol, r, views = setup()
    print("case autotune cell + launch", flush=True)
    base, bst, _ = r.renderBatch(views)
    t = r.autotune(views, flags=r.TUNE_CELL | r.TUNE_LAUNCH)
    print(f"tuned: cell {t.cell} batch {t.samples_per_batch} tile {t.tile_width} "
          f"group {t.tile_group}: {t.kernel_ms:.3f} ms (from {t.baseline_ms:.3f}, "
          f"{t.candidates} settings)")
    assert t.brick_size == 32 and 32 % t.cell == 0
    assert t.samples_per_batch in (2, 4, 8) and t.tile_width in (4, 8, 16)
    assert t.wg_waves in (4, 8)
    assert t.kernel_ms <= t.baseline_ms and t.candidates >= 8
    got, st, _ = r.renderBatch(views)
    for i in range(len(views)):
        assert np.array_equal(got[i], base[i]), f"view {i}"
        assert st[i].nominal_samples == bst[i].nominal_samples
    r.cleanup()


def test_brick_size_tuning_and_rebrick():
    w, vol, r, views = setup()
    print("case autotune brick size", flush=True)
    base, _, _ = r.renderBatch(views)
    t = r.autotune(views, flags=r.TUNE_BRICK | r.TUNE_CELL | r.TUNE_LAUNCH)
    print(f"tuned: brick {t.brick_size} cell {t.cell}: {t.kernel_ms:.3f} ms "
          f"(from {t.baseline_ms:.3f})")
    assert t.brick_size in (32, 64, 128)
    grid, B, a = r.getDataset().getBrickLayout(0)
    assert B == t.brick_size and a == 2
    got, _, _ = r.renderBatch(views)
    for i in range(len(views)):
        d, rmse = pixel_diff(got[i], base[i])
        assert d.max() <= 1 and rmse < 0.05, f"view {i}"
    # the re-bricked pool holds the same voxels
    from oracle import oracle as O
    assert np.array_equal(r.readBrick(0, 0, 0, 0), O.extract_brick(vol, B, 2, 0, 0, 0))
    r.cleanup()
