"""The allocation-free batched raycast (tvk_render_batch_noalloc) through
the C ABI: its images and sample counts equal tvk_render_batch's for the
same views, on two calls in a row (the second with fewer, smaller views),
while the device memory the library holds does not change; it refuses more
views than its scratch holds and TF slots that would need uploading into
new buffers."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_renderer
from imagevis3d_amd import TvkError, transfer, workloads
from imagevis3d_amd._lib import ViewStats, call
from imagevis3d_amd.renderer import Renderer, memman_stats

pytestmark = pytest.mark.gpu


def call_noalloc(r, views, ptrs):
    arr, _ = Renderer._views(views, False)
    stats = (ViewStats * len(views))()
    call("tvk_render_batch_noalloc", r._h, arr, len(views), (C.c_void_p * len(views))(*ptrs), stats)
    return list(stats)


def render_noalloc(r, views):
    """The views rendered into device buffers the caller owns: the device
    framebuffers of helper renderers of each view's size (device memory of
    the same runtime, read back with tvk_read_framebuffer)."""
    helpers = [gpu_renderer(width=v["width"], height=v["height"]) for v in views]
    try:
        ptrs = [h.deviceFramebuffer()[0] for h in helpers]
        before = memman_stats(0).gpu_allocated
        stats = call_noalloc(r, views, ptrs)
        assert memman_stats(0).gpu_allocated == before  # nothing allocated by the call
        return [h.readFramebuffer() for h in helpers], stats
    finally:
        for h in helpers:
            h.cleanup()


def test_noalloc_equals_render_batch_on_two_calls():
    w = workloads.medium(dims=(128, 128, 128), width=200, height=120)
    vol = w.volume()
    print("case allocation-free batch vs render_batch", flush=True)
    r = gpu_renderer(w)
    r.setVolume(vol, 32, 2, 1)
    r.set1DTransSlot(0, w.tf)
    r.set1DTransSlot(1, transfer.ramp_tf(4096, 1500, 3800))
    workloads.apply(r, w, w.cameras[0])
    views = [workloads.view_of(w.cameras[4 * i], 200, 120, i % 2) for i in range(8)]
    ref, rst, _ = r.renderBatch(views)  # puts both TF slots on the device
    got, gst = render_noalloc(r, views)
    for i in range(len(views)):
        assert np.array_equal(got[i], ref[i]), f"view {i}"
        assert gst[i].evaluated_samples == rst[i].evaluated_samples
        assert gst[i].nominal_samples == rst[i].nominal_samples > 0
    # a second call with fewer, smaller views (the scratch keeps the first
    # call's table: nothing of it may leak into this one)
    views2 = [workloads.view_of(w.cameras[3 + 9 * i], 96, 64, 1) for i in range(3)]
    ref2, rst2, _ = r.renderBatch(views2)
    got2, gst2 = render_noalloc(r, views2)
    for i in range(len(views2)):
        assert np.array_equal(got2[i], ref2[i]), f"second call view {i}"
        assert gst2[i].evaluated_samples == rst2[i].evaluated_samples
    # limits: more views than the scratch's job table, a TF slot not on the
    # device yet (both refused before any device work)
    h = gpu_renderer(width=96, height=64)
    fb = h.deviceFramebuffer()[0]
    with pytest.raises(TvkError, match="at most 64"):
        call_noalloc(r, [views2[0]] * 65, [fb] * 65)
    r.set1DTransSlot(5, w.tf)
    with pytest.raises(TvkError, match="(?i)render it once"):
        call_noalloc(r, [workloads.view_of(w.cameras[1], 96, 64, 5)], [fb])
    h.cleanup()
    r.cleanup()
