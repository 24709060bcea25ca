"""Render server: requests over several resident volumes (u8 and u16, 1D
and 2D TFs, mixed image sizes) are batched into shared launches, admitted
by output memory, and every image is the same alone and in the mix -- and
the same as the renderer's own capture of that view."""
import math

import numpy as np
import pytest

from conftest import gpu_renderer
from imagevis3d_amd import transfer, workloads
from imagevis3d_amd.renderer import VOL_BLOBS, VOL_FBM_SPHERES, RenderMode, generate_volume
from imagevis3d_amd.serving import RenderServer

pytestmark = pytest.mark.gpu


def make_volumes():
    vols = []
    r = gpu_renderer(width=64, height=64)
    r.setVolume(generate_volume(VOL_BLOBS, 1, (64, 64, 64), 8), 64, 2, 1)
    vols.append((r, 256))
    r = gpu_renderer(width=64, height=64)
    r.setGeneratedVolume(VOL_FBM_SPHERES, 5, (96, 96, 96), 16, 32, 2, 2)
    vols.append((r, 4096))
    r = gpu_renderer(width=64, height=64)
    r.setVolume(workloads.medium(dims=(80, 64, 48)).volume(), 32, 2, 1)
    vols.append((r, 4096))
    for i, (r, n) in enumerate(vols):
        r.set1DTransSlot(0, transfer.gaussian_bumps_tf(n, seed=10 + i))
        r.set1DTransSlot(1, transfer.ramp_tf(n, 0.2 * n, 0.9 * n))
        _, mg = r.histogram2D(256, 256)
        r.set2DTransSlot(2, transfer.gaussian_lobes_2d(256, 256, seed=20 + i), mg)
        r.setSampleRateModifier(1.25)
        r.setBGColors((0.1, 0.1, 0.2), (0.0, 0.0, 0.0))
    return vols


def random_requests(n, n_vol, seed):
    rng = np.random.default_rng(seed)
    sizes = [(64, 64), (128, 128), (200, 120), (96, 160)]
    out = []
    for _ in range(n):
        th, ph = rng.uniform(0, 2 * math.pi), rng.uniform(-1.2, 1.2)
        d = 1.6 + rng.uniform(0, 0.6)
        eye = (d * math.cos(ph) * math.sin(th), d * math.sin(ph), d * math.cos(ph) * math.cos(th))
        w, h = sizes[rng.integers(len(sizes))]
        mode = RenderMode.RM_2DTRANS if rng.uniform() < 0.35 else RenderMode.RM_1DTRANS
        slot = 2 if mode == RenderMode.RM_2DTRANS else int(rng.integers(2))
        view = dict(view_pos=eye, view_dir=tuple(-x for x in eye), up_dir=(0.0, 1.0, 0.0),
                    fov=float(rng.uniform(35, 60)), width=w, height=h, tf_slot=slot)
        out.append((int(rng.integers(n_vol)), view, mode))
    return out


def test_results_same_alone_and_in_the_mix():
    vols = make_volumes()
    print("case serving mix of 40 requests over 3 volumes", flush=True)
    srv = RenderServer(arena_bytes=64 << 20, lanes=3)
    ids = [srv.addVolume(r) for r, _ in vols]
    reqs = random_requests(40, len(vols), seed=7)
    tickets = [srv.submit(ids[v], view, mode) for v, view, mode in reqs]
    mixed = [srv.wait(t) for t in tickets]
    st = srv.stats()
    print(f"launches {st.launches} max jobs {st.max_jobs_per_launch} max lanes {st.max_lanes_busy}")
    assert st.completed == 40 and st.launches < 40 and st.max_jobs_per_launch > 1
    for i, (v, view, mode) in enumerate(reqs):
        img, res = srv.wait(srv.submit(ids[v], view, mode))
        assert res.status == 0 and res.launch_jobs == 1
        assert np.array_equal(img, mixed[i][0]), f"request {i}"
        assert res.evaluated_samples == mixed[i][1].evaluated_samples
        assert res.nominal_samples == mixed[i][1].nominal_samples > 0
    srv.close()
    # the served image is the renderer's own render of that view
    for i in range(8):
        v, view, mode = reqs[i]
        r = vols[v][0]
        r.setRenderMode(mode)
        imgs, _, _ = r.renderBatch([view])
        assert np.array_equal(imgs[0], mixed[i][0]), f"renderer's own view {i}"

    for r, _ in vols:
        r.cleanup()


def test_admission_waits_for_output_memory():
    vols = make_volumes()[:1]
    print("case serving admission with a 1 MiB arena", flush=True)
    srv = RenderServer(arena_bytes=1 << 20, lanes=4)
    vid = srv.addVolume(vols[0][0])
    view = dict(view_pos=(0.3, 0.4, 1.8), view_dir=(-0.3, -0.4, -1.8), fov=45.0, width=256,
                height=256, tf_slot=0)   # 256 KiB per image: at most 4 in flight
    tickets = [srv.submit(vid, view) for _ in range(12)]
    out = [srv.wait(t) for t in tickets]
    assert all(np.array_equal(o[0], out[0][0]) for o in out)
    assert srv.stats().completed == 12
    srv.close()
    vols[0][0].cleanup()
