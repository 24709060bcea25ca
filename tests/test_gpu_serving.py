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


def test_requests_in_parts_equal_whole_requests():
    """A request split over launches (its tile slots in parts, one part on
    the GPU at a time, light requests launched between its parts) gives the
    image and sample counts of the whole request."""
    vols = make_volumes()
    print("case serving: 40 requests in parts vs whole", flush=True)
    reqs = random_requests(40, len(vols), seed=9)
    out = {}
    for label, cost in (("whole", 2**64 - 1), ("parts", 100_000)):
        srv = RenderServer(arena_bytes=64 << 20, lanes=3)
        srv.setPartCost(cost)
        ids = [srv.addVolume(r) for r, _ in vols]
        tickets = [srv.submit(ids[v], view, mode) for v, view, mode in reqs]
        out[label] = [srv.wait(t) for t in tickets]
        st = srv.stats()
        print(f"{label}: launches {st.launches} split {st.split_requests} parts {st.parts_launched}",
              flush=True)
        assert st.completed == 40
        if label == "whole":
            assert st.split_requests == 0 and st.parts_launched == 40
        else:
            assert st.split_requests > 20 and st.parts_launched > 100
        srv.close()
    assert max(res.parts for _, res in out["parts"]) > 8
    for i in range(len(reqs)):
        (a, ra), (b, rb) = out["whole"][i], out["parts"][i]
        assert ra.parts == 1
        assert np.array_equal(a, b), f"request {i} ({rb.parts} parts)"
        assert ra.evaluated_samples == rb.evaluated_samples, f"request {i}"
        assert ra.nominal_samples == rb.nominal_samples > 0, f"request {i}"
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


def test_paged_volumes_served_and_admitted_by_pool_memory():
    """Two paged volumes and an in-core one behind one server.  The GPU
    budget leaves the pool room for one paged batch's working set but not
    two, so paged batches never hold more pool memory at once than the
    larger working set; every served image equals the renderer's own render
    of the view and, for paged volumes, the in-core render."""
    from imagevis3d_amd import brick_slot_bytes
    from imagevis3d_amd.renderer import generate_volume_gpu, memman_stats
    print("case serving paged volumes, pool admission", flush=True)
    sb = brick_slot_bytes(32, 2, 16)
    specs = [((256, 256, 256), 21, 512), ((192, 192, 192), 22, 216)]
    paged, incore = [], []
    for dims, seed, n in specs:
        vol = generate_volume_gpu(VOL_FBM_SPHERES, seed, dims, 16)
        tf = transfer.gaussian_bumps_tf(4096, seed=seed)
        tf2 = transfer.gaussian_lobes_2d(256, 256, seed=seed)
        for kind in ("paged", "incore"):
            r = gpu_renderer(width=64, height=64)
            if kind == "paged":
                r.setPagedFlatVolume(vol, 32, n * sb, n_lod=1)
            else:
                r.setVolume(vol, 32, 2, 1)
            r.set1DTransSlot(0, tf)
            r.set1DTransSlot(1, transfer.ramp_tf(4096, 1500, 3800))
            r.set2DTransSlot(2, tf2, 0.05)
            r.setSampleRateModifier(1.25)
            (paged if kind == "paged" else incore).append((r, vol))
    other = make_volumes()[0][0]
    srv = RenderServer(arena_bytes=64 << 20, lanes=2)
    ids = [srv.addVolume(r) for r, _ in paged] + [srv.addVolume(other)]
    need = [n * sb for _, _, n in specs]
    s = memman_stats(0)
    budget = s.gpu_budget
    paged[0][0].setMaxUsableGPUMem(s.gpu_allocated - s.pool_bytes + max(need) + (16 << 20))
    try:
        rng = np.random.default_rng(3)
        reqs = []
        for i in range(30):
            v = i % 3
            _, view, mode = workloads.serving_request(rng, 1, sizes=[(160, 96), (200, 120), (96, 96)])
            reqs.append((v, view, mode))
        tickets = [srv.submit(ids[v], view, mode) for v, view, mode in reqs]
        mixed = [srv.wait(t) for t in tickets]
        st = srv.stats()
        print(f"paged launches {st.paged_launches} pool waits {st.pool_waits} "
              f"max reserved {st.max_pool_reserved} need {need}")
        assert st.completed == 30 and st.paged_launches >= 2
        assert 0 < st.max_pool_reserved <= max(need)  # never two paged batches at once
    finally:
        paged[0][0].setMaxUsableGPUMem(budget)
    srv.close()
    for i, (v, view, mode) in enumerate(reqs):
        assert mixed[i][1].status == 0
        r = paged[v][0] if v < 2 else other
        r.setRenderMode(mode)
        alone, st_alone, _ = r.renderBatch([view])
        assert np.array_equal(alone[0], mixed[i][0]), f"request {i} (volume {v})"
        assert st_alone[0].evaluated_samples == mixed[i][1].evaluated_samples
        if v < 2:
            ic = incore[v][0]
            ic.setRenderMode(mode)
            ref, _, _ = ic.renderBatch([view])
            assert np.array_equal(ref[0], mixed[i][0]), f"request {i} vs in-core"
    for r, _ in paged + incore:
        r.cleanup()
    other.cleanup()


def test_serving_mix_at_full_size_same_alone_and_in_the_mix():
    """The serving workload's 7-volume pool (64^3 ... 2048^3, u8 / u16) with
    64 requests of 256^2 ... 3840x2160, 1D and 2D TFs, submitted at once:
    every image and sample count equals the same request served alone."""
    print("case serving mix, 7 volumes incl. 2048^3, 64 requests up to 4K", flush=True)
    rs = workloads.serving_pool()
    srv = RenderServer(arena_bytes=4 << 30, lanes=4)
    ids = [srv.addVolume(r) for r in rs]
    rng = np.random.default_rng(5)
    reqs = [workloads.serving_request(rng, len(ids)) for _ in range(64)]
    assert {v for v, _, _ in reqs} == set(range(len(ids)))
    assert {(view["width"], view["height"]) for _, view, _ in reqs} == set(workloads.SERVING_SIZES)
    tickets = [srv.submit(ids[v], view, mode) for v, view, mode in reqs]
    mixed = [srv.wait(t) for t in tickets]
    st = srv.stats()
    print(f"launches {st.launches} max jobs {st.max_jobs_per_launch} max lanes {st.max_lanes_busy}",
          flush=True)
    assert st.completed == 64 and st.launches < 64
    for i, (v, view, mode) in enumerate(reqs):
        img, res = srv.wait(srv.submit(ids[v], view, mode))
        assert res.status == 0 and res.launch_jobs == 1
        assert np.array_equal(img, mixed[i][0]), f"request {i}"
        assert res.evaluated_samples == mixed[i][1].evaluated_samples
    srv.close()
    for r in rs:
        r.cleanup()
