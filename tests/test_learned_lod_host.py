"""Host logic of the learned-LOD selector (no GPU): the octree cut honours
the error and brick budgets, forms a valid cut of the brick pyramid, and the
feature / bf16 helpers behave."""
import numpy as np

from imagevis3d_amd import learned_lod as LL


def test_octree_cut_budgets_and_validity():
    rng = np.random.default_rng(0)
    g, L = (8, 8, 8), 4
    y = np.log(rng.uniform(1e-5, 1e-3, (L - 1, 512)) * np.arange(1, L)[:, None])
    area = rng.uniform(0, 400, 512)
    for budget in (1e9, 5.0, 1.0, 0.0):
        lev, sse, mem = LL.select_octree(y, area, budget, g, 1, 0)
        assert sse <= budget + 1e-9 or (lev == 0).all()
        assert mem == LL.bricks_needed(lev, g)
        # a cut: every level-l brick's regions share level l
        lv = lev.reshape(8, 8, 8)
        for l in range(1, L):
            s = 1 << l
            for z in range(0, 8, s):
                for y0 in range(0, 8, s):
                    for x in range(0, 8, s):
                        blk = lv[z:z + s, y0:y0 + s, x:x + s]
                        if (blk == l).any():
                            assert (blk == l).all()
    lev, sse, mem = LL.select_octree(y, area, 0.0, g, 1000, 100_000)
    assert mem <= 100_000 and mem == 1000 * LL.bricks_needed(lev, g)
    # the error only falls as the brick budget grows
    errs = [LL.select_octree(y, area, 0.0, g, 1, m)[1] for m in (8, 64, 200, 512)]
    assert all(a >= b for a, b in zip(errs, errs[1:])) and errs[-1] == 0.0


def test_invisible_regions_stay_coarse():
    g, L = (4, 4, 4), 3
    y = np.full((L - 1, 64), np.log(1e-3))
    area = np.zeros(64)
    area[:8] = 100.0  # only the first row of regions is on screen
    lev, sse, mem = LL.select_octree(y, area, 0.0, g, 1, 0)
    # the bricks holding no visible region are never split
    assert (lev.reshape(4, 4, 4)[2:] > 0).all() and sse == 0.0


def test_bf16_round_trip_and_features():
    x = np.array([1.0, -2.5, 3.14159, 1e-3, 65504.0], np.float32)
    b = LL.to_bf16_bits(x)
    back = LL.from_bf16_bits(b)
    assert np.allclose(back, x, rtol=2 ** -8)
    assert LL.from_bf16_bits(LL.to_bf16_bits(np.float32([1.0 + 2 ** -9]))) == 1.0  # ties to even
    view = dict(view_pos=(0.0, 0.0, 2.0), view_dir=(0.0, 0.0, -1.0), fov=50.0, width=200,
                height=100)
    area, dist = LL.region_geometry((64, 64, 64), 16, view)
    assert area.shape == (64,) and (area > 0).all() and area.sum() < 200 * 100 * 4
    # nearer regions (larger z) project larger
    a = area.reshape(4, 4, 4)
    assert a[3].sum() > a[0].sum() and np.isclose(dist.min(), 2.0 - 0.375, atol=0.2)
    tf = np.zeros((256, 4), np.float32)
    tf[128:, 3] = 0.5
    h = np.zeros((2, 16))
    h[0, 0] = 1.0
    h[1, 15] = 1.0
    m = LL.tf_moments(h, tf)
    assert m[0, 0] == 0.0 and np.isclose(m[1, 0], 0.5)
    x = LL.features(2, area[:2], dist[:2], h, m)
    assert x.shape == (2, 32) and (x[:, 27:] == 0).all()


def _planner_case(seed, dims, edge, n_lod):
    rng = np.random.default_rng(seed)
    g = tuple(int(v) for v in np.ceil(np.asarray(dims) / edge))
    n_reg = int(np.prod(g))
    hist = rng.dirichlet(np.ones(16), n_reg)
    tf = np.zeros((256, 4), np.float32)
    tf[:, 3] = np.clip((np.arange(256) - 90) / 120.0, 0, 1)
    tf[:, :3] = rng.uniform(0, 1, 3)
    mom = LL.tf_moments(hist, tf)
    detail = rng.uniform(0, 0.2, (n_reg, n_lod))
    # coordinates exact in float32 (the library's view struct holds floats)
    view = dict(view_pos=(1.25, 0.5, -0.75), view_dir=(-1.25, -0.5, 0.75), up_dir=(0.0, 1.0, 0.0),
                fov=45.0, width=640, height=360)
    return g, n_reg, hist, mom, detail, view


def test_native_features_match_numpy_statement():
    """tvk_lod_features against region_geometry + features, with a ragged
    grid (dims not a multiple of the brick edge) and a view that puts some
    regions behind the camera."""
    for dims, edge, n_lod in (((96, 80, 72), 16, 4), ((128, 128, 128), 32, 3)):
        g, n_reg, hist, mom, detail, view = _planner_case(1, dims, edge, n_lod)
        for v in (view, dict(view, view_pos=(0.125, 0.0, 0.25))):
            x, area, dist = LL.plan_features(v, dims, edge, n_lod, hist, mom, detail)
            a_ref, d_ref = LL.region_geometry(dims, edge, v)
            np.testing.assert_allclose(area, a_ref, rtol=1e-12, atol=1e-9)
            np.testing.assert_allclose(dist, d_ref, rtol=1e-12)
            ref = np.concatenate([LL.features(l, a_ref, d_ref, hist, mom, detail)
                                  for l in range(1, n_lod)])
            np.testing.assert_allclose(x, ref, rtol=1e-6, atol=1e-7)
            assert (area == 0).any() or v is view
        x0, _, _ = LL.plan_features(view, dims, edge, n_lod, hist, mom, None)
        assert (x0[:, 27:] == 0).all()


def test_native_octree_cut_matches_numpy_statement():
    """tvk_lod_select_octree makes the same cut as select_octree on random
    predictions, with and without error and brick budgets."""
    rng = np.random.default_rng(3)
    for g, L in (((8, 8, 8), 4), ((6, 5, 7), 4), ((16, 16, 16), 5)):
        n = int(np.prod(g))
        y = np.log(rng.uniform(1e-6, 1e-3, (L - 1, n)) * np.arange(1, L)[:, None]).astype(np.float32)
        area = rng.uniform(0, 400, n) * (rng.uniform(size=n) > 0.2)
        for budget, bb, mb in ((1e9, 1, 0), (2.0, 1, 0), (0.0, 1, 0), (0.0, 1000, 1000 * n // 5),
                               (0.5, 7, 7 * n // 3)):
            lev, sse, mem = LL.plan_octree(y, area, budget, g, bb, mb)
            lr, sr, mr = LL.select_octree(y, area, budget, g, bb, mb)
            assert np.array_equal(lev, lr), (g, budget, mb)
            assert mem == mr and np.isclose(sse, sr, rtol=1e-9, atol=1e-12)
