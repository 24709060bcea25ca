"""Pins of the oracle's per-region LOD model (the learned-LOD render path):
closed forms that hold for any region -> level assignment, and the
special case where per-region segments coincide with the uniform LOD
render's samples."""
import numpy as np
import pytest

from oracle import oracle as O

ORTHO_Z = dict(view_pos=(0.0, 0.0, 1.5), view_dir=(0.0, 0.0, -1.0), up_dir=(0.0, 1.0, 0.0),
               ortho=True, ortho_half_height=0.5)


def tf_const(n, rgba):
    return np.tile(np.asarray(rgba, np.float32), (n, 1))


@pytest.mark.parametrize("rate", [1.0, 1.7])
def test_constant_volume_any_region_levels_is_beer_lambert(rate):
    """Segments restart per region and change length per level, but their
    lengths still sum to the chord: A = 1 - (1 - a)^L for any assignment."""
    n, E = 32, 8
    vol = np.full((n, n, n), 90, np.uint8)
    levels = O.lod_pyramid(vol, 3)
    rng = np.random.default_rng(1)
    rl = rng.integers(0, 3, (n // E,) * 3).astype(np.uint8)
    a = 0.04
    tf = tf_const(256, (0.7, 0.2, 0.5, a))
    p = O.make_params(vol, tf, width=24, height=24, rate=rate, term_alpha=2.0,
                      view_pos=(0.8, 0.5, 1.7), view_dir=(-0.8, -0.5, -1.7), fov=40.0)
    img, ns, _ = O.render_regions(p, levels, rl, E)
    # the uniform LOD-0 render obeys the same closed form (pinned against an
    # independent chord computation in test_oracle_render.py)
    ref, _, _ = O.render(p)
    assert np.allclose(img[..., :3], ref[..., :3], atol=1e-12)
    assert (ref[..., 0] > 0).sum() > 100 and ns.sum() > 0


def test_all_regions_at_one_level_match_uniform_lod_on_aligned_rays():
    """Orthographic rays along z at rate 1: the LOD-1 step is 2 voxels and a
    region (8 voxels) holds whole segments, so the per-region samples are
    the uniform LOD-1 samples and the images agree to rounding."""
    n, E = 32, 8
    rng = np.random.default_rng(4)
    vol = rng.integers(0, 256, (n, n, n)).astype(np.uint8)
    levels = O.lod_pyramid(vol, 3)
    tf = np.zeros((256, 4), np.float32)
    tf[:, 0] = np.linspace(0, 1, 256)
    tf[:, 2] = np.linspace(1, 0, 256)
    tf[:, 3] = np.linspace(0, 0.15, 256)
    kw = dict(width=32, height=32, rate=1.0, term_alpha=2.0, **ORTHO_Z)
    for lod in (1, 2):
        rl = np.full((n // E,) * 3, lod, np.uint8)
        img, ns, _ = O.render_regions(O.make_params(vol, tf, **kw), levels, rl, E)
        ref, ns_ref, _ = O.render(O.make_params(levels[lod], tf, dims0=(n, n, n), lod=lod, **kw))
        assert np.abs(img - ref).max() < 1e-12, f"lod {lod}"
        assert np.array_equal(ns, ns_ref)


def test_transparent_regions_are_skipped_exactly():
    n, E = 32, 8
    vol = np.zeros((n, n, n), np.uint8)
    vol[4:12, 8:24, 10:20] = 200
    levels = O.lod_pyramid(vol, 2)
    tf = np.zeros((256, 4), np.float32)
    tf[120:] = (1.0, 0.6, 0.2, 0.3)
    rl = np.zeros((n // E,) * 3, np.uint8)
    rl[2:] = 1
    p = O.make_params(vol, tf, width=40, height=30, view_pos=(1.1, 0.4, 1.3),
                      view_dir=(-1.1, -0.4, -1.3), fov=45.0)
    img, ns, _ = O.render_regions(p, levels, rl, E)
    # a TF with alpha everywhere: nothing can be skipped, the samples add up
    tf2 = tf.copy()
    tf2[:120, 3] = 1e-30
    p2 = O.make_params(vol, tf2, width=40, height=30, view_pos=(1.1, 0.4, 1.3),
                       view_dir=(-1.1, -0.4, -1.3), fov=45.0)
    img2, ns2, _ = O.render_regions(p2, levels, rl, E)
    assert np.abs(img - img2).max() < 1e-12
    assert ns.sum() < ns2.sum()
