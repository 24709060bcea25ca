"""Pins of the oracle's clip plane (setClipPlane, RenderWindow.cpp:1920-1927)
and blend precision (setBlendPrecision, RenderWindow.cpp:1618-1623) by
closed forms: Beer-Lambert on constant volumes for clipped chords, exact
half-images for planes through the centre, complementary half-spaces, and
the per-brick rounding of the composite computed brick by brick in numpy.
"""
import numpy as np
import pytest

from oracle import oracle as O

ORTHO_Z = dict(view_pos=(0.0, 0.0, 1.5), view_dir=(0.0, 0.0, -1.0), up_dir=(0.0, 1.0, 0.0),
               ortho=True, ortho_half_height=0.5)


def tf_const(n, rgba):
    return np.tile(np.asarray(rgba, np.float32), (n, 1))


def const_volume(n=32, v=100):
    return np.full((n, n, n), v, np.uint8)


def test_plane_through_centre_cuts_exactly_half_the_image():
    """Ortho view along -z of a face-filling box; the plane x = 0 (world)
    cuts away x > 0: the right half of the image is the background exactly,
    the left half is the unclipped render."""
    n, W = 32, 32
    vol = const_volume(n)
    tf = tf_const(256, (0.9, 0.5, 0.1, 0.05))
    kw = dict(width=W, height=W, term_alpha=2.0, bg_top=(0.2, 0.3, 0.4), bg_bottom=(0.2, 0.3, 0.4),
              **ORTHO_Z)
    full, ns_full, _ = O.render(O.make_params(vol, tf, **kw))
    cut, ns_cut, _ = O.render(O.make_params(vol, tf, clip_plane=(1.0, 0.0, 0.0, 0.0), **kw))
    assert np.array_equal(cut[:, :W // 2], full[:, :W // 2])
    assert np.array_equal(ns_cut[:, :W // 2], ns_full[:, :W // 2])
    bg = np.array([0.2, 0.3, 0.4, 1.0])
    assert np.all(cut[:, W // 2:] == bg)
    assert np.all(ns_cut[:, W // 2:] == 0)
    # nominal samples are the unclipped box chords
    assert O.nominal_samples(O.make_params(vol, tf, clip_plane=(1.0, 0.0, 0.0, 0.0), **kw)) == \
        O.nominal_samples(O.make_params(vol, tf, **kw))


@pytest.mark.parametrize("depth", [0.25, -0.1, 0.4])
def test_plane_across_the_view_shortens_chords_beer_lambert(depth):
    """Plane z = depth (world; normal +z, cut z > depth) seen along -z: every
    ray's chord becomes (depth + 0.5) * n voxels, and a constant volume gives
    A = 1 - (1 - a)^L exactly, whatever the sample spacing."""
    n, a = 32, 0.05
    vol = const_volume(n)
    tf = tf_const(256, (0.9, 0.5, 0.1, a))
    img, _, _ = O.render(O.make_params(vol, tf, width=8, height=8, term_alpha=2.0, rate=1.3,
                                       clip_plane=(0.0, 0.0, 1.0, -depth), **ORTHO_Z))
    L = (depth + 0.5) * n
    A = 1 - (1 - float(np.float32(a))) ** L
    assert np.allclose(img[..., 3], 1.0)
    assert np.allclose(img[..., 0], float(np.float32(0.9)) * A, atol=1e-12, rtol=0)


def test_complementary_planes_compose_to_the_full_chord():
    """(n, d) and (-n, -d) keep complementary pieces of every chord: on a
    constant volume (1 - A_P)(1 - A_-P) = 1 - A_full per pixel, and a plane
    scaled by k > 0 clips identically (perspective, oblique plane)."""
    vol = const_volume(24)
    tf = tf_const(256, (0.3, 0.6, 0.9, 0.04))
    kw = dict(width=24, height=18, term_alpha=2.0, view_pos=(0.8, 0.5, 1.7),
              view_dir=(-0.8, -0.5, -1.7), fov=45.0, rate=1.7, bg_alpha=0.0)
    plane = np.array([0.3, -0.5, 0.8, 0.07])
    full, _, _ = O.render(O.make_params(vol, tf, **kw))
    a, _, _ = O.render(O.make_params(vol, tf, clip_plane=tuple(plane), **kw))
    b, _, _ = O.render(O.make_params(vol, tf, clip_plane=tuple(-plane), **kw))
    k, _, _ = O.render(O.make_params(vol, tf, clip_plane=tuple(3.5 * plane), **kw))
    hit = full[..., 3] > 0.05
    assert hit.mean() > 0.3
    assert np.allclose((1 - a[..., 3]) * (1 - b[..., 3]), 1 - full[..., 3], atol=1e-12)
    assert np.allclose(k, a, atol=1e-12)
    # the plane crosses the volume in view: each side keeps part of some chords
    part = hit & (a[..., 3] > 1e-6) & (b[..., 3] > 1e-6)
    assert part.mean() > 0.1


def test_clip_follows_the_world_not_the_volume():
    """The plane lives in world space: rotating the volume by a region
    matrix leaves the clipped half of a symmetric scene where it was, and a
    translated volume is cut where the world plane crosses it."""
    n, W = 32, 32
    vol = const_volume(n)
    tf = tf_const(256, (0.9, 0.5, 0.1, 0.05))
    kw = dict(width=W, height=W, term_alpha=2.0, **ORTHO_Z)
    ref, _, _ = O.render(O.make_params(vol, tf, clip_plane=(1.0, 0.0, 0.0, 0.0), **kw))
    rot = np.eye(4)
    rot[:3, :3] = [[0, 0, -1], [0, 1, 0], [1, 0, 0]]  # 90 degrees about y (cube maps onto itself)
    got, _, _ = O.render(O.make_params(vol, tf, clip_plane=(1.0, 0.0, 0.0, 0.0), model=rot, **kw))
    assert np.allclose(got, ref, atol=1e-12)
    # volume moved by +0.25 in x: the plane x = 0 now cuts at a quarter of it
    tr = np.eye(4)
    tr[3, 0] = 0.25
    moved, _, _ = O.render(O.make_params(vol, tf, model=tr, **kw))
    cut, _, _ = O.render(O.make_params(vol, tf, clip_plane=(1.0, 0.0, 0.0, 0.0), model=tr, **kw))
    xs = (np.arange(W) + 0.5) / W - 0.5  # world x of each pixel column
    keep = xs < 0
    assert np.array_equal(cut[:, keep], moved[:, keep])
    assert np.all(cut[:, ~keep, 3] == 1.0) and np.all(cut[:, ~keep, :3] == 0.0)


def test_half_quantization_is_ieee_binary16():
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.random(2000), rng.random(200) * 1e-4, 2.0 ** -np.arange(10, 26),
                         (np.arange(1, 200) + 0.5) * 2.0 ** -11])  # ties
    got = np.array([O.blend_quantize(x, 16) for x in xs])
    assert np.array_equal(got, xs.astype(np.float16).astype(np.float64))
    assert O.blend_quantize(0.123456789, 32) == 0.123456789


def test_unorm8_quantization_properties():
    xs = np.concatenate([np.linspace(0, 1, 1001), (np.arange(255) + 0.5) / 255, [-0.1, 1.2]])
    got = np.array([O.blend_quantize(x, 8) for x in xs])
    k = got * 255
    assert np.allclose(k, np.round(k), atol=1e-9)          # multiples of 1/255
    inside = (xs >= 0) & (xs <= 1)
    assert np.all(np.abs(got[inside] - xs[inside]) <= 0.5 / 255 + 1e-15)
    assert np.array_equal(np.array([O.blend_quantize(x, 8) for x in got]), got)  # idempotent
    ties = (np.arange(255) + 0.5) / 255
    assert np.allclose([O.blend_quantize(x, 8) for x in ties], (np.arange(255) + 1) / 255)
    assert O.blend_quantize(-0.1, 8) == 0.0 and O.blend_quantize(1.2, 8) == 1.0


@pytest.mark.parametrize("bits", [8, 16])
def test_blend_rounds_per_brick_closed_form(bits):
    """Ortho rays along z at rate 1 through a constant 16^3-brick volume:
    each brick contributes 16 unit segments, A_b = 1 - (1-a)^16, and the
    composite is rounded after every brick; computed here brick by brick."""
    n, B, a = 64, 16, 0.05
    vol = const_volume(n)
    c = (0.9, 0.5, 0.1)
    tf = tf_const(256, (*c, a))
    kw = dict(width=4, height=4, term_alpha=2.0, **ORTHO_Z)
    img, _, _ = O.render(O.make_params(vol, tf, blend_bits=bits, blend_brick=B, **kw))
    q = (lambda x: np.float64(np.float16(x))) if bits == 16 else \
        (lambda x: np.floor(min(max(x, 0.0), 1.0) * 255 + 0.5) / 255)
    a32 = float(np.float32(a))
    Ab = 1 - (1 - a32) ** B
    C, A = np.zeros(3), 0.0
    for _ in range(n // B):
        w = (1 - A) * Ab
        C = np.array([q(C[i] + w * float(np.float32(c[i]))) for i in range(3)])
        A = q(A + w)
    assert np.allclose(img[..., 0], C[0], atol=1e-12) and np.allclose(img[..., 2], C[2], atol=1e-12)
    # 32 bit and one brick over the whole ray: only the end is rounded
    full, _, _ = O.render(O.make_params(vol, tf, **kw))
    one, _, _ = O.render(O.make_params(vol, tf, blend_bits=bits, blend_brick=n, **kw))
    assert np.allclose(one[..., 0], q(full[0, 0, 0]), atol=1e-12)
    assert not np.allclose(img[..., 0], full[..., 0], atol=1e-9)
