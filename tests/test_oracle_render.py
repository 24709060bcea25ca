"""Pins of the FP64 oracle's render model by closed forms and invariants.

Each test is chosen so that a plausible slip (wrong segment length, a
dropped opacity correction, a half-voxel shift, a transposed matrix,
a skipped sample) fails it without retyping the oracle's formulas.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

ORTHO_Z = dict(view_pos=(0.0, 0.0, 1.5), view_dir=(0.0, 0.0, -1.0), up_dir=(0.0, 1.0, 0.0),
               ortho=True, ortho_half_height=0.5)


def tf_const(n, rgba):
    return np.tile(np.asarray(rgba, np.float32), (n, 1))


def test_empty_volume_gives_background_exactly():
    vol = np.zeros((16, 16, 16), np.uint8)
    tf = tf_const(256, (1, 1, 1, 0))
    tf[1:, 3] = 1.0  # any non-zero value would be opaque
    p = O.make_params(vol, tf, width=20, height=10, bg_top=(0.2, 0.4, 0.6),
                      bg_bottom=(0.8, 0.6, 0.0), view_pos=(0.3, 0.2, 2.0),
                      view_dir=(-0.3, -0.2, -2.0))
    img, ns, _ = O.render(p)
    y = (np.arange(10) + 0.5) / 10
    wy = 1.0 - y
    for c, (t, b) in enumerate(zip((0.2, 0.4, 0.6), (0.8, 0.6, 0.0))):
        exp = b + (t - b) * wy
        assert np.allclose(img[:, :, c], exp[:, None], atol=1e-15, rtol=0)
    assert np.all(img[:, :, 3] == 1.0)


@pytest.mark.parametrize("rate", [1.0, 1.5, 0.7, 3.3])
def test_constant_volume_beer_lambert_any_rate(rate):
    """A = 1 - (1-a)^L for chord length L, independent of the rate because
    the segment lengths (including the partial last one) sum to L."""
    n = 24
    vol = np.full((n, n, n), 100, np.uint8)
    a = 0.05
    tf = tf_const(256, (0.9, 0.5, 0.1, a))
    W = H = 16
    p = O.make_params(vol, tf, width=W, height=H, rate=rate, term_alpha=2.0,
                      view_pos=(0.7, 0.4, 1.9), view_dir=(-0.7, -0.4, -1.9), fov=40.0)
    img, _, _ = O.render(p)
    # chord lengths from an independent ray/box computation in numpy
    f = np.array([-0.7, -0.4, -1.9]); f /= np.linalg.norm(f)
    r = np.cross(f, [0, 1, 0]); r /= np.linalg.norm(r)
    u = np.cross(r, f)
    th = math.tan(math.radians(20.0))
    for (y, x) in [(0, 0), (7, 8), (15, 3), (4, 12)]:
        ndx = 2 * (x + 0.5) / W - 1
        ndy = 1 - 2 * (y + 0.5) / H
        d = f + ndx * th * r + ndy * th * u
        d /= np.linalg.norm(d)
        o = np.array([0.7, 0.4, 1.9]) * n + n / 2
        dv = d * n
        dv /= np.linalg.norm(dv)
        with np.errstate(divide="ignore"):
            t1 = (0 - o) / dv
            t2 = (n - o) / dv
        tn, tf_ = np.max(np.minimum(t1, t2)), np.min(np.maximum(t1, t2))
        L = max(0.0, tf_ - max(tn, 0.0))
        a32 = float(np.float32(a))  # the TF is stored in float32
        A = 1 - (1 - a32) ** L
        assert img[y, x, 3] == pytest.approx(1.0)
        assert img[y, x, 0] == pytest.approx(float(np.float32(0.9)) * A, abs=1e-12)
        assert img[y, x, 2] == pytest.approx(float(np.float32(0.1)) * A, abs=1e-12)


def test_single_slab_nearest_is_one_sample():
    """Ortho rays along z at rate 1 sample voxel centres; a one-voxel slab
    is hit by exactly one sample: pixel = a*c + (1-a)*bg."""
    n = 16
    vol = np.zeros((n, n, n), np.uint8)
    vol[5] = 200
    tf = tf_const(256, (0, 0, 0, 0))
    tf[200] = (0.25, 0.5, 1.0, 0.3)
    p = O.make_params(vol, tf, width=n, height=n, rate=1.0, interp=0, bg_top=(1, 1, 1),
                      bg_bottom=(1, 1, 1), term_alpha=2.0, **ORTHO_Z)
    img, ns, fl = O.render(p)
    assert np.all(ns == n)
    a = float(np.float32(0.3))
    assert np.allclose(img[..., 0], a * 0.25 + (1 - a), atol=1e-14)
    assert np.allclose(img[..., 2], a * 1.0 + (1 - a), atol=1e-14)


def test_trilinear_reproduces_a_linear_ramp():
    """v = x; TF red = v/63, alpha 0.5: each ortho ray sees one value, the
    trilinear interpolant of the ramp at the pixel centre (clamped at the
    edges: texel centres at i + 0.5)."""
    n = 64
    vol = np.broadcast_to(np.arange(n, dtype=np.uint16)[None, None, :], (n, n, n)).copy()
    tf = np.zeros((n, 4), np.float32)
    tf[:, 0] = np.arange(n) / (n - 1)
    tf[:, 3] = 0.5
    W = 128
    p = O.make_params(vol, tf, width=W, height=4, rate=1.0, term_alpha=2.0,
                      view_pos=(0, 0, 1.5), view_dir=(0, 0, -1), up_dir=(0, 1, 0),
                      ortho=True, ortho_half_height=0.5 * 4 / W)
    img, _, _ = O.render(p)
    x = (np.arange(W) + 0.5) * n / W
    v = np.clip(x - 0.5, 0, n - 1)
    A = 1 - 0.5 ** n
    red = np.interp(v, np.arange(n), tf[:, 0].astype(np.float64))
    assert np.allclose(img[1, :, 0], red * A, atol=1e-12)


def test_early_termination_bound():
    n = 32
    rng = np.random.default_rng(0)
    vol = rng.integers(0, 256, (n, n, n)).astype(np.uint8)
    tf = np.zeros((256, 4), np.float32)
    tf[:, :3] = rng.uniform(0, 1, (256, 3))
    tf[:, 3] = np.linspace(0, 0.4, 256)
    kw = dict(width=24, height=24, view_pos=(0.5, 0.8, 1.6), view_dir=(-0.5, -0.8, -1.6),
              bg_top=(0.3, 0.3, 0.3), bg_bottom=(0.3, 0.3, 0.3))
    full, _, _ = O.render(O.make_params(vol, tf, term_alpha=2.0, **kw))
    term, ns_t, _ = O.render(O.make_params(vol, tf, term_alpha=0.99, **kw))
    _, ns_f, _ = O.render(O.make_params(vol, tf, term_alpha=2.0, **kw))
    assert np.max(np.abs(full - term)[..., :3]) <= 0.01 + 1e-12
    assert ns_t.sum() < ns_f.sum()


def test_empty_space_skipping_is_exact():
    n = 48
    vol = np.zeros((n, n, n), np.uint8)
    vol[20:30, 10:40, 5:25] = 180
    vol[2:6, 2:6, 40:46] = 90
    tf = np.zeros((256, 4), np.float32)
    tf[150:, :] = (1.0, 0.5, 0.2, 0.2)
    kw = dict(width=40, height=30, view_pos=(1.1, 0.3, 1.4), view_dir=(-1.1, -0.3, -1.4))
    a, ns_a, _ = O.render(O.make_params(vol, tf, brick_size=0, **kw))
    b, ns_b, _ = O.render(O.make_params(vol, tf, brick_size=8, apron=2, **kw))
    assert np.array_equal(a, b)
    assert ns_b.sum() < ns_a.sum()


def test_sampling_rate_converges():
    n = 32
    x = np.linspace(0, 1, n)
    vol = (120 + 100 * np.sin(6 * x)[None, None, :] * np.cos(4 * x)[None, :, None]).astype(np.uint8)
    vol = np.broadcast_to(vol, (n, n, n)).copy()
    tf = np.zeros((256, 4), np.float32)
    tf[:, 0] = np.linspace(0, 1, 256)
    tf[:, 3] = np.linspace(0, 0.08, 256) ** 1.5
    kw = dict(width=16, height=16, term_alpha=2.0, view_pos=(0.9, 0.5, 1.5),
              view_dir=(-0.9, -0.5, -1.5))
    ref, _, _ = O.render(O.make_params(vol, tf, rate=32.0, **kw))
    errs = []
    for rate in (0.5, 1.0, 2.0, 4.0):
        img, _, _ = O.render(O.make_params(vol, tf, rate=rate, **kw))
        errs.append(np.abs(img - ref).max())
    assert errs[0] > errs[1] > errs[2] > errs[3]
    assert errs[3] < 0.25 * errs[0]


def test_region_rotation_equals_inverse_camera_rotation():
    """Rotating the volume by R (row-vector convention p' = p R) looks the
    same as moving the camera by R^-1."""
    n = 20
    rng = np.random.default_rng(3)
    vol = rng.integers(0, 256, (n, n, n)).astype(np.uint8)
    vol[:, :, :10] //= 4  # make the volume asymmetric
    tf = np.zeros((256, 4), np.float32)
    tf[:, 0] = np.linspace(0, 1, 256)
    tf[:, 1] = np.linspace(1, 0, 256)
    tf[:, 3] = np.linspace(0, 0.2, 256)
    c, s = math.cos(0.6), math.sin(0.6)
    R = np.array([[c, 0, -s, 0], [0, 1, 0, 0], [s, 0, c, 0], [0, 0, 0, 1]], np.float64)
    eye = np.array([0.2, 0.3, 1.8])
    kw = dict(width=24, height=24)
    rot, _, _ = O.render(O.make_params(vol, tf, view_pos=eye, view_dir=-eye, model=R, **kw))
    Rinv = np.linalg.inv(R)[:3, :3]
    e2 = eye @ Rinv
    up2 = np.array([0, 1, 0]) @ Rinv
    cam, _, _ = O.render(O.make_params(vol, tf, view_pos=e2, view_dir=-e2, up_dir=up2, **kw))
    assert np.allclose(rot, cam, atol=1e-9)


def test_small_workload_nominal_samples():
    """64^3 ortho face-filling 512x512 at rate 1: 512^2 * 64 nominal samples."""
    vol = np.zeros((64, 64, 64), np.uint8)
    tf = tf_const(256, (0, 0, 0, 0))
    p = O.make_params(vol, tf, width=512, height=512, rate=1.0, **ORTHO_Z)
    assert O.nominal_samples(p) == 512 * 512 * 64


def test_fp32_baseline_tracks_fp64_oracle():
    from imagevis3d_amd import workloads as W
    w = W.small()
    vol = w.volume()
    w.width = w.height = 256
    p = O.make_params(vol, w.tf, **W.oracle_kwargs(w, w.cameras[0]))
    ref, ns, flags = O.render(p)
    img, n = O.baseline_render(p, 4)
    d = np.abs(O.to_rgba8(ref).astype(int) - img.astype(int))
    assert d[flags == 0].max() <= 2
    assert np.sqrt((d ** 2).mean()) <= 0.5
    assert abs(n - int(ns.sum())) <= 1e-3 * ns.sum()


@pytest.mark.parametrize("bits", [8, 16])
def test_fp32_baseline_skips_empty_spans_exactly(bits):
    """A sparse volume (content in a few 16^3 cells) seen in perspective:
    the baseline jumps across empty cells and still evaluates the samples
    the oracle evaluates (same brick-granular rule) with the same image."""
    n = 96
    rng = np.random.default_rng(bits)
    dt = np.uint8 if bits == 8 else np.uint16
    vol = np.zeros((n, n, n), dt)
    for _ in range(6):
        z, y, x = rng.integers(0, n - 20, 3)
        vol[z:z + 18, y:y + 14, x:x + 20] = rng.integers(60, 250, (18, 14, 20))
    tf = tf_const(256 if bits == 8 else 4096, (0.8, 0.5, 0.2, 0.0))
    tf[50 if bits == 8 else 60:, 3] = 0.08
    p = O.make_params(vol, tf, width=160, height=120, rate=1.7, view_pos=(0.9, 0.6, 1.4),
                      view_dir=(-0.9, -0.6, -1.4), fov=45.0, brick_size=16, apron=1)
    ref, ns, flags = O.render(p)
    img, cnt = O.baseline_render(p, 4)
    d = np.abs(O.to_rgba8(ref).astype(int) - img.astype(int))
    assert d.max() <= 2 and np.sqrt((d ** 2).mean()) <= 0.5
    assert abs(cnt - int(ns.sum())) <= 1e-3 * ns.sum() + 16
    assert 0 < ns.sum() < 0.5 * O.nominal_samples(p)  # most of each ray was skipped
