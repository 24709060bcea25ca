"""Closed-form pins of the oracle's 2D TF, gradient/Phong and isosurface
modes (see oracle.h for the stated model)."""
import math

import numpy as np
import pytest

from oracle import oracle as O

ORTHO_Z = dict(view_pos=(0.0, 0.0, 1.5), view_dir=(0.0, 0.0, -1.0), up_dir=(0.0, 1.0, 0.0),
               ortho=True, ortho_half_height=0.5)


def ramp(n, axis, dtype=np.uint8, scale=1):
    v = np.arange(n, dtype=np.int64) * scale
    shape = [1, 1, 1]
    shape[2 - axis] = n
    return np.broadcast_to(v.reshape(shape), (n, n, n)).astype(dtype).copy()


def test_2d_tf_with_rows_equal_to_a_1d_tf_matches_1d_mode():
    rng = np.random.default_rng(1)
    vol = rng.integers(0, 256, (24, 24, 24)).astype(np.uint8)
    tf1 = np.zeros((256, 4), np.float32)
    tf1[:, :3] = rng.uniform(0, 1, (256, 3))
    tf1[100:, 3] = np.linspace(0, 0.3, 156)
    tf2 = np.repeat(tf1[None], 16, axis=0)  # every gradient row the same
    kw = dict(width=20, height=20, view_pos=(0.6, 0.5, 1.7), view_dir=(-0.6, -0.5, -1.7),
              brick_size=8, apron=1)
    a, ns_a, _ = O.render(O.make_params(vol, tf1, **kw))
    b, ns_b, _ = O.render(O.make_params(vol, None, mode=1, tf2d=tf2, max_grad=0.5, **kw))
    assert np.allclose(a, b, atol=1e-12)
    assert np.array_equal(ns_a, ns_b)


def test_2d_tf_row_follows_gradient_magnitude():
    """On the ramp v = 3x (u8) the gradient is 3/255 per voxel inside: a 2D
    TF whose alpha is non-zero only on the row of that magnitude lights up
    the interior, and one on another row does not."""
    n = 32
    vol = ramp(n, 0, np.uint8, 3)
    g = 3.0 / 255.0
    H = 9
    max_grad = 8 * g / 4  # row of g = (H-1) * g / max_grad = 4
    tf2 = np.zeros((H, 256, 4), np.float32)
    tf2[4, :, :] = (1.0, 0.0, 0.0, 0.05)
    kw = dict(width=16, height=16, term_alpha=2.0, mode=1, max_grad=max_grad, **ORTHO_Z)
    img, _, _ = O.render(O.make_params(vol, None, tf2d=tf2, **kw))
    # interior pixels (away from the one-voxel border in x): the full chord
    x_int = img[:, 4:12, :]
    A = 1 - (1 - float(np.float32(0.05))) ** n
    assert np.allclose(x_int[..., 0], A, atol=1e-9)
    tf2b = np.zeros((H, 256, 4), np.float32)
    tf2b[7, :, :] = (1.0, 0.0, 0.0, 0.05)
    img2, _, _ = O.render(O.make_params(vol, None, tf2d=tf2b, **kw))
    assert np.all(img2[:, 4:12, 0] == 0.0)


def test_phong_closed_form_on_a_ramp():
    """v = 5x: gradient (5/255, 0, 0) inside; light along +x, view along -z:
    |N.L| = g/sqrt(g^2+eps^2), H = (L+V)/|L+V| with V = +z."""
    n = 24
    vol = ramp(n, 0, np.uint8, 5)
    tf = np.zeros((256, 4), np.float32)
    tf[:, :3] = (0.5, 0.25, 1.0)
    tf[:, 3] = 0.04
    ka, kd, ks, sh = (0.1, 0.2, 0.3), (0.6, 0.5, 0.4), (0.2, 0.3, 0.1), 8.0
    p = O.make_params(vol, tf, width=12, height=12, term_alpha=2.0, lighting=True, ambient=ka,
                      diffuse=kd, specular=ks, shininess=sh, light_dir=(1.0, 0.0, 0.0), **ORTHO_Z)
    img, _, _ = O.render(p)
    g = 5.0 / 255.0
    nl = g / math.sqrt(g * g + 1e-8)
    nh = nl / math.sqrt(2.0)
    A = 1 - (1 - float(np.float32(0.04))) ** n
    for c in range(3):
        col = float(tf[0, c]) * (ka[c] + kd[c] * nl) + ks[c] * nh ** sh
        assert np.allclose(img[2:10, 3:9, c], col * A, atol=1e-9)


def test_isosurface_on_a_z_ramp():
    """v = 4(31 - z), iso 50.6: ortho rays along -z (entering at z = 32, low
    values first) hit where the trilinear value reaches the iso value, in
    the interior; the normal is +-z and the light is +z, so every pixel is
    the iso colour times (ka + kd f) + ks f^4 with f = |N.L| = |N.H|."""
    n = 32
    vol = ramp(n, 2, np.uint8, 4)[::-1].copy()
    p = O.make_params(vol, None, width=16, height=16, mode=2, iso_value=50.6,
                      iso_color=(0.8, 0.6, 0.4), ambient=(0.1,) * 3, diffuse=(0.6,) * 3,
                      specular=(0.2,) * 3, shininess=4.0, light_dir=(0.0, 0.0, 1.0), **ORTHO_Z)
    img, ns, fl = O.render(p)
    g = 4.0 / 255.0
    f = g / math.sqrt(g * g + 1e-8)
    for c, ic in enumerate((0.8, 0.6, 0.4)):
        assert np.allclose(img[..., c], ic * (0.1 + 0.6 * f) + 0.2 * f ** 4.0, atol=1e-9)
    assert np.all(img[..., 3] == 1.0)
    # samples at q = 31.5, 30.5, ... have values 4 (q - 0.5) from the top:
    # 0, 4, ..., the first >= 50.6 is 52 at the 14th sample
    assert np.all(ns == 14)


def test_isosurface_sphere_coverage_and_skip_invariance():
    n = 48
    z, y, x = np.mgrid[0:n, 0:n, 0:n] + 0.5
    r = np.sqrt((x - 24) ** 2 + (y - 24) ** 2 + (z - 24) ** 2)
    vol = np.clip(200 - 10 * r, 0, 255).astype(np.uint8)  # 100 at r = 10
    kw = dict(width=64, height=64, mode=2, iso_value=100.0, bg_alpha=0.0, **ORTHO_Z)
    a, ns_a, _ = O.render(O.make_params(vol, None, **kw))
    b, ns_b, _ = O.render(O.make_params(vol, None, brick_size=8, apron=1, **kw))
    assert np.array_equal(a, b)
    assert ns_b.sum() < ns_a.sum()
    hit = a[..., 3] > 0.5
    frac = hit.mean()
    assert frac == pytest.approx(math.pi * 10 ** 2 / 48 ** 2, rel=0.05)
