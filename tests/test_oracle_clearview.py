"""The oracle's ClearView model (isosurface mode with a context and a focus
surface) pinned by reductions to plain isosurface renders and by the
analytic normals of nested spheres.  Tuvok's ClearView shaders are absent
from the reference, so these pin the model stated in oracle/oracle.h, not
the reference's shader."""
import math

import numpy as np

from oracle import oracle as O

N = 40
W = H = 48
ISO, CV_ISO = 100.0, 200.0


def nested_spheres():
    """u8 volume 250 - 200 d / R (clamped), d = distance from the centre:
    the 100 level set is a sphere of radius 0.75 R (context) around the 200
    level set of radius 0.25 R (focus)."""
    z, y, x = np.indices((N, N, N), dtype=np.float64)
    c = (N - 1) / 2.0
    d = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    R = 0.45 * N
    return np.clip(250.0 - 200.0 * d / R, 0, 255).astype(np.uint8), R


def render(vol, **kw):
    args = dict(view_pos=(0.0, 0.0, 2.2), view_dir=(0.0, 0.0, -1.0), fov=40.0, width=W, height=H,
                rate=2.0, mode=2, iso_value=ISO, iso_color=(0.9, 0.8, 0.7), bg_top=(0, 0, 0),
                bg_bottom=(0, 0, 0), bg_alpha=0.0, term_alpha=2.0)
    args.update(kw)
    img, ns, fl = O.render(O.make_params(vol, None, **args))
    return img, ns, fl


def cv(**kw):
    base = dict(iso_value=CV_ISO, color=(0.1, 0.3, 1.0), focus=(W / 2, H / 2), size=0.0,
                context_scale=0.0, border_scale=1.0)
    base.update(kw)
    return base


def test_focus_outside_the_image_is_the_isosurface():
    vol, _ = nested_spheres()
    plain, _, _ = render(vol)
    img, _, _ = render(vol, clearview=cv(focus=(-1e4, -1e4), size=1.0, context_scale=0.5))
    assert np.array_equal(img, plain)
    assert (plain[..., 3] == 1.0).sum() > 0.1 * W * H


def test_full_focus_without_context_is_the_focus_isosurface():
    """size 0 puts every pixel at the focus centre (w = 1); context scale 0
    makes the context fully transparent: the focus surface in the focus
    colour, exactly the plain isosurface render of the focus value."""
    vol, _ = nested_spheres()
    img, _, _ = render(vol, clearview=cv())
    focus_only, _, _ = render(vol, iso_value=CV_ISO, iso_color=(0.1, 0.3, 1.0))
    assert np.array_equal(img[..., :3], focus_only[..., :3])
    # pixels that hit the context but not the focus keep alpha 0 (ac = 0)
    ctx, _, _ = render(vol)
    ring = (ctx[..., 3] == 1.0) & (focus_only[..., 3] < 1.0)
    assert ring.sum() > 0
    assert np.all(img[ring][:, 3] == 0.0)


def test_uniform_ghost_scales_the_context():
    """border scale 0 makes the silhouette term 1: the context is kept with
    opacity = context scale everywhere; with no focus surface behind it
    (focus value above the data) the pixel is that fraction of the context."""
    vol, _ = nested_spheres()
    ctx, _, _ = render(vol)
    img, _, _ = render(vol, clearview=cv(iso_value=300.0, context_scale=0.35, border_scale=0.0))
    hit = ctx[..., 3] == 1.0
    np.testing.assert_allclose(img[hit][:, :3], 0.35 * ctx[hit][:, :3], rtol=0, atol=1e-12)
    np.testing.assert_allclose(img[hit][:, 3], 0.35, atol=1e-12)
    assert np.array_equal(img[~hit], ctx[~hit])


def test_focus_disc_blends_context_and_focus_linearly():
    """Context scale 0: ac = 1 - w, w = 1 - size * r (r in image heights):
    each pixel is (1 - w) context + w focus, both taken from separate plain
    isosurface renders; outside the disc the context alone."""
    vol, _ = nested_spheres()
    ctx, _, _ = render(vol)
    foc, _, _ = render(vol, iso_value=CV_ISO, iso_color=(0.1, 0.3, 1.0))
    fx, fy, size = 20.0, 27.0, 5.0
    img, _, _ = render(vol, clearview=cv(focus=(fx, fy), size=size))
    jj, ii = np.mgrid[0:H, 0:W]
    w = np.clip(1.0 - size * np.hypot(ii - fx, jj - fy) / H, 0.0, 1.0)
    both = (ctx[..., 3] == 1.0) & (foc[..., 3] == 1.0)
    exp = (1.0 - w[..., None]) * ctx[..., :3] + w[..., None] * foc[..., :3]
    np.testing.assert_allclose(img[both][:, :3], exp[both], atol=1e-12)
    outside = (w == 0.0) & (ctx[..., 3] == 1.0)
    assert outside.sum() > 0 and np.array_equal(img[outside], ctx[outside])
    assert (both & (w > 0.2) & (w < 0.8)).sum() > 10  # the blend is exercised


def test_silhouette_term_follows_the_analytic_normal():
    """Full focus, context scale 1, border exponent b: the context keeps
    opacity (1 - |N.V|)^b, N the sphere's radial normal at the hit, V the
    reverse ray (over the focus surface where it lies behind, else alone).  The trilinear central-difference gradient of the sampled
    distance field is within a few degrees of the analytic normal at
    grazing hits (0.06 on the opacity; |N.V|^b in place of (1 - |N.V|)^b
    misses by more than 0.3 over most of the sphere)."""
    vol, R = nested_spheres()
    ctx, _, _ = render(vol)
    foc, _, _ = render(vol, iso_value=CV_ISO, iso_color=(0.1, 0.3, 1.0))
    b = 1.5
    img, _, _ = render(vol, clearview=cv(context_scale=1.0, border_scale=b))
    # rays of the 40 degree camera at distance 2.2 (world box: extent 1 = N
    # voxels); the context is the sphere of radius 0.75 R voxels
    c = (N - 1) / 2.0 + 0.5  # voxel-centre convention: world 0 = voxel N/2
    rs = 0.75 * R / N  # context radius in world units
    cen = np.array([(c - N / 2) / N] * 3)
    th = math.tan(math.radians(20.0))
    checked = 0
    for j in range(H):
        for i in range(W):
            if ctx[j, i, 3] != 1.0:
                continue
            dx = (2 * (i + 0.5) / W - 1) * th
            dy = (1 - 2 * (j + 0.5) / H) * th
            d = np.array([dx, dy, -1.0])
            d /= np.linalg.norm(d)
            o = np.array([0.0, 0.0, 2.2])
            oc = o - cen
            bq = np.dot(oc, d)
            disc = bq * bq - (np.dot(oc, oc) - rs * rs)
            if disc <= 0:
                continue
            t = -bq - math.sqrt(disc)
            n = (o + t * d - cen) / rs
            ac = (1.0 - abs(np.dot(n, -d))) ** b
            if foc[j, i, 3] == 1.0:  # the focus surface shows through
                exp = ac * ctx[j, i, :3] + (1 - ac) * foc[j, i, :3]
            else:  # nothing behind the context: premultiplied by its opacity
                exp = ac * ctx[j, i, :3]
                assert abs(img[j, i, 3] - ac) < 0.06, (i, j)
            assert np.abs(img[j, i, :3] - exp).max() < 0.06, (i, j)
            checked += 1
    assert checked > 50
