"""The .2dt swatch rasteriser (transfer.rasterize_2dt) against closed forms:
linear and radial gradients, polygon coverage by pixel centre (even-odd),
source-over compositing in file order, 8-bit stop quantisation, and the
c60.2dt demo file.  Tuvok's own raster code is absent from the reference,
so these pin the stated rule, not Qt's painter."""
import os

import numpy as np

from imagevis3d_amd import transfer as T

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def tf(swatches, size=(64, 32)):
    return {"size": size, "tf1d": np.zeros((4, 4), np.float32), "swatches": swatches}


def square(x0=0.0, y0=0.0, x1=1.0, y1=1.0):
    return [(x0, y0), (x0, y1), (x1, y1), (x1, y0)]


def q(c):
    return int(c * 255) / 255.0


def test_linear_gradient_closed_form():
    W, H = 64, 32
    sw = {"radial": 0, "points": square(), "gradient": (0.0, 0.5, 1.0, 0.5),
          "stops": [(0.0, (1.0, 0.2, 0.0, 1.0)), (1.0, (0.0, 0.6, 1.0, 0.5))]}
    r = T.rasterize_2dt(tf([sw], (W, H)))
    t = (np.arange(W) + 0.5) / W
    a0, a1 = np.array([q(1.0), q(0.2), q(0.0), q(1.0)]), np.array([q(0.0), q(0.6), q(1.0), q(0.5)])
    exp = a0[None, :] + t[:, None] * (a1 - a0)[None, :]
    for row in (0, H // 2, H - 1):
        np.testing.assert_allclose(r[row], exp, atol=1e-6)


def test_radial_gradient_and_padding():
    W = H = 40
    sw = {"radial": 1, "points": square(), "gradient": (0.5, 0.5, 0.75, 0.5),
          "stops": [(0.0, (1.0, 1.0, 1.0, 1.0)), (1.0, (0.0, 0.0, 0.0, 1.0))]}
    r = T.rasterize_2dt(tf([sw], (W, H)))
    jj, ii = np.mgrid[0:H, 0:W]
    d = np.hypot(ii + 0.5 - 20, jj + 0.5 - 20) / 10.0
    exp = 1.0 - np.clip(d, 0, 1)  # beyond the radius the end stop repeats
    np.testing.assert_allclose(r[..., 0], exp, atol=1e-6)
    assert np.all(r[..., 3] == 1.0)


def test_coverage_by_pixel_centre_even_odd():
    W, H = 50, 40
    tri = [(0.1, 0.1), (0.9, 0.2), (0.3, 0.95)]
    sw = {"radial": 0, "points": tri, "gradient": (0, 0, 1, 0), "stops": [(0.5, (0, 0, 1, 1))]}
    r = T.rasterize_2dt(tf([sw], (W, H)))
    v = [(int(x * W), int(y * H)) for x, y in tri]
    jj, ii = np.mgrid[0:H, 0:W]
    px, py = ii + 0.5, jj + 0.5

    def side(a, b):
        return (b[0] - a[0]) * (py - a[1]) - (b[1] - a[1]) * (px - a[0])
    s = [side(v[0], v[1]), side(v[1], v[2]), side(v[2], v[0])]
    inside = ((s[0] > 0) & (s[1] > 0) & (s[2] > 0)) | ((s[0] < 0) & (s[1] < 0) & (s[2] < 0))
    # centres exactly on an edge follow the fill rule's boundary convention
    on_edge = (s[0] == 0) | (s[1] == 0) | (s[2] == 0)
    assert on_edge.sum() < 0.1 * inside.sum()
    assert np.array_equal((r[..., 3] > 0)[~on_edge], inside[~on_edge])
    # a pentagram: its centre lies inside the outline but outside by even-odd
    ang = np.pi / 2 + np.arange(5) * 4 * np.pi / 5
    star = [(0.5 + 0.45 * np.cos(a), 0.5 + 0.45 * np.sin(a)) for a in ang]
    sw2 = dict(sw, points=star)
    r2 = T.rasterize_2dt(tf([sw2], (W, W)))
    assert r2[W // 2, W // 2, 3] == 0.0
    # an arm (between the inner pentagon and a tip) is covered once
    assert r2[int(0.8 * W), W // 2, 3] == 1.0


def test_source_over_in_file_order():
    base = {"radial": 0, "gradient": (0, 0, 1, 0)}
    a = dict(base, points=square(0.0, 0.0, 0.75, 1.0), stops=[(0.5, (1.0, 0.0, 0.0, 1.0))])
    b = dict(base, points=square(0.25, 0.0, 1.0, 1.0), stops=[(0.5, (0.0, 0.0, 1.0, 0.4))])
    r = T.rasterize_2dt(tf([a, b], (8, 4)))
    # only a, both (b over a), only b
    np.testing.assert_allclose(r[0, 0], [1, 0, 0, 1], atol=1e-6)
    ab = q(0.4)
    np.testing.assert_allclose(r[0, 4], [1 - ab, 0, ab, 1], atol=1e-6)
    np.testing.assert_allclose(r[0, 7], [0, 0, 1, ab], atol=1e-6)
    # the later swatch wins when it is opaque
    c = dict(b, stops=[(0.5, (0.0, 1.0, 0.0, 1.0))])
    r2 = T.rasterize_2dt(tf([a, c], (8, 4)))
    np.testing.assert_allclose(r2[0, 4], [0, 1, 0, 1], atol=1e-6)


def test_c60_2dt():
    d = T.parse_2dt(os.path.join(GOLDEN, "c60.2dt"))
    assert d["size"] == (256, 256) and len(d["swatches"]) == 6
    r = T.rasterize_2dt(d)
    assert r.shape == (256, 256, 4) and r.dtype == np.float32
    assert np.all((r >= 0) & (r <= 1))
    # alpha only inside the swatches' bounding boxes
    box = np.zeros((256, 256), bool)
    for sw in d["swatches"]:
        xs = [int(x * 256) for x, _ in sw["points"]]
        ys = [int(y * 256) for _, y in sw["points"]]
        box[max(min(ys), 0):max(ys) + 1, max(min(xs), 0):max(xs) + 1] = True
    assert (r[..., 3] > 0).any() and not (r[..., 3] > 0)[~box].any()
    # the first swatch is a single red stop and nothing is drawn over it
    assert np.array_equal(r[int(0.82 * 256), int(0.22 * 256)], np.float32([1, 0, 0, 1]))
