"""Transfer functions: the reference's .1dt/.2dt files and the seeded TFs of
the benchmark workloads.

``.1dt`` (DemoData/c60.1dt): first line N, then N lines ``r g b a``
(TransferFunction1D, Q1DTransferFunction.cpp:573-589).  ``.2dt``
(DemoData/c60.2dt): ``W H``, an embedded 1D TF, then a swatch count and
swatches (TransferFunction2D, Q2DTransferFunction.cpp:426-470), rasterised
by ``rasterize_2dt``.
"""
from __future__ import annotations

import numpy as np


def load_1dt(path: str) -> np.ndarray:
    """Read a .1dt file -> (N, 4) float32 RGBA."""
    with open(path) as f:
        toks = f.read().split()
    n = int(toks[0])
    vals = np.asarray([float(t) for t in toks[1:1 + 4 * n]], np.float32)
    if vals.size != 4 * n:
        raise ValueError(f"{path}: expected {n} RGBA entries")
    return vals.reshape(n, 4)


def resample_1d(tf: np.ndarray, size: int) -> np.ndarray:
    """Linear resampling of a 1D TF to `size` entries (end points kept), the
    operation loadFromFileWithSize applies to match the histogram size.
    Entry j of the result is the table read at x = j (n-1)/(size-1),
    linearly between the two entries around x.  Tuvok's own resampling rule
    is not in the reference tree (parity unpinned); this one is pinned by
    closed forms in tests/test_transfer_1d.py."""
    tf = np.asarray(tf, np.float32)
    n = tf.shape[0]
    if size < 1 or n < 1:
        raise ValueError("a 1D TF needs at least one entry")
    if size == 1:
        return tf[:1].copy()
    if n == 1:
        return np.repeat(tf, size, axis=0)
    if n == size:
        return tf.copy()
    x = np.linspace(0.0, n - 1.0, size)
    out = np.empty((size, 4), np.float32)
    for c in range(4):
        out[:, c] = np.interp(x, np.arange(n), tf[:, c])
    return out


class TransferFunction1D:
    """The 1D TF proxy the editor talks to (Q1DTransferFunction.cpp): an
    (N, 4) RGBA float table with the Lua names the reference calls --
    ``getSize`` (:578), ``getColorData`` (:254-258; the returned array is the
    table itself, so edits through it change the TF, as the reference's
    shared colour vector does), ``setStdFunction`` (:441-455),
    ``loadFromFileWithSize`` (:573-589).  ``Renderer.set1DTrans(tf.getColorData())``
    is the upload the reference's ``tuvok.gpu.changed1DTrans`` makes (:468)."""

    def __init__(self, size: int = 256):
        if size < 1:
            raise ValueError("a 1D TF needs at least one entry")
        self._data = np.zeros((int(size), 4), np.float32)

    def getSize(self) -> int:
        return int(self._data.shape[0])

    def getColorData(self) -> np.ndarray:
        return self._data

    def setStdFunction(self, fCenterPoint: float, fInvGradient: float, iComponent: int,
                       bInvertedStep: bool = False) -> None:
        """Step function in one component: 0 below and 1 above a linear ramp
        whose width is fInvGradient (a fraction of the table) and whose
        middle sits at fCenterPoint (a fraction of the table); inverted, 1
        below and 0 above.  Both fractions are clamped to [0, 1].  The
        rounding of the ramp's end points in Tuvok's TransferFunction1D is
        not in the reference tree (parity unpinned); this ramp is exact."""
        if iComponent not in (0, 1, 2, 3):
            raise ValueError("component must be 0..3")
        n = self.getSize()
        c = min(max(float(fCenterPoint), 0.0), 1.0) * (n - 1)
        wdt = min(max(float(fInvGradient), 0.0), 1.0) * (n - 1)
        x = np.arange(n, dtype=np.float64)
        if wdt > 0.0:
            ramp = np.clip((x - (c - 0.5 * wdt)) / wdt, 0.0, 1.0)
        else:
            ramp = (x >= c).astype(np.float64)
        self._data[:, iComponent] = (1.0 - ramp if bInvertedStep else ramp).astype(np.float32)

    def loadFromFile(self, path: str) -> bool:
        try:
            self._data = load_1dt(path)
        except (OSError, ValueError):
            return False
        return True

    def loadFromFileWithSize(self, path: str, size: int) -> bool:
        """Load a .1dt and resample it to `size` entries (the histogram size,
        Q1DTransferFunction.cpp:573-589); size 0 keeps the file's size."""
        try:
            tf = load_1dt(path)
        except (OSError, ValueError):
            return False
        self._data = resample_1d(tf, int(size)) if size else tf
        return True


def parse_2dt(path: str) -> dict:
    """Read a .2dt file: raster size, embedded 1D TF and swatches."""
    with open(path) as f:
        toks = f.read().split()
    it = iter(toks)
    w, h = int(next(it)), int(next(it))
    n1 = int(next(it))
    tf1 = np.asarray([float(next(it)) for _ in range(4 * n1)], np.float32).reshape(n1, 4)
    ns = int(next(it))
    swatches = []
    for _ in range(ns):
        radial = int(next(it))
        npts = int(next(it))
        pts = [(float(next(it)), float(next(it))) for _ in range(npts)]
        grad = tuple(float(next(it)) for _ in range(4))
        nstops = int(next(it))
        stops = []
        for _ in range(nstops):
            pos = float(next(it))
            rgba = tuple(float(next(it)) for _ in range(4))
            stops.append((pos, rgba))
        swatches.append({"radial": radial, "points": pts, "gradient": grad, "stops": stops})
    return {"size": (w, h), "tf1d": tf1, "swatches": swatches}


def _stop_colours(stops):
    """Stops sorted by position, colours quantised as the swatch painter
    does (QColor(int(c * 255)), Q2DTransferFunction.cpp:459-463)."""
    st = sorted(stops, key=lambda s: s[0])
    pos = np.asarray([p for p, _ in st], np.float64)
    col = np.asarray([[int(c * 255) / 255.0 for c in rgba] for _, rgba in st], np.float64)
    return pos, col


def _even_odd(px, py, poly):
    """Even-odd fill rule: pixel centres (px, py) inside polygon `poly`."""
    inside = np.zeros(px.shape, bool)
    n = len(poly)
    for k in range(n):
        x0, y0 = poly[k]
        x1, y1 = poly[(k + 1) % n]
        if y0 == y1:
            continue
        cross = (y0 > py) != (y1 > py)
        xi = (x1 - x0) * (py - y0) / (y1 - y0) + x0
        inside ^= cross & (px < xi)
    return inside


def rasterize_2dt(tf2d: dict, size=None) -> np.ndarray:
    """The 2D TF raster of a parsed .2dt: (H, W, 4) float32 RGBA, column =
    value bin, row = gradient-magnitude bin (row 0 = gradient 0), the layout
    set2DTrans takes.

    Follows the swatch painter of Q2DTransferFunction::DrawSwatches
    (Q2DTransferFunction.cpp:426-470): swatch vertices and gradient end
    points go to integer raster positions (truncation of x*W, y*H); each
    swatch fills the pixels whose centres lie inside its polygon (even-odd
    rule, no antialiasing) with its gradient -- linear from the first to the
    second gradient point, or radial around the first with the distance to
    the second as radius; positions outside [0, 1] take the end stops -- at
    stop colours quantised to 8 bits, interpolated per component; swatches
    are composited in file order with source-over on a transparent canvas
    and the result is returned non-premultiplied.  Tuvok's own raster code
    (TransferFunction2D) is not in the reference tree: the interpolation
    space and the vertex snapping are this rule's choices."""
    W, H = size or tf2d["size"]
    jj, ii = np.mgrid[0:H, 0:W]
    px, py = ii + 0.5, jj + 0.5
    acc = np.zeros((H, W, 4), np.float64)  # premultiplied rgb, alpha
    for sw in tf2d["swatches"]:
        poly = [(float(int(x * W)), float(int(y * H))) for x, y in sw["points"]]
        inside = _even_odd(px, py, poly)
        if not inside.any() or not sw["stops"]:
            continue
        g = sw["gradient"]
        g0 = (float(int(g[0] * W)), float(int(g[1] * H)))
        g1 = (float(int(g[2] * W)), float(int(g[3] * H)))
        dx, dy = g1[0] - g0[0], g1[1] - g0[1]
        if sw["radial"]:
            r = np.hypot(dx, dy)
            t = np.hypot(px - g0[0], py - g0[1]) / r if r > 0 else np.ones(px.shape)
        else:
            ll = dx * dx + dy * dy
            t = ((px - g0[0]) * dx + (py - g0[1]) * dy) / ll if ll > 0 else np.zeros(px.shape)
        t = np.clip(t, 0.0, 1.0)
        pos, col = _stop_colours(sw["stops"])
        rgba = np.stack([np.interp(t, pos, col[:, c]) for c in range(4)], -1)
        a = rgba[..., 3:4]
        src = np.concatenate([rgba[..., :3] * a, a], -1)
        m = inside[..., None]
        acc = np.where(m, src + acc * (1.0 - a), acc)
    out = np.zeros((H, W, 4), np.float32)
    alpha = acc[..., 3]
    nz = alpha > 0
    out[..., 3] = alpha
    for c in range(3):
        out[..., c][nz] = acc[..., c][nz] / alpha[nz]
    return out


def ramp_tf(n: int = 256, lo: float = 40.0, hi: float = 220.0) -> np.ndarray:
    """Linear alpha ramp from value `lo` (alpha 0) to `hi` (alpha 1); colour
    ramp r = t, g = t^2, b = 1 - t with t = i / (n - 1)."""
    i = np.arange(n, dtype=np.float64)
    t = i / (n - 1)
    tf = np.empty((n, 4), np.float32)
    tf[:, 0] = t
    tf[:, 1] = t * t
    tf[:, 2] = 1.0 - t
    tf[:, 3] = np.clip((i - lo) / (hi - lo), 0.0, 1.0)
    return tf


def gaussian_bumps_tf(n: int = 4096, seed: int = 2, bumps: int = 3) -> np.ndarray:
    """`bumps` seeded Gaussian alpha bumps with compact support (+-3 sigma):
    centre U[0.15, 0.93]*n, sigma U[0.015, 0.05]*n, peak U[0.02, 0.08];
    colour = bump colours weighted by the bump profiles."""
    rng = np.random.default_rng(seed)
    mu = rng.uniform(0.15, 0.93, bumps) * n
    sig = rng.uniform(0.015, 0.05, bumps) * n
    amp = rng.uniform(0.02, 0.08, bumps)
    col = rng.uniform(0.2, 1.0, (bumps, 3))
    i = np.arange(n, dtype=np.float64)[:, None]
    prof = np.exp(-0.5 * ((i - mu) / sig) ** 2)
    prof[np.abs(i - mu) > 3.0 * sig] = 0.0
    alpha = np.clip((prof * amp).sum(1), 0.0, 1.0)
    wsum = prof.sum(1, keepdims=True)
    rgb = np.where(wsum > 0, (prof @ col) / np.maximum(wsum, 1e-30), 0.0)
    tf = np.concatenate([rgb, alpha[:, None]], 1).astype(np.float32)
    return tf


def gaussian_lobes_2d(w: int = 256, h: int = 256, seed: int = 3, lobes: int = 4) -> np.ndarray:
    """2D TF raster (h gradient rows x w value columns x RGBA) of `lobes`
    seeded Gaussian lobes with compact support (+-3 sigma per axis): centre
    U[0.15, 0.9] x U[0.1, 0.8], sigma U[0.03, 0.12] x U[0.05, 0.25] of the
    axis, peak alpha U[0.02, 0.1]; colour weighted by the lobe profiles."""
    rng = np.random.default_rng(seed)
    mx = rng.uniform(0.15, 0.9, lobes) * (w - 1)
    my = rng.uniform(0.1, 0.8, lobes) * (h - 1)
    sx = rng.uniform(0.03, 0.12, lobes) * w
    sy = rng.uniform(0.05, 0.25, lobes) * h
    amp = rng.uniform(0.02, 0.1, lobes)
    col = rng.uniform(0.2, 1.0, (lobes, 3))
    x = np.arange(w, dtype=np.float64)[None, :, None]
    y = np.arange(h, dtype=np.float64)[:, None, None]
    dx = (x - mx) / sx
    dy = (y - my) / sy
    prof = np.exp(-0.5 * (dx ** 2 + dy ** 2))
    prof[(np.abs(dx) > 3) | (np.abs(dy) > 3)] = 0.0
    alpha = np.clip((prof * amp).sum(-1), 0.0, 1.0)
    wsum = prof.sum(-1, keepdims=True)
    rgb = np.where(wsum > 0, (prof @ col) / np.maximum(wsum, 1e-30), 0.0)
    return np.concatenate([rgb, alpha[..., None]], -1).astype(np.float32)


def tubes_tf(n: int = 256, seed: int = 0) -> np.ndarray:
    """1D TF of the learned-LOD workload (noise + thin tubes): the tube
    values (top 22% of the range) nearly opaque in a seeded warm colour, a
    seeded low-alpha Gaussian bump (peak 0.01..0.04, sigma 4..10% of the
    range) over the noise values (centre 15..50% of the range) in a seeded
    cool colour."""
    rng = np.random.default_rng(seed)
    x = np.arange(n, dtype=np.float64) / (n - 1)
    tf = np.zeros((n, 4), np.float32)
    warm = np.array([rng.uniform(0.8, 1.0), rng.uniform(0.3, 0.7), rng.uniform(0.0, 0.3)])
    cool = np.array([rng.uniform(0.0, 0.3), rng.uniform(0.4, 0.8), rng.uniform(0.7, 1.0)])
    c, s, peak = rng.uniform(0.15, 0.5), rng.uniform(0.04, 0.10), rng.uniform(0.01, 0.04)
    bump = peak * np.exp(-0.5 * ((x - c) / s) ** 2)
    bump[np.abs(x - c) > 3 * s] = 0.0
    tube = np.clip((x - 0.78) / 0.04, 0.0, 1.0) * rng.uniform(0.7, 0.95)
    a = np.maximum(bump, tube)
    w = np.where(tube > bump, 1.0, 0.0)[:, None]
    tf[:, :3] = w * warm + (1 - w) * cool
    tf[:, 3] = a
    return tf


def constant_tf(n: int, rgba) -> np.ndarray:
    return np.tile(np.asarray(rgba, np.float32), (n, 1))
