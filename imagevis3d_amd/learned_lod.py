"""Learned per-brick LOD selection (Tuvok's frame planning, which picks one
LOD per frame from setLODLimits -- ImageVis3D/UI/RenderWindow.cpp:447-451 --
made per brick and learned).

A bf16 MLP (32 inputs -> 4 x 64 ReLU -> 1, evaluated on MFMA by
``tvk_lod_mlp_predict``) predicts, for every LOD-0 brick region and every
coarser level, the mean squared RGB error per screen pixel that region
causes when it is rendered from that level.  Inputs (29 used, padded to 32):

    level / 8, log2(1 + projected footprint in pixels) / 24,
    the region's 16-bin value histogram (fractions),
    8 TF moments over that histogram (mean alpha, mean alpha^2, alpha-
    weighted r, g, b, max alpha, alpha variance, alpha roughness),
    view distance (world units),
    and (slots 27-28) the region's detail at that level -- mean |LOD-0 voxel
    - level-l ancestor| (tvk_region_detail) -- plain and times the region's
    maximum TF alpha.

Training data come from the renderer itself: for each (view, TF) the frame
is rendered with every region at level 0 (recording per pixel the region
that added the most opacity) and with every region at level l; the per-pixel
squared error of level l is attributed to that region.  The selector then
cuts the brick octree: starting from the coarsest bricks it splits the
brick whose children reduce the predicted error (mse x footprint) the most
per added brick, until the predicted error fits the frame's budget (RMSE
1/255, scaled by a factor calibrated against full-resolution renders) or
the brick budget is spent; full resolution (level 0) is where splitting
ends at the finest level.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import call

N_IN = 32
HIDDEN = 64
MSE_FLOOR = 1e-6  # per-pixel squared RGB error below this (about 0.25/255) does not matter


# ------------------------------------------------------------- geometry
def camera_basis(view):
    f = np.asarray(view["view_dir"], np.float64)
    f = f / np.linalg.norm(f)
    up = np.asarray(view.get("up_dir", (0.0, 1.0, 0.0)), np.float64)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return f, r, u


def region_geometry(dims, edge, view):
    """Per LOD-0 region (x fastest): projected footprint in pixels (screen
    box of the region's 8 corners, clipped to the image; 0 when behind the
    camera) and distance of its centre from the eye, in world units (box
    centred at 0, largest extent 1)."""
    dims = np.asarray(dims, np.float64)
    g = np.ceil(dims / edge).astype(int)
    m = dims.max()
    zz, yy, xx = np.meshgrid(np.arange(g[2]), np.arange(g[1]), np.arange(g[0]), indexing="ij")
    lo = np.stack([xx, yy, zz], -1).reshape(-1, 3) * edge
    hi = np.minimum(lo + edge, dims)
    eye = np.asarray(view["view_pos"], np.float64)
    f, r, u = camera_basis(view)
    W, H = view["width"], view["height"]
    th = math.tan(math.radians(view.get("fov", 50.0)) / 2)
    asp = W / H
    xs, ys, ok = [], [], np.ones(len(lo), bool)
    for c in range(8):
        p = np.where([(c >> i) & 1 for i in range(3)], hi, lo)
        w = (p - dims / 2) / m - eye
        z = w @ f
        ok &= z > 1e-3
        zs = np.maximum(z, 1e-3)
        xs.append((w @ r) / (zs * th * asp))
        ys.append((w @ u) / (zs * th))
    xs, ys = np.stack(xs, 1), np.stack(ys, 1)
    x0 = np.clip(xs.min(1), -1, 1)
    x1 = np.clip(xs.max(1), -1, 1)
    y0 = np.clip(ys.min(1), -1, 1)
    y1 = np.clip(ys.max(1), -1, 1)
    area = (x1 - x0) * 0.5 * W * (y1 - y0) * 0.5 * H
    area = np.where(ok, np.maximum(area, 0.0), 0.0)
    centre = ((lo + hi) / 2 - dims / 2) / m
    dist = np.linalg.norm(centre - eye, axis=1)
    return area, dist


# ------------------------------------------------------------- features
def tf_moments(hist_frac: np.ndarray, tf: np.ndarray) -> np.ndarray:
    """8 TF moments per region from its 16-bin histogram and a 1D TF."""
    n = tf.shape[0]
    bins = hist_frac.shape[1]
    edges = np.linspace(0, n, bins + 1).astype(int)
    tb = np.stack([tf[edges[i]:max(edges[i + 1], edges[i] + 1)].mean(0) for i in range(bins)])
    a = tb[:, 3].astype(np.float64)
    p = hist_frac
    ma = p @ a
    m2 = p @ (a * a)
    cr, cg, cb = p @ (a * tb[:, 0]), p @ (a * tb[:, 1]), p @ (a * tb[:, 2])
    amax = np.where(p > 0, a[None, :], 0.0).max(1)
    var = np.maximum(m2 - ma * ma, 0.0)
    rough = p[:, :-1] @ np.abs(np.diff(a))
    return np.stack([ma, m2, cr, cg, cb, amax, var, rough], 1)


def features(level: int, area, dist, hist_frac, moments, detail=None) -> np.ndarray:
    n = len(area)
    x = np.zeros((n, N_IN), np.float32)
    x[:, 0] = level / 8.0
    x[:, 1] = np.log2(1.0 + area) / 24.0
    x[:, 2:18] = hist_frac
    x[:, 18:26] = moments
    x[:, 26] = dist
    if detail is not None:
        # the region's detail lost at this level, plain and opacity-weighted
        x[:, 27] = detail[:, level]
        x[:, 28] = detail[:, level] * moments[:, 5]
    return x


def plan_features(view, dims, edge, n_lod, hist_frac, moments, detail=None, out=None):
    """Predictor inputs of every region at levels 1..n_lod-1 (level-major
    rows, 32 columns as `features`), footprints and distances, computed by
    the library (tvk_lod_features) -- the path the renderer plans with.
    `region_geometry` and `features` are the numpy statement of the same.
    `out`, if given, is a float32 array of the inputs' shape to fill."""
    from .renderer import Renderer
    g = np.ceil(np.asarray(dims, np.float64) / edge).astype(np.int64)
    n_reg = int(np.prod(g))
    arr, _ = Renderer._views([view], False)
    d = (C.c_uint64 * 3)(*[int(v) for v in dims])
    hist = np.ascontiguousarray(hist_frac, np.float64)
    mom = np.ascontiguousarray(moments, np.float64)
    det = None if detail is None else np.ascontiguousarray(detail, np.float64)
    if hist.shape != (n_reg, 16) or mom.shape != (n_reg, 8):
        raise ValueError("hist_frac must be (regions, 16) and moments (regions, 8)")
    if det is not None and det.shape != (n_reg, n_lod):
        raise ValueError("detail must be (regions, n_lod)")
    shape = ((n_lod - 1) * n_reg, N_IN)
    if out is not None and (out.shape != shape or out.dtype != np.float32
                            or not out.flags.c_contiguous):
        raise ValueError(f"out must be a C-contiguous float32 array of shape {shape}")
    x = np.empty(shape, np.float32) if out is None else out
    area = np.empty(n_reg, np.float64)
    dist = np.empty(n_reg, np.float64)
    call("tvk_lod_features", arr, d, edge, n_lod, hist.ctypes.data, mom.ctypes.data,
         None if det is None else det.ctypes.data, x.ctypes.data, area.ctypes.data,
         dist.ctypes.data)
    return x, area, dist


def plan_octree(pred_log_mse: np.ndarray, area: np.ndarray, budget_sse: float, grid,
                brick_bytes: int = 1, mem_budget: int = 0):
    """`select_octree` computed by the library (tvk_lod_select_octree)."""
    pred = np.ascontiguousarray(pred_log_mse, np.float32)
    area = np.ascontiguousarray(area, np.float64)
    gx, gy, gz = (int(v) for v in grid)
    if pred.shape != (pred.shape[0], gx * gy * gz) or area.shape != (gx * gy * gz,):
        raise ValueError("pred must be (levels - 1, regions) and area (regions,)")
    lev = np.empty(gx * gy * gz, np.uint8)
    g = (C.c_uint32 * 3)(gx, gy, gz)
    sse, nbytes = C.c_double(), C.c_uint64()
    call("tvk_lod_select_octree", pred.ctypes.data, area.ctypes.data, g, pred.shape[0] + 1,
         float(budget_sse), int(brick_bytes), int(mem_budget), lev.ctypes.data, C.byref(sse),
         C.byref(nbytes))
    return lev, float(sse.value), int(nbytes.value)


def region_histograms(renderer) -> np.ndarray:
    h = renderer.brickHistograms(0, 16).astype(np.float64)
    return h / np.maximum(h.sum(1, keepdims=True), 1.0)


def region_detail(renderer) -> np.ndarray:
    """(n_regions, n_lod) detail per level, in units of the value range."""
    bits = renderer._bit_width()
    return renderer.regionDetail() * (16.0 / ((1 << bits) - 1))


# --------------------------------------------------------- training data
def training_rows(renderer, dims, edge, n_lod, views, tfs, hist_frac, detail=None):
    """Render every (view, TF) at all-level-0 and all-level-l and attribute
    the per-pixel squared RGB error of level l to the dominant region.
    Returns (inputs, log targets)."""
    g = np.ceil(np.asarray(dims) / edge).astype(int)
    n_reg = int(np.prod(g))
    X, Y = [], []
    renderer.setRegionOutput(True)
    for view in views:
        renderer.resize((view["width"], view["height"]))
        set_view(renderer, view)
        for tf in tfs:
            renderer.set1DTrans(tf)
            mom = tf_moments(hist_frac, tf)
            xs = plan_features(view, dims, edge, n_lod, hist_frac, mom, detail)[0]
            xs = xs.reshape(n_lod - 1, n_reg, N_IN)
            renderer.setRegionLODs(np.zeros(n_reg, np.uint8))
            base = renderer.capture().astype(np.float64) / 255.0
            ids = renderer.readRegionOutput().ravel()
            cnt = np.bincount(ids[ids >= 0], minlength=n_reg)
            seen = cnt > 0
            for lev in range(1, n_lod):
                renderer.setRegionLODs(np.full(n_reg, lev, np.uint8))
                img = renderer.capture().astype(np.float64) / 255.0
                e = ((img[..., :3] - base[..., :3]) ** 2).sum(-1).ravel()
                sse = np.bincount(ids[ids >= 0], weights=e[ids >= 0], minlength=n_reg)
                mse = sse / np.maximum(cnt, 1)
                X.append(xs[lev - 1][seen])
                Y.append(np.log(MSE_FLOOR + mse[seen]).astype(np.float32))
    renderer.setRegionOutput(False)
    renderer.setRegionLODs(None)
    return np.concatenate(X), np.concatenate(Y)


def set_view(r, view):
    r.setViewPos(view["view_pos"])
    r.setViewDir(view["view_dir"])
    r.setUpDir(view.get("up_dir", (0.0, 1.0, 0.0)))
    r.setFoV(view.get("fov", 50.0))
    r.setOrthoViewEnabled(False, 0.5)


# ---------------------------------------------------------------- model
def train(X: np.ndarray, Y: np.ndarray, steps: int = 3000, seed: int = 0, device: str = "cuda"):
    """Fit the 32 -> 4 x 64 -> 1 MLP (PyTorch, FP32) and return its weights
    rounded to bf16 in the predictor's layout, plus the training loss."""
    import torch

    torch.manual_seed(seed)
    dev = torch.device(device if torch.cuda.is_available() else "cpu")
    layers = []
    d = N_IN
    for _ in range(4):
        layers += [torch.nn.Linear(d, HIDDEN), torch.nn.ReLU()]
        d = HIDDEN
    layers.append(torch.nn.Linear(HIDDEN, 1))
    net = torch.nn.Sequential(*layers).to(dev)
    x = torch.from_numpy(X).to(dev)
    y = torch.from_numpy(Y).to(dev)
    mu = y.mean()
    with torch.no_grad():
        net[-1].bias.fill_(float(mu))
    opt = torch.optim.Adam(net.parameters(), lr=2e-3)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    g = torch.Generator(device="cpu").manual_seed(seed)
    loss = None
    for _ in range(steps):
        idx = torch.randint(0, x.shape[0], (4096,), generator=g).to(dev)
        loss = torch.nn.functional.mse_loss(net(x[idx]).squeeze(1), y[idx])
        opt.zero_grad()
        loss.backward()
        opt.step()
        sched.step()
    lin = [m for m in net if isinstance(m, torch.nn.Linear)]
    ws = [m.weight.detach().float().t().contiguous().cpu() for m in lin]  # [in][out]
    bs = [m.bias.detach().float().cpu() for m in lin]
    wbits = np.concatenate([to_bf16_bits(w.numpy()).ravel() for w in ws])
    biases = np.concatenate([b.numpy() for b in bs]).astype(np.float32)
    return wbits, biases, float(loss)


def to_bf16_bits(a: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even float32 -> bfloat16 bits (uint16)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)


def reference_forward(wbits: np.ndarray, biases: np.ndarray, x: np.ndarray) -> np.ndarray:
    """The predictor's arithmetic in numpy: bf16 weights and bf16-rounded
    activations between layers, FP32 sums (order aside)."""
    w = from_bf16_bits(wbits)
    h = from_bf16_bits(to_bf16_bits(x))
    off, boff = 0, 0
    d = N_IN
    for _ in range(4):
        W = w[off:off + d * HIDDEN].reshape(d, HIDDEN)
        off += d * HIDDEN
        h = np.maximum(h @ W + biases[boff:boff + HIDDEN], 0.0)
        h = from_bf16_bits(to_bf16_bits(h))
        boff += HIDDEN
        d = HIDDEN
    return h @ w[off:off + HIDDEN] + biases[boff]


class Predictor:
    """The bf16 MLP on the device (tvk_lod_mlp_*)."""

    def __init__(self, wbits: np.ndarray, biases: np.ndarray, device: int = 0):
        self._w = np.ascontiguousarray(wbits, np.uint16)
        self._b = np.ascontiguousarray(biases, np.float32)
        self._h = C.c_void_p()
        call("tvk_lod_mlp_create", int(device), self._w.ctypes.data, self._b.ctypes.data,
             C.byref(self._h))
        self.kernel_ms = 0.0

    def __call__(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty(len(x), np.float32)
        ms = C.c_float(0.0)
        call("tvk_lod_mlp_predict", self._h, x.ctypes.data, len(x), y.ctypes.data, C.byref(ms))
        self.kernel_ms = ms.value
        return y

    def close(self):
        if self._h:
            _lib.load().tvk_lod_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- selection
def bricks_needed(levels: np.ndarray, grid) -> int:
    """Distinct bricks a per-region selection reads (power-of-two pyramid:
    region c at level l lies in level-l brick c >> l)."""
    gz, gy, gx = (int(v) for v in grid[::-1])
    zz, yy, xx = np.meshgrid(np.arange(gz), np.arange(gy), np.arange(gx), indexing="ij")
    l = levels.reshape(gz, gy, gx).astype(np.int64)
    key = (l << 48) | ((zz >> l) << 32) | ((yy >> l) << 16) | (xx >> l)
    return int(np.unique(key).size)


def select_octree(pred_log_mse: np.ndarray, area: np.ndarray, budget_sse: float, grid,
                  brick_bytes: int = 1, mem_budget: int = 0):
    """Octree cut of the brick pyramid: start from the coarsest level's
    bricks and repeatedly split the brick whose children cut the predicted
    SSE the most per added brick, until the predicted SSE fits `budget_sse`
    or the next split would exceed `mem_budget` bytes (0: no memory cap).
    A level-l brick covers the LOD-0 regions c with c >> l equal to its
    index; its predicted SSE is the sum over those regions of mse_l x
    footprint.  Returns (levels per region, predicted SSE, brick bytes)."""
    import heapq

    L = pred_log_mse.shape[0] + 1
    gx, gy, gz = (int(v) for v in grid)
    mse = np.exp(pred_log_mse.astype(np.float64))
    cost = np.vstack([np.zeros((1, mse.shape[1])), mse * area[None, :]]).reshape(L, gz, gy, gx)
    # group sums per level: E[l][bz, by, bx]
    E = []
    for l in range(L):
        s = 1 << l
        ez, ey, ex = -(-gz // s), -(-gy // s), -(-gx // s)
        e = np.zeros((ez, ey, ex))
        np.add.at(e, tuple(np.meshgrid(np.arange(gz) >> l, np.arange(gy) >> l, np.arange(gx) >> l,
                                       indexing="ij")), cost[l])
        E.append(e)
    top = L - 1
    active = {(top, z, y, x) for z in range(E[top].shape[0]) for y in range(E[top].shape[1])
              for x in range(E[top].shape[2])}
    total = float(E[top].sum())
    n_bricks = len(active)

    def children(l, z, y, x):
        out = []
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    c = (l - 1, 2 * z + dz, 2 * y + dy, 2 * x + dx)
                    if c[1] < E[l - 1].shape[0] and c[2] < E[l - 1].shape[1] and c[3] < E[l - 1].shape[2]:
                        out.append(c)
        return out

    def gain(b):
        l = b[0]
        if l == 0:
            return None
        ch = children(*b)
        g = E[l][b[1], b[2], b[3]] - sum(E[c[0]][c[1], c[2], c[3]] for c in ch)
        return g / max(len(ch) - 1, 1), ch

    # bricks with error left are split in order of gain per added brick; a
    # split that gains nothing now may still open finer levels that do
    def push(b):
        gc = gain(b)
        if gc and E[b[0]][b[1], b[2], b[3]] > 0:
            heapq.heappush(heap, (-gc[0], b))

    heap = []
    for b in active:
        push(b)
    while heap and total > budget_sse:
        neg, b = heapq.heappop(heap)
        gc = gain(b)
        ch = gc[1]
        if mem_budget and (n_bricks + len(ch) - 1) * brick_bytes > mem_budget:
            break
        active.discard(b)
        total -= E[b[0]][b[1], b[2], b[3]]
        for c in ch:
            active.add(c)
            total += E[c[0]][c[1], c[2], c[3]]
            push(c)
        n_bricks += len(ch) - 1
    lev = np.zeros((gz, gy, gx), np.uint8)
    for (l, z, y, x) in active:
        s = 1 << l
        lev[z * s:(z + 1) * s, y * s:(y + 1) * s, x * s:(x + 1) * s] = l
    return lev.ravel(), max(total, 0.0), n_bricks * brick_bytes


def select_levels(pred_log_mse: np.ndarray, area: np.ndarray, budget_sse: float,
                  grid=None, brick_bytes: int = 0, mem_budget: int = 0):
    """pred_log_mse: (n_levels - 1, n_regions) predictions for levels 1..L-1.
    Per region, the coarsest level l whose predicted SSE (mse_l x
    footprint) and every finer level's stay <= tau; tau is the largest
    value whose summed predicted SSE fits `budget_sse` (bisection).  With a
    memory budget, tau is raised (coarser) until the bricks read fit.
    Returns (levels, predicted SSE, bytes of the bricks read)."""
    mse = np.exp(pred_log_mse.astype(np.float64))          # (L-1, R)
    cost = np.vstack([np.zeros((1, mse.shape[1])), mse * area[None, :]])  # (L, R)
    # a level is usable only if every finer level is too (monotone choice)
    run = np.maximum.accumulate(cost, axis=0)

    def choose(tau):
        ok = run <= tau
        lev = ok.shape[0] - 1 - np.argmax(ok[::-1], axis=0)
        return lev, cost[lev, np.arange(cost.shape[1])].sum()

    lo, hi = 0.0, float(run.max()) + 1.0
    lev, sse = choose(lo)
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        l2, s = choose(mid)
        if s <= budget_sse:
            lo, lev, sse = mid, l2, s
        else:
            hi = mid
    mem = bricks_needed(lev, grid) * brick_bytes if grid is not None else 0
    if mem_budget and mem > mem_budget:
        lo2, hi2 = lo, float(run.max()) + 1.0
        for _ in range(50):
            mid = 0.5 * (lo2 + hi2)
            l2, s = choose(mid)
            if bricks_needed(l2, grid) * brick_bytes <= mem_budget:
                hi2, lev, sse = mid, l2, s
            else:
                lo2 = mid
        mem = bricks_needed(lev, grid) * brick_bytes
    return lev.astype(np.uint8), float(sse), int(mem)


# ------------------------------------------------------------ rendering
def rmse255(a: np.ndarray, b: np.ndarray) -> float:
    """RMSE over the colour channels (RGB) in units of 1/255.  Alpha is left
    out: over an opaque background both frames hold 255 there, and counting
    it would dilute the colour error by sqrt(3/4)."""
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    return float(np.sqrt((d * d).mean()))


def full_resolution(renderer, dims, edge) -> np.ndarray:
    """The full-resolution frame of the per-region path (every region at
    level 0): the reference the learned frames' error is measured against,
    sampled like them."""
    n_reg = int(np.prod(np.ceil(np.asarray(dims) / edge)))
    renderer.setRegionLODs(np.zeros(n_reg, np.uint8))
    img = renderer.capture()
    renderer.setRegionLODs(None)
    return img


_PLAN_BUF: dict = {}  # the predictor-input buffer of choose_levels, reused per shape


def choose_levels(pred, view, dims, edge, n_lod, hist, moments, safety: float,
                  brick_bytes: int = 0, mem_budget: int = 0, detail=None):
    """Features -> MLP -> selection for one view.  The SSE budget is
    safety x (1/255)^2 x 4 x W x H (an RGBA RMSE of sqrt(safety)/255 if the
    predictions were exact)."""
    n_reg = int(np.prod(np.ceil(np.asarray(dims, np.float64) / edge)))
    shape = ((n_lod - 1) * n_reg, N_IN)
    buf = _PLAN_BUF.get(shape)
    if buf is None:
        _PLAN_BUF.clear()
        buf = _PLAN_BUF[shape] = np.empty(shape, np.float32)
    x, area, _ = plan_features(view, dims, edge, n_lod, hist, moments, detail, out=buf)
    y = pred(x).reshape(n_lod - 1, -1)
    budget = safety * (1.0 / 255.0) ** 2 * 4 * view["width"] * view["height"]
    grid = tuple(int(v) for v in np.ceil(np.asarray(dims) / edge))
    return plan_octree(y, area, budget, grid, max(brick_bytes, 1), mem_budget)


def calibrate_safety(renderer, pred, views, dims, edge, n_lod, hist, moments, target: float = 1.0,
                     brick_bytes: int = 0, mem_budget: int = 0, detail=None, refine: int = 4,
                     reduce: str = "max"):
    """Largest budget factor whose learned frames stay within `target` RMSE
    of full resolution over the calibration views -- on every view
    (reduce="max") or on average (reduce="mean"): doubled (or halved) from 4
    until the views cross the target, then bisected geometrically `refine`
    times.  Returns (factor, RMSE statistic at it); (0, statistic) when even
    the smallest factor misses the target (a binding brick budget)."""
    if reduce not in ("max", "mean"):
        raise ValueError("reduce must be 'max' or 'mean'")
    refs = []
    for v in views:
        renderer.resize((v["width"], v["height"]))
        set_view(renderer, v)
        refs.append(full_resolution(renderer, dims, edge))

    def worst_at(safety):
        errs = []
        for v, ref in zip(views, refs):
            renderer.resize((v["width"], v["height"]))
            set_view(renderer, v)
            lev, _, _ = choose_levels(pred, v, dims, edge, n_lod, hist, moments, safety,
                                      brick_bytes, mem_budget, detail)
            renderer.setRegionLODs(lev)
            errs.append(rmse255(renderer.capture(), ref))
            renderer.setRegionLODs(None)
        return max(errs) if reduce == "max" else float(np.mean(errs))

    s = 4.0
    w = worst_at(s)
    if w <= target:
        good, good_w, bad = s, w, None
        while bad is None and good < 2.0 ** 24:
            s = good * 2.0
            w = worst_at(s)
            if w <= target:
                good, good_w = s, w
            else:
                bad = s
        if bad is None:
            return good, good_w
    else:
        bad, good, good_w = s, None, w
        while good is None and bad > 2.0 ** -12:
            s = bad * 0.5
            w = worst_at(s)
            if w <= target:
                good, good_w = s, w
            else:
                bad = s
        if good is None:
            return 0.0, w
    for _ in range(refine):
        s = math.sqrt(good * bad)
        w = worst_at(s)
        if w <= target:
            good, good_w = s, w
        else:
            bad = s
    return good, good_w
