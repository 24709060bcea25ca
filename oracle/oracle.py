"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, as the checker or as the timed CPU baseline; the
product (``imagevis3d_amd``) never does.

Also holds ``read_uvf_v2``: an independent numpy reader of the UVF v2
container (doc/import.adoc:904-1102 of the reference) used to check the
library's own C++ reader and to pull the golden blocks out of c60.uvf.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    """Compile liboracle.so with oracle/Makefile (make rebuilds it when a
    source is newer than the library, so an edited oracle is never stale)."""
    subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


class RenderParams(C.Structure):
    _fields_ = [
        ("voxels", C.c_void_p), ("bits", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("n0x", C.c_int32), ("n0y", C.c_int32), ("n0z", C.c_int32),
        ("lod", C.c_int32), ("rescale", C.c_double * 3),
        ("tf", C.c_void_p), ("n_tf", C.c_int32), ("tf_scale", C.c_double),
        ("view_pos", C.c_double * 3), ("view_dir", C.c_double * 3),
        ("up_dir", C.c_double * 3), ("fov_deg", C.c_double),
        ("ortho", C.c_int32), ("ortho_half_height", C.c_double),
        ("model", C.c_double * 16),
        ("width", C.c_int32), ("height", C.c_int32),
        ("rate", C.c_double), ("term_alpha", C.c_double), ("interp", C.c_int32),
        ("bg_top", C.c_double * 3), ("bg_bottom", C.c_double * 3),
        ("bg_alpha", C.c_double),
        ("brick_size", C.c_int32), ("apron", C.c_int32),
        ("mode", C.c_int32), ("tf2d", C.c_void_p), ("tf2d_w", C.c_int32), ("tf2d_h", C.c_int32),
        ("max_grad", C.c_double), ("lighting", C.c_int32),
        ("ambient", C.c_double * 3), ("diffuse", C.c_double * 3), ("specular", C.c_double * 3),
        ("shininess", C.c_double), ("light_dir", C.c_double * 3),
        ("iso_value", C.c_double), ("iso_color", C.c_double * 3),
        ("cv_enabled", C.c_int32), ("cv_iso_value", C.c_double), ("cv_color", C.c_double * 3),
        ("cv_focus", C.c_double * 2), ("cv_size", C.c_double), ("cv_context_scale", C.c_double),
        ("cv_border_scale", C.c_double),
        ("clip_enabled", C.c_int32), ("clip_plane", C.c_double * 4),
        ("blend_bits", C.c_int32), ("blend_brick", C.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(RenderParams)
        L.oracle_render_1d.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_render_1d.restype = C.c_int
        L.oracle_render_regions.argtypes = [P, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_render_regions.restype = C.c_int
        L.oracle_nominal_samples.argtypes = [P]
        L.oracle_nominal_samples.restype = C.c_uint64
        L.baseline_render_1d_f32.argtypes = [P, C.c_void_p, C.c_int32]
        L.baseline_render_1d_f32.restype = C.c_uint64
        L.baseline_uses_avx512.argtypes = []
        L.baseline_uses_avx512.restype = C.c_int
        L.oracle_downsample.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_void_p]
        L.oracle_downsample_median.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                               C.c_int32, C.c_void_p]
        L.oracle_extract_brick.argtypes = [C.c_void_p, C.c_int32] + [C.c_int32] * 8 + [C.c_void_p]
        L.oracle_brick_minmax.argtypes = [C.c_void_p, C.c_int32] + [C.c_int32] * 8 + [
            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_histogram_1d.argtypes = [C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p]
        L.oracle_histogram_2d.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_uint32, C.c_uint32, C.c_void_p]
        L.oracle_histogram_2d.restype = C.c_float
        L.oracle_blend_quantize.argtypes = [C.c_double, C.c_int32]
        L.oracle_blend_quantize.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def bits_of(vol: np.ndarray) -> int:
    if vol.dtype == np.uint8:
        return 8
    if vol.dtype == np.uint16:
        return 16
    raise TypeError(f"unsupported voxel type {vol.dtype}")


# --------------------------------------------------------------- rendering
def make_params(vol: np.ndarray, tf: np.ndarray, *, dims0=None, lod=0, rescale=(1, 1, 1),
                tf_scale=1.0, view_pos=(0, 0, 2), view_dir=(0, 0, -1), up_dir=(0, 1, 0),
                fov=60.0, ortho=False, ortho_half_height=0.5, model=None,
                width=64, height=64, rate=1.0, term_alpha=0.99, interp=1,
                bg_top=(0, 0, 0), bg_bottom=(0, 0, 0), bg_alpha=1.0,
                brick_size=0, apron=2, mode=0, tf2d=None, max_grad=1.0, lighting=False,
                ambient=(0.1, 0.1, 0.1), diffuse=(0.9, 0.9, 0.9), specular=(0.3, 0.3, 0.3),
                shininess=16.0, light_dir=(0.0, 0.0, 1.0), iso_value=0.0,
                iso_color=(1.0, 1.0, 1.0), clearview=None, clip_plane=None, blend_bits=32,
                blend_brick=0):
    """Parameters of one render; ``vol`` is (nz, ny, nx) x-fastest.  For
    mode 1 ``tf2d`` is (H, W, 4) float32 (value along W) and ``tf_scale``
    maps values to its columns."""
    vol = np.ascontiguousarray(vol)
    if tf is None:
        tf = np.zeros((1, 4), np.float32)
    tf = np.ascontiguousarray(tf, dtype=np.float32)
    t2 = np.ascontiguousarray(tf2d if tf2d is not None else np.zeros((1, 1, 4)), np.float32)
    p = RenderParams()
    p.voxels = _ptr(vol)
    p.bits = bits_of(vol)
    p.nz, p.ny, p.nx = vol.shape
    d0 = dims0 if dims0 is not None else (p.nx, p.ny, p.nz)
    p.n0x, p.n0y, p.n0z = d0
    p.lod = lod
    p.rescale[:] = rescale
    p.tf = _ptr(tf)
    p.n_tf = tf.shape[0]
    p.tf_scale = tf_scale
    p.view_pos[:] = view_pos
    p.view_dir[:] = view_dir
    p.up_dir[:] = up_dir
    p.fov_deg = fov
    p.ortho = int(ortho)
    p.ortho_half_height = ortho_half_height
    m = np.eye(4) if model is None else np.asarray(model, dtype=np.float64)
    p.model[:] = m.reshape(-1).tolist()
    p.width, p.height = width, height
    p.rate = rate
    p.term_alpha = term_alpha
    p.interp = interp
    p.bg_top[:] = bg_top
    p.bg_bottom[:] = bg_bottom
    p.bg_alpha = bg_alpha
    p.brick_size = brick_size
    p.apron = apron
    p.mode = mode
    p.tf2d = _ptr(t2)
    p.tf2d_h, p.tf2d_w = t2.shape[:2]
    p.max_grad = max_grad
    p.lighting = int(lighting)
    p.ambient[:] = ambient
    p.diffuse[:] = diffuse
    p.specular[:] = specular
    p.shininess = shininess
    p.light_dir[:] = light_dir
    p.iso_value = iso_value
    p.iso_color[:] = iso_color
    # ClearView (mode 2): dict(iso_value, color, focus=(x, y), size,
    # context_scale, border_scale)
    if clearview is not None:
        p.cv_enabled = 1
        p.cv_iso_value = clearview["iso_value"]
        p.cv_color[:] = clearview.get("color", (1.0, 0.0, 0.0))
        p.cv_focus[:] = clearview.get("focus", (0.0, 0.0))
        p.cv_size = clearview.get("size", 5.5)
        p.cv_context_scale = clearview.get("context_scale", 1.0)
        p.cv_border_scale = clearview.get("border_scale", 60.0)
    # clip plane (a, b, c, d) in world space; blend precision 8 / 16 / 32
    # with blend_brick = the sampled level's brick payload
    if clip_plane is not None:
        p.clip_enabled = 1
        p.clip_plane[:] = clip_plane
    if blend_bits < 32 and blend_brick <= 0:
        raise ValueError("blend_bits < 32 needs the level's brick size (blend_brick)")
    p.blend_bits = blend_bits
    p.blend_brick = blend_brick
    # keep the buffers alive as long as the params
    p._keep = (vol, tf, t2)
    return p


def blend_quantize(x: float, bits: int) -> float:
    """The blend target's rounding of one composite value (8 or 16 bits)."""
    return lib().oracle_blend_quantize(float(x), int(bits))


def render(p: RenderParams):
    """FP64 render -> (rgba float64 HxWx4, samples uint32 HxW, flags uint8 HxW)."""
    out = np.zeros((p.height, p.width, 4), np.float64)
    ns = np.zeros((p.height, p.width), np.uint32)
    fl = np.zeros((p.height, p.width), np.uint8)
    rc = lib().oracle_render_1d(C.byref(p), _ptr(out), _ptr(ns), _ptr(fl))
    if rc != 0:
        raise RuntimeError("oracle_render_1d failed (singular region matrix?)")
    return out, ns, fl


def render_regions(p: RenderParams, levels, region_lod: np.ndarray, region_edge: int):
    """FP64 per-region LOD render (1D TF): `levels` = LOD pyramid (z, y, x
    arrays), `region_lod` = level per LOD-0 region of edge `region_edge`
    ((gz, gy, gx) or flat, x fastest).  Returns (rgba, samples, flags)."""
    levels = [np.ascontiguousarray(v) for v in levels]
    ptrs = (C.c_void_p * len(levels))(*[v.ctypes.data for v in levels])
    dims = np.array([[v.shape[2], v.shape[1], v.shape[0]] for v in levels], np.int32).ravel()
    rl = np.ascontiguousarray(region_lod, np.uint8).ravel()
    out = np.zeros((p.height, p.width, 4), np.float64)
    ns = np.zeros((p.height, p.width), np.uint32)
    fl = np.zeros((p.height, p.width), np.uint8)
    rc = lib().oracle_render_regions(C.byref(p), len(levels), ptrs, _ptr(dims), int(region_edge),
                                     _ptr(rl), _ptr(out), _ptr(ns), _ptr(fl))
    if rc != 0:
        raise RuntimeError("oracle_render_regions failed")
    return out, ns, fl


def to_rgba8(img: np.ndarray) -> np.ndarray:
    return np.floor(np.clip(img, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)


def nominal_samples(p: RenderParams) -> int:
    return int(lib().oracle_nominal_samples(C.byref(p)))


def baseline_render(p: RenderParams, threads: int = 0):
    """FP32 OpenMP CPU raycaster -> (rgba8 HxWx4, evaluated samples)."""
    out = np.zeros((p.height, p.width, 4), np.uint8)
    n = lib().baseline_render_1d_f32(C.byref(p), _ptr(out), threads)
    return out, int(n)


# ------------------------------------------------------------ bricking / LOD
def downsample(vol: np.ndarray, median: bool = False) -> np.ndarray:
    """One 2x2x2 reduction: integer mean, or the lower median (median=True)."""
    vol = np.ascontiguousarray(vol)
    nz, ny, nx = vol.shape
    out = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), vol.dtype)
    fn = lib().oracle_downsample_median if median else lib().oracle_downsample
    fn(_ptr(vol), bits_of(vol), nx, ny, nz, _ptr(out))
    return out


def lod_pyramid(vol: np.ndarray, n_lod: int):
    levels = [np.ascontiguousarray(vol)]
    for _ in range(1, n_lod):
        levels.append(downsample(levels[-1]))
    return levels


def extract_brick(vol: np.ndarray, B: int, apron: int, bx: int, by: int, bz: int) -> np.ndarray:
    vol = np.ascontiguousarray(vol)
    nz, ny, nx = vol.shape
    S = B + 2 * apron
    out = np.zeros((S, S, S), vol.dtype)
    lib().oracle_extract_brick(_ptr(vol), bits_of(vol), nx, ny, nz, B, apron, bx, by, bz, _ptr(out))
    return out


def brick_minmax(vol: np.ndarray, B: int, apron: int, bx: int, by: int, bz: int):
    vol = np.ascontiguousarray(vol)
    nz, ny, nx = vol.shape
    mn, mx = C.c_double(), C.c_double()
    lib().oracle_brick_minmax(_ptr(vol), bits_of(vol), nx, ny, nz, B, apron, bx, by, bz,
                              C.byref(mn), C.byref(mx))
    return mn.value, mx.value


def histogram_1d(vol: np.ndarray) -> np.ndarray:
    vol = np.ascontiguousarray(vol)
    b = bits_of(vol)
    h = np.zeros(1 << b, np.uint64)
    lib().oracle_histogram_1d(_ptr(vol), b, vol.size, _ptr(h))
    return h


def histogram_2d(vol: np.ndarray, n_value: int = 256, n_grad: int = 256):
    vol = np.ascontiguousarray(vol)
    nz, ny, nx = vol.shape
    h = np.zeros((n_value, n_grad), np.uint64)
    g = lib().oracle_histogram_2d(_ptr(vol), bits_of(vol), nx, ny, nz, n_value, n_grad, _ptr(h))
    return h, float(g)


def histogram_2d_exact(vol: np.ndarray, n_value: int = 256, n_grad: int = 256):
    """Exact-integer statement of the 2D histogram rule: squared gradient K in
    units of 1/(2*maxval) (central differences inside, doubled one-sided
    differences at the borders), bin = max{b : b^2 Kmax <= (n_grad-1)^2 K},
    value bin = v >> (bits - log2 n_value).  Returns (hist, max_grad f32).
    On c60.uvf it reproduces the stored block and max gradient exactly."""
    b = bits_of(vol)
    v = np.asarray(vol).astype(np.int64)

    def k(ax):
        g = np.zeros_like(v)
        if v.shape[ax] < 2:
            return g
        sl = lambda a, c: tuple(slice(a, c) if i == ax else slice(None) for i in range(3))
        g[sl(1, -1)] = v[sl(2, None)] - v[sl(None, -2)]
        g[sl(0, 1)] = 2 * (v[sl(1, 2)] - v[sl(0, 1)])
        g[sl(-1, None)] = 2 * (v[sl(-1, None)] - v[sl(-2, -1)])
        return g

    K = sum(k(a) ** 2 for a in range(3))
    km = int(K.max())
    F = n_grad - 1
    if km > 0:
        gb = np.floor(F * np.sqrt(K / km)).astype(np.int64)
        gb = np.where((gb + 1 <= F) & ((gb + 1) ** 2 * km <= F * F * K), gb + 1, gb)
        gb = np.where((gb > 0) & (gb ** 2 * km > F * F * K), gb - 1, gb)
    else:
        gb = np.zeros_like(K)
    shift = 0
    while (n_value << shift) < (1 << b):
        shift += 1
    h = np.zeros((n_value, n_grad), np.uint64)
    np.add.at(h, ((v >> shift).ravel(), gb.ravel()), 1)
    maxval = (1 << b) - 1
    return h, float(np.float32(np.sqrt(km) / (2.0 * maxval)))


# ------------------------------------------------------------------ UVF v2
def read_uvf_v2(path: str) -> dict:
    """Parse a UVF v2 file (little endian, uncompressed blocks) into a dict.

    Layout: doc/import.adoc:916-975 (master header), :977-1031 (block
    header), :1033-1102 (histogram, raster blocks); the raster header tail
    (element semantics, signedness, data offset) follows the bytes of
    DemoData/c60.uvf.
    """
    b = open(path, "rb").read()
    if b[:8] != b"UVF-DATA":
        raise ValueError("not a UVF file")
    if b[8] != 0:
        raise ValueError("big-endian UVF not supported")
    u64 = lambda o: struct.unpack_from("<Q", b, o)[0]
    out = {"version": u64(9), "checksum_type": u64(17)}
    ncs = u64(25)
    out["checksum"] = b[33:33 + ncs]
    o = 33 + ncs
    seek = u64(o)
    out["checksum_offset"] = o
    o = o + 8 + seek
    if out["checksum_type"] == 2:
        out["md5_ok"] = hashlib.md5(b[out["checksum_offset"]:]).digest() == out["checksum"]
    blocks = []
    while True:
        start = o
        n = u64(o)
        desc = b[o + 8:o + 8 + n].decode("latin-1")
        o += 8 + n
        sem, comp, nxt = u64(o), u64(o + 8), u64(o + 16)
        o += 24
        blk = {"offset": start, "desc": desc, "semantic": sem, "compression": comp}
        if sem == 1:  # raster
            nd = u64(o); o += 8
            blk["dim_semantics"] = [u64(o + 8 * i) for i in range(nd)]; o += 8 * nd
            blk["transform"] = np.frombuffer(b, np.float64, (nd + 1) ** 2, o).reshape(nd + 1, nd + 1); o += 8 * (nd + 1) ** 2
            for key in ("domain", "brick_size", "overlap", "lod_factor", "lod_groups"):
                blk[key] = [u64(o + 8 * i) for i in range(nd)]; o += 8 * nd
            ngroups = max(blk["lod_groups"]) + 1
            blk["lod_counts"] = [u64(o + 8 * i) for i in range(ngroups)]; o += 8 * ngroups
            nel = u64(o); o += 8
            el_size = [u64(o + 8 * i) for i in range(nel)]; o += 8 * nel
            elems = []
            for es in el_size:
                sems = [u64(o + 8 * i) for i in range(es)]; o += 8 * es
                bits = [u64(o + 8 * i) for i in range(es)]; o += 8 * es
                mant = [u64(o + 8 * i) for i in range(es)]; o += 8 * es
                signed = list(b[o:o + es]); o += es
                elems.append({"semantics": sems, "bits": bits, "mantissa": mant, "signed": signed})
            blk["elements"] = elems
            data_off = u64(o); o += 8 + data_off
            blk["data_offset"] = o
            d = blk["domain"]
            bpv = elems[0]["bits"][0] // 8
            dt = {1: np.uint8, 2: np.uint16}[bpv]
            nvox = int(np.prod(d))
            # a single-brick level-0 raster (no LOD): the payload is the
            # flat volume, x fastest (domain listed x, y, z in c60).
            if blk["lod_counts"] == [1] and all(di <= bs for di, bs in zip(d, blk["brick_size"])):
                blk["volume"] = np.frombuffer(b, dt, nvox, o).reshape(d[2], d[1], d[0])
        elif sem == 5:  # 1D histogram
            hn = u64(o)
            blk["histogram"] = np.frombuffer(b, np.uint64, hn, o + 8).copy()
        elif sem == 6:  # 2D histogram
            mg = struct.unpack_from("<f", b, o)[0]
            ex, ey = u64(o + 4), u64(o + 12)
            blk["max_grad"] = mg
            blk["histogram"] = np.frombuffer(b, np.uint64, ex * ey, o + 20).reshape(ex, ey).copy()
        elif sem == 7:  # max/min
            nb, nc = u64(o), u64(o + 8)
            vals = np.frombuffer(b, np.float64, nb * nc * 4, o + 16).reshape(nb, nc, 4)
            blk["minmax"] = vals.copy()
        blocks.append(blk)
        if nxt == 0:
            break
        o = start + nxt
    out["blocks"] = blocks
    return out
