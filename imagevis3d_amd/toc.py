"""TOC (UVF version 3) volume files: the bricked multi-resolution container
of Tuvok's TOCBlock (UVFReader/DataSource.h:475-491, queried as in
UVFReader/BlockInfo.h:38-96), built on the GPU by ``tvk_toc_build`` and read
back or paged into a renderer's brick pool.

Options follow uvfconvert (UVFReader/main.cpp:86-109) and the converter
settings (ImageVis3D_Settings.cpp:516-548): ``max_brick`` (overlap
included), ``overlap``, ``use_median``, ``clamp_to_edge``, ``compression``
(0 none, 1 zlib, 2 lzma, 3 lz4, 4 bzip2), ``level`` (1..10) and ``layout``
(0 scanline, 1 Morton, 2 Hilbert, 3 random).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import TocInfo, call

COMPRESSION = {"none": 0, "zlib": 1, "lzma": 2, "lz4": 3, "bzip2": 4}
LAYOUT = {"scanline": 0, "morton": 1, "hilbert": 2, "random": 3}


def _info_dict(i: TocInfo) -> dict:
    return {"dims": tuple(int(v) for v in i.dims), "bits": i.bits, "max_brick": i.max_brick,
            "overlap": i.overlap, "n_lod": i.n_lod, "use_median": bool(i.use_median),
            "clamp_to_edge": bool(i.clamp_to_edge), "compression": i.compression,
            "level": i.level, "layout": i.layout, "n_bricks": int(i.n_bricks),
            "stored_bytes": int(i.stored_bytes), "raw_bytes": int(i.raw_bytes)}


def build_toc(vol: np.ndarray, path: str, max_brick: int = 128, overlap: int = 2,
              use_median: bool = False, clamp_to_edge: bool = True, compression: int = 1,
              level: int = 1, layout: int = 0, device: int = 0) -> dict:
    """Write `vol` (z, y, x; uint8 or uint16) as a TOC volume file."""
    vol = np.ascontiguousarray(vol)
    bits = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16}[vol.dtype]
    nz, ny, nx = vol.shape
    info = TocInfo()
    call("tvk_toc_build", int(device), vol.ctypes.data, nx, ny, nz, bits, int(max_brick),
         int(overlap), int(bool(use_median)), int(bool(clamp_to_edge)), int(compression),
         int(level), int(layout), str(path).encode(), C.byref(info))
    return _info_dict(info)


def build_toc_raw(raw_path: str, header_bytes: int, dims, bits: int, path: str,
                  max_brick: int = 128, overlap: int = 2, use_median: bool = False,
                  clamp_to_edge: bool = True, compression: int = 1, level: int = 1,
                  layout: int = 0, device: int = 0) -> dict:
    """The same build from a raw file (x fastest, little endian) whose voxels
    start `header_bytes` into the file (uvfconvert's raw input)."""
    nx, ny, nz = (int(v) for v in dims)
    info = TocInfo()
    call("tvk_toc_build_raw", int(device), str(raw_path).encode(), int(header_bytes), nx, ny, nz,
         int(bits), int(max_brick), int(overlap), int(bool(use_median)), int(bool(clamp_to_edge)),
         int(compression), int(level), int(layout), str(path).encode(), C.byref(info))
    return _info_dict(info)


class TocVolume:
    """An open TOC volume file (tvk_toc_*)."""

    def __init__(self, path: str):
        h = C.c_void_p()
        call("tvk_toc_open", str(path).encode(), C.byref(h))
        self._h = h
        i = TocInfo()
        call("tvk_toc_get_info", self._h, C.byref(i))
        self.info = _info_dict(i)
        self._dtype = np.uint8 if i.bits == 8 else np.uint16

    def lod(self, lod: int):
        """(domain size, brick grid) of one LOD, x/y/z order."""
        d = (C.c_uint64 * 3)()
        g = (C.c_uint64 * 3)()
        call("tvk_toc_get_lod", self._h, int(lod), d, g)
        return tuple(int(v) for v in d), tuple(int(v) for v in g)

    def read_brick(self, lod: int, bx: int, by: int, bz: int):
        """(voxels (z, y, x), compression, min, max) of one stored brick."""
        s = self.info["max_brick"]
        out = np.empty((s, s, s), self._dtype)
        comp, mn, mx = _lib.u32(), C.c_float(), C.c_float()
        call("tvk_toc_read_brick", self._h, int(lod), int(bx), int(by), int(bz), out.ctypes.data,
             out.nbytes, C.byref(comp), C.byref(mn), C.byref(mx))
        return out, int(comp.value), float(mn.value), float(mx.value)

    def histogram1D(self) -> np.ndarray:
        """The file's Histogram1D block (u64 counts per raw value)."""
        n = _lib.u64(0)
        call("tvk_toc_get_histogram_1d", self._h, None, C.byref(n))
        out = np.zeros(n.value, np.uint64)
        call("tvk_toc_get_histogram_1d", self._h, out.ctypes.data, C.byref(n))
        return out

    def histogram2D(self):
        """The file's Histogram2D block: (counts [value, gradient], max gradient)."""
        nv, ng, mg = _lib.u64(0), _lib.u64(0), C.c_float()
        call("tvk_toc_get_histogram_2d", self._h, None, C.byref(nv), C.byref(ng), C.byref(mg))
        out = np.zeros((nv.value, ng.value), np.uint64)
        call("tvk_toc_get_histogram_2d", self._h, out.ctypes.data, C.byref(nv), C.byref(ng), C.byref(mg))
        return out, float(mg.value)

    def close(self) -> None:
        if self._h:
            _lib.load().tvk_toc_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass
