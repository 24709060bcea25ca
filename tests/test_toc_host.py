"""TOC (UVF version 3) volume files on the host, no GPU: the library's reader
(tvk_toc_open / tvk_toc_read_brick) against files written here by an
independent encoder of the documented byte layout, with Python's own zlib,
lzma (xz) and bz2 codecs; malformed files are rejected."""
import bz2
import lzma
import os
import struct
import zlib

import numpy as np
import pytest

from imagevis3d_amd import TvkError
from imagevis3d_amd.toc import TocVolume
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C60 = os.path.join(ROOT, "tests", "golden", "c60.uvf")

CODECS = {0: lambda b: b, 1: zlib.compress, 2: lambda b: lzma.compress(b, format=lzma.FORMAT_XZ),
          4: bz2.compress}


def toc_bricks(vol, max_brick, overlap, median=False, clamp=True):
    """LOD pyramid and overlap-inclusive bricks as the builder defines them."""
    P = max_brick - 2 * overlap
    levels = [vol]
    while max(levels[-1].shape) > P:
        levels.append(O.downsample(levels[-1], median=median))
    out = []
    for lv in levels:
        nz, ny, nx = lv.shape
        g = [-(-nx // P), -(-ny // P), -(-nz // P)]
        pad = [(overlap, g[2] * P - nz + overlap), (overlap, g[1] * P - ny + overlap),
               (overlap, g[0] * P - nx + overlap)]
        big = np.pad(lv, pad, mode="edge") if clamp else np.pad(lv, pad, mode="constant")
        bricks = []
        for bz in range(g[2]):
            for by in range(g[1]):
                for bx in range(g[0]):
                    bricks.append(np.ascontiguousarray(
                        big[bz * P:bz * P + max_brick, by * P:by * P + max_brick, bx * P:bx * P + max_brick]))
        out.append((lv.shape[::-1], tuple(g), bricks))
    return out


def write_toc(path, vol, max_brick, overlap, codec, order=None, version=3, magic=b"TVKTOC01",
              hists=None):
    """The documented layout: UVF master header, one block (semantic 9), the
    TOC payload, then the stored bricks (here in reverse scanline order);
    with `hists` = (h1, h2, max_grad), a Histogram1D and a Histogram2D block
    after the bricks (doc/import.adoc layouts)."""
    levels = toc_bricks(vol, max_brick, overlap)
    bits = 8 if vol.dtype == np.uint8 else 16
    name = b"test TOC volume"
    head = b"UVF-DATA" + b"\x00" + struct.pack("<QQQQ", version, 0, 0, 0)
    head += struct.pack("<Q", len(name)) + name + struct.pack("<QQQ", 9, 0, 0)
    next_at = len(head) - 8  # the TOC block's next-block field
    toc = magic + struct.pack("<II", bits, 1) + struct.pack("<QQQ", *vol.shape[::-1])
    toc += struct.pack("<IIII", max_brick, overlap, max_brick - 2 * overlap, len(levels))
    toc += struct.pack("<IIIIII", 0, 1, codec, 1, 0, 0)
    first = 0
    for dims, grid, bricks in levels:
        toc += struct.pack("<QQQQQQQ", *dims, *grid, first)
        first += len(bricks)
    flat = [b for _, _, bricks in levels for b in bricks]
    data_at = len(head) + len(toc) + 40 * len(flat)
    stored = [CODECS[codec](b.tobytes()) for b in flat]
    order = list(range(len(flat)))[::-1] if order is None else order
    offsets, pos = [0] * len(flat), data_at
    for i in order:
        offsets[i] = pos
        pos += len(stored[i])
    table = b"".join(struct.pack("<QQQIIff", offsets[i], len(stored[i]), flat[i].nbytes, codec, 0,
                                 float(flat[i].min()), float(flat[i].max())) for i in range(len(flat)))
    tail = b""
    if hists is not None:
        h1, h2, mg = hists
        n1 = b"1D Histogram"
        b1 = struct.pack("<Q", len(n1)) + n1
        b1_len = len(b1) + 24 + 8 + 8 * h1.size
        b1 += struct.pack("<QQQ", 5, 0, b1_len) + struct.pack("<Q", h1.size) + h1.astype("<u8").tobytes()
        n2 = b"2D Histogram"
        b2 = struct.pack("<Q", len(n2)) + n2 + struct.pack("<QQQ", 6, 0, 0)
        b2 += struct.pack("<fQQ", mg, *h2.shape) + h2.astype("<u8").tobytes()
        tail = b1 + b2
        toc_start = 9 + 32
        head = head[:next_at] + struct.pack("<Q", pos - toc_start) + head[next_at + 8:]
    with open(path, "wb") as f:
        f.write(head + toc + table)
        for i in order:
            f.write(stored[i])
        f.write(tail)
    return levels


@pytest.mark.parametrize("codec", [0, 1, 2, 4])
@pytest.mark.parametrize("dt", [np.uint8, np.uint16])
def test_reader_against_independent_writer(tmp_path, codec, dt):
    rng = np.random.default_rng(codec + dt().itemsize)
    vol = rng.integers(0, 200 if dt == np.uint8 else 4000, (16, 18, 20)).astype(dt)
    path = str(tmp_path / "v.uvf")
    levels = write_toc(path, vol, 12, 2, codec)
    t = TocVolume(path)
    info = t.info
    assert info["dims"] == (20, 18, 16) and info["max_brick"] == 12 and info["overlap"] == 2
    assert info["n_lod"] == len(levels) == 3 and info["compression"] == codec
    assert info["n_bricks"] == sum(len(b) for _, _, b in levels)
    for lod, (dims, grid, bricks) in enumerate(levels):
        assert t.lod(lod) == (tuple(dims), grid)
        i = 0
        for bz in range(grid[2]):
            for by in range(grid[1]):
                for bx in range(grid[0]):
                    got, comp, mn, mx = t.read_brick(lod, bx, by, bz)
                    assert np.array_equal(got, bricks[i]), (lod, bx, by, bz)
                    assert comp == codec and mn == bricks[i].min() and mx == bricks[i].max()
                    i += 1
    with pytest.raises(TvkError):
        t.read_brick(3, 0, 0, 0)
    with pytest.raises(TvkError):
        t.read_brick(0, levels[0][1][0], 0, 0)
    t.close()


def test_malformed_files_are_rejected(tmp_path):
    vol = np.arange(10 * 10 * 10, dtype=np.uint16).reshape(10, 10, 10)
    good = str(tmp_path / "good.uvf")
    write_toc(good, vol, 8, 2, 1)
    TocVolume(good).close()
    raw = open(good, "rb").read()
    cases = {
        "truncated": raw[:len(raw) // 2],
        "header_only": raw[:60],
        "not_uvf": b"NOT-UVF!" + raw[8:],
    }
    for name, data in cases.items():
        p = str(tmp_path / f"{name}.uvf")
        open(p, "wb").write(data)
        with pytest.raises(TvkError):
            TocVolume(p)
    v2 = str(tmp_path / "v2.uvf")
    write_toc(v2, vol, 8, 2, 1, version=2)
    with pytest.raises(TvkError):
        TocVolume(v2)
    badmagic = str(tmp_path / "m.uvf")
    write_toc(badmagic, vol, 8, 2, 1, magic=b"XXXXXXXX")
    with pytest.raises(TvkError):
        TocVolume(badmagic)
    # a version-2 RasterDataBlock file is not a TOC volume
    with pytest.raises(TvkError):
        TocVolume(C60)
    # a corrupt compressed brick fails on read, not on open
    blob = bytearray(raw)
    blob[-20:] = b"\xff" * 20
    p = str(tmp_path / "corrupt.uvf")
    open(p, "wb").write(bytes(blob))
    t = TocVolume(p)
    with pytest.raises(TvkError):
        t.read_brick(0, 0, 0, 0)  # reverse order: brick 0 is stored last
    t.close()


def test_histogram_blocks_are_read(tmp_path):
    vol = np.arange(12 * 12 * 12, dtype=np.uint8).reshape(12, 12, 12)
    h1 = np.bincount(vol.ravel(), minlength=256).astype(np.uint64)
    h2 = np.arange(256 * 256, dtype=np.uint64).reshape(256, 256)
    path = str(tmp_path / "h.uvf")
    write_toc(path, vol, 8, 2, 1, hists=(h1, h2, 0.25))
    t = TocVolume(path)
    assert np.array_equal(t.histogram1D(), h1)
    got, mg = t.histogram2D()
    assert np.array_equal(got, h2) and mg == 0.25
    # bricks still read after the extra blocks
    assert np.array_equal(t.read_brick(0, 0, 0, 0)[0], toc_bricks(vol, 8, 2)[0][2][0])
    t.close()
    # a file without histogram blocks reports their absence
    p2 = str(tmp_path / "n.uvf")
    write_toc(p2, vol, 8, 2, 1)
    t2 = TocVolume(p2)
    with pytest.raises(TvkError):
        t2.histogram1D()
    t2.close()


def test_crafted_headers_are_rejected(tmp_path):
    """Headers whose fields contradict each other: no payload between the
    overlaps, a level 0 that is not the volume, a level that does not halve
    the one before (the brick grid still matching its own dims)."""
    vol = np.arange(10 * 10 * 10, dtype=np.uint16).reshape(10, 10, 10)
    good = str(tmp_path / "good.uvf")
    levels = write_toc(good, vol, 8, 2, 0)
    assert [d for d, _, _ in levels] == [(10, 10, 10), (5, 5, 5), (3, 3, 3)]
    raw = bytearray(open(good, "rb").read())
    toc = raw.index(b"TVKTOC01")
    hdr = toc + 16                     # dims (3 x u64)
    brick_at = hdr + 24                # max_brick, overlap, payload, n_lod (u32)
    lvl = brick_at + 16 + 24           # per level: dims, grid (3 x u64 each), first (u64)

    def patched(name, edits):
        b = bytearray(raw)
        for off, fmt, val in edits:
            struct.pack_into(fmt, b, off, *val)
        p = str(tmp_path / f"{name}.uvf")
        open(p, "wb").write(bytes(b))
        return p

    cases = {
        "no_payload": [(brick_at, "<III", (4, 2, 0))],
        "level0_not_volume": [(hdr, "<Q", (11,))],
        "level1_not_half": [(lvl + 56, "<Q", (6,))],  # 6 still needs 2 bricks of 4
    }
    for name, edits in cases.items():
        with pytest.raises(TvkError, match="TOC"):
            TocVolume(patched(name, edits))
    TocVolume(patched("unchanged", [])).close()
