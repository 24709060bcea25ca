"""Bricking / LOD twins of the oracle against brute force and invariants.

The CombineAverage rounding of the reference lives in the absent Tuvok
source ("parity unpinned" against the reference); the rule implemented and
pinned here is floor(sum / count) over the existing children.
"""
import numpy as np
import pytest

from oracle import oracle as O


def brute_downsample(v):
    nz, ny, nx = v.shape
    out = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), v.dtype)
    for z in range(out.shape[0]):
        for y in range(out.shape[1]):
            for x in range(out.shape[2]):
                blk = v[2 * z:2 * z + 2, 2 * y:2 * y + 2, 2 * x:2 * x + 2].astype(np.int64)
                out[z, y, x] = blk.sum() // blk.size
    return out


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 5, 6), (9, 7, 8), (1, 2, 5), (2, 2, 2)])
@pytest.mark.parametrize("dt", [np.uint8, np.uint16])
def test_downsample_brute_force(shape, dt):
    rng = np.random.default_rng(sum(shape))
    v = rng.integers(0, np.iinfo(dt).max + 1, shape).astype(dt)
    assert np.array_equal(O.downsample(v), brute_downsample(v))


def brute_median(v):
    nz, ny, nx = v.shape
    out = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), v.dtype)
    for z in range(out.shape[0]):
        for y in range(out.shape[1]):
            for x in range(out.shape[2]):
                blk = np.sort(v[2 * z:2 * z + 2, 2 * y:2 * y + 2, 2 * x:2 * x + 2].ravel())
                out[z, y, x] = blk[(blk.size - 1) // 2]
    return out


@pytest.mark.parametrize("shape", [(3, 3, 3), (4, 5, 6), (9, 7, 8), (1, 2, 5), (2, 2, 2), (1, 1, 1)])
@pytest.mark.parametrize("dt", [np.uint8, np.uint16])
def test_downsample_median_brute_force(shape, dt):
    """The median rule: element (n-1)/2 of the sorted existing children, so
    the lower of the two middle values for 8 (and 2, 4) children."""
    rng = np.random.default_rng(3 + sum(shape))
    v = rng.integers(0, np.iinfo(dt).max + 1, shape).astype(dt)
    assert np.array_equal(O.downsample(v, median=True), brute_median(v))


def test_downsample_median_cases():
    # 8 distinct children 0..7 -> the 4th smallest (3), not the mean (3.5)
    w = np.arange(8, dtype=np.uint8).reshape(2, 2, 2)
    assert O.downsample(w, median=True)[0, 0, 0] == 3
    # a single outlier does not move the median, it does move the mean
    w = np.full((2, 2, 2), 10, np.uint16)
    w[1, 1, 1] = 60000
    assert O.downsample(w, median=True)[0, 0, 0] == 10
    assert O.downsample(w)[0, 0, 0] == (7 * 10 + 60000) // 8
    # constant volumes stay constant, values stay within the children's range
    c = np.full((7, 6, 5), 99, np.uint8)
    assert np.all(O.downsample(c, median=True) == 99)


def test_downsample_invariants():
    rng = np.random.default_rng(7)
    c = np.full((13, 11, 9), 4321, np.uint16)
    assert np.all(O.downsample(c) == 4321)
    v = rng.integers(0, 65536, (10, 12, 14)).astype(np.uint16)
    d = O.downsample(v)
    mx = np.maximum.reduce([v[a::2, b::2, c::2][:d.shape[0], :d.shape[1], :d.shape[2]]
                            for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    mn = np.minimum.reduce([v[a::2, b::2, c::2][:d.shape[0], :d.shape[1], :d.shape[2]]
                            for a in (0, 1) for b in (0, 1) for c in (0, 1)])
    assert np.all(mn <= d) and np.all(d <= mx)
    # exact case: the mean of {0, 1} x 8 voxels is floored
    w = np.zeros((2, 2, 2), np.uint8)
    w[0, 0, 0] = 7
    assert O.downsample(w)[0, 0, 0] == 0


def test_lod_pyramid_shapes():
    v = np.zeros((512 // 8, 512 // 8, 512 // 8), np.uint16)
    lv = O.lod_pyramid(v, 3)
    assert [x.shape for x in lv] == [(64, 64, 64), (32, 32, 32), (16, 16, 16)]
    odd = O.lod_pyramid(np.zeros((5, 9, 17), np.uint8), 4)
    assert [x.shape for x in odd] == [(5, 9, 17), (3, 5, 9), (2, 3, 5), (1, 2, 3)]


@pytest.mark.parametrize("dt", [np.uint8, np.uint16])
def test_extract_brick_matches_edge_padding(dt):
    rng = np.random.default_rng(11)
    v = rng.integers(0, 200, (21, 17, 30)).astype(dt)
    B, a = 8, 2
    pad = np.pad(v, a + B, mode="edge")
    grid = [(s + B - 1) // B for s in v.shape]
    for bz in range(grid[0]):
        for by in range(grid[1]):
            for bx in range(grid[2]):
                s = O.extract_brick(v, B, a, bx, by, bz)
                z0, y0, x0 = bz * B + B, by * B + B, bx * B + B
                exp = pad[z0:z0 + B + 2 * a, y0:y0 + B + 2 * a, x0:x0 + B + 2 * a]
                assert np.array_equal(s, exp)
                mn, mx = O.brick_minmax(v, B, a, bx, by, bz)
                assert (mn, mx) == (float(exp.min()), float(exp.max()))


def test_lod_count_rule():
    """RasterDataBlock LOD count: 1 + halvings until max dim <= brick size
    (UVFReader/DataSource.h:521-529)."""
    def count(m, b):
        n = 1
        while m > b:
            m //= 2
            n += 1
        return n
    assert count(128, 256) == 1   # c60: one level
    assert count(512, 256) == 2
    assert count(2048, 128) == 5
    assert count(4096, 128) == 6
