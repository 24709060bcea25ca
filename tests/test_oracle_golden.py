"""The oracle and the UVF reader against the known answers stored inside the
reference's DemoData/c60.uvf (MD5, 1D/2D histograms, max gradient, min/max)
and the reference's TF files."""
import hashlib
import os

import numpy as np

from conftest import GOLDEN
from oracle import oracle as O


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_uvf_header_and_md5(c60, golden):
    g = golden["c60"]
    assert c60["version"] == 2 and c60["checksum_type"] == 2
    assert c60["checksum"].hex() == g["md5"] == "e7a470483e4da953efbfb224619df22f"
    assert c60["md5_ok"]


def test_uvf_raster_block(c60, golden):
    ras = c60["blocks"][0]
    assert ras["semantic"] == 1 and ras["domain"] == [128, 128, 128]
    assert ras["brick_size"] == [256, 256, 256] and ras["overlap"] == [4, 4, 4]
    assert ras["lod_counts"] == [1] and ras["data_offset"] == 484
    assert sha(ras["volume"]) == golden["c60"]["volume_sha256"]


def test_md5_detects_corruption(tmp_path, c60_path):
    b = bytearray(open(c60_path, "rb").read())
    b[1000] ^= 0xFF
    p = tmp_path / "bad.uvf"
    p.write_bytes(bytes(b))
    assert not O.read_uvf_v2(str(p))["md5_ok"]


def test_histogram_1d_matches_stored_block(c60):
    vol = c60["blocks"][0]["volume"]
    stored = c60["blocks"][1]["histogram"]
    assert stored.sum() == 128 ** 3
    assert np.array_equal(O.histogram_1d(vol), stored)
    assert np.array_equal(np.bincount(vol.ravel(), minlength=256).astype(np.uint64), stored)


def test_histogram_2d_matches_stored_block(c60):
    vol = c60["blocks"][0]["volume"]
    blk = c60["blocks"][2]
    h, gmax = O.histogram_2d(vol, 256, 256)
    assert np.array_equal(h, blk["histogram"])
    # the stored max gradient is the f32 rounding of the exact maximum
    assert abs(gmax - blk["max_grad"]) <= 2e-8


def test_histogram_2d_exact_integer_rule(c60):
    """The exact-integer statement of the rule (the one the HIP kernel
    implements) reproduces the stored block and max gradient bit for bit."""
    blk = c60["blocks"][2]
    h, mg = O.histogram_2d_exact(c60["blocks"][0]["volume"], 256, 256)
    assert np.array_equal(h, blk["histogram"])
    assert np.float32(mg) == np.float32(blk["max_grad"])


def test_histogram_2d_twins_agree_on_u16():
    rng = np.random.default_rng(5)
    v = (rng.integers(0, 4096, (24, 40, 56)) * rng.integers(0, 2, (24, 40, 56))).astype(np.uint16)
    a, _ = O.histogram_2d(v, 256, 256)
    b, _ = O.histogram_2d_exact(v, 256, 256)
    assert np.array_equal(a, b)


def test_maxmin_block(c60):
    mm = c60["blocks"][3]["minmax"]
    vol = c60["blocks"][0]["volume"]
    assert mm.shape == (1, 1, 4)
    assert mm[0, 0, 0] == vol.min() == 0 and mm[0, 0, 1] == vol.max() == 255
    # the whole volume is one brick; with an apron beyond the volume the
    # slot range equals the volume range
    assert O.brick_minmax(vol, 128, 2, 0, 0, 0) == (0.0, 255.0)


def test_tf_files(golden):
    from imagevis3d_amd import transfer
    tf = transfer.load_1dt(os.path.join(GOLDEN, "c60.1dt"))
    assert tf.shape == (256, 4)
    nz = np.nonzero(tf[:, 3] > 0)[0]
    assert [nz[0], nz[-1]] == golden["c60_1dt"]["alpha_nonzero"] == [87, 255]
    t2 = transfer.parse_2dt(os.path.join(GOLDEN, "c60.2dt"))
    assert t2["size"] == (256, 256) and t2["tf1d"].shape == (256, 4)
    assert len(t2["swatches"]) == 6
    assert all(len(s["points"]) == 4 for s in t2["swatches"])


def test_oracle_c60_regression(golden):
    """FP64 oracle renders of a c60 crop are stable (fixture written by
    tests/golden/make_golden.py)."""
    import sys
    sys.path.insert(0, GOLDEN)
    import make_golden as mg
    u = O.read_uvf_v2(os.path.join(GOLDEN, "c60.uvf"))
    crop = np.ascontiguousarray(u["blocks"][0]["volume"][mg.CROP])
    tf = np.loadtxt(os.path.join(GOLDEN, "c60.1dt"), skiprows=1, dtype=np.float32)
    ref = np.load(os.path.join(GOLDEN, "oracle_c60_views.npz"))
    for i, v in enumerate(mg.VIEWS[:1]):
        p = O.make_params(crop, tf, width=96, height=96, fov=50.0, rate=1.0, brick_size=16,
                          apron=2, **v)
        img, ns, _ = O.render(p)
        assert np.array_equal(O.to_rgba8(img), ref["rgba8"][i])
        assert np.array_equal(ns, ref["samples"][i])
