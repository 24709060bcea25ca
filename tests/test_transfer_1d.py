"""1D TF ingestion (no GPU): .1dt resampling to the histogram size
(loadFromFileWithSize, Q1DTransferFunction.cpp:573-589) and the editor's
TF proxy (getColorData / setStdFunction, Q1DTransferFunction.cpp:254-258,
441-455), pinned by closed forms."""
import os

import numpy as np
import pytest

from imagevis3d_amd import transfer
from imagevis3d_amd.transfer import TransferFunction1D

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_resample_identity_at_equal_size():
    tf = transfer.load_1dt(os.path.join(GOLDEN, "c60.1dt"))
    out = transfer.resample_1d(tf, tf.shape[0])
    assert np.array_equal(out, tf) and out is not tf


@pytest.mark.parametrize("size", [2, 37, 256, 1000, 4096])
def test_resample_keeps_end_points_and_linear_ramps(size):
    n = 17
    a, b = np.array([0.1, 0.2, 0.3, 0.0]), np.array([0.9, 0.5, 0.1, 1.0])
    ramp = (a + (b - a) * (np.arange(n)[:, None] / (n - 1))).astype(np.float32)
    out = transfer.resample_1d(ramp, size)
    assert out.shape == (size, 4) and out.dtype == np.float32
    assert np.array_equal(out[0], ramp[0]) and np.array_equal(out[-1], ramp[-1])
    expect = a + (b - a) * (np.arange(size)[:, None] / (size - 1))
    assert np.allclose(out, expect, atol=1e-6)


def test_resample_constant_and_grid_points():
    const = np.tile(np.float32([0.25, 0.5, 0.75, 0.125]), (9, 1))
    assert np.all(transfer.resample_1d(const, 50) == const[0])
    # upsampling from n to (n-1) k + 1 entries keeps every original entry
    rng = np.random.default_rng(3)
    tf = rng.random((9, 4)).astype(np.float32)
    up = transfer.resample_1d(tf, 8 * 5 + 1)
    assert np.allclose(up[::5], tf, atol=1e-7)
    # and the entries in between lie between their neighbours
    lo = np.minimum(tf[:-1], tf[1:])
    hi = np.maximum(tf[:-1], tf[1:])
    mid = up[2::5][:8]
    assert np.all(mid >= lo - 1e-7) and np.all(mid <= hi + 1e-7)


def test_load_from_file_with_size(tmp_path):
    path = os.path.join(GOLDEN, "c60.1dt")
    tf = TransferFunction1D()
    assert tf.loadFromFileWithSize(path, 4096)
    assert tf.getSize() == 4096
    src = transfer.load_1dt(path)
    assert np.array_equal(tf.getColorData()[0], src[0])
    assert np.array_equal(tf.getColorData()[-1], src[-1])
    assert tf.loadFromFileWithSize(path, 0) and tf.getSize() == src.shape[0]
    assert not tf.loadFromFileWithSize(str(tmp_path / "missing.1dt"), 256)
    bad = tmp_path / "bad.1dt"
    bad.write_text("3\n1 0 0 1\n")
    assert not tf.loadFromFile(str(bad))


def test_color_data_is_the_table():
    tf = TransferFunction1D(8)
    cd = tf.getColorData()
    cd[3] = (1.0, 0.5, 0.25, 0.75)
    assert np.array_equal(tf.getColorData()[3], np.float32([1.0, 0.5, 0.25, 0.75]))


def test_std_function_step_closed_form():
    n = 101
    tf = TransferFunction1D(n)
    tf.setStdFunction(0.5, 0.2, 3)
    a = tf.getColorData()[:, 3]
    # ramp from entry 40 to entry 60 (width 0.2 * 100 around 50)
    assert np.all(a[:41] == 0.0) and np.all(a[60:] == 1.0)
    assert a[50] == pytest.approx(0.5)
    assert np.allclose(np.diff(a[40:61]), 1.0 / 20.0, atol=1e-6)
    assert np.all(tf.getColorData()[:, :3] == 0.0)  # other components untouched
    tf.setStdFunction(0.5, 0.2, 0, bInvertedStep=True)
    assert np.allclose(tf.getColorData()[:, 0], 1.0 - a, atol=1e-7)
    # zero width: a hard step at the centre; clamped arguments
    tf.setStdFunction(0.3, 0.0, 1)
    g = tf.getColorData()[:, 1]
    assert np.all(g[:30] == 0.0) and np.all(g[30:] == 1.0)
    tf.setStdFunction(2.0, -1.0, 2)
    assert np.all(tf.getColorData()[:-1, 2] == 0.0) and tf.getColorData()[-1, 2] == 1.0
    with pytest.raises(ValueError):
        tf.setStdFunction(0.5, 0.1, 4)
