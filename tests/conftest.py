import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtvkhip.so on device 0)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def c60_path():
    return os.path.join(GOLDEN, "c60.uvf")


@pytest.fixture(scope="session")
def c60(c60_path):
    from oracle import oracle as O
    return O.read_uvf_v2(c60_path)


@pytest.fixture(scope="session")
def c60_tf():
    from imagevis3d_amd import transfer
    return transfer.load_1dt(os.path.join(GOLDEN, "c60.1dt"))


def pixel_diff(a: np.ndarray, b: np.ndarray):
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    return d, float(np.sqrt((d.astype(np.float64) ** 2).mean()))


# Parity bar for RGBA8 images against the FP64 oracle: per-pixel max |d| <=
# 2/255 on every pixel and RMSE <= 0.5/255 over all pixels.  The oracle
# flags pixels whose outcome hangs on a comparison within rounding of its
# threshold (early termination within 1e-5 of alpha 0.99; an isosurface or
# ClearView sample within 1e-2 of its iso value; nearest sampling within
# 1e-4 of a voxel face).  Only isosurface-type renders may differ beyond the
# bar there -- a ray that grazes the surface can hit in one precision and
# miss in the other -- and only on at most `max_flip_frac` of the pixels.
MAX_DIFF = 2
MAX_RMSE = 0.5
ISO_FLIP_FRAC = 0.02


def assert_parity(img, ref8, flags, what="", max_flip_frac=0.0):
    d, rmse = pixel_diff(img, ref8)
    ok = flags == 0
    dp = d.max(axis=-1)
    m_ok = int(dp[ok].max()) if ok.any() else 0
    m_fl = int(dp[~ok].max()) if (~ok).any() else 0
    flips = float(((dp > MAX_DIFF) & ~ok).mean())
    print(f"{what}: max|d| {m_ok}/255 (unflagged), {m_fl}/255 on {int((~ok).sum())} flagged px, "
          f"{flips:.4%} beyond the bar, rmse {rmse:.4f}/255")
    assert m_ok <= MAX_DIFF, f"{what}: max |d| {m_ok} > {MAX_DIFF}"
    assert flips <= max_flip_frac, f"{what}: {flips:.4%} of the pixels beyond {MAX_DIFF}/255"
    assert rmse <= MAX_RMSE, f"{what}: rmse {rmse}"
    return d, rmse


def gpu_renderer(w=None, width=None, height=None):
    """A renderer on device 0 through the C ABI (tests marked gpu only)."""
    from imagevis3d_amd import tuvok
    r = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, False)
    ww = width or (w.width if w else 64)
    hh = height or (w.height if w else 64)
    assert r.initialize(tuvok.createContext(ww, hh, device=0)), r.lastError()
    r.setRendererTarget(tuvok.renderer.types.RT_Capture)
    return r
