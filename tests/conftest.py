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
# 2/255 on pixels whose early-termination decision is not within 1e-5 of
# the threshold (flagged by the oracle; there the two precisions may stop
# one sample apart, a difference bounded by 0.01 = 2.55/255 before
# rounding, so 4/255 is allowed), and RMSE <= 0.5/255 over all pixels.
MAX_DIFF = 2
MAX_DIFF_FLAGGED = 4
MAX_RMSE = 0.5


def assert_parity(img, ref8, flags, what="", flagged_max=MAX_DIFF_FLAGGED, max_flagged_frac=1.0):
    """Isosurface renders pass flagged_max=255: a ray whose samples come
    within 1e-2 of the iso value may hit in one precision and miss in the
    other; such pixels are counted (max_flagged_frac) instead of bounded."""
    d, rmse = pixel_diff(img, ref8)
    ok = flags == 0
    m_ok = int(d[ok].max()) if ok.any() else 0
    m_fl = int(d[~ok].max()) if (~ok).any() else 0
    print(f"{what}: max|d| {m_ok}/255 (unflagged), {m_fl}/255 on {int((~ok).sum())} flagged px, "
          f"rmse {rmse:.4f}/255")
    assert m_ok <= MAX_DIFF, f"{what}: max |d| {m_ok} > {MAX_DIFF}"
    assert m_fl <= flagged_max, f"{what}: flagged max |d| {m_fl}"
    assert (~ok).mean() <= max_flagged_frac, f"{what}: {(~ok).mean():.4f} of the pixels flagged"
    rmse_ok = float(np.sqrt((d[ok].astype(np.float64) ** 2).mean())) if ok.any() else 0.0
    assert (rmse if flagged_max <= MAX_DIFF_FLAGGED else rmse_ok) <= MAX_RMSE, f"{what}: rmse {rmse}"
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
