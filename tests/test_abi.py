"""The C-ABI library without a GPU: it loads, exports every symbol the header
declares, validates arguments, and its host-only generators are seeded."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tvk_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tvk_[a-z0-9_]+)\s*\(", src)))


def test_library_loads_and_exports_every_declared_symbol():
    from imagevis3d_amd import _lib
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)
    assert lib.tvk_version().startswith(b"tvk-hip")


def test_create_rejects_renderers_the_backend_does_not_provide():
    from imagevis3d_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.tvk_create(0, 0, 0, 0, 0, C.byref(h)) == -1  # OpenGL SBVR
    assert b"raycaster" in lib.tvk_last_error()
    assert lib.tvk_create(1, 0, 0, 0, 0, C.byref(h)) == 0
    # state errors before initialize: no device call is made
    assert lib.tvk_paint(h) == -3
    assert b"initialize" in lib.tvk_last_error()
    assert lib.tvk_set_fov(h, C.c_float(200.0)) == -1
    assert lib.tvk_set_sample_rate_modifier(h, C.c_float(0.0)) == -1
    assert lib.tvk_set_lod_limits(h, 2, 1) == -1
    assert lib.tvk_destroy(h) == 0
    assert lib.tvk_paint(None) == -1


def test_python_facade_error_behaviour():
    from imagevis3d_amd import TvkError, tuvok
    r = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, False)
    assert r.paint() is False and "initialize" in r.lastError()
    with pytest.raises(TvkError):
        r.setFoV(0.0)
    with pytest.raises(TvkError):
        tuvok.renderer.new(tuvok.renderer.types.OpenGL_SBVR, False, False, False, False)
    r.cleanup()


def test_generators_are_seeded_and_deterministic():
    from imagevis3d_amd.renderer import VOL_BLOBS, VOL_CONST, VOL_RANDOM, generate_volume
    a = generate_volume(VOL_BLOBS, 1, (64, 64, 64), 8)
    b = generate_volume(VOL_BLOBS, 1, (64, 64, 64), 8)
    c = generate_volume(VOL_BLOBS, 2, (64, 64, 64), 8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.max() > 150 and np.percentile(a, 50) < 10  # blobs over low noise
    assert np.all(generate_volume(VOL_CONST, 77, (5, 6, 7), 16) == 77)
    r = generate_volume(VOL_RANDOM, 3, (32, 32, 32), 16)
    assert r.std() > 15000


def test_medium_generator_statistics():
    from imagevis3d_amd.renderer import VOL_PERLIN_SPHERES, generate_volume
    v = generate_volume(VOL_PERLIN_SPHERES, 2, (96, 80, 64), 16)
    assert v.min() == 0 and v.max() <= 4095
    assert np.count_nonzero((v >= 3000) & (v <= 4000)) > 0
