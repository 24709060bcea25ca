"""ClearView (isosurface mode: a context surface that fades around a focus
position to reveal a focus surface) on the HIP raycaster: parity with the
FP64 oracle's ClearView model on nested spheres and on the medium volume,
the reductions that must hold byte for byte (focus outside the image =
plain isosurface; whole image in focus with no context = the focus
isosurface alone), batched launches, and the setters' checks."""
import numpy as np
import pytest

from conftest import ISO_FLIP_FRAC, assert_parity
from imagevis3d_amd import RenderMode, TvkError, transfer, tuvok, workloads
from imagevis3d_amd.workloads import Camera
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LIGHT = dict(ambient=(0.15, 0.1, 0.1), diffuse=(0.7, 0.8, 0.9), specular=(0.3, 0.3, 0.2),
             shininess=12.0, light_dir=(0.3, 0.8, 0.5))
ISO_COLOR = (0.9, 0.8, 0.6)
CV_COLOR = (0.2, 0.4, 1.0)


def nested_spheres(n=64):
    """u8 250 - 200 d / R: the 100 level set (context) encloses the 200
    level set (focus)."""
    z, y, x = np.indices((n, n, n), dtype=np.float64)
    c = (n - 1) / 2.0
    d = np.sqrt((x - c) ** 2 + (y - 0.9 * c) ** 2 + (z - c) ** 2)
    return np.clip(250.0 - 200.0 * d / (0.45 * n), 0, 255).astype(np.uint8)


def make(vol, w, iso, cv=None, brick=16):
    r = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, False)
    assert r.initialize(tuvok.createContext(w.width, w.height, device=0)), r.lastError()
    r.setRendererTarget(tuvok.renderer.types.RT_Capture)
    r.setVolume(vol, brick, 2, 1)
    r.set1DTrans(w.tf)
    r.setLightColors((*LIGHT["ambient"], 1.0), (*LIGHT["diffuse"], 1.0),
                     (*LIGHT["specular"], LIGHT["shininess"]), LIGHT["light_dir"])
    r.setRenderMode(RenderMode.RM_ISOSURFACE)
    r.setIsoValue(iso)
    r.setIsosurfaceColor(ISO_COLOR)
    if cv is not None:
        apply_cv(r, cv)
    return r


def apply_cv(r, cv):
    r.setClearViewEnabled(True)
    r.setCVIsoValue(cv["iso_value"])
    r.setCVColor(cv.get("color", CV_COLOR))
    r.setCVSize(cv["size"])
    r.setCVContextScale(cv["context_scale"])
    r.setCVBorderScale(cv["border_scale"])
    r.setCVFocusPos(None, cv["focus"])


def paint(r, w, cam):
    workloads.apply(r, w, cam)
    assert r.paint(), r.lastError()
    return r.capture()


def oracle(vol, w, cam, iso, cv=None, color=ISO_COLOR):
    kw = workloads.oracle_kwargs(w, cam)
    kw.update(mode=2, iso_value=iso, iso_color=color, **LIGHT)
    if cv is not None:
        kw["clearview"] = dict(cv, color=cv.get("color", CV_COLOR))
    img, ns, fl = O.render(O.make_params(vol, w.tf, **kw))
    return O.to_rgba8(img), fl


def sphere_workload(width=160, height=120):
    cams = [Camera((0.9, 0.6, 1.5), (-0.9, -0.6, -1.5), fov=40.0),
            Camera((-0.3, 1.6, 0.9), (0.3, -1.6, -0.9), (0.0, 0.0, 1.0), fov=35.0)]
    return workloads.Workload("cv", (64, 64, 64), 8, 0, 0, 16, 2, 1, transfer.ramp_tf(256, 20, 250),
                              width, height, 1.5, cams)


@pytest.mark.parametrize("view", [0, 1])
def test_clearview_parity_nested_spheres(view):
    vol = nested_spheres()
    w = sphere_workload()
    cv = dict(iso_value=200.0, focus=(70, 55), size=3.0, context_scale=0.8, border_scale=2.0)
    print(f"case clearview spheres view={view}", flush=True)
    r = make(vol, w, 100.0, cv)
    img = paint(r, w, w.cameras[view])
    ref8, fl = oracle(vol, w, w.cameras[view], 100.0, cv)
    plain, _ = oracle(vol, w, w.cameras[view], 100.0)
    assert (np.abs(ref8.astype(int) - plain.astype(int)).max(-1) > 30).sum() > 100  # focus visible
    assert_parity(img, ref8, fl, f"clearview spheres view {view}", max_flip_frac=ISO_FLIP_FRAC)
    r.setEmptySpaceSkipping(False)
    assert r.paint()
    assert np.array_equal(r.capture(), img)
    r.cleanup()


def test_clearview_parity_medium_u16():
    w = workloads.medium(dims=(128, 128, 128), width=256, height=144)
    w.brick = 32
    vol = w.volume()
    cam = w.cameras[9]
    cv = dict(iso_value=3300.0, focus=(128, 72), size=2.0, context_scale=1.5, border_scale=3.0)
    print("case clearview medium u16", flush=True)
    r = make(vol, w, 2700.0, cv, brick=32)
    img = paint(r, w, cam)
    ref8, fl = oracle(vol, w, cam, 2700.0, cv)
    assert_parity(img, ref8, fl, "clearview medium", max_flip_frac=ISO_FLIP_FRAC)
    r.cleanup()


def test_clearview_reductions_bit_exact():
    """Focus far outside the image: every pixel has focus weight 0, so the
    image is the plain isosurface.  Size 0 (whole image in focus) and
    context scale 0: the context is fully transparent and the image is the
    plain isosurface of the focus value in the focus colour."""
    vol = nested_spheres()
    w = sphere_workload()
    cam = w.cameras[0]
    print("case clearview reductions", flush=True)
    r = make(vol, w, 100.0)
    plain = paint(r, w, cam)
    apply_cv(r, dict(iso_value=200.0, focus=(-100000, -100000), size=1.0, context_scale=0.5,
                     border_scale=2.0))
    assert np.array_equal(paint(r, w, cam), plain)
    apply_cv(r, dict(iso_value=200.0, focus=(80, 60), size=0.0, context_scale=0.0, border_scale=2.0))
    full_focus = paint(r, w, cam)
    r.setClearViewEnabled(False)
    r.setIsoValue(200.0)
    r.setIsosurfaceColor(CV_COLOR)
    focus_only = paint(r, w, cam)
    assert np.array_equal(full_focus, focus_only)
    assert not np.array_equal(full_focus, plain)
    r.cleanup()


def test_clearview_batch_matches_paint():
    vol = nested_spheres()
    w = sphere_workload()
    cv = dict(iso_value=200.0, focus=(60, 70), size=2.5, context_scale=0.6, border_scale=1.0)
    r = make(vol, w, 100.0, cv)
    paints = [paint(r, w, c) for c in w.cameras]
    views = [dict(view_pos=c.view_pos, view_dir=c.view_dir, up_dir=c.up_dir, fov=c.fov,
                  width=w.width, height=w.height) for c in w.cameras]
    for _ in range(2):
        got, _, _ = r.renderBatch(views)
        for i in range(len(views)):
            assert np.array_equal(got[i], paints[i]), f"view {i}"
    r.cleanup()


def test_clearview_setters_and_getters():
    vol = nested_spheres(32)
    w = sphere_workload(64, 48)
    r = make(vol, w, 100.0)
    assert r.supportsClearView() and r.getClearViewDisabledReason() == ""
    assert not r.getClearViewEnabled()
    assert r.getCVSize() == 5.5 and r.getCVContextScale() == 1.0 and r.getCVBorderScale() == 60.0
    apply_cv(r, dict(iso_value=150.0, focus=(7, 9), size=2.0, context_scale=0.25, border_scale=4.0,
                     color=(0.5, 0.25, 1.0)))
    assert r.getClearViewEnabled() and r.getCVIsoValue() == 150.0
    assert r.getCVSize() == 2.0 and r.getCVContextScale() == 0.25 and r.getCVBorderScale() == 4.0
    assert r.getCVColor() == (0.5, 0.25, 1.0) and r.getCVFocusPos() == (7, 9)
    for bad in (lambda: r.setCVSize(-1.0), lambda: r.setCVContextScale(float("nan")),
                lambda: r.setCVBorderScale(-0.5), lambda: r.setCVIsoValue(float("inf"))):
        with pytest.raises(TvkError):
            bad()
    r.cleanup()
