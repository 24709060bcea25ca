"""Python mirror of Tuvok's Lua renderer interface over the HIP backend.

``tuvok.renderer.new(...)`` returns a :class:`Renderer` whose methods carry
the Lua names the ImageVis3D tree calls (``LuaScripts/BatchRenderer.lua``,
``ImageVis3D/UI/RenderWindow.cpp``), so a batch script reads like
BatchRenderer.lua:

    r = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster,
                           False, False, False, False, False)
    r.loadDataset("c60.uvf")
    ctx = tuvok.createContext(640, 480, 32, 24, 8, True, False)
    r.initialize(ctx); r.resize((640, 480))
    r.setRendererTarget(tuvok.renderer.types.RT_Headless)
    r.paint()
    r.captureSingleFrame("render.png", True)
    r.cleanup(); deleteClass(r)

Error behaviour follows the reference: ``paint``, ``loadDataset`` and
``initialize`` return ``False`` on failure (the message is in
``lastError()``), setters raise :class:`TvkError` (Tuvok's LuaError).
Every call goes through the C ABI in ``include/tvk_hip.h``.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import FrameStats, MemStats, RegionCheck, Tuning, TvkError, View, ViewStats, call


class _Types:
    """tuvok.renderer.types (SettingsDlg.cpp:486-517, DebugScriptWindow.cpp:219-221)."""
    OpenGL_SBVR = 0
    OpenGL_Raycaster = 1
    DirectX_SBVR = 2
    DirectX_Raycaster = 3
    OpenGL_2DSBVR = 4
    DirectX_2DSBVR = 5
    OpenGL_GridLeaper = 6
    DirectX_GridLeaper = 7
    RT_Interactive = 0
    RT_Capture = 1
    RT_Headless = 2


class RenderMode:
    """AbstrRenderer::ERenderMode (RenderWindow.cpp:2051-2059)."""
    RM_1DTRANS = 0
    RM_2DTRANS = 1
    RM_ISOSURFACE = 2


class Interpolant:
    NearestNeighbor = 0
    Linear = 1


class BlendPrecision:
    """AbstrRenderer::EBlendPrecision (SettingsDlg.cpp:247-251)."""
    BP_8BIT = 0
    BP_16BIT = 1
    BP_32BIT = 2


class Context:
    """What tuvok.createContext returns: the HIP device the renderer uses
    (BatchRenderer/main.cpp:49-65 makes a GL context instead)."""

    def __init__(self, width: int, height: int, device: int = 0):
        self.width, self.height, self.device = width, height, device


def _f4(v: Sequence[float]):
    a = np.ascontiguousarray(np.asarray(v, np.float32).reshape(-1))
    return a, a.ctypes.data


def _mat(m) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(m, np.float32).reshape(16))
    return a


class Dataset:
    """The dataset proxy of ``renderer.getDataset()`` (ImageVis3D_Progress.cpp:104-117)."""

    def __init__(self, r: "Renderer", path: str | None):
        self._r, self._path = r, path

    def getLODLevelCount(self) -> int:
        n = C.c_uint32()
        call("tvk_get_lod_level_count", self._r._h, C.byref(n))
        return n.value

    def getDomainSize(self, lod: int = 0, ts: int = 0):
        d = (C.c_uint64 * 3)()
        call("tvk_get_domain_size", self._r._h, lod, d)
        return tuple(int(x) for x in d)

    def getBitWidth(self) -> int:
        return self._r._bit_width()

    def getBrickLayout(self, lod: int):
        g = (C.c_uint32 * 3)()
        b, a = C.c_uint32(), C.c_uint32()
        call("tvk_get_brick_layout", self._r._h, lod, g, C.byref(b), C.byref(a))
        return tuple(int(x) for x in g), b.value, a.value

    def fullpath(self) -> str | None:
        return self._path


class Renderer:
    """One ``tuvok.renderer`` instance (AbstrRenderer) on the HIP backend."""

    def __init__(self, rtype: int, pow2: bool = False, to8bit: bool = False,
                 disable_border: bool = False, bias_scale_tf: bool = False, *_extra):
        h = C.c_void_p()
        call("tvk_create", int(rtype), int(pow2), int(to8bit), int(disable_border),
             int(bias_scale_tf), C.byref(h))
        self._h = h
        self._pending_dataset: tuple | None = None
        self._path = None
        self._bits = 0
        self._size = (0, 0)
        self.lastErrorMessage = ""

    # ------------------------------------------------------------ lifecycle
    def _bool(self, name, *args) -> bool:
        rc = getattr(_lib.load(), name)(self._h, *args)
        if rc < 0:
            self.lastErrorMessage = _lib.last_error()
            return False
        return True

    def lastError(self) -> str:
        return self.lastErrorMessage

    def addShaderPath(self, path: str) -> None:
        """No shaders: the raycaster is compiled HIP code (kept for script compatibility)."""

    def loadDataset(self, path: str, verify: bool = True) -> bool:
        """Before initialize() the load is deferred, as BatchRenderer.lua does it."""
        self._path = path
        if not self._initialized():
            self._pending_dataset = ("uvf", path, verify)
            return True
        ok = self._bool("tvk_load_dataset", path.encode(), int(verify))
        if ok:
            self._bits = self._bit_width()
        return ok

    def setVolume(self, vol: np.ndarray, brick_size: int, apron: int = 2, n_lod: int = 0) -> None:
        """Flat (nz, ny, nx) uint8/uint16 volume -> bricked device pool + LOD pyramid."""
        vol = np.ascontiguousarray(vol)
        bits = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16}[vol.dtype]
        nz, ny, nx = vol.shape
        call("tvk_set_volume", self._h, vol.ctypes.data, nx, ny, nz, bits, brick_size, apron, n_lod)
        self._bits = bits

    def setGeneratedVolume(self, kind: int, seed: int, dims, bits: int, brick_size: int,
                           apron: int = 2, n_lod: int = 0) -> None:
        """setVolume with a volume generated straight into device memory."""
        nx, ny, nz = (int(d) for d in dims)
        call("tvk_set_generated_volume", self._h, int(kind), int(seed), nx, ny, nz, int(bits),
             int(brick_size), int(apron), int(n_lod))
        self._bits = int(bits)
        self._path = None

    def setPagedFlatVolume(self, vol: np.ndarray, brick_size: int, pool_bytes: int,
                           apron: int = 2, n_lod: int = 0) -> None:
        """Out-of-core volume: LOD 0 stays in this host array (kept alive by
        the renderer), bricks are paged into a pool of `pool_bytes`."""
        vol = np.ascontiguousarray(vol)
        bits = {np.dtype(np.uint8): 8, np.dtype(np.uint16): 16}[vol.dtype]
        nz, ny, nx = vol.shape
        call("tvk_set_paged_flat_volume", self._h, vol.ctypes.data, nx, ny, nz, bits, brick_size,
             apron, n_lod, int(pool_bytes))
        self._paged_keep = vol
        self._bits = bits

    def setTOCVolume(self, toc, pool_bytes: int) -> None:
        """Page a TOC volume (imagevis3d_amd.toc.TocVolume, kept open by the
        renderer) into a brick pool of `pool_bytes`: bricks of the TOC's
        payload with its overlap as apron, its LODs and brick ranges."""
        call("tvk_set_toc_volume", self._h, toc._h, int(pool_bytes))
        self._paged_keep = toc
        self._bits = toc.info["bits"]

    def setPagedVolume(self, dims, bits: int, brick_size: int, source, pool_bytes: int,
                       brick_ranges=None, apron: int = 2, n_lod: int = 0) -> None:
        """Out-of-core volume from a brick source:
        ``source(lod, bx, by, bz) -> ndarray`` of the (B+2a)^3 stored brick."""
        def fill(user, lod, bx, by, bz, slot, nbytes):
            try:
                b = np.ascontiguousarray(source(lod, bx, by, bz))
                C.memmove(slot, b.ctypes.data, min(b.nbytes, nbytes))
                return 0
            except Exception:  # noqa: BLE001 -- reported to the C side as a failure
                return -1
        cb = _lib.BRICK_SOURCE(fill)
        rng = None if brick_ranges is None else np.ascontiguousarray(brick_ranges, np.float32)
        nx, ny, nz = (int(d) for d in dims)
        call("tvk_set_paged_volume", self._h, nx, ny, nz, bits, brick_size, apron, n_lod,
             C.cast(cb, C.c_void_p), None, None if rng is None else rng.ctypes.data,
             int(pool_bytes))
        self._paged_keep = (cb, rng)
        self._bits = bits

    def _bit_width(self) -> int:
        b = C.c_uint32()
        call("tvk_get_bit_width", self._h, C.byref(b))
        return b.value

    def _initialized(self) -> bool:
        return getattr(self, "_init", False)

    def initialize(self, context: Context | int = 0) -> bool:
        dev = context.device if isinstance(context, Context) else int(context)
        ok = self._bool("tvk_initialize", dev)
        self._init = ok
        if ok and self._pending_dataset is not None:
            _, path, verify = self._pending_dataset
            self._pending_dataset = None
            ok = self._bool("tvk_load_dataset", path.encode(), int(verify))
            if ok:
                self._bits = self._bit_width()
        if ok and isinstance(context, Context):
            self.resize((context.width, context.height))
        return ok

    def resize(self, size: Sequence[int]) -> None:
        w, h = int(size[0]), int(size[1])
        call("tvk_resize", self._h, w, h)
        self._size = (w, h)

    def setRendererTarget(self, t: int) -> None:
        call("tvk_set_renderer_target", self._h, int(t))

    def getRendererTarget(self) -> int:
        t = C.c_int()
        call("tvk_get_renderer_target", self._h, C.byref(t))
        return t.value

    def cleanup(self) -> None:
        if self._h:
            _lib.load().tvk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass

    # ------------------------------------------------------------ rendering
    def paint(self) -> bool:
        return self._bool("tvk_paint")

    def checkForRedraw(self) -> bool:
        return bool(_lib.load().tvk_check_for_redraw(self._h))

    def scheduleCompleteRedraw(self) -> None:
        call("tvk_schedule_complete_redraw", self._h)

    def capture(self, preserve_transparency: bool = False) -> np.ndarray:
        """Finish the frame and return it as an (H, W, 4) uint8 array."""
        w, h = self._size
        out = np.empty((h, w, 4), np.uint8)
        call("tvk_capture_single_frame", self._h, out.ctypes.data, out.nbytes,
             int(preserve_transparency))
        return out

    def captureSingleFrame(self, filename: str, preserve_transparency: bool = False) -> bool:
        try:
            img = self.capture(preserve_transparency)
        except TvkError as e:
            self.lastErrorMessage = str(e)
            return False
        try:
            from PIL import Image
            Image.fromarray(img, "RGBA").save(filename)
        except ImportError:
            img.tofile(filename)
        return True

    def synchronize(self) -> None:
        call("tvk_synchronize", self._h)

    def frameStats(self) -> FrameStats:
        s = FrameStats()
        call("tvk_get_frame_stats", self._h, C.byref(s))
        return s

    def deviceFramebuffer(self):
        p, s = C.c_void_p(), C.c_void_p()
        call("tvk_get_device_framebuffer", self._h, C.byref(p), C.byref(s))
        return p.value, s.value

    # progress counters (RenderWindow.cpp:453-486)
    def _progress(self):
        sf, wsf, bc, wb, ml = C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        call("tvk_get_progress", self._h, C.byref(sf), C.byref(wsf), C.byref(bc), C.byref(wb),
             C.byref(ml))
        return sf.value, wsf.value, bc.value, wb.value, ml.value

    def getCurrentSubFrameCount(self) -> int:
        return self._progress()[0]

    def getWorkingSubFrame(self) -> int:
        return self._progress()[1]

    def getCurrentBrickCount(self) -> int:
        return self._progress()[2]

    def getWorkingBrick(self) -> int:
        return self._progress()[3]

    def getMinLODIndex(self) -> int:
        return self._progress()[4]

    # ------------------------------------------------------- mode and TFs
    def setRenderMode(self, mode: int) -> None:
        call("tvk_set_render_mode", self._h, int(mode))

    def set1DTrans(self, rgba: np.ndarray) -> None:
        """Upload a 1D TF (n x RGBA floats): TransferFunction1D + tuvok.gpu.changed1DTrans."""
        a = np.ascontiguousarray(np.asarray(rgba, np.float32).reshape(-1, 4))
        call("tvk_set_1d_trans", self._h, a.ctypes.data, a.shape[0])

    def set2DTrans(self, rgba: np.ndarray, max_grad: float) -> None:
        """Upload a 2D TF raster (H gradient rows x W value columns x RGBA);
        `max_grad` is the gradient magnitude of the last row (the 2D
        histogram's maximum)."""
        a = np.ascontiguousarray(np.asarray(rgba, np.float32))
        h, w = a.shape[:2]
        call("tvk_set_2d_trans", self._h, a.ctypes.data, w, h, float(max_grad))

    def setUseMIP(self, use_mip: bool) -> None:
        """Maximum intensity projection in 1D TF mode (RenderWindow.cpp:281-289)."""
        call("tvk_set_use_mip", self._h, int(bool(use_mip)))

    def getUseMIP(self) -> bool:
        v = C.c_int()
        call("tvk_get_use_mip", self._h, C.byref(v))
        return bool(v.value)

    def setMIPRotationAngle(self, degrees: float) -> None:
        """MIP views turned about the up vector (RenderWindow.cpp:1687-1703)."""
        call("tvk_set_mip_rotation_angle", self._h, float(degrees))

    def setIsoValue(self, v: float) -> None:
        call("tvk_set_iso_value", self._h, float(v))

    def setIsosurfaceColor(self, c: Sequence[float]) -> None:
        call("tvk_set_isosurface_color", self._h, *[float(x) for x in c[:3]])

    # ------------------------------------------------------------ ClearView
    def supportsClearView(self) -> bool:
        return True

    def getClearViewDisabledReason(self) -> str:
        return ""

    def setClearViewEnabled(self, on: bool) -> None:
        """ClearView in isosurface mode (RenderWindow.cpp:2348-2356)."""
        call("tvk_set_clearview_enabled", self._h, int(bool(on)))

    def setCVIsoValue(self, v: float) -> None:
        call("tvk_set_cv_iso_value", self._h, float(v))

    def setCVColor(self, c: Sequence[float]) -> None:
        call("tvk_set_cv_color", self._h, *[float(x) for x in c[:3]])

    def setCVSize(self, v: float) -> None:
        call("tvk_set_cv_size", self._h, float(v))

    def setCVContextScale(self, v: float) -> None:
        call("tvk_set_cv_context_scale", self._h, float(v))

    def setCVBorderScale(self, v: float) -> None:
        call("tvk_set_cv_border_scale", self._h, float(v))

    def setCVFocusPos(self, *args) -> None:
        """setCVFocusPos(region, (x, y)) as the reference calls it
        (RenderWindow.cpp:2359-2370), or setCVFocusPos((x, y)): the focus in
        pixels, row 0 at the top.  The region is the raycast view."""
        x, y = args[-1]
        call("tvk_set_cv_focus_pos", self._h, int(x), int(y))

    def _clearview(self) -> _lib.ClearView:
        cv = _lib.ClearView()
        call("tvk_get_clearview", self._h, C.byref(cv))
        return cv

    def getClearViewEnabled(self) -> bool:
        return bool(self._clearview().enabled)

    def getCVIsoValue(self) -> float:
        return self._clearview().iso_value

    def getCVSize(self) -> float:
        return self._clearview().size

    def getCVContextScale(self) -> float:
        return self._clearview().context_scale

    def getCVBorderScale(self) -> float:
        return self._clearview().border_scale

    def getCVColor(self):
        return tuple(self._clearview().color)

    def getCVFocusPos(self):
        return tuple(self._clearview().focus)

    # --------------------------------------------------------------- camera
    def setViewPos(self, v: Sequence[float]) -> None:
        call("tvk_set_view_pos", self._h, *[float(x) for x in v])

    def setViewDir(self, v: Sequence[float]) -> None:
        call("tvk_set_view_dir", self._h, *[float(x) for x in v])

    def setUpDir(self, v: Sequence[float]) -> None:
        call("tvk_set_up_dir", self._h, *[float(x) for x in v])

    def setFoV(self, deg: float) -> None:
        call("tvk_set_fov", self._h, float(deg))

    def setOrthoViewEnabled(self, on: bool, half_height: float = 0.5) -> None:
        call("tvk_set_ortho_view", self._h, int(on), float(half_height))

    def setRotation4x4(self, m) -> None:
        a = _mat(m)
        call("tvk_set_rotation_4x4", self._h, a.ctypes.data)

    def setTranslation4x4(self, m) -> None:
        a = _mat(m)
        call("tvk_set_translation_4x4", self._h, a.ctypes.data)

    def setRescaleFactors(self, s: Sequence[float]) -> None:
        call("tvk_set_rescale_factors", self._h, *[float(x) for x in s])

    # ------------------------------------------------------ quality and LOD
    def setSampleRateModifier(self, r: float) -> None:
        call("tvk_set_sample_rate_modifier", self._h, float(r))

    def setInterpolant(self, i: int) -> None:
        call("tvk_set_interpolant", self._h, int(i))

    def setLODLimits(self, lim: Sequence[int]) -> None:
        call("tvk_set_lod_limits", self._h, int(lim[0]), int(lim[1]))

    def setBGColors(self, top: Sequence[float], bottom: Sequence[float]) -> None:
        t, tp = _f4(top)
        b, bp = _f4(bottom)
        call("tvk_set_bg_colors", self._h, tp, bp)

    def setLightingEnabled(self, on: bool) -> None:
        call("tvk_set_lighting_enabled", self._h, int(on))

    def setLightColors(self, ambient, diffuse, specular, light_dir) -> None:
        a, ap = _f4(ambient)
        d, dp = _f4(diffuse)
        s, sp = _f4(specular)
        l, lp = _f4(light_dir)
        call("tvk_set_light_colors", self._h, ap, dp, sp, lp)

    def set1DTransSlot(self, slot: int, rgba: np.ndarray) -> None:
        """1D TF of TF slot `slot` (0 = set1DTrans) for batched views."""
        a = np.ascontiguousarray(np.asarray(rgba, np.float32).reshape(-1, 4))
        call("tvk_set_1d_trans_slot", self._h, int(slot), a.ctypes.data, a.shape[0])

    def set2DTransSlot(self, slot: int, rgba: np.ndarray, max_grad: float) -> None:
        """2D TF of TF slot `slot` (0 = set2DTrans) for batched views."""
        a = np.ascontiguousarray(np.asarray(rgba, np.float32))
        h, w = a.shape[:2]
        call("tvk_set_2d_trans_slot", self._h, int(slot), a.ctypes.data, w, h, float(max_grad))

    @staticmethod
    def _views(views, readback: bool):
        n = len(views)
        arr = (View * max(n, 1))()
        imgs = []
        for i, v in enumerate(views):
            g = v.get if isinstance(v, dict) else (lambda k, d=None, v=v: getattr(v, k, d))
            x = arr[i]
            x.view_pos[:] = [float(t) for t in g("view_pos")]
            x.view_dir[:] = [float(t) for t in g("view_dir")]
            x.up_dir[:] = [float(t) for t in g("up_dir", (0.0, 1.0, 0.0))]
            x.fov_deg = float(g("fov", 50.0))
            x.ortho = int(bool(g("ortho", False)))
            x.ortho_half_height = float(g("ortho_half_height", 0.5))
            x.width, x.height = int(g("width")), int(g("height"))
            x.tf_slot = int(g("tf_slot", 0))
            if readback:
                img = np.empty((x.height, x.width, 4), np.uint8)
                imgs.append(img)
                x.rgba = img.ctypes.data
            else:
                x.rgba = None
        return arr, imgs

    TUNE_BRICK, TUNE_CELL, TUNE_LAUNCH = 1, 2, 4

    def autotune(self, views, flags: int = 7) -> Tuning:
        """Pick brick size / empty-space cell / launch shape for these views
        (fastest summed kernel time) and leave them applied."""
        arr, _ = self._views(views, False)
        t = Tuning()
        call("tvk_autotune", self._h, arr, len(views), int(flags), C.byref(t))
        return t

    def renderBatch(self, views, readback: bool = True):
        """Render many views in one launch.  `views` holds objects with
        view_pos, view_dir, up_dir, fov, ortho, ortho_half_height, width,
        height and tf_slot (a workloads.Camera plus size/slot, or dicts).
        Returns (images or None, [ViewStats], kernel_ms)."""
        n = len(views)
        arr, imgs = self._views(views, readback)
        stats = (ViewStats * max(n, 1))()
        kms = C.c_float(0.0)
        call("tvk_render_batch", self._h, arr, n, stats, C.byref(kms))
        return (imgs if readback else None), list(stats)[:n], kms.value

    def batchFramebuffer(self, index: int) -> int:
        """Device pointer of view `index` of the last batch."""
        p = C.c_void_p()
        call("tvk_get_batch_framebuffer", self._h, int(index), C.byref(p))
        return p.value

    def setRegionLODs(self, levels) -> None:
        """Per-region LOD: one level per LOD-0 brick ((gz, gy, gx) or flat);
        None turns it off."""
        if levels is None:
            call("tvk_set_region_lods", self._h, None, 0)
            return
        a = np.ascontiguousarray(np.asarray(levels, np.uint8).ravel())
        call("tvk_set_region_lods", self._h, a.ctypes.data, a.size)

    def brickHistograms(self, lod: int = 0, bins: int = 16) -> np.ndarray:
        """(n_bricks, bins) uint32 value histograms of a level's bricks."""
        grid, _, _ = self.getDataset().getBrickLayout(lod)
        out = np.empty((grid[0] * grid[1] * grid[2], bins), np.uint32)
        call("tvk_brick_histograms", self._h, int(lod), int(bins), out.ctypes.data, out.size)
        return out

    def regionDetail(self) -> np.ndarray:
        """(n_regions, n_lod) mean |LOD-0 voxel - level-l ancestor| per region."""
        grid, _, _ = self.getDataset().getBrickLayout(0)
        nl = self.getDataset().getLODLevelCount()
        out = np.empty((grid[0] * grid[1] * grid[2], nl), np.float32)
        call("tvk_region_detail", self._h, out.ctypes.data, out.size)
        return out

    def setRegionLODCheck(self, tile_sample: int, max_rmse: float = 1.0) -> None:
        """Per-frame check of per-region frames: one ray tile in each run of
        `tile_sample` rendered at level 0, the frame's RGB RMSE estimated
        from them (1/255 units) and, above max_rmse, the frame rendered at
        level 0; 0 turns it off."""
        call("tvk_set_region_check", self._h, int(tile_sample), float(max_rmse))

    def regionLODCheck(self) -> RegionCheck:
        """The last checked frame: rmse_estimate, fell_back, check_ms."""
        c = RegionCheck()
        call("tvk_get_region_check", self._h, C.byref(c))
        return c

    def setRegionOutput(self, on: bool) -> None:
        call("tvk_set_region_output", self._h, int(bool(on)))

    def readRegionOutput(self) -> np.ndarray:
        """(H, W) int32: per pixel the LOD-0 region adding the most opacity."""
        w, h = self._size
        out = np.empty((h, w), np.int32)
        call("tvk_read_region_output", self._h, out.ctypes.data, out.size)
        return out

    def setLaunchGraphs(self, on: bool) -> None:
        call("tvk_set_launch_graphs", self._h, int(bool(on)))

    def setRayTileShape(self, tile_width: int) -> None:
        """Ray tile of one wavefront: tile_width x 64/tile_width (no image change)."""
        call("tvk_set_ray_tile_shape", self._h, int(tile_width))

    def setEmptySpaceCell(self, cell: int) -> None:
        """Empty-space cell edge for the next setVolume (no image change)."""
        call("tvk_set_empty_space_cell", self._h, int(cell))

    def setEarlyRayTermination(self, alpha: float) -> None:
        call("tvk_set_early_ray_termination", self._h, float(alpha))

    def setEmptySpaceSkipping(self, on: bool) -> None:
        call("tvk_set_empty_space_skipping", self._h, int(on))

    def setClipPlane(self, plane) -> None:
        """World-space plane (a, b, c, d): points with a*x + b*y + c*z + d > 0
        are cut away while the plane is enabled (RenderWindow.cpp:1920-1927)."""
        a = np.ascontiguousarray(np.asarray(plane, np.float32).reshape(4))
        call("tvk_set_clip_plane", self._h, a.ctypes.data)

    def getClipPlane(self) -> tuple:
        a = np.zeros(4, np.float32)
        call("tvk_get_clip_plane", self._h, a.ctypes.data)
        return tuple(float(x) for x in a)

    def enableClipPlane(self, on: bool) -> None:
        call("tvk_enable_clip_plane", self._h, int(on))

    def isClipPlaneEnabled(self) -> bool:
        v = C.c_int32(0)
        call("tvk_is_clip_plane_enabled", self._h, C.byref(v))
        return bool(v.value)

    def setBlendPrecision(self, precision: int) -> None:
        """BlendPrecision.BP_8BIT / BP_16BIT / BP_32BIT (RenderWindow.cpp:1618-1623)."""
        call("tvk_set_blend_precision", self._h, int(precision))

    def getBlendPrecision(self) -> int:
        v = C.c_int32(0)
        call("tvk_get_blend_precision", self._h, C.byref(v))
        return v.value

    def setPerfMeasures(self, minFramerate: int, lowResIntermediate: bool = False,
                        screenResDecFactor: float = 2.0, sampleDecFactor: float = 2.0,
                        LODDelay: int = 0) -> None:
        """The interactive frame-time governor (RenderWindow.cpp:1625-1640);
        minFramerate 0 turns it off."""
        call("tvk_set_perf_measures", self._h, int(minFramerate), int(lowResIntermediate),
             float(screenResDecFactor), float(sampleDecFactor), int(LODDelay))

    def readFramebuffer(self) -> np.ndarray:
        """The current subframe as (H, W, 4) uint8, without painting further."""
        w, h = self._size
        out = np.empty((h, w, 4), np.uint8)
        call("tvk_read_framebuffer", self._h, out.ctypes.data, out.nbytes)
        return out

    def setKernelTuning(self, samples_per_batch: int = 8, tile_group: int = 4,
                        xcd_split: int = 0) -> None:
        """Launch tuning picked by the autotuner (no image change)."""
        call("tvk_set_kernel_tuning", self._h, int(samples_per_batch), int(tile_group),
             int(xcd_split))

    def setWorkGroupWaves(self, waves: int) -> None:
        """Waves per work-group (4 or 8) of the raycast variants without an
        LDS TF (no image change)."""
        call("tvk_set_work_group_waves", self._h, int(waves))

    def setBrickStaging(self, on: bool) -> None:
        """1D march over LDS-staged 8^3 slot cells (no image change)."""
        call("tvk_set_brick_staging", self._h, int(bool(on)))

    def setTFGrouping(self, on: bool) -> None:
        """Gradient-mode launches walk views that differ only in their TF
        slot once per ray, compositing every view's TF (default on; no
        image change)."""
        call("tvk_set_tf_grouping", self._h, int(bool(on)))

    # --------------------------------------------------------------- data
    def getDataset(self) -> Dataset:
        return Dataset(self, self._path)

    def readLOD(self, lod: int) -> np.ndarray:
        nx, ny, nz = self.getDataset().getDomainSize(lod)
        dt = np.uint8 if self._bits == 8 else np.uint16
        out = np.empty((nz, ny, nx), dt)
        call("tvk_read_lod", self._h, lod, out.ctypes.data, out.nbytes)
        return out

    def readBrick(self, lod: int, bx: int, by: int, bz: int) -> np.ndarray:
        _, B, a = self.getDataset().getBrickLayout(lod)
        S = B + 2 * a
        dt = np.uint8 if self._bits == 8 else np.uint16
        out = np.empty((S, S, S), dt)
        call("tvk_read_brick", self._h, lod, bx, by, bz, out.ctypes.data, out.nbytes)
        return out

    def brickMinMax(self, lod: int):
        g, _, _ = self.getDataset().getBrickLayout(lod)
        n = g[0] * g[1] * g[2]
        mn = np.empty(n, np.float32)
        mx = np.empty(n, np.float32)
        call("tvk_get_brick_minmax", self._h, lod, mn.ctypes.data, mx.ctypes.data, n)
        return mn.reshape(g[2], g[1], g[0]), mx.reshape(g[2], g[1], g[0])

    def histogram1D(self) -> np.ndarray:
        h = np.empty(1 << self._bits, np.uint64)
        call("tvk_compute_histogram_1d", self._h, h.ctypes.data, h.size)
        return h

    def histogram2D(self, n_value: int = 256, n_grad: int = 256):
        h = np.empty((n_value, n_grad), np.uint64)
        mg = C.c_float()
        call("tvk_compute_histogram_2d", self._h, h.ctypes.data, n_value, n_grad, C.byref(mg))
        return h, mg.value

    # ------------------------------------------------------------- memory
    def setMaxUsableGPUMem(self, nbytes: int) -> None:
        call("tvk_set_max_usable_gpu_mem", self._h, int(nbytes))

    def setMaxUsableCPUMem(self, nbytes: int) -> None:
        call("tvk_set_max_usable_cpu_mem", self._h, int(nbytes))

    def memStats(self) -> MemStats:
        s = MemStats()
        call("tvk_get_mem_stats", self._h, C.byref(s))
        return s

    def setBrickSourceThreads(self, n: int) -> None:
        call("tvk_set_brick_source_threads", self._h, int(n))

    def pagingStats(self) -> tuple:
        """(bytes of bricks streamed in, host ms spent paging) since the volume was set."""
        b, ms = C.c_uint64(0), C.c_double(0.0)
        call("tvk_get_paging_stats", self._h, C.byref(b), C.byref(ms))
        return b.value, ms.value


def memman_stats(device: int = 0):
    """The device-wide memory manager's stats (MemManStats)."""
    from ._lib import MemManStats
    s = MemManStats()
    call("tvk_get_memman_stats", int(device), C.byref(s))
    return s


class _RendererNS:
    types = _Types

    @staticmethod
    def new(rtype: int, *flags) -> Renderer:
        return Renderer(rtype, *flags)


class tuvok:  # noqa: N801 -- the Lua namespace name
    """The `tuvok.*` names BatchRenderer.lua uses."""
    renderer = _RendererNS

    @staticmethod
    def createContext(width: int, height: int, color_bits: int = 32, depth_bits: int = 24,
                      stencil_bits: int = 8, double_buffer: bool = True, visible: bool = False,
                      device: int = 0) -> Context:
        return Context(width, height, device)


def deleteClass(r: Renderer) -> None:  # noqa: N802 -- Lua name
    r.cleanup()


def generate_volume(kind: int, seed: int, dims: Sequence[int], bits: int) -> np.ndarray:
    """Seeded synthetic volume from the library's host generators, (nz, ny, nx)."""
    nx, ny, nz = (int(d) for d in dims)
    out = np.empty((nz, ny, nx), np.uint8 if bits == 8 else np.uint16)
    call("tvk_generate_volume", int(kind), int(seed), nx, ny, nz, int(bits), out.ctypes.data,
         out.nbytes)
    return out


def brick_slot_bytes(brick: int, apron: int = 2, bits: int = 16) -> int:
    """Device bytes of one brick-pool slot (tvk_brick_slot_bytes): pools of
    pool_bytes hold pool_bytes // this many bricks."""
    n = int(_lib.load().tvk_brick_slot_bytes(int(brick), int(apron), int(bits)))
    if n == 0:
        raise ValueError("invalid brick size, apron or bit width")
    return n


def generate_volume_gpu(kind: int, seed: int, dims: Sequence[int], bits: int,
                        device: int = 0, out: np.ndarray | None = None) -> np.ndarray:
    """Device-generated synthetic volume (kinds VOL_FBM_PLANES, VOL_NOISE_TUBES,
    VOL_FBM_SPHERES) copied into host memory, (nz, ny, nx)."""
    nx, ny, nz = (int(d) for d in dims)
    if out is None:
        out = np.empty((nz, ny, nx), np.uint8 if bits == 8 else np.uint16)
    call("tvk_generate_volume_gpu", int(device), int(kind), int(seed), nx, ny, nz, int(bits),
         out.ctypes.data, out.nbytes)
    return out


VOL_CONST, VOL_RANDOM, VOL_SPHERE, VOL_BLOBS, VOL_PERLIN_SPHERES = 0, 1, 2, 3, 4
VOL_FBM_PLANES, VOL_NOISE_TUBES, VOL_FBM_SPHERES = 5, 6, 7
