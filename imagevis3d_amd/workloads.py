"""The benchmark workloads: seeded volume, TF, cameras and render settings.

``small``  -- 64^3 uint8, 60 Gaussian blobs on a sphere + noise, one 64^3
              brick (apron 2, no LOD), 256-entry ramp TF (alpha 0 at 40 to
              1 at 220), 512x512 orthographic face-filling view, rate 1.0,
              early termination at alpha >= 0.99.
``medium`` -- 512^3 uint16, 5-octave Perlin fBm in [0,4095] + 20 spheres,
              256^3 bricks with a 2-voxel apron, 3 LOD levels, 4096-entry TF
              with three seeded Gaussian bumps, 1920x1080 perspective (60
              degree vertical FoV), 36-view orbit (10 degree azimuth steps,
              20 degree elevation, bounding sphere filling the FoV), rate 1.5,
              min/max empty-space skipping.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import transfer
from .renderer import VOL_BLOBS, VOL_PERLIN_SPHERES, generate_volume


@dataclass
class Camera:
    view_pos: tuple
    view_dir: tuple
    up_dir: tuple = (0.0, 1.0, 0.0)
    fov: float = 60.0
    ortho: bool = False
    ortho_half_height: float = 0.5


@dataclass
class Workload:
    name: str
    dims: tuple            # (nx, ny, nz)
    bits: int
    volume_kind: int
    seed: int
    brick: int
    apron: int
    n_lod: int
    tf: np.ndarray
    width: int
    height: int
    rate: float
    cameras: list = field(default_factory=list)
    term_alpha: float = 0.99
    skip: bool = True
    bg_top: tuple = (0.0, 0.0, 0.0)
    bg_bottom: tuple = (0.0, 0.0, 0.0)

    def volume(self) -> np.ndarray:
        return generate_volume(self.volume_kind, self.seed, self.dims, self.bits)


def orbit_cameras(n_views: int = 36, fov: float = 60.0, elevation_deg: float = 20.0,
                  extent=(1.0, 1.0, 1.0)) -> list:
    """Cameras on a circle around the volume centre, looking at it, at the
    distance where the bounding sphere exactly fills the vertical FoV."""
    r = 0.5 * math.sqrt(sum(e * e for e in extent))
    d = r / math.sin(math.radians(fov) / 2.0)
    el = math.radians(elevation_deg)
    cams = []
    for k in range(n_views):
        az = math.radians(10.0 * k * 36.0 / n_views)
        pos = (d * math.cos(el) * math.sin(az), d * math.sin(el), d * math.cos(el) * math.cos(az))
        cams.append(Camera(pos, tuple(-p for p in pos), (0.0, 1.0, 0.0), fov))
    return cams


def small(seed: int = 1) -> Workload:
    return Workload(
        name="small-64^3-u8-ortho512", dims=(64, 64, 64), bits=8, volume_kind=VOL_BLOBS,
        seed=seed, brick=64, apron=2, n_lod=1, tf=transfer.ramp_tf(256, 40, 220),
        width=512, height=512, rate=1.0,
        cameras=[Camera((0.0, 0.0, 1.5), (0.0, 0.0, -1.0), (0.0, 1.0, 0.0), 60.0, True, 0.5)])


def medium(seed: int = 2, n_views: int = 36, width: int = 1920, height: int = 1080,
           dims=(512, 512, 512)) -> Workload:
    return Workload(
        name="medium-512^3-u16-persp1080p-36views", dims=tuple(dims), bits=16,
        volume_kind=VOL_PERLIN_SPHERES, seed=seed, brick=256, apron=2, n_lod=3,
        tf=transfer.gaussian_bumps_tf(4096, seed, 3), width=width, height=height, rate=1.5,
        cameras=orbit_cameras(n_views, 60.0, 20.0))


def apply(r, w: Workload, cam: Camera) -> None:
    """Push a workload's settings and one camera into a renderer."""
    r.setSampleRateModifier(w.rate)
    r.setEarlyRayTermination(w.term_alpha)
    r.setEmptySpaceSkipping(w.skip)
    r.setBGColors(w.bg_top, w.bg_bottom)
    set_camera(r, cam)


def set_camera(r, cam: Camera) -> None:
    r.setViewPos(cam.view_pos)
    r.setViewDir(cam.view_dir)
    r.setUpDir(cam.up_dir)
    r.setFoV(cam.fov)
    r.setOrthoViewEnabled(cam.ortho, cam.ortho_half_height)


def skip_cell(brick: int) -> int:
    """Edge of the renderer's empty-space cells for a brick size."""
    return 16 if brick % 16 == 0 else brick


def oracle_kwargs(w: Workload, cam: Camera, lod: int = 0, level_dims=None) -> dict:
    """Keyword arguments of oracle.make_params for the same render.  The
    oracle counts skipped samples on the renderer's empty-space cells: edge
    skip_cell(brick), value range over the cell plus one voxel (apron 1)."""
    return dict(dims0=w.dims, lod=lod, view_pos=cam.view_pos, view_dir=cam.view_dir,
                up_dir=cam.up_dir, fov=cam.fov, ortho=cam.ortho,
                ortho_half_height=cam.ortho_half_height, width=w.width, height=w.height,
                rate=w.rate, term_alpha=w.term_alpha, interp=1, bg_top=w.bg_top,
                bg_bottom=w.bg_bottom, bg_alpha=1.0,
                brick_size=skip_cell(w.brick) if w.skip else 0, apron=1)


SERVING_POOL = [((64, 64, 64), 8, 64), ((64, 64, 64), 16, 64), ((256, 256, 256), 8, 128),
                ((256, 256, 256), 16, 128), ((512, 512, 512), 8, 128), ((512, 512, 512), 16, 128),
                ((2048, 2048, 2048), 16, 128)]
SERVING_SIZES = [(256, 256), (512, 512), (1920, 1080), (3840, 2160)]


def serving_pool(pool=SERVING_POOL, seed: int = 100):
    """The serving workload's volumes: device-generated fBm + spheres at
    64^3, 256^3 and 512^3 in uint8 and uint16 and one 2048^3 uint16, each on
    its own renderer with two 1D TF slots (Gaussian bumps, a ramp) and one
    2D TF slot (Gaussian lobes), sample rate 1."""
    from .renderer import VOL_FBM_SPHERES, tuvok
    rs = []
    for i, (dims, bits, brick) in enumerate(pool):
        # the volumes span the whole 8 / 16-bit range: TF entries are spread
        # over it (biasAndScaleTF, LuaScripts/BatchRenderer.lua:19-20)
        r = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, True)
        assert r.initialize(tuvok.createContext(256, 256, device=0)), r.lastError()
        r.setGeneratedVolume(VOL_FBM_SPHERES, seed + i, dims, bits, brick, 2, 1)
        n = 256 if bits == 8 else 4096
        r.set1DTransSlot(0, transfer.gaussian_bumps_tf(n, seed=seed + 100 + i))
        r.set1DTransSlot(1, transfer.ramp_tf(n, 0.45 * n, 0.95 * n))
        _, mg = r.histogram2D(256, 256)
        r.set2DTransSlot(2, transfer.gaussian_lobes_2d(256, 256, seed=seed + 200 + i), mg)
        r.setSampleRateModifier(1.0)
        rs.append(r)
    return rs


BATCHED_DIMS, BATCHED_BRICK, BATCHED_LODS = (2048, 2048, 1024), 128, 5


def batched_renderer(vol, max_grad: float, pool_bytes: int, brick: int = BATCHED_BRICK,
                     n_lod: int = BATCHED_LODS):
    """The out-of-core batched workload's renderer: `vol` (u16, fBm +
    layered slabs) paged from host memory through a device pool of
    `pool_bytes`, 4 seeded 256x256 Gaussian-lobe 2D TFs over value x
    gradient magnitude (value axis over the data range), Phong lighting,
    sample rate 1, finest LOD."""
    from .renderer import RenderMode, tuvok
    r = tuvok.renderer.new(tuvok.renderer.types.OpenGL_Raycaster, False, False, False, True)
    assert r.initialize(tuvok.createContext(64, 64, device=0)), r.lastError()
    r.setPagedFlatVolume(vol, brick, pool_bytes, n_lod=n_lod)
    r.setRenderMode(RenderMode.RM_2DTRANS)
    for k in range(4):
        r.set2DTransSlot(k, transfer.gaussian_lobes_2d(256, 256, seed=40 + k), max_grad)
    r.setLightingEnabled(True)
    r.setSampleRateModifier(1.0)
    r.setLODLimits((0, 0))
    return r


def batched_views(dims=BATCHED_DIMS, width: int = 1920, height: int = 1080, n_cams: int = 8):
    """8 orbit cameras (60 degree FoV, 25 degree elevation) x 4 TF slots."""
    ext = tuple(d / max(dims) for d in dims)
    cams = orbit_cameras(n_cams, fov=60.0, elevation_deg=25.0, extent=ext)
    return [dict(view_pos=c.view_pos, view_dir=c.view_dir, up_dir=c.up_dir, fov=c.fov,
                 width=width, height=height, tf_slot=k) for k in range(4) for c in cams]


def serving_request(rng, n_vol: int, sizes=SERVING_SIZES):
    """One random request of the serving workload: (volume, view, mode) with
    a camera looking at the volume from distance U(1.5, 2.4) (50 degree FoV),
    an image size from `sizes`, and with probability 1/2 the 2D TF (slot 2),
    else a 1D TF (slot 0 or 1)."""
    from .renderer import RenderMode
    th, ph = rng.uniform(0, 2 * math.pi), rng.uniform(-1.3, 1.3)
    d = rng.uniform(1.5, 2.4)
    eye = (d * math.cos(ph) * math.sin(th), d * math.sin(ph), d * math.cos(ph) * math.cos(th))
    w, h = sizes[rng.integers(len(sizes))]
    two_d = rng.uniform() < 0.5
    view = dict(view_pos=eye, view_dir=tuple(-x for x in eye), fov=50.0, width=w, height=h,
                tf_slot=2 if two_d else int(rng.integers(2)))
    return int(rng.integers(n_vol)), view, RenderMode.RM_2DTRANS if two_d else RenderMode.RM_1DTRANS


def view_of(cam: Camera, width: int, height: int, tf_slot: int = 0) -> dict:
    """One entry of a renderBatch / serving view list for a camera."""
    return dict(view_pos=cam.view_pos, view_dir=cam.view_dir, up_dir=cam.up_dir, fov=cam.fov,
                ortho=cam.ortho, ortho_half_height=cam.ortho_half_height, width=width,
                height=height, tf_slot=tf_slot)
