"""Raymarch host: orbit-camera snapshot, volume bounds and VolumeRaymarchPass.

Mirrors src/engine/scene/camera/createOrbitCameraController.ts:30-41, 120-162
(snapshot basis), src/engine/render/volumetrics/volumeWorldBounds.ts:14-31 and
src/engine/render/passes/createVolumeRaymarchPass.ts:8-303 (host packing,
light sanitising, parameter clamps). The shader itself runs in libfenix.
"""
import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np

from . import abi

MAX_RENDER_LIGHTS = 3  # renderLight.ts:10
DEFAULT_LIGHT_DIRECTION = (-0.34, 0.88, 0.31)  # createVolumeRaymarchPass.ts:36


@dataclass
class RenderLight:  # src/engine/render/lighting/renderLight.ts:3-9
    type: str = "directional"
    direction: Tuple[float, float, float] = DEFAULT_LIGHT_DIRECTION
    position: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    color: Tuple[float, float, float] = (1.0, 0.95, 0.88)
    intensity: float = 1.25


DEFAULT_LIGHTS = (RenderLight(),)  # createVolumeRaymarchPass.ts:38-46


@dataclass
class OrbitCameraSnapshot:
    position: Tuple[float, float, float]
    target: Tuple[float, float, float]
    forward: Tuple[float, float, float]
    right: Tuple[float, float, float]
    up: Tuple[float, float, float]
    aspect: float
    tanHalfFovY: float


def _normalize(v):
    n = math.hypot(*v)
    if n <= 1e-6:  # vector3.ts:51-59
        return (0.0, 0.0, 0.0)
    return tuple(c / n for c in v)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def orbit_camera_snapshot(yaw=-0.62, pitch=0.31, radius=40.0, target=(0.0, 16.0, 0.0), aspect=1.0) -> OrbitCameraSnapshot:
    """getSnapshot() of the orbit camera (createOrbitCameraController.ts:120-162)."""
    cp = math.cos(pitch)
    position = (target[0] + math.cos(yaw) * cp * radius, target[1] + math.sin(pitch) * radius,
                target[2] + math.sin(yaw) * cp * radius)
    forward = _normalize(tuple(t - p for t, p in zip(target, position)))
    right = _normalize(_cross(forward, (0.0, 1.0, 0.0)))
    up = _normalize(_cross(right, forward))
    return OrbitCameraSnapshot(position, tuple(target), forward, right, up, aspect if aspect > 0 else 1.0,
                               math.tan(48 * math.pi / 180))


def get_volume_world_bounds(resolution):
    """getVolumeWorldBounds (volumeWorldBounds.ts:14-31): (center, halfExtents)."""
    w, h, d = resolution
    scale = 13.5
    m = max(w, d)
    half = (2.05 * scale * (w / m), 2.05 * scale * (h / m), 2.05 * scale * (d / m))
    return (0.0, half[1] - 0.25, 0.0), half


def default_step_count(resolution) -> int:
    """createVolumeRaymarchPass.ts:86-88."""
    w, h, d = resolution
    n = w * h * d
    hi = 720 if n >= 4_000_000 else 560
    return int(max(180, min(hi, _js_round(max(w, h, d) * 1.5))))


def _js_round(v):
    return int(math.floor(v + 0.5))


def sanitize_render_lights(lights: Optional[Sequence[RenderLight]]):
    """createVolumeRaymarchPass.ts:48-76."""
    src = DEFAULT_LIGHTS if lights is None else lights
    out = []
    for light in list(src)[:MAX_RENDER_LIGHTS]:
        n = math.hypot(*light.direction)
        direction = DEFAULT_LIGHT_DIRECTION if n <= 0.0001 else tuple(c / n for c in light.direction)
        out.append(RenderLight(light.type, direction, tuple(light.position), tuple(max(c, 0.0) for c in light.color),
                               max(light.intensity, 0.0)))
    return out


def camera_struct(cam: OrbitCameraSnapshot) -> abi.Camera:
    c = abi.Camera()
    c.position[:] = cam.position
    c.target[:] = cam.target
    c.forward[:] = cam.forward
    c.right[:] = cam.right
    c.up[:] = cam.up
    c.aspect = cam.aspect
    c.tan_half_fov_y = cam.tanHalfFovY
    return c


class VolumeRaymarchPass:
    """VolumeRaymarchPass (createVolumeRaymarchPass.ts:15-29) rendering into rgba16f."""

    DISPLAY_MODES = {"temperature": 0, "density": 1, "fuel": 2}  # RP:194

    def __init__(self, resolution):
        self.resolution = tuple(resolution)
        self.stepCount = default_step_count(self.resolution)
        self.lights = sanitize_render_lights(None)
        self.scatteringForward = 0.32
        self.scatteringBack = -0.18

    def setRenderParams(self, stepCount=None, lights=None, scatteringForward=None, scatteringBack=None):
        """RP:284-297."""
        if stepCount is not None:
            self.stepCount = max(32, _js_round(stepCount))
        if lights is not None:
            self.lights = sanitize_render_lights(lights)
        if scatteringForward is not None:
            self.scatteringForward = min(max(scatteringForward, 0.0), 0.98)
        if scatteringBack is not None:
            self.scatteringBack = min(max(scatteringBack, -0.98), 0.0)

    def params_struct(self, displayMode="temperature", elapsedSeconds=0.0) -> abi.RenderParams:
        p = abi.RenderParams()
        p.step_count = self.stepCount
        p.light_count = len(self.lights)
        for i, l in enumerate(self.lights):
            p.lights[i].type = 1 if l.type == "point" else 0
            p.lights[i].direction[:] = l.direction
            p.lights[i].position[:] = l.position
            p.lights[i].color[:] = l.color
            p.lights[i].intensity = l.intensity
        p.scattering_forward = self.scatteringForward
        p.scattering_back = self.scatteringBack
        p.display_mode = self.DISPLAY_MODES[displayMode]
        p.elapsed_s = elapsedSeconds
        return p

    def render(self, sim, camera: OrbitCameraSnapshot, displayMode="temperature", renderWidth=256, renderHeight=256,
               elapsedSeconds=0.0) -> np.ndarray:
        """render(encoder, view, buffers, camera, displayMode, w, h, elapsed) -> float16[h, w, 4]."""
        out = np.zeros((renderHeight, renderWidth, 4), dtype=np.float16)
        cam = camera_struct(camera)
        p = self.params_struct(displayMode, elapsedSeconds)
        dst = out.ctypes.data_as(C.POINTER(C.c_uint16))
        if hasattr(sim, "raymarch_call"):  # a scene of a brick-queue batch
            sim.raymarch_call("fx_scene_raymarch", cam, p, renderWidth, renderHeight, dst)
        else:
            sim._check(sim._lib.fx_raymarch(sim.handle, C.byref(cam), C.byref(p), renderWidth, renderHeight, dst))
        return out

    def render_device(self, sim, camera, device_ptr, displayMode="temperature", renderWidth=256, renderHeight=256,
                      elapsedSeconds=0.0):
        cam = camera_struct(camera)
        p = self.params_struct(displayMode, elapsedSeconds)
        if hasattr(sim, "raymarch_call"):  # a scene of a brick-queue batch
            sim.raymarch_call("fx_scene_raymarch_device", cam, p, renderWidth, renderHeight, C.c_void_p(device_ptr))
        else:
            sim._check(sim._lib.fx_raymarch_device(sim.handle, C.byref(cam), C.byref(p), renderWidth, renderHeight,
                                                   C.c_void_p(device_ptr)))


class BloomPass:
    """BloomPass (src/engine/render/passes/createBloomPass.ts:26-146) over libfenix:
    the raymarch's RGBA float16 image -> the canvas image (RGBA8 unorm)."""

    def __init__(self, lib=None):
        self._lib = lib or abi.load()

    @staticmethod
    def params_struct(enabled=True, threshold=0.6, strength=0.9, radius=0.4) -> abi.BloomParams:
        p = abi.BloomParams()
        p.enabled, p.threshold, p.strength, p.radius = int(bool(enabled)), threshold, strength, radius
        return p

    def render(self, sim, src_rgba_f16: np.ndarray, params: Optional[abi.BloomParams] = None) -> np.ndarray:
        """Host image in, uint8[h, w, 4] out (synchronous)."""
        src = np.ascontiguousarray(np.asarray(src_rgba_f16, dtype=np.float16))
        h, w = src.shape[:2]
        out = np.zeros((h, w, 4), dtype=np.uint8)
        p = params or self.params_struct()
        rc = self._lib.fx_bloom(sim._h, C.byref(p), w, h, src.ctypes.data_as(C.POINTER(C.c_uint16)),
                                out.ctypes.data_as(C.POINTER(C.c_uint8)))
        abi.check(self._lib, sim._h, rc)
        return out

    def render_device(self, sim, src_ptr, dst_ptr, width, height, params: Optional[abi.BloomParams] = None) -> None:
        """Device images (asynchronous, on the render stream after the raymarch)."""
        p = params or self.params_struct()
        rc = self._lib.fx_bloom_device(sim._h, C.byref(p), int(width), int(height), C.c_void_p(src_ptr), C.c_void_p(dst_ptr))
        abi.check(self._lib, sim._h, rc)
