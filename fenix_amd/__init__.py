"""fenix_amd — MI355X (gfx950) implementation of fenix's per-step combustion-volume
pipeline behind a C ABI (include/fenix.h, libfenix.so) with a Python host that
mirrors the reference's CombustionVolumeSimulation / VolumeRaymarchPass API."""
from . import abi
from .emitters import (BurstEmitter, IgniterEmitter, ScalarEmitter, VelocityEmitter, pack_emitter_sources)
from .render import (RenderLight, VolumeRaymarchPass, get_volume_world_bounds, orbit_camera_snapshot)
from .simulation import (CombustionVolumeSimulation, VolumeResolution, createCombustionVolumeSimulation, make_desc)

__all__ = [
    "abi", "BurstEmitter", "IgniterEmitter", "ScalarEmitter", "VelocityEmitter", "pack_emitter_sources",
    "RenderLight", "VolumeRaymarchPass", "get_volume_world_bounds", "orbit_camera_snapshot",
    "CombustionVolumeSimulation", "VolumeResolution", "createCombustionVolumeSimulation", "make_desc",
]
