"""Host mirror of the reference's CombustionVolumeSimulation over libfenix.

Reference: src/engine/simulation/runtime/createCombustionVolumeSimulation.ts
(ORCH below) and combustion-volume-simulation/types.ts (CVS/types.ts). Method
names, argument meaning, defaults and clamps follow the reference; GPU buffers
are replaced by device pointers owned by the library plus readback helpers.
"""
import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np

from . import abi
from .emitters import EmitterSource, pack_emitter_sources

_ADVECTION_MODES = {"maccormack": abi.FX_ADVECT_MACCORMACK, "semi-lagrangian": abi.FX_ADVECT_SEMI_LAGRANGIAN}
_ADVECTION_NAMES = {v: k for k, v in _ADVECTION_MODES.items()}

# SimulationRuntimeParams keys (CVS/types.ts:46-58) -> (struct field, present bit)
_RUNTIME_KEYS = {
    "wind": ("wind", abi.FX_RP_WIND),
    "windStrength": ("wind_strength", abi.FX_RP_WIND_STRENGTH),
    "gravity": ("gravity", abi.FX_RP_GRAVITY),
    "gravityStrength": ("gravity_strength", abi.FX_RP_GRAVITY_STRENGTH),
    "buoyancy": ("buoyancy", abi.FX_RP_BUOYANCY),
    "vorticityStrength": ("vorticity_strength", abi.FX_RP_VORTICITY_STRENGTH),
    "vorticityConstantMask": ("vorticity_constant_mask", abi.FX_RP_VORTICITY_CONSTANT_MASK),
    "vorticityVelocityMask": ("vorticity_velocity_mask", abi.FX_RP_VORTICITY_VELOCITY_MASK),
    "vorticityHeatMask": ("vorticity_heat_mask", abi.FX_RP_VORTICITY_HEAT_MASK),
    "vorticityDensityMask": ("vorticity_density_mask", abi.FX_RP_VORTICITY_DENSITY_MASK),
    "worldSize": ("world_size", abi.FX_RP_WORLD_SIZE),
}
# SimulationQualitySettings keys (CVS/types.ts:6-14)
_QUALITY_KEYS = {
    "pressureInterval": "pressure_interval",
    "vorticityInterval": "vorticity_interval",
    "finePreIterations": "fine_pre_iterations",
    "finePostIterations": "fine_post_iterations",
    "midPreIterations": "mid_pre_iterations",
    "midPostIterations": "mid_post_iterations",
    "coarseIterations": "coarse_iterations",
}


@dataclass(frozen=True)
class VolumeResolution:  # src/engine/simulation/common/volumeResolution.ts:1-5
    width: int
    height: int
    depth: int

    @property
    def voxel_count(self):
        return self.width * self.height * self.depth


def make_desc(resolution=(128, 256, 128), **kw) -> abi.CreateDesc:
    """fx_create_desc with the reference defaults (ORCH:54, 194-204), overridable by
    keyword: any CreateDesc field, plus `quality=dict`, `runtime=dict`, `vcycle=dict`."""
    d = abi.CreateDesc()
    d.width, d.height, d.depth = resolution
    d.advection_mode = abi.FX_ADVECT_MACCORMACK
    rt = d.runtime
    rt.wind[:] = (0.0, 0.0, 0.0)
    rt.wind_strength = 0.0
    rt.gravity[:] = (0.0, -1.0, 0.0)
    rt.gravity_strength = 0.45
    rt.buoyancy = 3.6
    rt.vorticity_strength = 2.15
    rt.vorticity_constant_mask = 1.0
    rt.vorticity_velocity_mask = 0.4
    rt.vorticity_heat_mask = 0.8
    rt.vorticity_density_mask = 0.35
    rt.world_size = 10.0
    q = d.quality
    q.pressure_interval, q.vorticity_interval = 1, 1
    q.fine_pre_iterations, q.fine_post_iterations = 2, 3
    q.mid_pre_iterations, q.mid_post_iterations = 2, 3
    q.coarse_iterations = 10
    d.domain_mode = abi.FX_DOMAIN_DENSE
    d.activation_rule = abi.FX_ACTIVATION_REFERENCE
    d.solver = abi.FX_SOLVER_REFERENCE
    d.precision = abi.FX_PRECISION_FP32
    v = d.vcycle
    v.levels, v.pre_sweeps, v.post_sweeps, v.cycles, v.coarse_sweeps, v.omega = 4, 2, 2, 4, 16, 2.0 / 3.0
    d.ignition_lo, d.ignition_hi, d.burn_rate = 0.1, 0.55, 1.0
    d.activation_density, d.activation_speed = 1e-4, 1e-3
    d.debug_fields = -1
    d.device = 0
    for key, value in kw.items():
        if key == "quality":
            for k, val in value.items():
                setattr(d.quality, _QUALITY_KEYS.get(k, k), int(val))
        elif key == "runtime":
            for k, val in value.items():
                name = _RUNTIME_KEYS[k][0] if k in _RUNTIME_KEYS else k
                if name in ("wind", "gravity"):
                    getattr(d.runtime, name)[:] = tuple(val)
                else:
                    setattr(d.runtime, name, float(val))
        elif key == "vcycle":
            for k, val in value.items():
                setattr(d.vcycle, k, val)
        else:
            setattr(d, key, value)
    return d


class DeviceField:
    """A field of the simulation: borrowed device pointer + readback (row-major, shape (D, H, W))."""

    def __init__(self, sim, field_id):
        self._sim = sim
        self.field_id = field_id

    @property
    def ptr(self):
        return self._sim._lib.fx_field_device_ptr(self._sim._h, self.field_id)

    def read(self):
        return self._sim.read_field(self.field_id)


class CombustionVolumeSimulation:
    """CombustionVolumeSimulation (CVS/types.ts:62-76) on one MI355X."""

    def __init__(self, desc: abi.CreateDesc, lib=None, borrowed_handle=None):
        self._lib = lib or abi.load()
        self._owned = borrowed_handle is None
        if borrowed_handle is not None:  # a scene of a CombustionVolumeBatch
            self._h = C.c_void_p(borrowed_handle)
        else:
            self._h = C.c_void_p()
            rc = self._lib.fx_create(C.byref(desc), C.byref(self._h))
            if rc != abi.FX_OK:
                msg = self._lib.fx_last_error(None)
                raise abi.FenixError(rc, msg.decode() if msg else "")
        self.desc = desc
        self.resolution = VolumeResolution(desc.width, desc.height, desc.depth)
        self.brick_counts = tuple((n + 7) // 8 for n in (desc.width, desc.height, desc.depth))
        self.brick_count = self.brick_counts[0] * self.brick_counts[1] * self.brick_counts[2]

    # -- internals -----------------------------------------------------------
    def _check(self, rc):
        return abi.check(self._lib, self._h, rc)

    @property
    def handle(self):
        return self._h

    # -- CVS/types.ts:64 -----------------------------------------------------
    def step(self, encoder, elapsedSeconds: float, stepSeconds: float) -> None:
        """Records one step on the simulation stream (ORCH:230-305). `encoder` is
        accepted for signature parity and ignored: work is stream-ordered."""
        self._check(self._lib.fx_step(self._h, float(elapsedSeconds), float(stepSeconds)))

    def synchronize(self):
        self._check(self._lib.fx_synchronize(self._h))

    def getRenderBuffers(self):  # ORCH:306-312
        return {
            "density": DeviceField(self, abi.FIELDS["density"]),
            "temperature": DeviceField(self, abi.FIELDS["temperature"]),
            "fuel": DeviceField(self, abi.FIELDS["fuel"]),
            "reaction": DeviceField(self, abi.FIELDS["reaction"]),
            "activeBrickFlags": self.read_flags,
            "activeBrickInfo": self.brick_info(),
        }

    def getDebugBuffers(self):  # ORCH:313-326
        """The render buffers plus the debug fields; like the reference, the call also requests
        the debug pass for the next step (debugRequested, ORCH:316), whatever the grid size."""
        self._check(self._lib.fx_request_debug_fields(self._h))
        bufs = self.getRenderBuffers()
        for name in ("velocity_magnitude", "divergence", "pressure", "vorticity_magnitude", "confinement_magnitude"):
            bufs[name] = DeviceField(self, abi.FIELDS[name])
        return bufs

    def getScalarAdvectionMode(self) -> str:  # ORCH:327-329
        out = C.c_int32()
        self._check(self._lib.fx_get_advection_mode(self._h, C.byref(out)))
        return _ADVECTION_NAMES[out.value]

    def setScalarAdvectionMode(self, mode: str) -> None:  # ORCH:330-337
        self._check(self._lib.fx_set_advection_mode(self._h, _ADVECTION_MODES[mode]))

    def getRuntimeParams(self) -> Dict:  # ORCH:338-352
        rp = abi.RuntimeParams()
        self._check(self._lib.fx_get_runtime_params(self._h, C.byref(rp)))
        out = {}
        for key, (name, _) in _RUNTIME_KEYS.items():
            val = getattr(rp, name)
            out[key] = tuple(val) if name in ("wind", "gravity") else float(val)
        return out

    def setRuntimeParams(self, params: Dict) -> None:  # ORCH:353-365 (Partial<...>)
        rp = abi.RuntimeParams()
        mask = 0
        for key, value in params.items():
            if value is None:
                continue
            name, bit = _RUNTIME_KEYS[key]
            if name in ("wind", "gravity"):
                getattr(rp, name)[:] = tuple(float(v) for v in value)
            else:
                setattr(rp, name, float(value))
            mask |= bit
        self._check(self._lib.fx_set_runtime_params(self._h, C.byref(rp), mask))

    def getQualitySettings(self) -> Dict:  # ORCH:366-368
        q = abi.Quality()
        self._check(self._lib.fx_get_quality(self._h, C.byref(q)))
        return {k: int(getattr(q, v)) for k, v in _QUALITY_KEYS.items()}

    def setQualitySettings(self, settings: Dict) -> None:  # ORCH:369-377 (clamped, CVS/types.ts:30-44)
        merged = self.getQualitySettings()
        merged.update(settings)
        q = abi.Quality()
        for k, v in _QUALITY_KEYS.items():
            # Math.round semantics are applied in the library (floor(v + 0.5)).
            setattr(q, v, int(np.floor(float(merged[k]) + 0.5)))
        self._check(self._lib.fx_set_quality(self._h, C.byref(q)))

    def updateSources(self, sources: Sequence[EmitterSource]) -> None:  # ORCH:378-380
        self.update_packed_sources(pack_emitter_sources(sources), len(sources))

    def update_packed_sources(self, packed: np.ndarray, count: int) -> None:
        packed = np.ascontiguousarray(packed, dtype=np.float32)
        ptr = packed.ctypes.data_as(C.POINTER(C.c_float)) if count else None
        self._check(self._lib.fx_update_sources(self._h, ptr, int(count)))

    def reset(self) -> None:  # ORCH:381-383
        self._check(self._lib.fx_reset(self._h))

    def dispose(self) -> None:  # ORCH:384-400
        if self._h and self._owned:
            self._lib.fx_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass

    # -- readback / test helpers ----------------------------------------------
    def read_field(self, field) -> np.ndarray:
        fid = abi.FIELDS[field] if isinstance(field, str) else int(field)
        r = self.resolution
        out = np.empty((r.depth, r.height, r.width), dtype=np.float32)
        self._check(self._lib.fx_read_field(self._h, fid, out.ctypes.data_as(C.POINTER(C.c_float)), out.nbytes))
        return out

    def write_field(self, field, values: np.ndarray) -> None:
        fid = abi.FIELDS[field] if isinstance(field, str) else int(field)
        r = self.resolution
        arr = np.ascontiguousarray(values, dtype=np.float32).reshape(r.depth, r.height, r.width)
        self._check(self._lib.fx_write_field(self._h, fid, arr.ctypes.data_as(C.POINTER(C.c_float)), arr.nbytes))

    def read_flags(self, which=abi.FX_FLAGS_RENDER) -> np.ndarray:
        out = np.empty(self.brick_count, dtype=np.uint32)
        self._check(self._lib.fx_read_flags(self._h, which, out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size))
        return out

    def read_active_list(self) -> np.ndarray:
        out = np.empty(self.brick_count, dtype=np.uint32)
        n = C.c_uint32()
        self._check(self._lib.fx_read_active_list(self._h, out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size, C.byref(n)))
        return out[: n.value].copy()

    def read_page_table(self) -> np.ndarray:
        out = np.empty(self.brick_count, dtype=np.int32)
        self._check(self._lib.fx_read_page_table(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), out.size))
        return out

    def level_dims(self, level):
        d = (C.c_uint32 * 3)()
        self._check(self._lib.fx_level_dims(self._h, level, d))
        return tuple(d)

    def read_level(self, level, which) -> np.ndarray:
        w, h, dd = self.level_dims(level)
        out = np.empty((dd, h, w), dtype=np.float32)
        self._check(self._lib.fx_read_level(self._h, level, which, out.ctypes.data_as(C.POINTER(C.c_float)), out.nbytes))
        return out

    def solver_residuals(self) -> np.ndarray:
        buf = (C.c_double * 64)()
        n = C.c_uint32()
        self._check(self._lib.fx_solver_residuals(self._h, buf, 64, C.byref(n)))
        return np.array(buf[: min(n.value, 64)], dtype=np.float64)

    def active_brick_count(self) -> int:
        n = C.c_uint32()
        self._check(self._lib.fx_active_brick_count(self._h, C.byref(n)))
        return n.value

    def run_pass(self, name_or_id, params: abi.SimParams) -> None:
        pid = abi.PASSES[name_or_id] if isinstance(name_or_id, str) else int(name_or_id)
        self._check(self._lib.fx_run_pass(self._h, pid, C.byref(params)))

    def set_vcycle_plan(self, down_split: int, up_split: int) -> None:
        """Per-level V-cycle kernel plan (bit l: split legs at level l); results are plan-independent."""
        self._check(self._lib.fx_set_vcycle_plan(self._h, int(down_split), int(up_split)))

    def vcycle_plan(self):
        d, u = C.c_uint32(), C.c_uint32()
        self._check(self._lib.fx_get_vcycle_plan(self._h, C.byref(d), C.byref(u)))
        return d.value, u.value

    def autotune_vcycle(self, repeats: int = 4):
        """Times the per-level leg plans on the current state and keeps the fastest; returns (down, up)."""
        d, u = C.c_uint32(), C.c_uint32()
        self._check(self._lib.fx_autotune_vcycle(self._h, int(repeats), C.byref(d), C.byref(u)))
        return d.value, u.value

    def time_fine_sweep(self, repeats: int = 50):
        """(mean ms per launch, voxels per launch) of the finest level's pressure sweep on the current state."""
        ms, vox = C.c_double(), C.c_uint64()
        self._check(self._lib.fx_time_fine_sweep(self._h, int(repeats), C.byref(ms), C.byref(vox)))
        return ms.value, vox.value

    def set_stage_timing(self, enable: bool) -> None:
        self._check(self._lib.fx_set_stage_timing(self._h, 1 if enable else 0))

    def stage_times(self):
        """{stage name: (total ms, calls)} since timing was enabled."""
        ms = (C.c_double * abi.FX_TIME_COUNT)()
        calls = (C.c_uint32 * abi.FX_TIME_COUNT)()
        self._check(self._lib.fx_read_stage_times(self._h, ms, calls, abi.FX_TIME_COUNT))
        return {abi.TIME_NAMES[i]: (ms[i], calls[i]) for i in range(abi.FX_TIME_COUNT) if calls[i]}

    def brick_info(self):
        """packSparseBrickLayout (src/engine/simulation/sparse/createSparseBrickLayout.ts:28-39)."""
        bx, by, bz = self.brick_counts
        return np.array([bx, by, bz, self.brick_count, 8, 0, 0, 0], dtype=np.uint32)


class CombustionVolumeBatch:
    """Many independent scenes stepped together (fx_batch_*): each scene keeps its
    own streams and step graphs, so the GPU runs them concurrently. `scenes[i]` is a
    CombustionVolumeSimulation over the borrowed handle of scene i."""

    def __init__(self, descs: Sequence[abi.CreateDesc], lib=None):
        self._lib = lib or abi.load()
        arr = (abi.CreateDesc * len(descs))(*descs)
        self._h = C.c_void_p()
        rc = self._lib.fx_batch_create(arr, len(descs), C.byref(self._h))
        if rc != abi.FX_OK:
            msg = self._lib.fx_last_error(None)
            raise abi.FenixError(rc, msg.decode() if msg else "")
        self.scenes = [CombustionVolumeSimulation(d, self._lib, borrowed_handle=self._lib.fx_batch_scene(self._h, i))
                       for i, d in enumerate(descs)]

    def __len__(self):
        return len(self.scenes)

    def step(self, elapsedSeconds: float, stepSeconds: float) -> None:
        rc = self._lib.fx_batch_step(self._h, float(elapsedSeconds), float(stepSeconds))
        if rc != abi.FX_OK:
            for sim in self.scenes:  # the failing scene carries the message
                msg = self._lib.fx_last_error(sim.handle)
                if msg:
                    raise abi.FenixError(rc, msg.decode())
            raise abi.FenixError(rc, "batch step failed")

    def synchronize(self) -> None:
        rc = self._lib.fx_batch_synchronize(self._h)
        if rc != abi.FX_OK:
            raise abi.FenixError(rc, "batch synchronize failed")

    def dispose(self) -> None:
        if self._h:
            for sim in self.scenes:
                sim.dispose()
            self._lib.fx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass


class QueueScene:
    """Scene `index` of a CombustionVolumeQueueBatch: the per-scene calls (fx_scene_*) with the
    method names of CombustionVolumeSimulation."""

    def __init__(self, batch, index: int, desc: abi.CreateDesc):
        self._batch = batch
        self.index = index
        self.desc = desc
        self.resolution = VolumeResolution(desc.width, desc.height, desc.depth)
        self.brick_counts = tuple((n + 7) // 8 for n in (desc.width, desc.height, desc.depth))
        self.brick_count = self.brick_counts[0] * self.brick_counts[1] * self.brick_counts[2]

    def _call(self, name, *args):
        b = self._batch
        return b._check(getattr(b._lib, name)(b._h, self.index, *args))

    def updateSources(self, sources: Sequence[EmitterSource]) -> None:  # ORCH:378-380
        self.update_packed_sources(pack_emitter_sources(sources), len(sources))

    def update_packed_sources(self, packed: np.ndarray, count: int) -> None:
        packed = np.ascontiguousarray(packed, dtype=np.float32)
        self._call("fx_scene_update_sources", packed.ctypes.data_as(C.POINTER(C.c_float)) if count else None, int(count))

    def getRuntimeParams(self) -> Dict:  # ORCH:338-352
        rp = abi.RuntimeParams()
        self._call("fx_scene_get_runtime_params", C.byref(rp))
        return {key: (tuple(getattr(rp, name)) if name in ("wind", "gravity") else float(getattr(rp, name)))
                for key, (name, _) in _RUNTIME_KEYS.items()}

    def setRuntimeParams(self, params: Dict) -> None:  # ORCH:353-365 (Partial<...>)
        rp = abi.RuntimeParams()
        mask = 0
        for key, value in params.items():
            if value is None:
                continue
            name, bit = _RUNTIME_KEYS[key]
            if name in ("wind", "gravity"):
                getattr(rp, name)[:] = tuple(float(v) for v in value)
            else:
                setattr(rp, name, float(value))
            mask |= bit
        self._call("fx_scene_set_runtime_params", C.byref(rp), mask)

    def read_field(self, field) -> np.ndarray:
        fid = abi.FIELDS[field] if isinstance(field, str) else int(field)
        r = self.resolution
        out = np.empty((r.depth, r.height, r.width), dtype=np.float32)
        self._call("fx_scene_read_field", fid, out.ctypes.data_as(C.POINTER(C.c_float)), out.nbytes)
        return out

    def write_field(self, field, values: np.ndarray) -> None:
        fid = abi.FIELDS[field] if isinstance(field, str) else int(field)
        r = self.resolution
        arr = np.ascontiguousarray(values, dtype=np.float32).reshape(r.depth, r.height, r.width)
        self._call("fx_scene_write_field", fid, arr.ctypes.data_as(C.POINTER(C.c_float)), arr.nbytes)

    def read_flags(self, which=abi.FX_FLAGS_RENDER) -> np.ndarray:
        out = np.empty(self.brick_count, dtype=np.uint32)
        self._call("fx_scene_read_flags", which, out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size)
        return out

    def read_active_list(self) -> np.ndarray:
        out = np.empty(self.brick_count, dtype=np.uint32)
        n = C.c_uint32()
        self._call("fx_scene_read_active_list", out.ctypes.data_as(C.POINTER(C.c_uint32)), out.size, C.byref(n))
        return out[: n.value].copy()

    def read_page_table(self) -> np.ndarray:
        out = np.empty(self.brick_count, dtype=np.int32)
        self._call("fx_scene_read_page_table", out.ctypes.data_as(C.POINTER(C.c_int32)), out.size)
        return out

    def active_brick_count(self) -> int:
        n = C.c_uint32()
        self._call("fx_scene_active_brick_count", C.byref(n))
        return n.value

    def raymarch_call(self, name, cam, params, width, height, dst):
        """fx_scene_raymarch / fx_scene_raymarch_device of this scene (VolumeRaymarchPass uses it)."""
        self._call(name, C.byref(cam), C.byref(params), int(width), int(height), dst)


class CombustionVolumeQueueBatch:
    """A brick-queue batch (fx_create_scenes): scenes of any sizes whose bricks share every
    level's storage, so that each pass of a step is one launch over the active bricks of all
    scenes and the coarsest V-cycle level one workgroup per scene. `scenes[i]` reads and
    writes scene i; step, reset and the solver controls apply to the whole batch."""

    def __init__(self, descs: Sequence[abi.CreateDesc], lib=None):
        self._lib = lib or abi.load()
        arr = (abi.CreateDesc * len(descs))(*descs)
        self._h = C.c_void_p()
        rc = self._lib.fx_create_scenes(arr, len(descs), C.byref(self._h))
        if rc != abi.FX_OK:
            msg = self._lib.fx_last_error(None)
            raise abi.FenixError(rc, msg.decode() if msg else "")
        self.scenes = [QueueScene(self, i, d) for i, d in enumerate(descs)]

    def _check(self, rc):
        return abi.check(self._lib, self._h, rc)

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return len(self.scenes)

    def step(self, elapsedSeconds: float, stepSeconds: float) -> None:
        self._check(self._lib.fx_step(self._h, float(elapsedSeconds), float(stepSeconds)))

    def synchronize(self) -> None:
        self._check(self._lib.fx_synchronize(self._h))

    def reset(self) -> None:
        self._check(self._lib.fx_reset(self._h))

    def active_brick_count(self) -> int:
        n = C.c_uint32()
        self._check(self._lib.fx_active_brick_count(self._h, C.byref(n)))
        return n.value

    def set_vcycle_plan(self, down_split: int, up_split: int) -> None:
        self._check(self._lib.fx_set_vcycle_plan(self._h, int(down_split), int(up_split)))

    def vcycle_plan(self):
        d, u = C.c_uint32(), C.c_uint32()
        self._check(self._lib.fx_get_vcycle_plan(self._h, C.byref(d), C.byref(u)))
        return d.value, u.value

    def autotune_vcycle(self, repeats: int = 4):
        d, u = C.c_uint32(), C.c_uint32()
        self._check(self._lib.fx_autotune_vcycle(self._h, int(repeats), C.byref(d), C.byref(u)))
        return d.value, u.value

    def set_stage_timing(self, enable: bool) -> None:
        self._check(self._lib.fx_set_stage_timing(self._h, 1 if enable else 0))

    def stage_times(self):
        return CombustionVolumeSimulation.stage_times(self)

    def dispose(self) -> None:
        if self._h:
            self._lib.fx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass


def createCombustionVolumeSimulation(device=None, options: Optional[Dict] = None, **build_knobs) -> CombustionVolumeSimulation:
    """createCombustionVolumeSimulation(device, options) — ORCH:50-226.

    `device` is a HIP device ordinal (None -> 0). `options` follows
    CombustionVolumeSimulationOptions (ORCH:39-48): scalarAdvectionMode,
    resolution, wind, windStrength, gravity, gravityStrength, worldSize,
    qualitySettings. `build_knobs` are extra fx_create_desc fields (domain_mode,
    solver, vcycle, ...).
    """
    options = dict(options or {})
    res = options.get("resolution", {"width": 128, "height": 256, "depth": 128})
    if isinstance(res, VolumeResolution):
        res = (res.width, res.height, res.depth)
    elif isinstance(res, dict):
        res = (res["width"], res["height"], res["depth"])
    runtime = {}
    for key in ("wind", "windStrength", "gravity", "gravityStrength", "worldSize"):
        if options.get(key) is not None:
            runtime[key] = options[key]
    kw = dict(build_knobs)
    if runtime:
        kw["runtime"] = {**runtime, **kw.get("runtime", {})}
    if options.get("qualitySettings"):
        kw["quality"] = {**options["qualitySettings"], **kw.get("quality", {})}
    if options.get("scalarAdvectionMode"):
        kw["advection_mode"] = _ADVECTION_MODES[options["scalarAdvectionMode"]]
    if device is not None:
        kw["device"] = int(device)
    return CombustionVolumeSimulation(make_desc(tuple(res), **kw))
