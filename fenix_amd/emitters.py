"""Emitter sources and their 32-float packing.

Mirrors src/engine/simulation/emitters/{scalarEmitter,velocityEmitter,igniterEmitter,
burstEmitter}.ts (the EmitterSource union) and packEmitterSources.ts:25-121.
"""
from dataclasses import dataclass, field
from typing import Sequence, Tuple, Union

import numpy as np

FLOATS_PER_EMITTER = 32  # packEmitterSources.ts:55
KIND_SCALAR, KIND_VELOCITY, KIND_IGNITER, KIND_BURST = 0, 1, 2, 3
VELOCITY_MODE_INDEX = {"radial": 0, "directional": 1, "turbulent": 2}  # packEmitterSources.ts:52

Vec3 = Tuple[float, float, float]


@dataclass
class ScalarEmitter:  # scalarEmitter.ts:2-21
    position: Vec3
    radius: float
    startTime: float = 0.0
    duration: float = 9999.0
    densityRate: float = 1.0
    heatRate: float = 0.0
    fuelRate: float = 0.0
    noiseScale: float = 0.0
    noiseMix: float = 0.0
    seed: float = 0.0
    kind: str = field(default="scalar", init=False)


@dataclass
class VelocityEmitter:  # velocityEmitter.ts:1-21
    position: Vec3
    radius: float
    mode: str = "directional"
    speed: float = 0.0
    direction: Vec3 = (0.0, 1.0, 0.0)
    falloff: float = 0.0
    tightness: float = 0.0
    startTime: float = 0.0
    duration: float = 9999.0
    seed: float = 0.0
    kind: str = field(default="velocity", init=False)


@dataclass
class IgniterEmitter:  # igniterEmitter.ts:2-12
    position: Vec3
    radius: float
    intensity: float = 1.0
    startTime: float = 0.0
    duration: float = 0.2
    seed: float = 0.0
    kind: str = field(default="igniter", init=False)


@dataclass
class BurstEmitter:  # burstEmitter.ts:2-24
    position: Vec3
    radius: float
    startTime: float = 0.0
    duration: float = 0.5
    densityAmount: float = 0.0
    heatAmount: float = 0.0
    fuelAmount: float = 0.0
    reactionAmount: float = 0.0
    expansionSpeed: float = 0.0
    liftSpeed: float = 0.0
    turbulenceSpeed: float = 0.0
    falloff: float = 0.5
    noiseScale: float = 0.0
    noiseMix: float = 0.0
    seed: float = 0.0
    kind: str = field(default="burst", init=False)


EmitterSource = Union[ScalarEmitter, VelocityEmitter, IgniterEmitter, BurstEmitter]


def pack_emitter_sources(sources: Sequence[EmitterSource]) -> np.ndarray:
    """packEmitterSources(sources) -> float32[len*32] (packEmitterSources.ts:69-121)."""
    data = np.zeros(len(sources) * FLOATS_PER_EMITTER, dtype=np.float32)
    for i, src in enumerate(sources):
        o = i * FLOATS_PER_EMITTER
        data[o + 0:o + 3] = src.position
        data[o + 3] = src.radius
        data[o + 4] = src.startTime
        data[o + 5] = src.duration
        data[o + 29] = src.seed
        if src.kind == "scalar":
            data[o + 8] = src.densityRate
            data[o + 9] = src.heatRate
            data[o + 10] = src.fuelRate
            data[o + 20] = src.noiseScale
            data[o + 21] = src.noiseMix
            data[o + 28] = KIND_SCALAR
        elif src.kind == "velocity":
            data[o + 12] = src.speed
            data[o + 13] = VELOCITY_MODE_INDEX[src.mode]
            data[o + 14] = src.falloff
            data[o + 15] = src.tightness if src.tightness is not None else 0.0
            data[o + 16:o + 19] = src.direction
            data[o + 28] = KIND_VELOCITY
        elif src.kind == "burst":
            data[o + 8] = src.densityAmount
            data[o + 9] = src.heatAmount
            data[o + 10] = src.fuelAmount
            data[o + 12] = src.expansionSpeed
            data[o + 13] = src.liftSpeed
            data[o + 14] = src.falloff
            data[o + 15] = src.turbulenceSpeed
            data[o + 20] = src.noiseScale
            data[o + 21] = src.noiseMix
            data[o + 22] = src.reactionAmount
            data[o + 28] = KIND_BURST
        else:
            data[o + 22] = src.intensity
            data[o + 28] = KIND_IGNITER
    return data
