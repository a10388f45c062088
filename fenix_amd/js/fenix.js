'use strict'
// fenix.js — JS host of libfenix with the reference's simulation API.
//
// createCombustionVolumeSimulation(device, options) returns an object with the
// members of CombustionVolumeSimulation
// (src/engine/simulation/runtime/combustion-volume-simulation/types.ts:62-76):
// step(encoder, elapsedSeconds, stepSeconds), getRenderBuffers(), getDebugBuffers(),
// get/setScalarAdvectionMode, get/setRuntimeParams, get/setQualitySettings,
// updateSources(sources), reset(), dispose(). It runs on the MI355X through the
// Node-API addon fenix_napi.node (built by fenix_amd/build.py), which calls the
// C ABI of include/fenix.h. GPU buffers are replaced by device-resident fields
// with read() methods; `device` is the HIP device ordinal (0 when not a number).
// Errors are thrown as Error(message) with the library's message, as the
// reference throws.

const path = require('path')

const native = require(process.env.FENIX_NAPI || path.join(__dirname, 'fenix_napi.node'))

const ADVECTION = { maccormack: 0, 'semi-lagrangian': 1 } // FX_ADVECT_*
const ADVECTION_NAMES = ['maccormack', 'semi-lagrangian']
const FIELD = {
  density: 0, temperature: 1, fuel: 2, reaction: 3,
  velocityX: 4, velocityY: 5, velocityZ: 6,
  velocityMagnitude: 15, divergence: 16, pressure: 17,
  vorticityMagnitude: 13, confinementForceMagnitude: 14,
}
const FLAGS_RENDER = 0
const DISPLAY_MODES = { temperature: 0, density: 1, fuel: 2 } // volumeDisplayMode.ts
// Storage of density, temperature and fuel (FX_PRECISION_*), for options.native.precision;
// UF16 is the 16-bit format that stays within 1e-3 of FP64.
const PRECISION = { FP32: 0, BF16: 1, FP16: 2, BF16_RNE: 3, FP16_RNE: 4, U16: 5, U16_RNE: 6, UF16: 7, UF16_RNE: 8 }

// Emitter packing, src/engine/simulation/emitters/packEmitterSources.ts:25-121:
// 32 floats per source in the layout of the EmitterSource struct of the inject shaders.
const FLOATS_PER_EMITTER = 32
const KIND = { scalar: 0, velocity: 1, igniter: 2, burst: 3 }
const VELOCITY_MODE_INDEX = { radial: 0, directional: 1, turbulent: 2 }

function packEmitterSources (sources) {
  const data = new Float32Array(sources.length * FLOATS_PER_EMITTER)
  sources.forEach((src, i) => {
    const o = i * FLOATS_PER_EMITTER
    data[o + 0] = src.position[0]
    data[o + 1] = src.position[1]
    data[o + 2] = src.position[2]
    data[o + 3] = src.radius
    data[o + 4] = src.startTime
    data[o + 5] = src.duration
    data[o + 29] = src.seed
    if (src.kind === 'scalar') {
      data[o + 8] = src.densityRate
      data[o + 9] = src.heatRate
      data[o + 10] = src.fuelRate
      data[o + 20] = src.noiseScale
      data[o + 21] = src.noiseMix
    } else if (src.kind === 'velocity') {
      data[o + 12] = src.speed
      data[o + 13] = VELOCITY_MODE_INDEX[src.mode]
      data[o + 14] = src.falloff
      data[o + 15] = (src.tightness == null ? 0 : src.tightness)
      data[o + 16] = src.direction[0]
      data[o + 17] = src.direction[1]
      data[o + 18] = src.direction[2]
    } else if (src.kind === 'burst') {
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
    } else if (src.kind === 'igniter') {
      data[o + 22] = src.intensity
    } else {
      throw new Error(`Unknown emitter kind: ${src.kind}`)
    }
    data[o + 28] = KIND[src.kind]
  })
  return data
}

// A device-resident field of the simulation (in place of a GPUBuffer): read()
// synchronises and returns the row-major Float32Array (x + W (y + H z)).
function deviceField (handle, id, voxelCount) {
  return { fieldId: id, read: () => native.readField(handle(), id, voxelCount) }
}

function createCombustionVolumeSimulation (device, options = {}) {
  const resolution = options.resolution || { width: 128, height: 256, depth: 128 } // ORCH:54
  const desc = {
    width: resolution.width,
    height: resolution.height,
    depth: resolution.depth,
    advectionMode: ADVECTION[options.scalarAdvectionMode || 'maccormack'],
    runtime: {},
    quality: options.qualitySettings || {},
    device: typeof device === 'number' ? device : 0,
    ...(options.native || {}), // library knobs without a reference counterpart (domainMode, solver, ...)
  }
  for (const key of ['wind', 'windStrength', 'gravity', 'gravityStrength', 'worldSize']) {
    if (options[key] !== undefined) desc.runtime[key] = options[key]
  }
  let h = native.create(desc)
  const handle = () => {
    if (h === null) throw new Error('CombustionVolumeSimulation used after dispose()')
    return h
  }
  const voxelCount = resolution.width * resolution.height * resolution.depth
  const bricks = [Math.ceil(resolution.width / 8), Math.ceil(resolution.height / 8), Math.ceil(resolution.depth / 8)]
  const brickCount = bricks[0] * bricks[1] * bricks[2]
  const renderBuffers = () => ({
    density: deviceField(handle, FIELD.density, voxelCount),
    temperature: deviceField(handle, FIELD.temperature, voxelCount),
    fuel: deviceField(handle, FIELD.fuel, voxelCount),
    reaction: deviceField(handle, FIELD.reaction, voxelCount),
    activeBrickFlags: { read: () => native.readFlags(handle(), FLAGS_RENDER, brickCount) },
    // packSparseBrickLayout (src/engine/simulation/sparse/createSparseBrickLayout.ts:28-39)
    activeBrickInfo: new Uint32Array([bricks[0], bricks[1], bricks[2], brickCount, 8, 0, 0, 0]),
  })
  return {
    resolution,
    // The encoder is accepted for signature parity: work is ordered on the library's stream.
    step (_encoder, elapsedSeconds, stepSeconds) { native.step(handle(), elapsedSeconds, stepSeconds) },
    getRenderBuffers: renderBuffers,
    getDebugBuffers () {
      native.requestDebugFields(handle()) // debugRequested, ORCH:316
      return {
        ...renderBuffers(),
        velocityMagnitude: deviceField(handle, FIELD.velocityMagnitude, voxelCount),
        divergence: deviceField(handle, FIELD.divergence, voxelCount),
        pressure: deviceField(handle, FIELD.pressure, voxelCount),
        vorticityMagnitude: deviceField(handle, FIELD.vorticityMagnitude, voxelCount),
        confinementForceMagnitude: deviceField(handle, FIELD.confinementForceMagnitude, voxelCount),
      }
    },
    getScalarAdvectionMode () { return ADVECTION_NAMES[native.getAdvectionMode(handle())] },
    setScalarAdvectionMode (mode) {
      if (!(mode in ADVECTION)) throw new Error(`Unknown scalar advection mode: ${mode}`)
      native.setAdvectionMode(handle(), ADVECTION[mode])
    },
    getRuntimeParams () { return native.getRuntimeParams(handle()) },
    setRuntimeParams (params) { native.setRuntimeParams(handle(), params) },
    getQualitySettings () { return native.getQualitySettings(handle()) },
    setQualitySettings (settings) { native.setQualitySettings(handle(), settings) },
    updateSources (sources) { native.updateSources(handle(), packEmitterSources(sources), sources.length) },
    reset () { native.reset(handle()) },
    dispose () {
      if (h !== null) native.destroy(h)
      h = null
    },
    // Not in the reference interface: readback helpers of this host.
    synchronize () { native.synchronize(handle()) },
    activeBrickCount () { return native.activeBrickCount(handle()) },
    _handle: handle,
  }
}

// VolumeRaymarchPass.render (src/engine/render/passes/createVolumeRaymarchPass.ts:183-283) over the
// simulation's current render buffers: camera = OrbitCameraSnapshot, params = the render
// parameters after setRenderParams' clamps plus sanitised lights ({type: 0 directional | 1 point,
// direction, position, color, intensity}). Returns width x height RGBA binary16 (Uint16Array).
function renderVolume (simulation, camera, params, displayMode, width, height, elapsedSeconds = 0) {
  return native.raymarch(simulation._handle(), camera,
    { ...params, displayMode: DISPLAY_MODES[displayMode || 'temperature'], elapsedSeconds }, width, height)
}

module.exports = { createCombustionVolumeSimulation, packEmitterSources, renderVolume, FLOATS_PER_EMITTER, PRECISION, native }
