// Types of the JS host (fenix.js), mirroring the reference's simulation interface
// (src/engine/simulation/runtime/combustion-volume-simulation/types.ts:4-76) and
// EmitterSource (src/engine/simulation/emitters/emitterSource.ts).

export type ScalarAdvectionMode = 'semi-lagrangian' | 'maccormack'

export interface VolumeResolution { width: number; height: number; depth: number }

export interface SimulationQualitySettings {
  pressureInterval: number
  vorticityInterval: number
  finePreIterations: number
  finePostIterations: number
  midPreIterations: number
  midPostIterations: number
  coarseIterations: number
}

export interface SimulationRuntimeParams {
  wind: readonly [number, number, number]
  windStrength: number
  gravity: readonly [number, number, number]
  gravityStrength: number
  buoyancy: number
  vorticityStrength: number
  vorticityConstantMask?: number
  vorticityVelocityMask?: number
  vorticityHeatMask?: number
  vorticityDensityMask?: number
  worldSize: number
}

type Vec3 = readonly [number, number, number]
interface EmitterBase { position: Vec3; radius: number; startTime: number; duration: number; seed: number }
export interface ScalarEmitterSource extends EmitterBase {
  kind: 'scalar'; densityRate: number; heatRate: number; fuelRate: number; noiseScale: number; noiseMix: number
}
export interface VelocityEmitterSource extends EmitterBase {
  kind: 'velocity'; mode: 'radial' | 'directional' | 'turbulent'; speed: number; direction: Vec3; falloff: number
  tightness?: number
}
export interface IgniterEmitterSource extends EmitterBase { kind: 'igniter'; intensity: number }
export interface BurstEmitterSource extends EmitterBase {
  kind: 'burst'; densityAmount: number; heatAmount: number; fuelAmount: number; reactionAmount: number
  expansionSpeed: number; liftSpeed: number; turbulenceSpeed: number; falloff: number; noiseScale: number; noiseMix: number
}
export type EmitterSource = ScalarEmitterSource | VelocityEmitterSource | IgniterEmitterSource | BurstEmitterSource

/** A field resident on the GPU (in place of a GPUBuffer); read() synchronises. */
export interface DeviceField { fieldId: number; read(): Float32Array }

export interface CombustionVolumeRenderBuffers {
  density: DeviceField
  temperature: DeviceField
  fuel: DeviceField
  reaction: DeviceField
  activeBrickFlags: { read(): Uint32Array }
  activeBrickInfo: Uint32Array
}

export interface SimulationDebugBuffers extends CombustionVolumeRenderBuffers {
  velocityMagnitude: DeviceField
  divergence: DeviceField
  pressure: DeviceField
  vorticityMagnitude: DeviceField
  confinementForceMagnitude: DeviceField
}

export interface CombustionVolumeSimulation {
  readonly resolution: VolumeResolution
  step(encoder: unknown, elapsedSeconds: number, stepSeconds: number): void
  getRenderBuffers(): CombustionVolumeRenderBuffers
  getDebugBuffers(): SimulationDebugBuffers
  getScalarAdvectionMode(): ScalarAdvectionMode
  setScalarAdvectionMode(mode: ScalarAdvectionMode): void
  getRuntimeParams(): SimulationRuntimeParams
  setRuntimeParams(params: Partial<SimulationRuntimeParams>): void
  getQualitySettings(): SimulationQualitySettings
  setQualitySettings(settings: Partial<SimulationQualitySettings>): void
  updateSources(sources: readonly EmitterSource[]): void
  reset(): void
  dispose(): void
  synchronize(): void
  activeBrickCount(): number
}

export interface CombustionVolumeSimulationOptions {
  scalarAdvectionMode?: ScalarAdvectionMode
  resolution?: VolumeResolution
  wind?: Vec3
  windStrength?: number
  gravity?: Vec3
  gravityStrength?: number
  worldSize?: number
  qualitySettings?: Partial<SimulationQualitySettings>
  /** Knobs of fx_create_desc with no reference counterpart (domainMode, solver, vcycle, precision, ...). */
  native?: Record<string, unknown>
}

/** `device`: the HIP device ordinal (the reference takes a GPUDevice). */
export function createCombustionVolumeSimulation(device: number | unknown, options?: CombustionVolumeSimulationOptions):
  CombustionVolumeSimulation

export function packEmitterSources(sources: readonly EmitterSource[]): Float32Array

/** Storage formats of density, temperature and fuel for `native.precision` (FX_PRECISION_*). */
export const PRECISION: {
  FP32: 0; BF16: 1; FP16: 2; BF16_RNE: 3; FP16_RNE: 4; U16: 5; U16_RNE: 6; UF16: 7; UF16_RNE: 8
}

export interface RenderLight { type: 0 | 1; direction: Vec3; position: Vec3; color: Vec3; intensity: number }
export interface OrbitCameraSnapshot {
  position: Vec3; target: Vec3; forward: Vec3; right: Vec3; up: Vec3; aspect: number; tanHalfFovY: number
}
export interface RaymarchRenderParams {
  stepCount: number; lights: readonly RenderLight[]; scatteringForward: number; scatteringBack: number
}
/** VolumeRaymarchPass.render into width x height RGBA binary16 (row 0 = top). */
export function renderVolume(simulation: CombustionVolumeSimulation, camera: OrbitCameraSnapshot, params: RaymarchRenderParams,
  displayMode: 'temperature' | 'density' | 'fuel', width: number, height: number, elapsedSeconds?: number): Uint16Array

export const FLOATS_PER_EMITTER: 32
