/*
 * fenix.h — C ABI of libfenix, the MI355X (gfx950) implementation of fenix's
 * per-step combustion-volume pipeline (simulation step + emission–absorption
 * raymarch).
 *
 * Every entry point replaces one member of the reference's TypeScript API; the
 * reference interface is cited next to each declaration (paths relative to the
 * reference repository root). Plain C types only: pointers, sizes, PODs.
 *
 * Conventions
 *   - Every call returns an int status: FX_OK (0) or a negative FX_ERR_* code;
 *     fx_last_error() returns the message of the last failure on that handle.
 *   - Field reads/writes use the reference's dense row-major FP32 layout
 *     (index = x + W*(y + H*z), src/engine/simulation/shaders/common/indexing.wgsl.ts:2-4).
 *     Device storage is brick-major; conversion happens inside the call.
 *   - fx_step and fx_raymarch_device enqueue work on the simulation's HIP
 *     stream and return without waiting; reads synchronise.
 *   - The library owns all device memory. Inputs are copied during the call.
 */
#ifndef FENIX_H
#define FENIX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define FX_OK 0
#define FX_ERR_INVALID_ARG (-1)
#define FX_ERR_OUT_OF_MEMORY (-2)
#define FX_ERR_DEVICE (-3)
#define FX_ERR_UNSUPPORTED (-4)
#define FX_ERR_STATE (-5)

/* ---- enums -------------------------------------------------------------- */
/* ScalarAdvectionMode (src/engine/simulation/runtime/combustion-volume-simulation/types.ts:4) */
#define FX_ADVECT_MACCORMACK 0
#define FX_ADVECT_SEMI_LAGRANGIAN 1

/* Domain: DENSE runs every stage over every brick with the reference's
 * activity-flag semantics; SPARSE runs every stage over the active-brick list
 * only, with all fields defined as zero outside the active bricks. */
#define FX_DOMAIN_DENSE 0
#define FX_DOMAIN_SPARSE 1

/* Activation: REFERENCE = compute-active-bricks.wgsl.ts:24-79 (energy samples,
 * conditional dilation); THRESHOLD = density > a or |u| > b anywhere in the
 * brick; bricks with density > a also activate their 26 neighbours. */
#define FX_ACTIVATION_REFERENCE 0
#define FX_ACTIVATION_THRESHOLD 1

/* Pressure: REFERENCE = the 3-level schedule of PressureSolvePass.ts:68-84;
 * VCYCLE = L-level weighted-Jacobi V-cycle with h^2-scaled residual restriction. */
#define FX_SOLVER_REFERENCE 0
#define FX_SOLVER_VCYCLE 1

/* Storage precision of density, temperature and fuel (reaction, velocity and
 * pressure always FP32; all arithmetic FP32). BF16/FP16 round every value a pass
 * stores stochastically: up with probability equal to the dropped fraction of an
 * ulp, decided by an integer hash of (voxel, step, pass, field), so the rounding is
 * unbiased and reproducible; the _RNE variants round to nearest even. Values
 * uploaded by fx_write_field round to nearest even in every mode. */
#define FX_PRECISION_FP32 0
#define FX_PRECISION_BF16 1
#define FX_PRECISION_FP16 2
#define FX_PRECISION_BF16_RNE 3
#define FX_PRECISION_FP16_RNE 4
/* 16-bit unsigned fixed point with 16 fraction bits (value = u / 65536, a uniform
 * step of 2^-16 over [0, 1 - 2^-16]) for these three fields, which every pass
 * clamps to [0, 1]: stored values are clamped to that range. U16 rounds
 * stochastically (unbiased to within 2^-24 of a step), U16_RNE to nearest even. */
#define FX_PRECISION_U16 5
#define FX_PRECISION_U16_RNE 6
/* Unsigned 16-bit float, 4 exponent and 12 mantissa bits, for values in [0, 1]
 * (stored values are clamped to that range): normal codes (1 + m/4096) * 2^(E-15),
 * E = 1..15, and steps of 2^-26 below 2^-14; every value is held to 2^-13 relative,
 * a quarter of a half's step, and 1.0 exactly. UF16 rounds stochastically, UF16_RNE
 * to nearest even. */
#define FX_PRECISION_UF16 7
#define FX_PRECISION_UF16_RNE 8

/* Field ids for fx_read_field / fx_write_field / fx_field_device_ptr. */
#define FX_FIELD_DENSITY 0          /* render set (getRenderBuffers) */
#define FX_FIELD_TEMPERATURE 1
#define FX_FIELD_FUEL 2
#define FX_FIELD_REACTION 3
#define FX_FIELD_VELOCITY_X 4       /* velocityCurrent */
#define FX_FIELD_VELOCITY_Y 5
#define FX_FIELD_VELOCITY_Z 6
#define FX_FIELD_SCRATCH_X 7        /* velocityScratch */
#define FX_FIELD_SCRATCH_Y 8
#define FX_FIELD_SCRATCH_Z 9
#define FX_FIELD_VORTICITY_X 10
#define FX_FIELD_VORTICITY_Y 11
#define FX_FIELD_VORTICITY_Z 12
#define FX_FIELD_VORTICITY_MAG 13
#define FX_FIELD_CONFINEMENT_MAG 14
#define FX_FIELD_VELOCITY_MAG 15
#define FX_FIELD_DIVERGENCE 16      /* fine-level divergence */
#define FX_FIELD_PRESSURE 17        /* fine-level pressure last used by projection */
#define FX_FIELD_ADV_DENSITY 18     /* advected set (scalar advection output, combustion input) */
#define FX_FIELD_ADV_TEMPERATURE 19
#define FX_FIELD_ADV_FUEL 20
#define FX_FIELD_ADV_REACTION 21
#define FX_FIELD_COUNT 22

/* Pressure-level buffers for fx_read_level. */
#define FX_LEVEL_DIVERGENCE 0
#define FX_LEVEL_PRESSURE_A 1
#define FX_LEVEL_PRESSURE_B 2

/* Flag sets for fx_read_flags. */
#define FX_FLAGS_RENDER 0           /* flags returned by getRenderBuffers */
#define FX_FLAGS_INJECTED 1         /* flags computed right after injection */

/* Single passes for fx_run_pass (test hook; the reference records these
 * inside step(), ORCH = src/engine/simulation/runtime/createCombustionVolumeSimulation.ts:270-301). */
#define FX_PASS_INJECT_SCALAR 1     /* inject-scalar.wgsl.ts:32-70 into the render set */
#define FX_PASS_INJECT_VELOCITY 2   /* inject-velocity.wgsl.ts:46-96 */
#define FX_PASS_INJECT_IGNITER 3    /* inject-igniter.wgsl.ts:32-66 */
#define FX_PASS_ACTIVATE_RENDER 4   /* compute-active-bricks over the render set -> injected flags */
#define FX_PASS_ACTIVATE_ADVECTED 5 /* compute-active-bricks over the advected set -> render flags */
#define FX_PASS_ADVECT_VELOCITY 6   /* advect-velocity-semilagrangian.wgsl.ts:23-41 */
#define FX_PASS_ADVECT_SCALARS 7    /* advect-scalar-{maccormack,semilagrangian}: render -> advected */
#define FX_PASS_COMBUSTION 8        /* combustion-update.wgsl.ts:42-93: advected -> render (read-old) */
#define FX_PASS_BUOYANCY 9          /* apply-buoyancy.wgsl.ts:37-110 */
#define FX_PASS_VORTICITY 10        /* compute-vorticity.wgsl.ts:23-50 */
#define FX_PASS_CONFINEMENT 11      /* apply-vorticity-confinement.wgsl.ts:30-66 */
#define FX_PASS_DIVERGENCE 12       /* compute-divergence.wgsl.ts:22-42 */
#define FX_PASS_PRESSURE 13         /* PressureSolvePass.clear + solve */
#define FX_PASS_PROJECTION 14       /* project-velocity.wgsl.ts:21-43 */
#define FX_PASS_DEBUG_FIELDS 15     /* compute-debug-fields.wgsl.ts:18-25 */

/* Raymarch display modes (src/engine/render/volumetrics/volumeDisplayMode.ts:1). */
#define FX_DISPLAY_TEMPERATURE 0
#define FX_DISPLAY_DENSITY 1
#define FX_DISPLAY_FUEL 2

#define FX_FLOATS_PER_EMITTER 32    /* packEmitterSources.ts:55 */
#define FX_MAX_RENDER_LIGHTS 3      /* renderLight.ts:10 */

/* present_mask bits for fx_set_runtime_params (Partial<SimulationRuntimeParams>). */
#define FX_RP_WIND (1u << 0)
#define FX_RP_WIND_STRENGTH (1u << 1)
#define FX_RP_GRAVITY (1u << 2)
#define FX_RP_GRAVITY_STRENGTH (1u << 3)
#define FX_RP_BUOYANCY (1u << 4)
#define FX_RP_VORTICITY_STRENGTH (1u << 5)
#define FX_RP_VORTICITY_CONSTANT_MASK (1u << 6)
#define FX_RP_VORTICITY_VELOCITY_MASK (1u << 7)
#define FX_RP_VORTICITY_HEAT_MASK (1u << 8)
#define FX_RP_VORTICITY_DENSITY_MASK (1u << 9)
#define FX_RP_WORLD_SIZE (1u << 10)
#define FX_RP_ALL 0x7ffu

/* ---- PODs --------------------------------------------------------------- */

/* SimulationQualitySettings, CVS = src/engine/simulation/runtime/combustion-volume-simulation/;
 * fields and clamps follow CVS/types.ts:6-44. */
typedef struct fx_quality {
  int32_t pressure_interval;    /* [1,16] */
  int32_t vorticity_interval;   /* [1,16] */
  int32_t fine_pre_iterations;  /* [0,32] */
  int32_t fine_post_iterations; /* [0,32] */
  int32_t mid_pre_iterations;   /* [0,32] */
  int32_t mid_post_iterations;  /* [0,32] */
  int32_t coarse_iterations;    /* [0,64] */
} fx_quality;

/* SimulationRuntimeParams (CVS/types.ts:46-58); defaults ORCH:194-204. */
typedef struct fx_runtime_params {
  float wind[3];
  float wind_strength;
  float gravity[3];
  float gravity_strength;
  float buoyancy;
  float vorticity_strength;
  float vorticity_constant_mask;
  float vorticity_velocity_mask;
  float vorticity_heat_mask;
  float vorticity_density_mask;
  float world_size;
} fx_runtime_params;

/* The 80-byte per-step uniform (src/engine/simulation/shaders/common/simulation-params.wgsl.ts:2-14,
 * packed at ORCH:241-266). Used by fx_run_pass. */
typedef struct fx_sim_params {
  float time;
  float delta_time;
  float previous_time;
  float scalar_set_index;
  float wind[4];          /* xyz direction, w strength */
  float gravity[4];       /* xyz direction, w strength */
  float buoyancy;
  float vorticity_strength;
  float dx;
  float world_size;
  float vorticity_mask[4];
} fx_sim_params;

/* V-cycle configuration (FX_SOLVER_VCYCLE only). */
typedef struct fx_vcycle {
  int32_t levels;        /* >= 2 */
  int32_t pre_sweeps;
  int32_t post_sweeps;
  int32_t cycles;
  int32_t coarse_sweeps;
  float omega;           /* weighted-Jacobi factor */
} fx_vcycle;

/* CombustionVolumeSimulationOptions (ORCH:39-48) plus this library's knobs. */
typedef struct fx_create_desc {
  uint32_t width, height, depth;  /* resolution; ORCH:54 default 128x256x128 */
  int32_t advection_mode;         /* FX_ADVECT_* */
  fx_runtime_params runtime;      /* wind/gravity/worldSize options + ORCH:198-203 defaults */
  fx_quality quality;             /* qualitySettings (clamped like CVS/types.ts:30-44) */
  int32_t domain_mode;            /* FX_DOMAIN_* */
  int32_t activation_rule;        /* FX_ACTIVATION_* */
  int32_t solver;                 /* FX_SOLVER_* */
  int32_t precision;              /* FX_PRECISION_* */
  fx_vcycle vcycle;
  float ignition_lo;              /* combustion ignition window; reference 0.1 .. 0.55 */
  float ignition_hi;
  float burn_rate;                /* multiplier of the reference burn; reference 1 */
  float activation_density;       /* THRESHOLD rule: density threshold */
  float activation_speed;         /* THRESHOLD rule: |u| threshold */
  int32_t debug_fields;           /* -1: reference rule (N < 1.8M, ORCH:193); 0 off; 1 on */
  int32_t device;                 /* HIP device ordinal */
} fx_create_desc;

/* OrbitCameraSnapshot (src/engine/scene/camera/createOrbitCameraController.ts:11-19). */
typedef struct fx_camera {
  float position[3];
  float target[3];
  float forward[3];
  float right[3];
  float up[3];
  float aspect;
  float tan_half_fov_y;
} fx_camera;

/* RenderLight after sanitizeRenderLights (createVolumeRaymarchPass.ts:62-76). */
typedef struct fx_light {
  int32_t type;          /* 0 directional, 1 point */
  float direction[3];
  float position[3];
  float color[3];
  float intensity;
} fx_light;

/* RaymarchRenderParams after setRenderParams clamps (createVolumeRaymarchPass.ts:8-13, 284-297)
 * plus the per-call render() arguments. */
typedef struct fx_render_params {
  int32_t step_count;
  int32_t light_count;    /* 0..3 */
  fx_light lights[FX_MAX_RENDER_LIGHTS];
  float scattering_forward;
  float scattering_back;
  int32_t display_mode;   /* FX_DISPLAY_* */
  float elapsed_s;
} fx_render_params;

typedef struct fx_sim fx_sim;

/* ---- lifecycle ---------------------------------------------------------- */

/* Fills *desc with the reference defaults (ORCH:54, 194-204; CVS/types.ts:16-24). */
void fx_default_create_desc(fx_create_desc* desc);

/* createCombustionVolumeSimulation(device, options) — ORCH:50-226. */
int fx_create(const fx_create_desc* desc, fx_sim** out);

/* dispose() — ORCH:384-400. Accepts NULL. */
void fx_destroy(fx_sim* sim);

/* reset() — ORCH:206-226, 381-383. */
int fx_reset(fx_sim* sim);

/* ---- step --------------------------------------------------------------- */

/* step(encoder, elapsedSeconds, stepSeconds) — CVS/types.ts:64, ORCH:230-305. */
int fx_step(fx_sim* sim, double elapsed_s, double step_s);

/* Waits for all work queued on the simulation's stream. */
int fx_synchronize(fx_sim* sim);

/* ---- batches (no reference counterpart) ---------------------------------- */

/* A batch of independent scenes (many scenes of different sizes and active sets).
 * Every scene is a full fx_sim with its own simulation and render streams and its
 * own captured step graphs, so one batch step runs the scenes concurrently on the
 * GPU: the latency-bound launches of one scene fill the CUs another leaves idle. */
typedef struct fx_batch fx_batch;

/* Creates n scenes from descs[0..n-1]. On failure nothing is left allocated. */
int fx_batch_create(const fx_create_desc* descs, uint32_t n, fx_batch** out);
/* Destroys every scene of the batch. Accepts NULL. */
void fx_batch_destroy(fx_batch* batch);
uint32_t fx_batch_size(const fx_batch* batch);
/* Borrowed handle of scene i (NULL if out of range): every fx_* call applies to it,
 * except fx_destroy. */
fx_sim* fx_batch_scene(fx_batch* batch, uint32_t i);
/* fx_step on every scene with the same clock; stops at the first failure and
 * returns its code (the message is on that scene's fx_last_error). */
int fx_batch_step(fx_batch* batch, double elapsed_s, double step_s);
/* Waits for every scene's streams. */
int fx_batch_synchronize(fx_batch* batch);

/* A brick-queue batch: n scenes of any sizes in one simulation. Every level's
 * storage holds the bricks of all scenes one after another, and the active-brick
 * lists, the page tables and every stage of the step run over all scenes at once:
 * each pass of fx_step is one launch over the (scene, brick) queue (the scene of a
 * brick selects its grid, step constants and emitters), the coarsest V-cycle
 * smoother runs one workgroup per scene, and the step is captured as one pair of
 * hipGraphs. Each scene's results are those of the same scene simulated alone.
 * The scenes must use the sparse domain, the V-cycle solver and threshold
 * activation, and share precision, advection mode, quality, V-cycle settings,
 * combustion and activation constants and device; their sizes, runtime
 * parameters and emitters are free. Scenes render one at a time
 * (fx_scene_raymarch*); a batch runs no debug pass. fx_step, fx_reset, fx_synchronize, fx_destroy, the quality, advection,
 * V-cycle plan, autotuner and stage-timing calls apply to the whole batch; the
 * calls that address one scene's fields, flags, emitters or runtime parameters
 * take the fx_scene_* forms below (FX_ERR_UNSUPPORTED otherwise). */
int fx_create_scenes(const fx_create_desc* descs, uint32_t n, fx_sim** out);
/* Scenes of the simulation (1 for fx_create's). */
uint32_t fx_scene_count(const fx_sim* sim);
/* The per-scene forms of fx_update_sources, fx_get/set_runtime_params,
 * fx_read_field, fx_write_field, fx_read_flags, fx_read_active_list,
 * fx_read_page_table and fx_active_brick_count: same arguments and meaning on
 * scene `scene` (its own grid; brick indices and list positions are the scene's).
 * On a simulation of one scene, scene 0 is the simulation itself. */
int fx_scene_update_sources(fx_sim* sim, uint32_t scene, const float* packed, uint32_t count);
int fx_scene_get_runtime_params(fx_sim* sim, uint32_t scene, fx_runtime_params* out);
int fx_scene_set_runtime_params(fx_sim* sim, uint32_t scene, const fx_runtime_params* params, uint32_t present_mask);
int fx_scene_read_field(fx_sim* sim, uint32_t scene, int32_t field, float* host, size_t bytes);
int fx_scene_write_field(fx_sim* sim, uint32_t scene, int32_t field, const float* host, size_t bytes);
int fx_scene_read_flags(fx_sim* sim, uint32_t scene, int32_t which, uint32_t* host, size_t count);
int fx_scene_read_active_list(fx_sim* sim, uint32_t scene, uint32_t* host, size_t capacity, uint32_t* count);
int fx_scene_read_page_table(fx_sim* sim, uint32_t scene, int32_t* host, size_t count);
int fx_scene_active_brick_count(fx_sim* sim, uint32_t scene, uint32_t* out);
/* fx_raymarch / fx_raymarch_device of one scene: its grid, world bounds and flags
 * in the step's render snapshot (which every scene of a batch shares). */
int fx_scene_raymarch(fx_sim* sim, uint32_t scene, const fx_camera* camera, const fx_render_params* params, uint32_t width,
                      uint32_t height, uint16_t* host_rgba16f);
int fx_scene_raymarch_device(fx_sim* sim, uint32_t scene, const fx_camera* camera, const fx_render_params* params,
                             uint32_t width, uint32_t height, void* device_rgba16f);

/* ---- parameters --------------------------------------------------------- */

/* getRuntimeParams / setRuntimeParams(Partial<...>) — ORCH:338-365. */
int fx_get_runtime_params(fx_sim* sim, fx_runtime_params* out);
int fx_set_runtime_params(fx_sim* sim, const fx_runtime_params* params, uint32_t present_mask);

/* getQualitySettings / setQualitySettings — ORCH:366-377 (clamps CVS/types.ts:30-44). */
int fx_get_quality(fx_sim* sim, fx_quality* out);
int fx_set_quality(fx_sim* sim, const fx_quality* quality);

/* getScalarAdvectionMode / setScalarAdvectionMode — ORCH:327-337. */
int fx_get_advection_mode(fx_sim* sim, int32_t* out);
int fx_set_advection_mode(fx_sim* sim, int32_t mode);

/* updateSources(sources) — ORCH:378-380, packed by packEmitterSources.ts:69-121
 * (32 floats per source). count may be 0. */
int fx_update_sources(fx_sim* sim, const float* packed, uint32_t count);

/* ---- outputs ------------------------------------------------------------ */

/* getRenderBuffers()/getDebugBuffers() readback — ORCH:306-326. `bytes` must be
 * W*H*D*4. Synchronises the stream. */
int fx_read_field(fx_sim* sim, int32_t field_id, float* host, size_t bytes);

/* Uploads a dense row-major FP32 field (test and scene-initialisation hook; the
 * reference writes buffers only through its passes). In FX_DOMAIN_SPARSE the
 * bricks holding non-zero values join the active set. */
int fx_write_field(fx_sim* sim, int32_t field_id, const float* host, size_t bytes);

/* Brick flags (u32 per brick, ORCH:71-81): FX_FLAGS_RENDER is activeBrickFlags of
 * getRenderBuffers (ORCH:309). `count` must equal the brick count. */
int fx_read_flags(fx_sim* sim, int32_t which, uint32_t* host, size_t count);

/* Active-brick list (ascending brick indices with flag != 0) of the render flags;
 * *count receives the list length. No reference counterpart (the reference has
 * flags only); pinned by the flags. */
int fx_read_active_list(fx_sim* sim, uint32_t* host, size_t capacity, uint32_t* count);

/* Page table: brick index -> position in the active list, or -1. */
int fx_read_page_table(fx_sim* sim, int32_t* host, size_t count);

/* Borrowed device pointer of a field's brick-major storage (valid until the next
 * fx_step / fx_reset / fx_destroy). Element type follows the precision. */
void* fx_field_device_ptr(fx_sim* sim, int32_t field_id);

/* getDebugBuffers() also sets debugRequested (ORCH:313-326): the next fx_step runs the
 * debug pass (velocity magnitude, ORCH:298-301) whatever the grid size, then clears
 * the request. Without a request the pass runs every step only below 1.8M voxels
 * (ORCH:193) or as fx_create_desc.debug_fields says. */
int fx_request_debug_fields(fx_sim* sim);

/* Pressure-hierarchy buffers (PressureSolvePass.ts:93-127): level 0 = fine. */
int fx_read_level(fx_sim* sim, int32_t level, int32_t which, float* host, size_t bytes);
int fx_level_dims(fx_sim* sim, int32_t level, uint32_t dims[3]);

/* FP64 residual norms ||div - L p||_2 of the fine level, one per V-cycle of the
 * last solve (FX_SOLVER_VCYCLE); *count receives how many were written. */
int fx_solver_residuals(fx_sim* sim, double* host, size_t capacity, uint32_t* count);

/* Number of bricks in the last computed active list (sparse) or all bricks (dense). */
int fx_active_brick_count(fx_sim* sim, uint32_t* out);

/* ---- raymarch ----------------------------------------------------------- */

/* VolumeRaymarchPass.render(...) — createVolumeRaymarchPass.ts:183-283, shader :573-724,
 * over the current render buffers. Writes width*height RGBA float16 pixels
 * (row 0 = top, premultiplied, cleared to 0 where the reference discards). */
int fx_raymarch(fx_sim* sim, const fx_camera* camera, const fx_render_params* params,
                uint32_t width, uint32_t height, uint16_t* host_rgba_f16);

/* Same, into a caller-owned device buffer of width*height*8 bytes; asynchronous. */
int fx_raymarch_device(fx_sim* sim, const fx_camera* camera, const fx_render_params* params,
                       uint32_t width, uint32_t height, void* device_rgba_f16);

/* ---- bloom and composite ------------------------------------------------ */

/* BloomParams, src/engine/render/passes/createBloomPass.ts:1-6. */
typedef struct fx_bloom_params {
  int32_t enabled;
  float threshold, strength, radius;
} fx_bloom_params;

/* BloomPass.render(encoder, sourceView, targetView, w, h, params) — createBloomPass.ts:99-129:
 * the raymarch's w x h RGBA float16 image (device) -> the w x h canvas image as RGBA8
 * unorm (R in the lowest byte of each 32-bit pixel, row 0 = top). Enabled: soft
 * luminance threshold into half resolution, 9-tap horizontal and vertical blur,
 * composite with pow(hdr, 0.92); disabled: the blit with pow(hdr, 0.92). Enqueued
 * on the render stream after the raymarch; the library keeps the half-resolution
 * images (resized when w, h change). */
int fx_bloom_device(fx_sim* sim, const fx_bloom_params* params, uint32_t width, uint32_t height,
                    const void* device_src_rgba_f16, void* device_dst_rgba8);
/* Same from and to host memory (synchronous). */
int fx_bloom(fx_sim* sim, const fx_bloom_params* params, uint32_t width, uint32_t height,
             const uint16_t* host_src_rgba_f16, uint8_t* host_dst_rgba8);

/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): fx_bloom_device without any allocation. The three half-resolution
 * images live in a static device array of 24 MiB per GPU (held until the process
 * ends), so the canvas is limited to 1024 x 1024 half-resolution pixels (up to
 * 2048 x 2048); larger sizes fail with FX_ERR_INVALID_ARG. Every call on one GPU
 * shares that array: calls that overlap on a GPU (several simulations rendering
 * concurrently on different streams) race and must be serialised by the caller.
 * Results equal fx_bloom_device's. Enqueued on the render stream; asynchronous. */
int fx_bloom_device_noalloc(fx_sim* sim, const fx_bloom_params* params, uint32_t width, uint32_t height,
                            const void* device_src_rgba_f16, void* device_dst_rgba8);

/* ---- stage timing ------------------------------------------------------- */

/* Timing ids: the FX_PASS_* ids (whole pass; FX_PASS_ACTIVATE_RENDER includes the
 * sparse list compaction), plus: */
#define FX_TIME_RAYMARCH 16        /* fx_raymarch / fx_raymarch_device */
#define FX_TIME_JACOBI_FINE 17     /* each fine-level pressure sweep (also inside FX_PASS_PRESSURE) */
#define FX_TIME_COUNT 18

/* Enables/disables HIP-event timing of every stage on the simulation's stream
 * (no reference counterpart; the reference times only its render pass through
 * stats-gl, src/engine/render/renderer/createViewportRuntime.ts:236-268).
 * Enabling clears the accumulators. */
int fx_set_stage_timing(fx_sim* sim, int32_t enable);

/* Synchronises, then returns per timing id the accumulated milliseconds and call
 * counts since timing was enabled. `n` entries are written (<= FX_TIME_COUNT). */
int fx_read_stage_times(fx_sim* sim, double* total_ms, uint32_t* calls, size_t n);

/* ---- V-cycle kernel plan (no reference counterpart) ---------------------- */

/* Kernel plan of the V-cycle legs, one bit per level (bit l = level l):
 * down_split: pre-sweeps as one-sweep kernels + residual + restriction instead of
 * one fused LDS kernel (tiles of the brick plus a pre_sweeps + 2 halo);
 * up_split: prolongation + one-sweep kernels instead of one fused LDS kernel
 * (halo post_sweeps).
 * Bit 8 + l of either mask: that split leg of level l runs its sweeps two per
 * launch (LDS tiles of halo 2) instead of one per launch; ignored where the leg
 * is fused.
 * Bit 16 + l: that split leg fuses the sweep next to the level change into one
 * LDS kernel: the last pre-sweep with the residual and the restriction (down,
 * tiles of halo 3), or the prolongation with the first post-sweep (up, halo 1).
 * Bits 24-25 of down_split: sweeps per launch of the coarsest level when its
 * active voxels exceed the one-workgroup smoother's capacity (0: one, 1: two,
 * 2: four, in LDS tiles of halo 2 or 4).
 * Bits 20 + 2l .. 21 + 2l of up_split: the compute tile of the split sweeps of
 * level l (both legs): 0 the 8^3 brick, 1 its 4^3 octants (one wave each), 2 a
 * 2x2x2 group of bricks (16^3, one workgroup); the pair bit of the leg gives the
 * halo (1 or 2 sweeps per launch, staged in LDS).
 * Every plan gives bit-identical results; only speed differs. */
int fx_set_vcycle_plan(fx_sim* sim, uint32_t down_split, uint32_t up_split);
int fx_get_vcycle_plan(fx_sim* sim, uint32_t* down_split, uint32_t* up_split);

/* Autotuner: times each level's leg plans (fused, split, split with paired
 * sweeps, for the down and the up leg; greedily, finest level first), then, where
 * a leg of the level is split, its sweeps' compute tile (4^3, 8^3 or 16^3, with
 * the chosen halo of 1 or 2); then the coarsest level's sweeps per launch when it
 * runs multi-workgroup launches; the bit 16 + l legs are not timed, only pinned by
 * fx_set_vcycle_plan)
 * with `repeats` graph-captured pressure solves of the current state and keeps
 * the fastest; returns the chosen masks. Solves start from p = 0 and every plan
 * is bit-identical, so the simulation's results do not depend on the outcome.
 * Synchronises; V-cycle solver only (FX_ERR_UNSUPPORTED otherwise). */
int fx_autotune_vcycle(fx_sim* sim, uint32_t repeats, uint32_t* down_split, uint32_t* up_split);

/* Kernel clock (no reference counterpart): times `repeats` back-to-back launches
 * of the finest level's weighted-Jacobi sweep on the current active set (reading
 * the last pressure solution, writing the other pressure buffer, so the solution
 * the projection uses is left intact) between one pair of HIP events on the
 * simulation stream, and returns the mean launch time in milliseconds and the
 * voxel count one launch updates. Synchronises; V-cycle solver only. */
int fx_time_fine_sweep(fx_sim* sim, uint32_t repeats, double* avg_ms, uint64_t* voxels);

/* ---- differentiable smoke step (no reference counterpart) ---------------- */

/* Inflow-control optimisation needs gradients of a rollout with respect to a
 * source control; the reference computes none. fx_diff runs a differentiable
 * model of the step on a dense grid: state density d and velocity u (voxel
 * units, 0 initially), and per step
 *   source       d += dt S_rho, u += dt S_u on the floor slab (interior voxels with
 *                1 <= y <= slab), S the cell-centred trilinear upsampling of the
 *                control over the slab box;
 *   advection    u_a = u(x - u(x) dt) (semi-Lagrangian, 0 on the boundary faces),
 *                d' = MacCormack(d, u_a) with the min/max(original, predicted)
 *                limiter of the scalar advection, without its decay terms;
 *   buoyancy     u_y += dt buoyancy d';
 *   confinement  u += dt confinement (N x w), w = curl u, N = grad|w| / max(|grad|w||, 1e-4);
 *   projection   u -= grad P(div u), P the multigrid V-cycle of fx_vcycle.
 * The control holds 4 x control_n^3 floats, row-major [channel][z][y][x], channels
 * (u_x, u_y, u_z, rho). The loss of a rollout is 0.5 sum (d_T - target)^2 over the
 * grid (FP64 reduction). fx_diff_gradient runs the rollout storing (d, u) every
 * checkpoint_every steps, then walks the segments backwards: each segment is
 * recomputed from its checkpoint with the per-step intermediates kept, and the
 * adjoint kernels run it in reverse (transposed trilinear gathers are float
 * atomics, so gradients vary in the last bits from run to run; the forward pass is
 * deterministic). The pressure projection is its own adjoint because the V-cycle
 * is symmetric (equal pre- and post-sweeps, restriction proportional to the
 * transposed prolongation), so pre_sweeps must equal post_sweeps. */
typedef struct fx_diff fx_diff;
typedef struct fx_diff_desc {
  uint32_t width, height, depth; /* dense grid, each >= 8 */
  uint32_t steps;                /* steps per rollout */
  uint32_t checkpoint_every;     /* steps per recomputed segment */
  uint32_t control_n;            /* control edge, >= 2 */
  uint32_t slab;                 /* source slab height (voxels above the floor face) */
  float dt, buoyancy, confinement;
  fx_vcycle vcycle;
  int32_t device;
} fx_diff_desc;

/* 64^3, 64 steps, checkpoint 8, control 32, slab 4, dt 1, buoyancy 0.02,
 * confinement 0.02, V-cycle 4 levels 2/2 x 4, 16 coarse sweeps, omega 2/3.
 * Confinement amplifies small perturbations: with strong confinement over long
 * rollouts (0.3 over 64 steps at 128^3) the loss is chaotic and its gradient,
 * exact for the discrete model, no longer predicts finite changes. */
void fx_diff_default_desc(fx_diff_desc* desc);
int fx_diff_create(const fx_diff_desc* desc, fx_diff** out);
void fx_diff_destroy(fx_diff* diff);
/* Target density, row-major [z][y][x] floats (bytes = 4 W H D). Zero until set. */
int fx_diff_set_target(fx_diff* diff, const float* host, size_t bytes);
/* One rollout from the zero state under `control` (host, 4 control_n^3 floats);
 * writes the loss. The final state stays readable with fx_diff_read_field. */
int fx_diff_forward(fx_diff* diff, const float* control, double* loss);
/* Loss and its gradient with respect to the control (host, 4 control_n^3 floats). */
int fx_diff_gradient(fx_diff* diff, const float* control, double* loss, float* grad);
/* Final state of the last rollout: field 0 density, 1..3 velocity x, y, z;
 * row-major [z][y][x] floats (bytes = 4 W H D). */
int fx_diff_read_field(fx_diff* diff, int32_t field, float* host, size_t bytes);
const char* fx_diff_last_error(fx_diff* diff);

/* ---- test hook ---------------------------------------------------------- */

/* Runs one pass of the step with explicit uniforms (no reference counterpart:
 * the reference only records passes inside step()). For FX_PASS_PRESSURE the
 * current quality/solver settings are used. */
int fx_run_pass(fx_sim* sim, int32_t pass_id, const fx_sim_params* params);

/* Message of the last failure on this handle (or of the last fx_create when sim is NULL). */
const char* fx_last_error(fx_sim* sim);

#ifdef __cplusplus
}
#endif

#endif /* FENIX_H */
