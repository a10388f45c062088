// fenix_api.cpp — host side of libfenix: the C ABI declared in include/fenix.h.
//
// Owns every device buffer of one simulation, records the reference's pass
// sequence (src/engine/simulation/runtime/createCombustionVolumeSimulation.ts:230-305,
// "ORCH" below) on the simulation's HIP stream and converts between the
// reference's dense row-major layout and the brick-major device layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fenix_internal.h"

using namespace fx;

namespace {

thread_local std::string g_create_error;

constexpr int kStageSlots = 8;
// The step constants and the candidate list share one staging block (pinned on
// the host, device-resident on the GPU): the constants at offset 0, the list at
// kConstBytes, so that a step uploads both with one copy.
constexpr size_t kConstBytes = (sizeof(StepConst) + 255) & ~size_t(255);
constexpr size_t kBlockSmoothVoxels = 65536;  // coarsest levels up to this size run in one workgroup

struct Level {
  Grid g;
  size_t nb = 0, nv = 0;
  float* div = nullptr;
  float* p[2] = {nullptr, nullptr};  // pressureA, pressureB
  uint32_t* flags = nullptr;         // sparse: level activity flags
  BrickList L{};                     // sparse: active list; dense: identity
  uint64_t* bits = nullptr;          // voxel activity words, 8 per listed brick (V-cycle)
  uint32_t* nbr = nullptr;           // fine level, sparse V-cycle: listed face neighbours per listed brick
  void* tab = nullptr;               // coarsest level: slot table of the block smoother
  uint32_t* tab_n = nullptr;
};

struct StageSlot {
  StepConst* c = nullptr;      // pinned staging block (const_bytes + 4 nb bytes)
  uint32_t* cand = nullptr;    // inside the block at const_bytes, capacity nb
  hipEvent_t done = nullptr;
  bool used = false;
};

}  // namespace

struct fx_sim {
  fx_create_desc desc{};
  int device = 0;
  hipStream_t stream = nullptr;
  Grid g{};
  size_t nb = 0, nv = 0;
  int blocks = 1;          // persistent grid for brick kernels
  int cus = 1;             // compute units of the device
  bool blocks_env = false; // FENIX_SIM_BLOCKS_PER_CU set
  bool sparse = false;
  bool stochastic = false; // 16-bit stores round stochastically (FX_PRECISION_BF16/FP16/U16)

  void* S[2][4] = {};               // d, t, f in the storage precision; r FP32
  float* u[3] = {};
  float* us[3] = {};
  float* w[3] = {};
  float* wmag = nullptr;
  float* cmag = nullptr;
  float* umag = nullptr;
  uint32_t* flagsInj = nullptr;
  uint32_t* flagsRender = nullptr;  // aliases flagsInj when sparse
  uint32_t* touched = nullptr;
  std::vector<Level> lv;
  uint32_t* identity = nullptr;     // 0..nb-1
  uint32_t* identity_count = nullptr;

  StepConst* d_const = nullptr;     // device copy of the step constants (start of the staging block)
  uint32_t* d_cand = nullptr;       // device candidate list (injection + host-written bricks), in the same block
  float* d_src = nullptr;
  size_t src_cap = 0;               // floats
  std::vector<float> h_src;         // packed, >= 1 source
  uint32_t nsrc = 0;
  StageSlot slots[kStageSlots];
  int slot_next = 0;
  double* d_norms = nullptr;
  int norm_cap = 0;
  int norm_count = 0;

  fx_runtime_params rt{};
  fx_quality q{};
  int mode = 0;

  // Brick-queue batch (fx_create_scenes): every scene's bricks share each level's
  // storage (scene k owns bricks [base, base + its bricks) of every level, in
  // scene order), and every stage is one launch over the listed bricks of all
  // scenes. g and lv[l].g are then selectors whose tables live on the device;
  // the per-scene grids, runtime parameters and emitters are kept here. The
  // step constants are one StepConst per scene (const_bytes apart from the
  // candidate list), the emitters one buffer with per-scene offsets.
  bool multi = false;
  struct Scene {
    Grid g[FX_MAX_LEVELS_DEV] = {};  // this scene's grid of every level (base and index set)
    size_t nv = 0;                    // voxels of the finest level
    fx_runtime_params rt{};
    std::vector<float> h_src;         // packed, >= 1 source
    uint32_t nsrc = 0;
  };
  std::vector<Scene> scenes;
  Grid* d_gtab = nullptr;             // [level][scene] grids
  uint32_t* d_owner = nullptr;        // scene of every brick, level after level
  uint32_t* d_scene_count = nullptr;  // active fine bricks per scene (end of the last step)
  uint32_t* h_scene_count = nullptr;  // pinned copy
  size_t coarse_nv_max = 0;           // largest coarsest-level storage of one scene
  size_t const_bytes = 0;             // bytes of the step constants before the candidate list
  bool initialized = false;
  double startS = 0, lastS = 0;
  uint64_t stepIndex = 0;
  int lastPressure = 0;
  int curSet = 0;
  std::vector<uint32_t> pending;    // sparse: bricks written by fx_write_field since the last step
  std::string err;

  CompactArgs cargs{};               // sparse: list compaction of every level
  LevelFlags lflags{};               // sparse: coarse flags marked by activation (n = 0 when dense)
  uint32_t* d_compact = nullptr;     // per-block counts and offsets
  uint32_t* d_bits = nullptr;        // packed render flags for the raymarch

  // V-cycle kernel plan per level (bit l set: split legs of lean one-sweep
  // kernels instead of the fused LDS legs). Part of the graph key.
  uint32_t split_down = 0, split_up = 0;
  // Coarsest level: the one-workgroup LDS smoother holds kCoarseCap active
  // voxels (beyond that it falls back to a slow one-workgroup global path);
  // coarse_split runs the coarsest sweeps as multi-workgroup launches instead.
  // Decided per step from an upper bound of the coarsest active voxel count:
  // exact when dense, from the fine active-brick count of an earlier step when
  // sparse (copied into pinned memory at the end of every step and read once
  // that step's event has completed, so no step waits for it). Results never
  // depend on the choice. Part of the graph key.
  bool coarse_split = false;
  uint32_t* h_fine_count = nullptr;  // pinned
  hipEvent_t ev_count = nullptr;
  bool count_pending = false;

  // hipGraphs of the step body, split where the scalars and render flags are
  // final (after combustion); rebuilt when the recorded sequence changes.
  bool use_graph = true;
  hipGraph_t graph = nullptr, graph_b = nullptr;
  hipGraphExec_t graph_exec = nullptr, graph_exec_b = nullptr;

  // Rendering runs on its own stream, overlapped with the simulation: after
  // combustion (where density, temperature, fuel, reaction and the render flags
  // are final for the step) the simulation stream packs them into a snapshot
  // (d_render, d_bits) and records ev_scalars; the raymarch reads only the
  // snapshot, so it overlaps the rest of the step and the first half of the
  // next one. ev_render marks the last raymarch done; the next pack waits for it.
  hipStream_t rstream = nullptr;
  hipEvent_t ev_scalars = nullptr, ev_render = nullptr;
  // A batch of scenes renders them on up to kRenderLanes streams (scene k on lane
  // k % rlanes; lane 0 is rstream / ev_render), as concurrently as separate
  // simulations would; each lane has its own tile counter (d_rqueue + lane).
  static constexpr int kRenderLanes = 4;
  int rlanes = 1;
  hipStream_t rl_stream[kRenderLanes] = {};
  hipEvent_t rl_ev[kRenderLanes] = {};
  bool rl_pending[kRenderLanes] = {};
  void* d_render = nullptr;      // float4 per voxel
  bool dense_tables = false;  // dense V-cycle: level activity words and coarsest table built
  // Bloom: three half-resolution rgba16f images (bright, horizontal and vertical
  // blur) for the last size, and staging for the host-memory variant.
  void* d_bloom = nullptr;
  uint32_t bloom_w = 0, bloom_h = 0;
  void* d_bloom_io = nullptr;
  size_t bloom_io_bytes = 0;
  bool in_step = false;          // recording the step's pass sequence
  bool debug_requested = false;  // getDebugBuffers(): the next step runs the debug pass (ORCH:298-301, 316)
  uint32_t* d_rqueue = nullptr;  // raymarch tile counter
  int ray_blocks = 1;            // resident raymarch workgroups
  bool render_overlap = true;    // false: the raymarch runs on the simulation stream
  bool render_pending = false;
  bool scalars_marked = false;  // the snapshot matches all simulation work enqueued so far
  bool render_full = true;      // the next pack rewrites every brick (lists changed outside fx_step)
  bool render_wanted = true;    // a raymarch read the last snapshot: the next step packs one
  hipStream_t last_render_stream = nullptr;  // where the last raymarch was enqueued
  std::vector<int64_t> graph_key;
  int graph_last_pressure = 0;

  // Stage timing (fx_set_stage_timing): event pairs recorded around stages.
  bool timing = false;
  struct Rec { int id; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double acc_ms[FX_TIME_COUNT] = {};
  uint32_t acc_calls[FX_TIME_COUNT] = {};
};

namespace {

int fail(fx_sim* s, int code, const std::string& msg) {
  if (s) s->err = msg; else g_create_error = msg;
  return code;
}

#define HIP_OK(sim, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(sim, FX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

// Calls that address one scene's fields, flags, emitters or image: a batch of
// scenes takes the fx_scene_* forms.
#define ONE_SCENE(sim)                                                                                       \
  do {                                                                                                       \
    if ((sim)->multi) return fail(sim, FX_ERR_UNSUPPORTED, "a batch of scenes is addressed per scene (fx_scene_*)"); \
  } while (0)

int clamp_int(double v, int lo, int hi) {  // CVS/types.ts:26-28
  double r = std::floor(v + 0.5);
  return int(std::max(double(lo), std::min(double(hi), r)));
}
fx_quality clamp_quality(const fx_quality& q) {  // CVS/types.ts:30-44
  fx_quality o;
  o.pressure_interval = clamp_int(q.pressure_interval, 1, 16);
  o.vorticity_interval = clamp_int(q.vorticity_interval, 1, 16);
  o.fine_pre_iterations = clamp_int(q.fine_pre_iterations, 0, 32);
  o.fine_post_iterations = clamp_int(q.fine_post_iterations, 0, 32);
  o.mid_pre_iterations = clamp_int(q.mid_pre_iterations, 0, 32);
  o.mid_post_iterations = clamp_int(q.mid_post_iterations, 0, 32);
  o.coarse_iterations = clamp_int(q.coarse_iterations, 0, 64);
  return o;
}
uint32_t halve(uint32_t n) { return std::max<uint32_t>(4u, n / 2u); }  // CVS/buffers.ts:10-16

template <class T>
int dalloc(fx_sim* s, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return fail(s, FX_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  *p = static_cast<T*>(q);
  return FX_OK;
}

Scalars set_of(fx_sim* s, int k) {
  return Scalars{s->S[k][0], s->S[k][1], s->S[k][2], static_cast<float*>(s->S[k][3]), s->desc.precision};
}
int field_esz(const fx_sim* s, int f) { return f < 3 ? int(precision_bytes(s->desc.precision)) : 4; }
Vec3 vec_of(float* const* v) { return Vec3{v[0], v[1], v[2]}; }

int grid_for(const fx_sim* s, size_t bricks) { return int(std::max<size_t>(1, std::min<size_t>(size_t(s->blocks), bricks))); }

// Fine-level brick list used by every stage.
const uint32_t* fine_list(const fx_sim* s) { return s->sparse ? s->lv[0].L.list : s->identity; }
const uint32_t* fine_count(const fx_sim* s) { return s->sparse ? s->lv[0].L.count : s->identity_count; }
const uint32_t* level_list(const fx_sim* s, int l) { return s->lv[l].L.list; }
const uint32_t* level_count(const fx_sim* s, int l) { return s->lv[l].L.count; }

// Voxel activity of level l (sparse). Level 0 needs it too: fused V-cycle
// tiles compute halo voxels that may lie in inactive bricks.
LvMask mask_of(const fx_sim* s, int l) {
  if (!s->sparse) return LvMask{nullptr, s->g, 0u};
  return LvMask{s->flagsInj, s->g, uint32_t(l)};
}
LvBits bits_of(const fx_sim* s, int l) {
  return LvBits{s->lv[l].bits, s->sparse ? s->lv[l].L.page : nullptr, l == 0 ? s->lv[0].nbr : nullptr};
}

struct Buf { void* p; int esz; };
std::vector<Buf> fine_buffers(fx_sim* s) {
  std::vector<Buf> b;
  for (int k = 0; k < 2; ++k) for (int f = 0; f < 4; ++f) b.push_back(Buf{s->S[k][f], field_esz(s, f)});
  for (int c = 0; c < 3; ++c) { b.push_back(Buf{s->u[c], 4}); b.push_back(Buf{s->us[c], 4}); b.push_back(Buf{s->w[c], 4}); }
  for (float* p : {s->wmag, s->cmag, s->umag, s->lv[0].div, s->lv[0].p[0], s->lv[0].p[1]}) b.push_back(Buf{p, 4});
  return b;
}

// ORCH:236-266 — computed in double, stored as f32; `rt` and `g` are those of the
// scene the constants are for.
fx_sim_params make_params(const fx_sim* s, const fx_runtime_params& rt, const Grid& g, double elapsed, double stepS) {
  fx_sim_params p;
  double eff = std::min(stepS, 1.0 / 15.0);
  p.time = float(elapsed - s->startS);
  p.delta_time = float(eff);
  p.previous_time = float(s->lastS - s->startS);
  p.scalar_set_index = float(s->curSet);
  p.wind[0] = rt.wind[0]; p.wind[1] = rt.wind[1]; p.wind[2] = rt.wind[2]; p.wind[3] = rt.wind_strength;
  p.gravity[0] = rt.gravity[0]; p.gravity[1] = rt.gravity[1]; p.gravity[2] = rt.gravity[2];
  p.gravity[3] = rt.gravity_strength;
  p.buoyancy = rt.buoyancy;
  p.vorticity_strength = rt.vorticity_strength;
  p.dx = float(double(rt.world_size) / double(std::max(g.W, g.D)));
  p.world_size = rt.world_size;
  p.vorticity_mask[0] = rt.vorticity_constant_mask; p.vorticity_mask[1] = rt.vorticity_velocity_mask;
  p.vorticity_mask[2] = rt.vorticity_heat_mask; p.vorticity_mask[3] = rt.vorticity_density_mask;
  return p;
}
fx_sim_params make_params(const fx_sim* s, double elapsed, double stepS) { return make_params(s, s->rt, s->g, elapsed, stepS); }

// Sparse injection bricks: every brick overlapping the voxel bounding box (+1
// voxel) of a gate-active source's influence sphere (burst: outer limit,
// inject-source-common.wgsl.ts:70-77).
// g: the scene's fine grid (brick ids are offset by its base); src/nsrc its emitters.
void injection_bricks(const Grid& g, const std::vector<float>& src, uint32_t nsrc, const fx_sim_params& p,
                      std::vector<uint8_t>& mark, std::vector<uint32_t>& out) {
  uint32_t ns = std::max<uint32_t>(nsrc, 1);
  for (uint32_t i = 0; i < ns; ++i) {
    const float* e = &src[size_t(i) * FX_FLOATS_PER_EMITTER];
    float age = p.time - e[4];
    if (!(age >= 0.f && e[5] >= age)) continue;
    float r = e[3];
    if (e[28] == 3.f) {
      float vw = p.dx / std::max(p.world_size, 0.0001f);
      r = r + std::max(r * 0.24f, vw * 3.f);
    }
    if (!(r > 0.f)) continue;
    const uint32_t n[3] = {g.W, g.H, g.D};
    int lo[3], hi[3];
    bool empty = false;
    for (int a = 0; a < 3; ++a) {
      double l = std::floor((double(e[a]) - r) * n[a] - 0.5) - 1.0;
      double h = std::ceil((double(e[a]) + r) * n[a] - 0.5) + 1.0;
      l = std::max(l, 0.0); h = std::min(h, double(n[a]) - 1.0);
      if (l > h) { empty = true; break; }
      lo[a] = int(l) >> 3; hi[a] = int(h) >> 3;
    }
    if (empty) continue;
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          uint32_t b = g.base + uint32_t(x) + g.BX * (uint32_t(y) + g.BY * uint32_t(z));
          if (!mark[b]) { mark[b] = 1; out.push_back(b); }
        }
  }
}

// Uploads the step constants (and, sparse, the candidate list) through a pinned
// staging slot; the slot is reused only after its previous copy completed.
// p: the constants of every scene (one for a single scene). Scene 0's inj_count
// and cand_count are the lengths of the whole candidate list.
int upload_step(fx_sim* s, const std::vector<fx_sim_params>& p, const std::vector<uint32_t>* cand, uint32_t inj_count) {
  StageSlot& slot = s->slots[s->slot_next];
  s->slot_next = (s->slot_next + 1) % kStageSlots;
  if (slot.used) HIP_OK(s, hipEventSynchronize(slot.done));
  uint32_t src_off = 0;
  for (size_t k = 0; k < p.size(); ++k) {
    StepConst& c = slot.c[k];
    const uint32_t nsrc = s->multi ? s->scenes[k].nsrc : s->nsrc;
    c.p = p[k];
    c.nsrc = std::max<uint32_t>(nsrc, 1);
    c.ignition_lo = s->desc.ignition_lo; c.ignition_hi = s->desc.ignition_hi; c.burn_rate = s->desc.burn_rate;
    c.act_density = s->desc.activation_density; c.act_speed = s->desc.activation_speed;
    c.omega = s->desc.vcycle.omega;
    c.inj_count = inj_count;
    c.cand_count = cand ? uint32_t(cand->size()) : 0u;
    c.step_key = uint32_t(s->stepIndex);
    c.sr = s->stochastic ? 1u : 0u;
    c.src_off = src_off;
    src_off += c.nsrc * FX_FLOATS_PER_EMITTER;
  }
  size_t bytes = sizeof(StepConst) * p.size();
  if (cand && !cand->empty()) {
    std::memcpy(slot.cand, cand->data(), cand->size() * 4);
    bytes = s->const_bytes + cand->size() * 4;
  }
  HIP_OK(s, hipMemcpyAsync(s->d_const, slot.c, bytes, hipMemcpyHostToDevice, s->stream));
  HIP_OK(s, hipEventRecord(slot.done, s->stream));
  slot.used = true;
  return FX_OK;
}

bool debug_each_frame(const fx_sim* s) {
  if (s->multi || s->desc.debug_fields == 0) return false;
  if (s->desc.debug_fields == 1) return true;
  return s->nv < 1800000ull;  // ORCH:193 (voxelCount)
}

// Brackets the launches of one stage with timing events (when enabled).
struct StageTimer {
  fx_sim* s;
  int id;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  static hipEvent_t take(fx_sim* s) {
    if (!s->pool.empty()) { hipEvent_t e = s->pool.back(); s->pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  StageTimer(fx_sim* sim, int stage, hipStream_t on = nullptr) : s(sim), id(stage), st(on ? on : sim->stream) {
    if (!s->timing) return;
    a = take(s); b = take(s);
    if (a) (void)hipEventRecord(a, st);
  }
  ~StageTimer() {
    if (!s->timing || !a || !b) return;
    (void)hipEventRecord(b, st);
    s->recs.push_back(fx_sim::Rec{id, a, b});
  }
};

// Folds completed event pairs into the accumulators (stream must be idle).
void drain_timing(fx_sim* s) {
  for (auto& r : s->recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess && r.id >= 0 && r.id < FX_TIME_COUNT) {
      s->acc_ms[r.id] += ms;
      s->acc_calls[r.id] += 1;
    }
    s->pool.push_back(r.a);
    s->pool.push_back(r.b);
  }
  s->recs.clear();
}

// ------------------------------------------------------------- passes --
void pass_activate_render(fx_sim* s) {
  StageTimer timer(s, FX_PASS_ACTIVATE_RENDER);
  if (s->sparse) {
    // Candidates: last step's active list + injection/host-written bricks. Inside a
    // step the injection launch cleared the flags.
    if (!s->in_step) launch_clear_u32(s->stream, s->flagsInj, s->nb, 0u);
    launch_activate(s->stream, grid_for(s, s->nb), s->lv[0].L.list, s->lv[0].L.count, s->d_cand, &s->d_const->cand_count,
                    s->g, s->d_const, s->desc.activation_rule, set_of(s, 0), vec_of(s->u), s->flagsInj, s->lflags);
    launch_compact(s->stream, s->cargs);
    RetireArgs R{};
    R.nlevels = int(s->lv.size());
    int nb = 0;
    for (int l = 0; l < R.nlevels; ++l) {
      R.list[l] = s->lv[l].L.retire;
      R.count[l] = s->lv[l].L.retire_count;
      R.first[l] = nb;
      if (l == 0) {
        for (const Buf& b : fine_buffers(s)) { R.esz[nb] = uint8_t(b.esz); R.b[nb++] = b.p; }
      } else {
        for (float* p : {s->lv[l].div, s->lv[l].p[0], s->lv[l].p[1]}) { R.esz[nb] = 4; R.b[nb++] = p; }
      }
    }
    R.first[R.nlevels] = nb;
    launch_retire(s->stream, grid_for(s, s->nb), R);
  } else {
    if (!s->in_step) launch_clear_u32(s->stream, s->flagsInj, s->nb, 0u);  // ORCH:271 (in a step: by the injection launch)
    launch_activate(s->stream, grid_for(s, s->nb), s->identity, s->identity_count, nullptr, nullptr, s->g, s->d_const,
                    s->desc.activation_rule, set_of(s, 0), vec_of(s->u), s->flagsInj, s->lflags);
  }
}
void pass_activate_advected(fx_sim* s) {  // ORCH:280-281 (dense only)
  StageTimer timer(s, FX_PASS_ACTIVATE_ADVECTED);
  launch_clear_u32(s->stream, s->flagsRender, s->nb, 0u);
  launch_activate(s->stream, grid_for(s, s->nb), s->identity, s->identity_count, nullptr, nullptr, s->g, s->d_const,
                  s->desc.activation_rule, set_of(s, 1), vec_of(s->us), s->flagsRender, s->lflags);
}
void pass_inject(fx_sim* s, int parts) {
  StageTimer timer(s, parts == 2 ? FX_PASS_INJECT_VELOCITY : parts == 4 ? FX_PASS_INJECT_IGNITER : FX_PASS_INJECT_SCALAR);
  if (s->sparse)
    launch_inject(s->stream, grid_for(s, s->nb), s->d_cand, &s->d_const->inj_count, s->g, s->d_const, s->d_src, set_of(s, 0),
                  vec_of(s->u), s->touched, parts, s->in_step ? s->flagsInj : nullptr,
                  uint32_t(s->nb));
  else
    launch_inject(s->stream, grid_for(s, s->nb), s->identity, s->identity_count, s->g, s->d_const, s->d_src, set_of(s, 0),
                  vec_of(s->u), nullptr, parts, s->in_step ? s->flagsInj : nullptr,
                  uint32_t(s->nb));
}
void zero_level_p(fx_sim* s, int l, int which) {
  float* b[1] = {s->lv[l].p[which]};
  launch_zero_bricks(s->stream, grid_for(s, s->lv[l].nb), level_list(s, l), level_count(s, l), b, 1);
}
void clear_pressure(fx_sim* s) {  // PressureSolvePass.ts:62-66
  for (size_t l = 0; l < s->lv.size(); ++l) zero_level_p(s, int(l), 0);
}
int smooth_ref(fx_sim* s, int l, int iters, int start) {  // PressureSolvePass.ts:129-143
  int a = start;
  Level& L = s->lv[l];
  for (int k = 0; k < iters; ++k) {
    StageTimer timer(s, l == 0 ? FX_TIME_JACOBI_FINE : -1);
    launch_jacobi(s->stream, grid_for(s, L.nb), level_list(s, l), level_count(s, l), L.g, mask_of(s, l), s->d_const, L.div, L.p[a], L.p[1 - a], 0);
    a = 1 - a;
  }
  return a;
}
int solve_reference(fx_sim* s) {  // PressureSolvePass.ts:68-84 with its buffer aliasing
  Level& F = s->lv[0]; Level& M = s->lv[1]; Level& C = s->lv[2];
  int finePre = smooth_ref(s, 0, s->q.fine_pre_iterations, 0);
  launch_residual(s->stream, grid_for(s, F.nb), level_list(s, 0), level_count(s, 0), F.g, mask_of(s, 0), F.div, F.p[finePre], F.p[1 - finePre], nullptr);
  launch_restrict(s->stream, grid_for(s, M.nb), level_list(s, 1), level_count(s, 1), F.g, M.g, mask_of(s, 1), F.div, M.div, 0);
  int midPre = smooth_ref(s, 1, s->q.mid_pre_iterations, 0);
  launch_residual(s->stream, grid_for(s, M.nb), level_list(s, 1), level_count(s, 1), M.g, mask_of(s, 1), M.div, M.p[midPre], M.p[1 - midPre], nullptr);
  launch_restrict(s->stream, grid_for(s, C.nb), level_list(s, 2), level_count(s, 2), M.g, C.g, mask_of(s, 2), M.div, C.div, 0);
  int co = smooth_ref(s, 2, s->q.coarse_iterations, 0);
  launch_prolong(s->stream, grid_for(s, M.nb), level_list(s, 1), level_count(s, 1), C.g, M.g, mask_of(s, 1), C.p[co], M.p[0]);
  int mid = smooth_ref(s, 1, s->q.mid_post_iterations, midPre);
  launch_prolong(s->stream, grid_for(s, F.nb), level_list(s, 0), level_count(s, 0), M.g, F.g, mask_of(s, 0), M.p[mid], F.p[0]);
  return smooth_ref(s, 0, s->q.fine_post_iterations, finePre);
}
// Weighted-Jacobi V-cycle; cur[l] = buffer holding the current iterate.
int solve_vcycle(fx_sim* s) {
  const int L = int(s->lv.size());
  const fx_vcycle& v = s->desc.vcycle;
  std::vector<int> cur(L, 0);
  {
    LevelMaskArgs A{};
    A.nlevels = L;
    for (int l = 0; l < L; ++l) {
      A.list[l] = level_list(s, l);
      A.count[l] = level_count(s, l);
      A.g[l] = s->lv[l].g;
      A.m[l] = mask_of(s, l);
      A.bits[l] = s->lv[l].bits;
    }
    // the FP64 residual-norm accumulators of this solve start at zero
    const int ncyc = std::min(v.cycles, s->norm_cap);
    A.zero = ncyc > 0 ? reinterpret_cast<uint32_t*>(s->d_norms) : nullptr;
    A.zero_words = size_t(std::max(ncyc, 0)) * kNormLanes * 2;
    A.nbr0 = s->lv[0].nbr;
    A.page0 = s->sparse ? s->lv[0].L.page : nullptr;
    // A dense domain's levels list every brick and activity is geometric, so the
    // activity words and the coarsest slot table (whose inactive voxels the
    // sweeps never write) are built once; later solves only clear the norms.
    if (s->sparse || !s->dense_tables)
      launch_level_masks(s->stream, grid_for(s, s->lv[0].nb), A);
    else if (A.zero)
      launch_clear_u32(s->stream, A.zero, A.zero_words, 0u);
  }
  Level& LC = s->lv[L - 1];
  const uint32_t nsc = s->multi ? uint32_t(s->scenes.size()) : 1u;
  const size_t coarse_nv = s->multi ? s->coarse_nv_max : LC.nv;  // one scene's coarsest storage
  const bool block_smoother = L > 1 && coarse_nv <= kBlockSmoothVoxels && !s->coarse_split;
  if (block_smoother && (s->sparse || !s->dense_tables))
    launch_coarse_setup(s->stream, level_list(s, L - 1), level_count(s, L - 1), LC.g, bits_of(s, L - 1), LC.tab, LC.tab_n, LC.p[0],
                        LC.p[1], nsc);
  if (!s->sparse) s->dense_tables = true;
  // n sweeps at level l, `per` (1, 2 or 4) per launch; zero_start: the first sweep
  // starts from p = 0 (a coarse level's correction), so neither a clear nor a read
  // of p is needed.
  auto smoothW = [&](int l, int n, bool zero_start, int per) {
    Level& V = s->lv[l];
    if (n == 0) {
      if (zero_start) { cur[l] = 0; zero_level_p(s, l, 0); }
      return;
    }
    if (zero_start) cur[l] = 0;
    if (l == L - 1 && block_smoother) {
      launch_smooth_block(s->stream, level_list(s, l), level_count(s, l), V.g, mask_of(s, l), s->d_const, V.div, V.p[cur[l]],
                          V.p[1 - cur[l]], n, 1, zero_start ? 1 : 0, V.tab, V.tab_n, nsc);
      if (n & 1) cur[l] = 1 - cur[l];
      return;
    }
    // compute tile of the level's split sweeps (plan bits kPlanTileShift + 2l of the up mask)
    const uint32_t tile = l < L - 1 ? (s->split_up >> (kPlanTileShift + 2 * l)) & 3u : 0u;
    for (int k = 0; k < n; ++k) {
      const float* src = (zero_start && k == 0) ? nullptr : V.p[cur[l]];
      if (tile == 1u || tile == 2u) {
        const int ns = (per >= 2 && k + 1 < n) ? 2 : 1;
        {
          StageTimer timer(s, l == 0 ? FX_TIME_JACOBI_FINE : -1);
          launch_sweep_tile(s->stream, grid_for(s, V.nb), tile == 1u ? 4 : 16, ns, level_list(s, l), level_count(s, l), V.g,
                            bits_of(s, l), s->d_const, V.div, src, V.p[1 - cur[l]]);
        }
        cur[l] = 1 - cur[l];
        k += ns - 1;
        continue;
      }
      if (per >= 4 && k + 3 < n) {
        launch_sweep_quad(s->stream, grid_for(s, V.nb), level_list(s, l), level_count(s, l), V.g, bits_of(s, l), s->d_const, V.div,
                          src, V.p[1 - cur[l]]);
        cur[l] = 1 - cur[l];
        k += 3;
        continue;
      }
      if (per >= 2 && k + 1 < n) {
        launch_sweep_pair(s->stream, grid_for(s, V.nb), level_list(s, l), level_count(s, l), V.g, bits_of(s, l), s->d_const, V.div,
                          src, V.p[1 - cur[l]]);
        cur[l] = 1 - cur[l];
        ++k;
        continue;
      }
      // fine sweeps reading a pressure are timed by the kernel's own start/end timestamps
      const bool exact = s->timing && l == 0 && src != nullptr && !s->multi;
      hipEvent_t ka = exact ? StageTimer::take(s) : nullptr, kb = exact ? StageTimer::take(s) : nullptr;
      {
        StageTimer timer(s, (l == 0 && !(ka && kb)) ? FX_TIME_JACOBI_FINE : -1);
        launch_sweep(s->stream, grid_for(s, V.nb), level_list(s, l), level_count(s, l), V.g, bits_of(s, l), s->d_const, V.div, src,
                     V.p[1 - cur[l]], l == 0, ka, kb);
      }
      if (ka && kb) s->recs.push_back(fx_sim::Rec{FX_TIME_JACOBI_FINE, ka, kb});
      cur[l] = 1 - cur[l];
    }
  };
  s->norm_count = std::min(v.cycles, s->norm_cap);
  for (int c = 0; c < v.cycles; ++c) {
    for (int l = 0; l < L - 1; ++l) {
      Level& V = s->lv[l];
      Level& W = s->lv[l + 1];
      double* norm = (l == 0 && c < s->norm_cap) ? s->d_norms + size_t(c) * kNormLanes : nullptr;
      // Level 0 starts each solve from p = 0; coarse levels start every cycle from 0.
      const bool zero = l > 0 || c == 0;
      if (zero) cur[l] = 0;
      const bool split = (s->split_down >> l) & 1u;
      if (split && ((s->split_down >> (kPlanFuseShift + l)) & 1u) && v.pre_sweeps >= 2) {
        // split leg whose last pre-sweep is fused with the residual and the restriction
        smoothW(l, v.pre_sweeps - 1, zero, ((s->split_down >> (kPlanPairShift + l)) & 1u) ? 2 : 1);
        if (launch_vdown(s->stream, grid_for(s, V.nb), 1, level_list(s, l), level_count(s, l), V.g, bits_of(s, l), W.g,
                         mask_of(s, l + 1), s->d_const, V.div, V.p[cur[l]], V.p[1 - cur[l]], W.div, norm)) {
          cur[l] = 1 - cur[l];
          continue;
        }
      }
      if (!split && launch_vdown(s->stream, grid_for(s, V.nb), v.pre_sweeps, level_list(s, l), level_count(s, l), V.g,
                                 bits_of(s, l), W.g, mask_of(s, l + 1), s->d_const, V.div, zero ? nullptr : V.p[cur[l]],
                                 V.p[1 - cur[l]], W.div, norm)) {
        cur[l] = 1 - cur[l];
        continue;
      }
      // split leg: lean sweeps, residual into the other pressure buffer, restriction
      smoothW(l, v.pre_sweeps, zero, ((s->split_down >> (kPlanPairShift + l)) & 1u) ? 2 : 1);
      if (zero && v.pre_sweeps == 0) zero_level_p(s, l, cur[l]);
      launch_residual_lean(s->stream, grid_for(s, V.nb), level_list(s, l), level_count(s, l), V.g, bits_of(s, l), V.div,
                           V.p[cur[l]], V.p[1 - cur[l]], norm);
      launch_restrict_lean(s->stream, grid_for(s, W.nb), level_list(s, l + 1), level_count(s, l + 1), V.g, W.g, bits_of(s, l + 1),
                           V.p[1 - cur[l]], W.div);
    }
    // coarsest: the one-workgroup smoother, or (beyond its capacity) multi-workgroup
    // launches of 1, 2 or 4 sweeps each
    const uint32_t cper = (s->split_down >> kPlanCoarseShift) & 3u;
    smoothW(L - 1, v.coarse_sweeps, true, cper == 2 ? 4 : cper == 1 ? 2 : 1);
    for (int l = L - 2; l >= 0; --l) {
      Level& V = s->lv[l];
      const uint32_t* lst = level_list(s, l);
      const uint32_t* cnt = level_count(s, l);
      const float* e = s->lv[l + 1].p[cur[l + 1]];
      if (((s->split_up >> l) & 1u) && ((s->split_up >> (kPlanFuseShift + l)) & 1u) && v.post_sweeps >= 1 &&
          launch_vup(s->stream, grid_for(s, V.nb), 1, lst, cnt, s->lv[l + 1].g, V.g, bits_of(s, l), s->d_const, e, V.div,
                     V.p[cur[l]], V.p[1 - cur[l]])) {
        // split leg whose prolongation is fused with the first post-sweep, then lean sweeps
        cur[l] = 1 - cur[l];
        smoothW(l, v.post_sweeps - 1, false, ((s->split_up >> (kPlanPairShift + l)) & 1u) ? 2 : 1);
        continue;
      }
      if ((s->split_up >> l) & 1u) {
        // split leg: in-place prolongation, then lean sweeps
        launch_prolong_add(s->stream, grid_for(s, V.nb), lst, cnt, s->lv[l + 1].g, V.g, bits_of(s, l), e, V.p[cur[l]]);
        smoothW(l, v.post_sweeps, false, ((s->split_up >> (kPlanPairShift + l)) & 1u) ? 2 : 1);
        continue;
      }
      if (launch_vup(s->stream, grid_for(s, V.nb), v.post_sweeps, lst, cnt, s->lv[l + 1].g, V.g, bits_of(s, l), s->d_const, e,
                     V.div, V.p[cur[l]], V.p[1 - cur[l]])) {
        cur[l] = 1 - cur[l];
        continue;
      }
      launch_prolong(s->stream, grid_for(s, V.nb), lst, cnt, s->lv[l + 1].g, V.g, mask_of(s, l), e, V.p[cur[l]]);
      smoothW(l, v.post_sweeps, false, 1);
    }
  }
  return cur[0];
}
// PressureSolvePass.solve (the clear is recorded separately, ORCH:291).
int pressure_solve(fx_sim* s) {
  StageTimer timer(s, FX_PASS_PRESSURE);
  if (s->desc.solver == FX_SOLVER_VCYCLE) return solve_vcycle(s);
  return solve_reference(s);
}

void run_stage(fx_sim* s, int pass) {
  StageTimer timer(s, pass);
  const uint32_t* list = fine_list(s);
  const uint32_t* count = fine_count(s);
  const int blocks = grid_for(s, s->nb);
  switch (pass) {
    case FX_PASS_ADVECT_VELOCITY:
      launch_advect_velocity(s->stream, blocks, list, count, s->g, s->d_const, vec_of(s->u), vec_of(s->us));
      break;
    case FX_PASS_ADVECT_SCALARS:
      launch_advect_scalars(s->stream, blocks, list, count, s->g, s->d_const, s->mode, set_of(s, 0), vec_of(s->us), set_of(s, 1),
                            s->flagsInj);
      break;
    case FX_PASS_COMBUSTION:
      launch_combustion(s->stream, blocks, list, count, s->g, s->d_const, set_of(s, 1), set_of(s, 0), s->flagsRender);
      break;
    case FX_PASS_BUOYANCY:
      launch_buoyancy(s->stream, blocks, list, count, s->g, s->d_const, set_of(s, 0), vec_of(s->us), s->flagsRender);
      break;
    case FX_PASS_VORTICITY:
      launch_vorticity(s->stream, blocks, list, count, s->g, vec_of(s->us), vec_of(s->w), s->wmag);
      break;
    case FX_PASS_CONFINEMENT:
      launch_confinement(s->stream, blocks, list, count, s->g, s->d_const, vec_of(s->w), s->wmag, vec_of(s->us), s->cmag, set_of(s, 0));
      break;
    case FX_PASS_DIVERGENCE:
      launch_divergence(s->stream, blocks, list, count, s->g, vec_of(s->us), s->lv[0].div);
      break;
    case FX_PASS_PROJECTION:
      launch_projection(s->stream, blocks, list, count, s->g, vec_of(s->us), s->lv[0].p[s->lastPressure], vec_of(s->u));
      break;
    case FX_PASS_DEBUG_FIELDS:
      launch_debug(s->stream, blocks, list, count, s->g, vec_of(s->u), s->umag);
      break;
    default:
      break;
  }
}

// Candidate list for the sparse activation: injection bricks then host-written bricks.
void build_candidates(fx_sim* s, const std::vector<fx_sim_params>& p, std::vector<uint32_t>& cand, uint32_t& inj_count) {
  std::vector<uint8_t> mark(s->nb, 0);
  if (s->multi) {
    for (size_t k = 0; k < s->scenes.size(); ++k)
      injection_bricks(s->scenes[k].g[0], s->scenes[k].h_src, s->scenes[k].nsrc, p[k], mark, cand);
  } else {
    injection_bricks(s->g, s->h_src, s->nsrc, p[0], mark, cand);
  }
  inj_count = uint32_t(cand.size());
  for (uint32_t b : s->pending)
    if (!mark[b]) { mark[b] = 1; cand.push_back(b); }
  s->pending.clear();
}

// Device buffer of a field and its element size in bytes (4, or 2 for d/t/f in 16-bit storage).
void* field_ptr(fx_sim* s, int id, int* esz = nullptr) {
  if (esz) *esz = 4;
  switch (id) {
    case FX_FIELD_DENSITY: case FX_FIELD_TEMPERATURE: case FX_FIELD_FUEL: case FX_FIELD_REACTION:
      if (esz) *esz = field_esz(s, id);
      return s->S[0][id];
    case FX_FIELD_VELOCITY_X: case FX_FIELD_VELOCITY_Y: case FX_FIELD_VELOCITY_Z: return s->u[id - FX_FIELD_VELOCITY_X];
    case FX_FIELD_SCRATCH_X: case FX_FIELD_SCRATCH_Y: case FX_FIELD_SCRATCH_Z: return s->us[id - FX_FIELD_SCRATCH_X];
    case FX_FIELD_VORTICITY_X: case FX_FIELD_VORTICITY_Y: case FX_FIELD_VORTICITY_Z: return s->w[id - FX_FIELD_VORTICITY_X];
    case FX_FIELD_VORTICITY_MAG: return s->wmag;
    case FX_FIELD_CONFINEMENT_MAG: return s->cmag;
    case FX_FIELD_VELOCITY_MAG: return s->umag;
    case FX_FIELD_DIVERGENCE: return s->lv[0].div;
    case FX_FIELD_PRESSURE: return s->lv[0].p[s->lastPressure];
    case FX_FIELD_ADV_DENSITY: case FX_FIELD_ADV_TEMPERATURE: case FX_FIELD_ADV_FUEL: case FX_FIELD_ADV_REACTION:
      if (esz) *esz = field_esz(s, id - FX_FIELD_ADV_DENSITY);
      return s->S[1][id - FX_FIELD_ADV_DENSITY];
    default: return nullptr;
  }
}

// brick-major <-> row-major
void gather_rowmajor(const Grid& g, const float* bm, float* rm) {
  for (uint32_t z = 0; z < g.D; ++z)
    for (uint32_t y = 0; y < g.H; ++y) {
      float* row = rm + (size_t(z) * g.H + y) * g.W;
      for (uint32_t x = 0; x < g.W; ++x) {
        size_t b = (x >> 3) + size_t(g.BX) * ((y >> 3) + size_t(g.BY) * (z >> 3));
        row[x] = bm[(b << 9) | ((z & 7u) << 6) | ((y & 7u) << 3) | (x & 7u)];
      }
    }
}
void scatter_brickmajor(const Grid& g, const float* rm, float* bm) {
  for (uint32_t z = 0; z < g.D; ++z)
    for (uint32_t y = 0; y < g.H; ++y) {
      const float* row = rm + (size_t(z) * g.H + y) * g.W;
      for (uint32_t x = 0; x < g.W; ++x) {
        size_t b = (x >> 3) + size_t(g.BX) * ((y >> 3) + size_t(g.BY) * (z >> 3));
        bm[(b << 9) | ((z & 7u) << 6) | ((y & 7u) << 3) | (x & 7u)] = row[x];
      }
    }
}

// Host conversions of the 16-bit storage formats (round to nearest even).
float bf16_to_f32(uint16_t h) { uint32_t u = uint32_t(h) << 16; float f; std::memcpy(&f, &u, 4); return f; }
uint16_t f32_to_bf16(float f) {
  uint32_t u; std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
float f16_to_f32(uint16_t h) {
  uint32_t sign = uint32_t(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu, u;
  if (e == 0) {
    if (m == 0) u = sign;
    else { e = 127 - 14; while (!(m & 0x400u)) { m <<= 1; --e; } m &= 0x3ffu; u = sign | (e << 23) | (m << 13); }
  } else if (e == 31) u = sign | 0x7f800000u | (m << 13);
  else u = sign | ((e + 112) << 23) | (m << 13);
  float f; std::memcpy(&f, &u, 4); return f;
}
uint16_t f32_to_f16(float f) {
  uint32_t u; std::memcpy(&u, &f, 4);
  uint32_t sign = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return uint16_t(sign | 0x7e00u);
  if (a >= 0x477ff000u) return uint16_t(sign | 0x7c00u);         // rounds to infinity
  if (a < 0x38800000u) {                                          // subnormal or zero half
    float af; std::memcpy(&af, &a, 4);
    float scaled = af * 16777216.f;                               // 2^24: half subnormal units
    uint32_t q = uint32_t(std::nearbyint(scaled));                // round half to even
    return uint16_t(sign | q);
  }
  uint32_t mant = a & 0x7fffffu, exp = (a >> 23) - 112;
  uint32_t half = (exp << 10) | (mant >> 13);
  uint32_t rest = mant & 0x1fffu;
  if (rest > 0x1000u || (rest == 0x1000u && (half & 1u))) ++half;
  return uint16_t(sign | half);
}

// 16-bit unsigned fixed point: value = u * 2^-16, clamped to [0, 1 - 2^-16].
float u16_to_f32(uint16_t h) { return float(h) * (1.f / 65536.f); }
uint16_t f32_to_u16(float f) {
  const float c = f > 0.f ? (f < 1.f ? f : 1.f) : 0.f;  // NaN -> 0
  return uint16_t(std::fmin(std::nearbyint(c * 65536.f), 65535.f));
}
// Unsigned float, 4 exponent / 12 mantissa bits: codes >= 4096 are the f32 bits
// (c + 0x70000) << 11, smaller codes c * 2^-26; values clamped to [0, 1].
float uf16_to_f32(uint16_t h) {
  if (h < 4096u) return float(h) * (1.f / 67108864.f);
  const uint32_t u = (uint32_t(h) + 0x70000u) << 11;
  float f; std::memcpy(&f, &u, 4); return f;
}
uint16_t f32_to_uf16(float f) {
  const float c = f > 0.f ? (f < 1.f ? f : 1.f) : 0.f;  // NaN -> 0
  uint32_t b; std::memcpy(&b, &c, 4);
  if (b < 0x38800000u) return uint16_t(std::nearbyint(c * 67108864.f));
  b += 0x3ffu + ((b >> 11) & 1u);
  return uint16_t((b >> 11) - 0x70000u);
}
// The 16-bit storage format of a simulation, decoded / encoded.
float load16(int prec, uint16_t h) {
  return prec == FX_PRECISION_BF16 ? bf16_to_f32(h) : prec == FX_PRECISION_U16 ? u16_to_f32(h)
       : prec == FX_PRECISION_UF16 ? uf16_to_f32(h) : f16_to_f32(h);
}
uint16_t store16(int prec, float f) {
  return prec == FX_PRECISION_BF16 ? f32_to_bf16(f) : prec == FX_PRECISION_U16 ? f32_to_u16(f)
       : prec == FX_PRECISION_UF16 ? f32_to_uf16(f) : f32_to_f16(f);
}

int read_brickmajor_typed(fx_sim* s, const Grid& g, const void* dev, int esz, float* host, size_t bytes) {
  size_t nv = size_t(g.W) * g.H * g.D;
  if (bytes != nv * 4) return fail(s, FX_ERR_INVALID_ARG, "buffer size mismatch");
  const size_t n = storage_count(g);
  std::vector<float> tmp(n);
  HIP_OK(s, hipStreamSynchronize(s->stream));
  if (esz == 4) {
    HIP_OK(s, hipMemcpy(tmp.data(), dev, n * 4, hipMemcpyDeviceToHost));
  } else {
    std::vector<uint16_t> h(n);
    HIP_OK(s, hipMemcpy(h.data(), dev, n * 2, hipMemcpyDeviceToHost));
    const int prec = s->desc.precision;
    for (size_t i = 0; i < n; ++i) tmp[i] = load16(prec, h[i]);
  }
  gather_rowmajor(g, tmp.data(), host);
  return FX_OK;
}

int read_brickmajor(fx_sim* s, const Grid& g, const float* dev, float* host, size_t bytes) {
  size_t nv = size_t(g.W) * g.H * g.D;
  if (bytes != nv * 4) return fail(s, FX_ERR_INVALID_ARG, "buffer size mismatch");
  std::vector<float> tmp(storage_count(g));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  HIP_OK(s, hipMemcpy(tmp.data(), dev, tmp.size() * 4, hipMemcpyDeviceToHost));
  gather_rowmajor(g, tmp.data(), host);
  return FX_OK;
}

// The pass sequence of one step (ORCH:270-301) on the simulation stream, in
// two parts: (a) up to combustion, after which the render scalars and flags are
// final, and (b) the rest, which reads them but writes only velocity, vorticity,
// pressure and diagnostics.
void record_step_a(fx_sim* s) {
  s->in_step = true;
  pass_inject(s, 7);                                   // ORCH:270
  pass_activate_render(s);                             // ORCH:271-272
  run_stage(s, FX_PASS_ADVECT_VELOCITY);               // ORCH:273
  run_stage(s, FX_PASS_ADVECT_SCALARS);                // ORCH:274-279
  if (!s->sparse) pass_activate_advected(s);           // ORCH:280-281
  run_stage(s, FX_PASS_COMBUSTION);                    // ORCH:282
  s->in_step = false;
}
void record_step_b(fx_sim* s, bool vort, bool pres, bool dbg) {
  run_stage(s, FX_PASS_BUOYANCY);                      // ORCH:283
  if (vort) {                                          // ORCH:285-288
    run_stage(s, FX_PASS_VORTICITY);
    run_stage(s, FX_PASS_CONFINEMENT);
  }
  if (pres) {                                          // ORCH:290-294
    // PressureSolvePass.clear (ORCH:291); the V-cycle starts from p = 0 in its
    // first sweep and needs no clear.
    if (s->desc.solver != FX_SOLVER_VCYCLE) clear_pressure(s);
    run_stage(s, FX_PASS_DIVERGENCE);
    s->lastPressure = pressure_solve(s);
  }
  run_stage(s, FX_PASS_PROJECTION);                    // ORCH:296
  if (dbg) run_stage(s, FX_PASS_DEBUG_FIELDS);         // ORCH:298-301
  if (s->multi && s->h_scene_count) {                  // for the next steps' coarsest-level plan
    launch_scene_counts(s->stream, s->lv[0].L.list, s->lv[0].L.count, s->lv[0].g.tab, uint32_t(s->scenes.size()),
                        s->d_scene_count);
    (void)hipMemcpyAsync(s->h_scene_count, s->d_scene_count, s->scenes.size() * 4, hipMemcpyDeviceToHost, s->stream);
  } else if (s->sparse && s->h_fine_count) {
    (void)hipMemcpyAsync(s->h_fine_count, s->lv[0].L.count, 4, hipMemcpyDeviceToHost, s->stream);
  }
}

// Upper bound of the coarsest level's active voxels: a level-l voxel J lies in
// fine brick (J << l) >> 3, so a fine brick holds 8^(3-l) level-l voxels for
// l <= 3 and at least one for l > 3.
uint64_t coarse_active_bound(const fx_sim* s, uint64_t fine_bricks) {
  const int l = int(s->lv.size()) - 1;
  uint64_t per = 1;
  for (int k = l; k < 3; ++k) per *= 8;
  return fine_bricks * per;
}
// Chooses the coarsest-level plan before a step (see fx_sim::coarse_split).
void update_coarse_plan(fx_sim* s) {
  if (s->desc.solver != FX_SOLVER_VCYCLE || s->lv.size() < 2) return;
  const Level& LC = s->lv.back();
  if (!s->sparse) {  // every interior voxel is active
    const uint64_t W = LC.g.W > 2 ? LC.g.W - 2 : 0, H = LC.g.H > 2 ? LC.g.H - 2 : 0, D = LC.g.D > 2 ? LC.g.D - 2 : 0;
    s->coarse_split = W * H * D > kCoarseCap;
    return;
  }
  if (s->count_pending && hipEventQuery(s->ev_count) == hipSuccess) {
    s->count_pending = false;
    uint64_t fine = s->multi ? 0 : *s->h_fine_count;  // the largest scene's count bounds every scene's smoother
    if (s->multi)
      for (size_t k = 0; k < s->scenes.size(); ++k) fine = std::max<uint64_t>(fine, s->h_scene_count[k]);
    const uint64_t bound = coarse_active_bound(s, fine);
    // hysteresis: a set hovering near the cap does not recapture the graph every step
    if (bound > kCoarseCap) s->coarse_split = true;
    else if (bound * 8 < uint64_t(kCoarseCap) * 7) s->coarse_split = false;
  }
}

// Captures the two step parts into hipGraphs (kernel nodes only: every size the
// kernels need is read from device memory, so one pair serves every step).
int capture_step(fx_sim* s, bool vort, bool pres, bool dbg) {
  if (s->graph_exec) { (void)hipGraphExecDestroy(s->graph_exec); s->graph_exec = nullptr; }
  if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
  if (s->graph_exec_b) { (void)hipGraphExecDestroy(s->graph_exec_b); s->graph_exec_b = nullptr; }
  if (s->graph_b) { (void)hipGraphDestroy(s->graph_b); s->graph_b = nullptr; }
  HIP_OK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  record_step_a(s);
  hipGraph_t g = nullptr;
  HIP_OK(s, hipStreamEndCapture(s->stream, &g));
  s->graph = g;
  HIP_OK(s, hipGraphInstantiateWithFlags(&s->graph_exec, s->graph, 0));
  HIP_OK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  record_step_b(s, vort, pres, dbg);
  g = nullptr;
  HIP_OK(s, hipStreamEndCapture(s->stream, &g));
  s->graph_b = g;
  HIP_OK(s, hipGraphInstantiateWithFlags(&s->graph_exec_b, s->graph_b, 0));
  s->graph_last_pressure = s->lastPressure;
  return FX_OK;
}
// The simulation stream waits for the last raymarch, then rewrites the render
// snapshot: the listed bricks and this step's retired ones, or every brick.
int pack_render(fx_sim* s, bool full) {
  if (s->render_pending) {
    HIP_OK(s, hipStreamWaitEvent(s->stream, s->ev_render, 0));
    s->render_pending = false;
  }
  for (int r = 1; r < s->rlanes; ++r)
    if (s->rl_pending[r]) {
      HIP_OK(s, hipStreamWaitEvent(s->stream, s->rl_ev[r], 0));
      s->rl_pending[r] = false;
    }
  {
    StageTimer timer(s, FX_TIME_RAYMARCH);
    const bool all = full || !s->sparse;
    launch_render_pack(s->stream, grid_for(s, s->nb), all ? s->identity : s->lv[0].L.list,
                       all ? s->identity_count : s->lv[0].L.count, set_of(s, 0), all ? nullptr : s->lv[0].L.retire,
                       all ? nullptr : s->lv[0].L.retire_count, s->d_render, s->flagsRender, uint32_t(s->nb), s->d_bits);
  }
  HIP_OK(s, hipEventRecord(s->ev_scalars, s->stream));
  s->render_full = false;
  s->scalars_marked = true;
  return FX_OK;
}
// Simulation work outside fx_step: the snapshot is stale and lists may have changed.
void snapshot_stale(fx_sim* s) {
  s->scalars_marked = false;
  s->render_full = true;
}
// The render snapshot of a step, taken between its two graphs when the previous
// snapshot was rendered (or none was taken yet); a simulation that is not being
// rendered skips it, and a raymarch after such steps repacks every brick first.
int step_snapshot(fx_sim* s) {
  if (!s->render_wanted) {
    snapshot_stale(s);
    return FX_OK;
  }
  s->render_wanted = false;
  return pack_render(s, s->render_full);
}

}  // namespace

// ====================================================================== ABI ==
extern "C" {

void fx_default_create_desc(fx_create_desc* d) {
  if (!d) return;
  std::memset(d, 0, sizeof(*d));
  d->width = 128; d->height = 256; d->depth = 128;           // ORCH:54
  d->advection_mode = FX_ADVECT_MACCORMACK;                  // ORCH:181
  d->runtime.gravity[1] = -1.f;                              // ORCH:194-204
  d->runtime.gravity_strength = 0.45f;
  d->runtime.buoyancy = 3.6f;
  d->runtime.vorticity_strength = 2.15f;
  d->runtime.vorticity_constant_mask = 1.0f;
  d->runtime.vorticity_velocity_mask = 0.4f;
  d->runtime.vorticity_heat_mask = 0.8f;
  d->runtime.vorticity_density_mask = 0.35f;
  d->runtime.world_size = 10.f;
  d->quality = fx_quality{1, 1, 2, 3, 2, 3, 10};             // CVS/types.ts:16-24
  d->domain_mode = FX_DOMAIN_DENSE;
  d->activation_rule = FX_ACTIVATION_REFERENCE;
  d->solver = FX_SOLVER_REFERENCE;
  d->precision = FX_PRECISION_FP32;
  d->vcycle = fx_vcycle{4, 2, 2, 4, 16, 2.f / 3.f};
  d->ignition_lo = 0.1f; d->ignition_hi = 0.55f; d->burn_rate = 1.f;  // SH/combustion-update.wgsl.ts:73-74
  d->activation_density = 1e-4f; d->activation_speed = 1e-3f;
  d->debug_fields = -1;
  d->device = 0;
}

namespace {
// Checks one creation descriptor (fx_create's rules).
int check_desc(const fx_create_desc* desc) {
  if (desc->width < 4 || desc->height < 4 || desc->depth < 4)
    return fail(nullptr, FX_ERR_INVALID_ARG, "resolution must be at least 4 in every axis");
  if (uint64_t(desc->width) * desc->height * desc->depth > (1ull << 31))
    return fail(nullptr, FX_ERR_INVALID_ARG, "resolution exceeds 2^31 voxels");
  if (desc->precision < FX_PRECISION_FP32 || desc->precision > FX_PRECISION_UF16_RNE)
    return fail(nullptr, FX_ERR_INVALID_ARG, "unknown precision");
  if (desc->solver == FX_SOLVER_VCYCLE && (desc->vcycle.levels < 2 || desc->vcycle.levels > FX_MAX_LEVELS_DEV))
    return fail(nullptr, FX_ERR_INVALID_ARG, "vcycle.levels must be in [2, 8]");
  if (desc->advection_mode != FX_ADVECT_MACCORMACK && desc->advection_mode != FX_ADVECT_SEMI_LAGRANGIAN)
    return fail(nullptr, FX_ERR_INVALID_ARG, "unknown advection mode");
  return FX_OK;
}
// Scenes of one brick-queue batch share every setting that shapes the recorded
// pass sequence or the kernels: domain, solver, activation, precision, advection,
// quality, V-cycle, combustion and activation constants, device.
int check_batch_desc(const fx_create_desc& d, const fx_create_desc& d0) {
  if (d.domain_mode != FX_DOMAIN_SPARSE || d.solver != FX_SOLVER_VCYCLE || d.activation_rule != FX_ACTIVATION_THRESHOLD)
    return fail(nullptr, FX_ERR_UNSUPPORTED, "a batch of scenes runs the sparse domain, the V-cycle solver and threshold activation");
  const fx_quality q = clamp_quality(d.quality), q0 = clamp_quality(d0.quality);
  if (d.precision != d0.precision || d.advection_mode != d0.advection_mode || d.device != d0.device ||
      std::memcmp(&q, &q0, sizeof(q)) != 0 || std::memcmp(&d.vcycle, &d0.vcycle, sizeof(d.vcycle)) != 0 ||
      d.ignition_lo != d0.ignition_lo || d.ignition_hi != d0.ignition_hi || d.burn_rate != d0.burn_rate ||
      d.activation_density != d0.activation_density || d.activation_speed != d0.activation_speed)
    return fail(nullptr, FX_ERR_INVALID_ARG,
                "the scenes of a batch share precision, advection mode, quality, V-cycle, combustion and activation constants and device");
  return FX_OK;
}

int create_impl(const fx_create_desc* descs, uint32_t nsc, bool multi, fx_sim** out) {
  if (!descs || !out) return fail(nullptr, FX_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (nsc == 0) return fail(nullptr, FX_ERR_INVALID_ARG, "a batch needs at least one scene");
  if (nsc > 65535) return fail(nullptr, FX_ERR_INVALID_ARG, "a batch holds at most 65535 scenes");
  for (uint32_t k = 0; k < nsc; ++k) {
    int r = check_desc(&descs[k]);
    if (r == FX_OK && multi) r = check_batch_desc(descs[k], descs[0]);
    if (r != FX_OK) return r;
  }
  const fx_create_desc* desc = &descs[0];
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, FX_ERR_DEVICE, "no HIP device");
  if (desc->device < 0 || desc->device >= ndev) return fail(nullptr, FX_ERR_INVALID_ARG, "device ordinal out of range");

  fx_sim* s = new (std::nothrow) fx_sim();
  if (!s) return fail(nullptr, FX_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->desc = *desc;
  s->multi = multi;
  if (multi) s->desc.debug_fields = 0;  // a batch records no debug pass
  // storage type in desc.precision; the rounding mode apart
  s->stochastic = desc->precision == FX_PRECISION_BF16 || desc->precision == FX_PRECISION_FP16 ||
                  desc->precision == FX_PRECISION_U16 || desc->precision == FX_PRECISION_UF16;
  if (desc->precision == FX_PRECISION_BF16_RNE) s->desc.precision = FX_PRECISION_BF16;
  if (desc->precision == FX_PRECISION_FP16_RNE) s->desc.precision = FX_PRECISION_FP16;
  if (desc->precision == FX_PRECISION_U16_RNE) s->desc.precision = FX_PRECISION_U16;
  if (desc->precision == FX_PRECISION_UF16_RNE) s->desc.precision = FX_PRECISION_UF16;
  s->device = desc->device;
  s->sparse = desc->domain_mode == FX_DOMAIN_SPARSE;
  s->rt = desc->runtime;
  s->q = clamp_quality(desc->quality);
  s->mode = desc->advection_mode;
  const int nlev = desc->solver == FX_SOLVER_VCYCLE ? desc->vcycle.levels : 3;  // PressureSolvePass.ts:32-35
  // Level grids; several scenes: every scene's grid of every level at its base
  // in the level's storage, and the level totals.
  std::vector<size_t> lev_nb(size_t(nlev), 0);
  if (multi) {
    s->scenes.resize(nsc);
    for (uint32_t k = 0; k < nsc; ++k) {
      fx_sim::Scene& sc = s->scenes[k];
      sc.rt = descs[k].runtime;
      sc.nv = size_t(descs[k].width) * descs[k].height * descs[k].depth;
      sc.h_src.assign(FX_FLOATS_PER_EMITTER, 0.f);
      Grid lg = make_grid(descs[k].width, descs[k].height, descs[k].depth);
      for (int l = 0; l < nlev; ++l) {
        lg.base = uint32_t(lev_nb[size_t(l)]);
        lg.scene = k;
        sc.g[l] = lg;
        lev_nb[size_t(l)] += brick_count(lg);
        if (l == nlev - 1) s->coarse_nv_max = std::max(s->coarse_nv_max, storage_count(lg));
        lg = make_grid(halve(lg.W), halve(lg.H), halve(lg.D));
      }
      s->nv += sc.nv;
    }
    if (lev_nb[0] * 512 > (size_t(1) << 32))
      return fail(nullptr, FX_ERR_INVALID_ARG, "the scenes of a batch exceed 2^23 bricks");
    for (uint32_t k = 0; k < nsc; ++k)
      for (int l = 0; l < nlev; ++l) s->scenes[k].g[l].span = uint32_t(lev_nb[size_t(l)]);
    s->g = s->scenes[0].g[0];
    s->nb = lev_nb[0];
  } else {
    s->g = make_grid(desc->width, desc->height, desc->depth);
    s->nb = brick_count(s->g);
    s->nv = size_t(desc->width) * desc->height * desc->depth;
    Grid lg = s->g;
    for (int l = 0; l < nlev; ++l) {
      lev_nb[size_t(l)] = brick_count(lg);
      lg = make_grid(halve(lg.W), halve(lg.H), halve(lg.D));
    }
  }
  s->const_bytes = (sizeof(StepConst) * nsc + 255) & ~size_t(255);
  s->h_src.assign(FX_FLOATS_PER_EMITTER, 0.f);
  auto bail = [&](int code) { std::string e = s->err; fx_destroy(s); g_create_error = e; return code; };

  if (hipSetDevice(s->device) != hipSuccess) return bail(fail(s, FX_ERR_DEVICE, "hipSetDevice failed"));
  // the static scratch of fx_bloom_device_noalloc: its address on this device, once per device
  if (static_scratch_setup(s->device, nullptr) != FX_OK)
    return bail(fail(s, FX_ERR_DEVICE, "hipGetSymbolAddress of the static scratch failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, s->device) != hipSuccess) return bail(fail(s, FX_ERR_DEVICE, "hipGetDeviceProperties failed"));
  {
    // persistent brick-kernel workgroups (512 threads each): 3 per CU, which a
    // CU holds beside a raymarch workgroup (128 VGPRs) for kernels of at most 64
    // VGPRs; 4 would leave a quarter of the grid waiting while the overlapped
    // raymarch runs (plume step 1.045 -> 1.018 ms). FENIX_SIM_BLOCKS_PER_CU
    // overrides; the grid stays a multiple of the 8 XCDs through the CU count.
    const char* b = std::getenv("FENIX_SIM_BLOCKS_PER_CU");
    s->cus = std::max(1, prop.multiProcessorCount);
    s->blocks_env = b != nullptr;
    s->blocks = s->cus * (b ? std::max(1, std::atoi(b)) : 3);
  }
  {
    // resident raymarch workgroups (256 threads each): one per CU;
    // FENIX_RAY_BLOCKS_PER_CU or the absolute FENIX_RAY_BLOCKS override
    const char* r = std::getenv("FENIX_RAY_BLOCKS_PER_CU");
    const int per_cu = r ? std::max(1, std::atoi(r)) : 1;
    s->ray_blocks = std::max(1, prop.multiProcessorCount * per_cu);
    if (const char* a = std::getenv("FENIX_RAY_BLOCKS")) s->ray_blocks = std::max(1, std::atoi(a));
    const char* o = std::getenv("FENIX_RENDER_OVERLAP");
    s->render_overlap = !(o && o[0] == '0');
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&s->rstream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_scalars, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_render, hipEventDisableTiming) != hipSuccess)
    return bail(fail(s, FX_ERR_DEVICE, "stream creation failed"));
  if (s->sparse && desc->solver == FX_SOLVER_VCYCLE) {
    if (hipHostMalloc(reinterpret_cast<void**>(&s->h_fine_count), 4, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s->ev_count, hipEventDisableTiming) != hipSuccess)
      return bail(fail(s, FX_ERR_OUT_OF_MEMORY, "pinned count allocation failed"));
    *s->h_fine_count = 0;
  }

  int rc = FX_OK;
  if (multi) {
    // device grid tables and brick owners of every level
    std::vector<Grid> gt(size_t(nlev) * nsc);
    std::vector<uint32_t> own;
    std::vector<size_t> own_at(size_t(nlev), 0);
    for (int l = 0; l < nlev; ++l) {
      own_at[size_t(l)] = own.size();
      for (uint32_t k = 0; k < nsc; ++k) {
        gt[size_t(l) * nsc + k] = s->scenes[k].g[l];
        own.insert(own.end(), brick_count(s->scenes[k].g[l]), k);
      }
    }
    rc = dalloc(s, &s->d_gtab, gt.size());
    if (rc == FX_OK) rc = dalloc(s, &s->d_owner, own.size());
    if (rc == FX_OK) rc = dalloc(s, &s->d_scene_count, nsc);
    if (rc == FX_OK && hipHostMalloc(reinterpret_cast<void**>(&s->h_scene_count), size_t(nsc) * 4, hipHostMallocDefault) != hipSuccess)
      rc = fail(s, FX_ERR_OUT_OF_MEMORY, "pinned count allocation failed");
    if (rc == FX_OK && (hipMemcpy(s->d_gtab, gt.data(), gt.size() * sizeof(Grid), hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(s->d_owner, own.data(), own.size() * 4, hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(s, FX_ERR_DEVICE, "grid table upload failed");
    if (rc != FX_OK) return bail(rc);
    for (uint32_t k = 0; k < nsc; ++k) s->h_scene_count[k] = 0;
    // selectors: the scene-0 grid, pointing at the level's table
    s->g.tab = s->d_gtab;
    s->g.owner = s->d_owner;
    for (int l = 0; l < nlev; ++l) {
      s->scenes[0].g[l].tab = s->d_gtab + size_t(l) * nsc;
      s->scenes[0].g[l].owner = s->d_owner + own_at[size_t(l)];
    }
    for (uint32_t k = 1; k < nsc; ++k)
      for (int l = 0; l < nlev; ++l) { s->scenes[k].g[l].tab = nullptr; s->scenes[k].g[l].owner = nullptr; }
  }

  const size_t nsv = s->nb * 512;
  for (int k = 0; k < 2 && rc == FX_OK; ++k)
    for (int f = 0; f < 4 && rc == FX_OK; ++f) {
      uint8_t* p = nullptr;
      rc = dalloc(s, &p, nsv * size_t(field_esz(s, f)));
      s->S[k][f] = p;
    }
  for (int c = 0; c < 3 && rc == FX_OK; ++c) {
    rc = dalloc(s, &s->u[c], nsv);
    if (rc == FX_OK) rc = dalloc(s, &s->us[c], nsv);
    if (rc == FX_OK) rc = dalloc(s, &s->w[c], nsv);
  }
  if (rc == FX_OK) rc = dalloc(s, &s->wmag, nsv);
  if (rc == FX_OK) rc = dalloc(s, &s->cmag, nsv);
  if (rc == FX_OK) rc = dalloc(s, &s->umag, nsv);
  if (rc == FX_OK) rc = dalloc(s, &s->flagsInj, s->nb);
  if (rc == FX_OK) rc = s->sparse ? FX_OK : dalloc(s, &s->flagsRender, s->nb);
  if (rc == FX_OK) rc = dalloc(s, &s->touched, s->nb);
  if (rc == FX_OK) rc = dalloc(s, &s->identity, s->nb);
  if (rc == FX_OK) rc = dalloc(s, &s->identity_count, 1);
  uint8_t* stage = nullptr;
  if (rc == FX_OK) rc = dalloc(s, &stage, s->const_bytes + s->nb * 4);
  if (rc == FX_OK) {
    s->d_const = reinterpret_cast<StepConst*>(stage);
    s->d_cand = reinterpret_cast<uint32_t*>(stage + s->const_bytes);
  }
  if (rc == FX_OK) rc = dalloc(s, &s->d_bits, (s->nb + 31) / 32);
  if (rc != FX_OK) return bail(rc);
  if (s->sparse) s->flagsRender = s->flagsInj;

  Grid lg = s->g;
  for (int l = 0; l < nlev && rc == FX_OK; ++l) {
    Level L;
    L.g = multi ? s->scenes[0].g[l] : lg;
    L.nb = lev_nb[size_t(l)];
    L.nv = L.nb * 512;
    rc = dalloc(s, &L.div, L.nv);
    if (rc == FX_OK) rc = dalloc(s, &L.p[0], L.nv);
    if (rc == FX_OK) rc = dalloc(s, &L.p[1], L.nv);
    if (rc == FX_OK) rc = dalloc(s, &L.L.list, L.nb);
    if (rc == FX_OK) rc = dalloc(s, &L.L.count, 1);
    if (rc == FX_OK) rc = dalloc(s, &L.L.page, L.nb);
    if (rc == FX_OK) rc = dalloc(s, &L.L.retire, L.nb);
    if (rc == FX_OK) rc = dalloc(s, &L.L.retire_count, 1);
    if (rc == FX_OK) rc = l == 0 ? FX_OK : dalloc(s, &L.flags, L.nb);
    if (rc == FX_OK && desc->solver == FX_SOLVER_VCYCLE) rc = dalloc(s, &L.bits, L.nb * 8);
    if (rc == FX_OK && desc->solver == FX_SOLVER_VCYCLE && l == 0 && s->sparse) rc = dalloc(s, &L.nbr, L.nb);
    if (rc == FX_OK && desc->solver == FX_SOLVER_VCYCLE && l > 0) {
      // one slot table per scene (only the coarsest level's is used)
      uint8_t* t = nullptr;
      rc = dalloc(s, &t, kCoarseTableBytes * nsc);
      L.tab = t;
      if (rc == FX_OK) rc = dalloc(s, &L.tab_n, nsc);
    }
    if (l == 0) L.flags = s->flagsInj;
    s->lv.push_back(L);
    lg = make_grid(halve(lg.W), halve(lg.H), halve(lg.D));
  }
  if (rc != FX_OK) return bail(rc);
  if (s->sparse) {
    CompactArgs& A = s->cargs;
    A.nlevels = int(s->lv.size());
    uint32_t blocks = 0;
    for (int l = 0; l < A.nlevels; ++l) {
      A.nb[l] = uint32_t(s->lv[l].nb);
      A.flags[l] = s->lv[l].flags;
      A.L[l] = s->lv[l].L;
      A.block_start[l] = blocks;
      blocks += uint32_t((s->lv[l].nb + 1023) / 1024);
    }
    A.block_start[A.nlevels] = blocks;
    A.touched = s->touched;
    s->lflags.n = A.nlevels - 1;
    for (int l = 1; l < A.nlevels; ++l) {
      s->lflags.f[l] = s->lv[l].flags;
      s->lflags.g[l] = s->lv[l].g;
    }
    if (dalloc(s, &s->d_compact, size_t(blocks) * 4) != FX_OK) return bail(FX_ERR_OUT_OF_MEMORY);
    A.counts = s->d_compact;
    A.offsets = s->d_compact + size_t(blocks) * 2;
  }
  {
    const char* g = std::getenv("FENIX_GRAPH");
    s->use_graph = !(g && g[0] == '0');
  }
  s->norm_cap = std::max(1, desc->vcycle.cycles);
  if (dalloc(s, &s->d_norms, size_t(s->norm_cap) * kNormLanes) != FX_OK) return bail(FX_ERR_OUT_OF_MEMORY);
  if (dalloc(s, &s->d_rqueue, fx_sim::kRenderLanes) != FX_OK) return bail(FX_ERR_OUT_OF_MEMORY);
  if (multi) {  // render lanes 1.. of a batch
    s->rlanes = int(std::min<uint32_t>(nsc, fx_sim::kRenderLanes));
    for (int r = 1; r < s->rlanes; ++r)
      if (hipStreamCreateWithFlags(&s->rl_stream[r], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&s->rl_ev[r], hipEventDisableTiming) != hipSuccess)
        return bail(fail(s, FX_ERR_DEVICE, "stream creation failed"));
  }
  {
    float* q = nullptr;
    if (dalloc(s, &q, nsv * 4) != FX_OK) return bail(FX_ERR_OUT_OF_MEMORY);
    s->d_render = q;
  }
  s->src_cap = size_t(FX_FLOATS_PER_EMITTER) * nsc;
  if (dalloc(s, &s->d_src, s->src_cap) != FX_OK) return bail(FX_ERR_OUT_OF_MEMORY);
  for (auto& slot : s->slots) {
    if (hipHostMalloc(reinterpret_cast<void**>(&slot.c), s->const_bytes + s->nb * 4, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&slot.done, hipEventDisableTiming) != hipSuccess)
      return bail(fail(s, FX_ERR_OUT_OF_MEMORY, "pinned staging allocation failed"));
    slot.cand = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(slot.c) + s->const_bytes);
  }

  // Identity lists (dense) and zeroed state.
  std::vector<uint32_t> ident(s->nb);
  for (size_t b = 0; b < s->nb; ++b) ident[b] = uint32_t(b);
  uint32_t nbc = uint32_t(s->nb);
  if (hipMemcpy(s->identity, ident.data(), s->nb * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->identity_count, &nbc, 4, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(s, FX_ERR_DEVICE, "identity upload failed"));
  for (auto& L : s->lv) {
    if (!s->sparse) {
      std::vector<uint32_t> id(L.nb);
      for (size_t b = 0; b < L.nb; ++b) id[b] = uint32_t(b);
      uint32_t n = uint32_t(L.nb);
      if (hipMemcpy(L.L.list, id.data(), L.nb * 4, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(L.L.count, &n, 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(s, FX_ERR_DEVICE, "level list upload failed"));
    }
  }
  {
    std::vector<float> zero(s->src_cap, 0.f);
    if (hipMemcpy(s->d_src, zero.data(), zero.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(s, FX_ERR_DEVICE, "source upload failed"));
  }
  int r = fx_reset(s);
  if (r != FX_OK) return bail(r);
  if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(s, FX_ERR_DEVICE, "initial clear failed"));
  update_coarse_plan(s);  // dense: fixed by the grid
  *out = s;
  return FX_OK;
}
}  // namespace

int fx_create(const fx_create_desc* desc, fx_sim** out) { return create_impl(desc, 1, false, out); }

int fx_create_scenes(const fx_create_desc* descs, uint32_t n, fx_sim** out) { return create_impl(descs, n, true, out); }

void fx_destroy(fx_sim* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->rstream) (void)hipStreamSynchronize(s->rstream);
  for (int r = 1; r < fx_sim::kRenderLanes; ++r)
    if (s->rl_stream[r]) (void)hipStreamSynchronize(s->rl_stream[r]);
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  for (int k = 0; k < 2; ++k) for (int f = 0; f < 4; ++f) F(s->S[k][f]);
  for (int c = 0; c < 3; ++c) { F(s->u[c]); F(s->us[c]); F(s->w[c]); }
  F(s->wmag); F(s->cmag); F(s->umag);
  F(s->flagsInj);
  if (!s->sparse) F(s->flagsRender);
  F(s->touched); F(s->d_bits); F(s->d_rqueue); F(s->d_render); F(s->d_bloom); F(s->d_bloom_io); F(s->identity); F(s->identity_count); F(s->d_const); F(s->d_src); F(s->d_norms);
  for (size_t l = 0; l < s->lv.size(); ++l) {
    Level& L = s->lv[l];
    F(L.div); F(L.p[0]); F(L.p[1]); F(L.bits); F(L.nbr); F(L.tab); F(L.tab_n); F(L.L.list); F(L.L.count); F(L.L.page); F(L.L.retire); F(L.L.retire_count);
    if (l > 0) F(L.flags);
  }
  for (auto& slot : s->slots) {
    if (slot.c) (void)hipHostFree(slot.c);
    if (slot.done) (void)hipEventDestroy(slot.done);
  }
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->graph_exec_b) (void)hipGraphExecDestroy(s->graph_exec_b);
  if (s->graph_b) (void)hipGraphDestroy(s->graph_b);
  F(s->d_compact); F(s->d_gtab); F(s->d_owner); F(s->d_scene_count);
  if (s->h_scene_count) (void)hipHostFree(s->h_scene_count);
  drain_timing(s);
  for (hipEvent_t e : s->pool) (void)hipEventDestroy(e);
  if (s->ev_scalars) (void)hipEventDestroy(s->ev_scalars);
  if (s->ev_render) (void)hipEventDestroy(s->ev_render);
  if (s->ev_count) (void)hipEventDestroy(s->ev_count);
  if (s->h_fine_count) (void)hipHostFree(s->h_fine_count);
  if (s->rstream) (void)hipStreamDestroy(s->rstream);
  for (int r = 1; r < fx_sim::kRenderLanes; ++r) {
    if (s->rl_ev[r]) (void)hipEventDestroy(s->rl_ev[r]);
    if (s->rl_stream[r]) (void)hipStreamDestroy(s->rl_stream[r]);
  }
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int fx_reset(fx_sim* s) {  // ORCH:206-226
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  HIP_OK(s, hipSetDevice(s->device));
  snapshot_stale(s);
  s->curSet = 0; s->stepIndex = 0; s->lastPressure = 0; s->initialized = false;
  s->startS = s->lastS;
  s->pending.clear();
  const size_t nsv = s->nb * 512;
  for (const Buf& b : fine_buffers(s)) launch_clear_u32(s->stream, static_cast<uint32_t*>(b.p), nsv * size_t(b.esz) / 4, 0u);
  launch_clear_u32(s->stream, s->flagsInj, s->nb, 0u);
  if (!s->sparse) launch_clear_u32(s->stream, s->flagsRender, s->nb, 0u);
  launch_clear_u32(s->stream, s->touched, s->nb, 0u);
  for (size_t l = 0; l < s->lv.size(); ++l) {
    Level& L = s->lv[l];
    if (l > 0) {
      launch_clear_f32(s->stream, L.div, L.nv);
      launch_clear_f32(s->stream, L.p[0], L.nv);
      launch_clear_f32(s->stream, L.p[1], L.nv);
      launch_clear_u32(s->stream, L.flags, L.nb, 0u);
    }
    if (s->sparse) {
      launch_clear_u32(s->stream, reinterpret_cast<uint32_t*>(L.L.page), L.nb, 0xffffffffu);
      launch_clear_u32(s->stream, L.L.count, 1, 0u);
      launch_clear_u32(s->stream, L.L.retire_count, 1, 0u);
    }
  }
  HIP_OK(s, hipGetLastError());
  return FX_OK;
}

int fx_step(fx_sim* s, double elapsed_s, double step_s) {  // ORCH:230-305
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  if (!std::isfinite(elapsed_s) || !std::isfinite(step_s)) return fail(s, FX_ERR_INVALID_ARG, "non-finite time");
  HIP_OK(s, hipSetDevice(s->device));
  if (!s->initialized) { s->initialized = true; s->startS = elapsed_s; s->lastS = elapsed_s; }
  std::vector<fx_sim_params> p;
  if (s->multi)
    for (const auto& sc : s->scenes) p.push_back(make_params(s, sc.rt, sc.g[0], elapsed_s, step_s));
  else
    p.push_back(make_params(s, elapsed_s, step_s));
  std::vector<uint32_t> cand;
  uint32_t inj = 0;
  if (s->sparse) build_candidates(s, p, cand, inj);
  int rc = upload_step(s, p, s->sparse ? &cand : nullptr, inj);
  if (rc != FX_OK) return rc;

  const bool vort = s->stepIndex % uint64_t(s->q.vorticity_interval) == 0;
  const bool pres = s->stepIndex % uint64_t(s->q.pressure_interval) == 0;
  const bool dbg = debug_each_frame(s) || s->debug_requested;
  s->debug_requested = false;
  update_coarse_plan(s);
  // Persistent brick-kernel workgroups: 3 per CU fit beside an overlapped raymarch
  // workgroup; while the snapshots go unrendered no raymarch shares the CUs, and 4
  // per CU are faster (C3 batch 6.61 -> 6.46 ms per step). FENIX_SIM_BLOCKS_PER_CU fixes it.
  if (!s->blocks_env) s->blocks = s->cus * (s->render_wanted ? 3 : 4);
  if (s->use_graph && !s->timing) {
    std::vector<int64_t> key = {vort, pres, dbg, s->mode, pres ? -1 : s->lastPressure, s->q.fine_pre_iterations,
                                s->q.fine_post_iterations, s->q.mid_pre_iterations, s->q.mid_post_iterations,
                                s->q.coarse_iterations, s->split_down, s->split_up, s->coarse_split, s->blocks};
    if (!s->graph_exec || key != s->graph_key) {
      rc = capture_step(s, vort, pres, dbg);
      if (rc != FX_OK) return rc;
      s->graph_key = key;
    } else if (pres) {
      s->lastPressure = s->graph_last_pressure;
    }
    HIP_OK(s, hipGraphLaunch(s->graph_exec, s->stream));
    rc = step_snapshot(s);
    if (rc != FX_OK) return rc;
    HIP_OK(s, hipGraphLaunch(s->graph_exec_b, s->stream));
  } else {
    record_step_a(s);
    rc = step_snapshot(s);
    if (rc != FX_OK) return rc;
    record_step_b(s, vort, pres, dbg);
  }
  if (s->sparse && s->ev_count && !s->count_pending) {  // the count copied at the end of this step
    HIP_OK(s, hipEventRecord(s->ev_count, s->stream));
    s->count_pending = true;
  }
  s->curSet = s->curSet == 0 ? 1 : 0;
  s->lastS = elapsed_s;
  s->stepIndex += 1;
  HIP_OK(s, hipGetLastError());
  return FX_OK;
}

int fx_run_pass(fx_sim* s, int32_t pass, const fx_sim_params* params) {
  if (!s || !params) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  HIP_OK(s, hipSetDevice(s->device));
  if (s->multi) return fail(s, FX_ERR_UNSUPPORTED, "single passes run on a simulation of one scene");
  std::vector<uint32_t> cand;
  uint32_t inj = 0;
  const std::vector<fx_sim_params> pv(1, *params);
  if (s->sparse) build_candidates(s, pv, cand, inj);
  snapshot_stale(s);
  int rc = upload_step(s, pv, s->sparse ? &cand : nullptr, inj);
  if (rc != FX_OK) return rc;
  update_coarse_plan(s);
  switch (pass) {
    case FX_PASS_INJECT_SCALAR: pass_inject(s, 1); break;
    case FX_PASS_INJECT_VELOCITY: pass_inject(s, 2); break;
    case FX_PASS_INJECT_IGNITER: pass_inject(s, 4); break;
    case FX_PASS_ACTIVATE_RENDER: pass_activate_render(s); break;
    case FX_PASS_ACTIVATE_ADVECTED:
      if (s->sparse) return fail(s, FX_ERR_UNSUPPORTED, "sparse domains activate once per step");
      pass_activate_advected(s);
      break;
    case FX_PASS_PRESSURE: clear_pressure(s); s->lastPressure = pressure_solve(s); break;
    case FX_PASS_ADVECT_VELOCITY: case FX_PASS_ADVECT_SCALARS: case FX_PASS_COMBUSTION: case FX_PASS_BUOYANCY:
    case FX_PASS_VORTICITY: case FX_PASS_CONFINEMENT: case FX_PASS_DIVERGENCE: case FX_PASS_PROJECTION:
    case FX_PASS_DEBUG_FIELDS:
      run_stage(s, pass);
      break;
    default: return fail(s, FX_ERR_INVALID_ARG, "unknown pass id");
  }
  HIP_OK(s, hipGetLastError());
  return FX_OK;
}

int fx_request_debug_fields(fx_sim* s) {  // getDebugBuffers() sets debugRequested, ORCH:316
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  ONE_SCENE(s);
  s->debug_requested = true;
  return FX_OK;
}

int fx_synchronize(fx_sim* s) {
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  HIP_OK(s, hipStreamSynchronize(s->rstream));
  for (int r = 1; r < s->rlanes; ++r) HIP_OK(s, hipStreamSynchronize(s->rl_stream[r]));
  return FX_OK;
}

int fx_get_runtime_params(fx_sim* s, fx_runtime_params* out) {
  if (!s || !out) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  *out = s->rt;
  return FX_OK;
}
}  // extern "C"
namespace {
// The fields of p selected by the present mask m (ORCH:353-365).
void apply_runtime(fx_runtime_params& r, const fx_runtime_params* p, uint32_t m) {
  if (m & FX_RP_WIND) { r.wind[0] = p->wind[0]; r.wind[1] = p->wind[1]; r.wind[2] = p->wind[2]; }
  if (m & FX_RP_WIND_STRENGTH) r.wind_strength = p->wind_strength;
  if (m & FX_RP_GRAVITY) { r.gravity[0] = p->gravity[0]; r.gravity[1] = p->gravity[1]; r.gravity[2] = p->gravity[2]; }
  if (m & FX_RP_GRAVITY_STRENGTH) r.gravity_strength = p->gravity_strength;
  if (m & FX_RP_BUOYANCY) r.buoyancy = p->buoyancy;
  if (m & FX_RP_VORTICITY_STRENGTH) r.vorticity_strength = p->vorticity_strength;
  if (m & FX_RP_VORTICITY_CONSTANT_MASK) r.vorticity_constant_mask = p->vorticity_constant_mask;
  if (m & FX_RP_VORTICITY_VELOCITY_MASK) r.vorticity_velocity_mask = p->vorticity_velocity_mask;
  if (m & FX_RP_VORTICITY_HEAT_MASK) r.vorticity_heat_mask = p->vorticity_heat_mask;
  if (m & FX_RP_VORTICITY_DENSITY_MASK) r.vorticity_density_mask = p->vorticity_density_mask;
  if (m & FX_RP_WORLD_SIZE) r.world_size = p->world_size;
}
}  // namespace
extern "C" {
int fx_set_runtime_params(fx_sim* s, const fx_runtime_params* p, uint32_t m) {  // ORCH:353-365
  if (!s || !p) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  apply_runtime(s->rt, p, m);
  return FX_OK;
}
int fx_get_quality(fx_sim* s, fx_quality* out) {
  if (!s || !out) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  *out = s->q;
  return FX_OK;
}
int fx_set_quality(fx_sim* s, const fx_quality* q) {
  if (!s || !q) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  s->q = clamp_quality(*q);
  return FX_OK;
}
int fx_get_advection_mode(fx_sim* s, int32_t* out) {
  if (!s || !out) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  *out = s->mode;
  return FX_OK;
}
int fx_set_advection_mode(fx_sim* s, int32_t m) {
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  if (m != FX_ADVECT_MACCORMACK && m != FX_ADVECT_SEMI_LAGRANGIAN) return fail(s, FX_ERR_INVALID_ARG, "unknown advection mode");
  s->mode = m;
  return FX_OK;
}

int fx_update_sources(fx_sim* s, const float* packed, uint32_t count) {  // SourceInjectionPass.ts:71-83
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  ONE_SCENE(s);
  if (count && !packed) return fail(s, FX_ERR_INVALID_ARG, "null source data");
  HIP_OK(s, hipSetDevice(s->device));
  const size_t nf = size_t(std::max<uint32_t>(count, 1)) * FX_FLOATS_PER_EMITTER;
  std::vector<float> h(nf, 0.f);
  if (count) std::memcpy(h.data(), packed, size_t(count) * FX_FLOATS_PER_EMITTER * 4);
  HIP_OK(s, hipStreamSynchronize(s->stream));
  if (nf > s->src_cap) {
    HIP_OK(s, hipFree(s->d_src));
    s->d_src = nullptr;
    if (dalloc(s, &s->d_src, nf) != FX_OK) return FX_ERR_OUT_OF_MEMORY;
    s->src_cap = nf;
  }
  HIP_OK(s, hipMemcpy(s->d_src, h.data(), nf * 4, hipMemcpyHostToDevice));
  s->h_src = std::move(h);
  s->nsrc = count;
  return FX_OK;
}

int fx_read_field(fx_sim* s, int32_t id, float* host, size_t bytes) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  int esz = 4;
  void* f = field_ptr(s, id, &esz);
  if (!f) return fail(s, FX_ERR_INVALID_ARG, "unknown field id");
  HIP_OK(s, hipSetDevice(s->device));
  return read_brickmajor_typed(s, s->g, f, esz, host, bytes);
}

int fx_write_field(fx_sim* s, int32_t id, const float* host, size_t bytes) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  int esz = 4;
  void* f = field_ptr(s, id, &esz);
  if (!f) return fail(s, FX_ERR_INVALID_ARG, "unknown field id");
  if (bytes != s->nv * 4) return fail(s, FX_ERR_INVALID_ARG, "buffer size mismatch");
  HIP_OK(s, hipSetDevice(s->device));
  snapshot_stale(s);
  std::vector<float> tmp(storage_count(s->g), 0.f);
  scatter_brickmajor(s->g, host, tmp.data());
  HIP_OK(s, hipStreamSynchronize(s->stream));
  if (esz == 4) {
    HIP_OK(s, hipMemcpy(f, tmp.data(), tmp.size() * 4, hipMemcpyHostToDevice));
  } else {
    const int prec = s->desc.precision;
    std::vector<uint16_t> h(tmp.size());
    for (size_t i = 0; i < tmp.size(); ++i) h[i] = store16(prec, tmp[i]);
    HIP_OK(s, hipMemcpy(f, h.data(), h.size() * 2, hipMemcpyHostToDevice));
  }
  if (s->sparse) {
    // Bricks holding non-zero data join the next activation's candidates.
    std::vector<uint32_t> nz;
    for (size_t b = 0; b < s->nb; ++b) {
      const float* p = &tmp[b << 9];
      for (int k = 0; k < 512; ++k)
        if (p[k] != 0.f) { nz.push_back(uint32_t(b)); break; }
    }
    if (!nz.empty()) {
      HIP_OK(s, hipMemcpy(s->d_cand, nz.data(), nz.size() * 4, hipMemcpyHostToDevice));
      launch_mark(s->stream, s->d_cand, uint32_t(nz.size()), s->touched);
      HIP_OK(s, hipStreamSynchronize(s->stream));
      std::vector<uint8_t> m(s->nb, 0);
      for (uint32_t b : s->pending) m[b] = 1;
      for (uint32_t b : nz) if (!m[b]) s->pending.push_back(b);
    }
  }
  return FX_OK;
}

int fx_read_flags(fx_sim* s, int32_t which, uint32_t* host, size_t count) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  if (count != s->nb) return fail(s, FX_ERR_INVALID_ARG, "count must equal the brick count");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  const uint32_t* f = which == FX_FLAGS_INJECTED ? s->flagsInj : s->flagsRender;
  HIP_OK(s, hipMemcpy(host, f, count * 4, hipMemcpyDeviceToHost));
  return FX_OK;
}

int fx_read_active_list(fx_sim* s, uint32_t* host, size_t capacity, uint32_t* count) {
  if (!s || !count) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  if (s->sparse) {
    uint32_t n = 0;
    HIP_OK(s, hipMemcpy(&n, s->lv[0].L.count, 4, hipMemcpyDeviceToHost));
    *count = n;
    if (host && capacity) HIP_OK(s, hipMemcpy(host, s->lv[0].L.list, std::min<size_t>(n, capacity) * 4, hipMemcpyDeviceToHost));
  } else {
    std::vector<uint32_t> f(s->nb);
    HIP_OK(s, hipMemcpy(f.data(), s->flagsRender, s->nb * 4, hipMemcpyDeviceToHost));
    uint32_t n = 0;
    for (size_t b = 0; b < s->nb; ++b)
      if (f[b]) { if (host && n < capacity) host[n] = uint32_t(b); ++n; }
    *count = n;
  }
  return FX_OK;
}

int fx_read_page_table(fx_sim* s, int32_t* host, size_t count) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  if (count != s->nb) return fail(s, FX_ERR_INVALID_ARG, "count must equal the brick count");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  if (s->sparse) {
    HIP_OK(s, hipMemcpy(host, s->lv[0].L.page, count * 4, hipMemcpyDeviceToHost));
  } else {
    std::vector<uint32_t> f(s->nb);
    HIP_OK(s, hipMemcpy(f.data(), s->flagsRender, s->nb * 4, hipMemcpyDeviceToHost));
    int32_t n = 0;
    for (size_t b = 0; b < s->nb; ++b) host[b] = f[b] ? n++ : -1;
  }
  return FX_OK;
}

void* fx_field_device_ptr(fx_sim* s, int32_t id) {
  if (!s || s->multi) return nullptr;
  return field_ptr(s, id);
}

int fx_read_level(fx_sim* s, int32_t level, int32_t which, float* host, size_t bytes) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  if (level < 0 || level >= int(s->lv.size())) return fail(s, FX_ERR_INVALID_ARG, "level out of range");
  Level& L = s->lv[level];
  const float* f = which == FX_LEVEL_DIVERGENCE ? L.div : which == FX_LEVEL_PRESSURE_A ? L.p[0] : L.p[1];
  HIP_OK(s, hipSetDevice(s->device));
  return read_brickmajor(s, L.g, f, host, bytes);
}
int fx_level_dims(fx_sim* s, int32_t level, uint32_t dims[3]) {
  if (!s || !dims) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  if (level < 0 || level >= int(s->lv.size())) return fail(s, FX_ERR_INVALID_ARG, "level out of range");
  dims[0] = s->lv[level].g.W; dims[1] = s->lv[level].g.H; dims[2] = s->lv[level].g.D;
  return FX_OK;
}
int fx_solver_residuals(fx_sim* s, double* host, size_t capacity, uint32_t* count) {
  if (!s || !count) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  uint32_t n = s->desc.solver == FX_SOLVER_VCYCLE ? uint32_t(s->norm_count) : 0u;
  *count = n;
  if (n && host && capacity) {
    std::vector<double> t(size_t(n) * kNormLanes);
    HIP_OK(s, hipMemcpy(t.data(), s->d_norms, t.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < std::min<size_t>(n, capacity); ++i) {
      double sum = 0.0;
      for (int k = 0; k < kNormLanes; ++k) sum += t[i * kNormLanes + k];
      host[i] = std::sqrt(sum);
    }
  }
  return FX_OK;
}
int fx_active_brick_count(fx_sim* s, uint32_t* out) {
  if (!s || !out) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  if (s->sparse) HIP_OK(s, hipMemcpy(out, s->lv[0].L.count, 4, hipMemcpyDeviceToHost));
  else *out = uint32_t(s->nb);
  return FX_OK;
}

// ------------------------------------------------ brick-queue batches --
uint32_t fx_scene_count(const fx_sim* s) { return s ? (s->multi ? uint32_t(s->scenes.size()) : 1u) : 0u; }

}  // extern "C"
namespace {
// The scene's fine grid (one scene: the simulation's), or nullptr after a failure.
const Grid* scene_grid(fx_sim* s, uint32_t k) {
  if (!s->multi) {
    if (k != 0) { fail(s, FX_ERR_INVALID_ARG, "scene index out of range"); return nullptr; }
    return &s->g;
  }
  if (k >= s->scenes.size()) { fail(s, FX_ERR_INVALID_ARG, "scene index out of range"); return nullptr; }
  return &s->scenes[k].g[0];
}
// The scene's bricks of the ascending fine list: list positions [k0, k0 + n).
int scene_list_range(fx_sim* s, const Grid& g, std::vector<uint32_t>& list, uint32_t& k0, uint32_t& n) {
  uint32_t total = 0;
  HIP_OK(s, hipStreamSynchronize(s->stream));
  HIP_OK(s, hipMemcpy(&total, s->lv[0].L.count, 4, hipMemcpyDeviceToHost));
  list.resize(total);
  if (total) HIP_OK(s, hipMemcpy(list.data(), s->lv[0].L.list, size_t(total) * 4, hipMemcpyDeviceToHost));
  const uint32_t lo = g.base, hi = g.base + uint32_t(brick_count(g));
  k0 = uint32_t(std::lower_bound(list.begin(), list.end(), lo) - list.begin());
  n = uint32_t(std::lower_bound(list.begin(), list.end(), hi) - list.begin()) - k0;
  return FX_OK;
}
}  // namespace
extern "C" {

int fx_scene_update_sources(fx_sim* s, uint32_t scene, const float* packed, uint32_t count) {
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  if (!s->multi) return scene == 0 ? fx_update_sources(s, packed, count) : fail(s, FX_ERR_INVALID_ARG, "scene index out of range");
  if (scene >= s->scenes.size()) return fail(s, FX_ERR_INVALID_ARG, "scene index out of range");
  if (count && !packed) return fail(s, FX_ERR_INVALID_ARG, "null source data");
  HIP_OK(s, hipSetDevice(s->device));
  fx_sim::Scene& sc = s->scenes[scene];
  sc.h_src.assign(size_t(std::max<uint32_t>(count, 1)) * FX_FLOATS_PER_EMITTER, 0.f);
  if (count) std::memcpy(sc.h_src.data(), packed, size_t(count) * FX_FLOATS_PER_EMITTER * 4);
  sc.nsrc = count;
  // every scene's emitters, in scene order (the step constants carry the offsets)
  std::vector<float> all;
  for (const auto& x : s->scenes) all.insert(all.end(), x.h_src.begin(), x.h_src.end());
  HIP_OK(s, hipStreamSynchronize(s->stream));
  if (all.size() > s->src_cap) {
    HIP_OK(s, hipFree(s->d_src));
    s->d_src = nullptr;
    if (dalloc(s, &s->d_src, all.size()) != FX_OK) return FX_ERR_OUT_OF_MEMORY;
    s->src_cap = all.size();
  }
  HIP_OK(s, hipMemcpy(s->d_src, all.data(), all.size() * 4, hipMemcpyHostToDevice));
  return FX_OK;
}
int fx_scene_get_runtime_params(fx_sim* s, uint32_t scene, fx_runtime_params* out) {
  if (!s || !out) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  if (!scene_grid(s, scene)) return FX_ERR_INVALID_ARG;
  *out = s->multi ? s->scenes[scene].rt : s->rt;
  return FX_OK;
}
int fx_scene_set_runtime_params(fx_sim* s, uint32_t scene, const fx_runtime_params* p, uint32_t m) {
  if (!s || !p) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  if (!scene_grid(s, scene)) return FX_ERR_INVALID_ARG;
  apply_runtime(s->multi ? s->scenes[scene].rt : s->rt, p, m);
  return FX_OK;
}
int fx_scene_read_field(fx_sim* s, uint32_t scene, int32_t id, float* host, size_t bytes) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  const Grid* g = scene_grid(s, scene);
  if (!g) return FX_ERR_INVALID_ARG;
  int esz = 4;
  const uint8_t* f = static_cast<const uint8_t*>(field_ptr(s, id, &esz));
  if (!f) return fail(s, FX_ERR_INVALID_ARG, "unknown field id");
  HIP_OK(s, hipSetDevice(s->device));
  return read_brickmajor_typed(s, *g, f + size_t(g->base) * 512 * size_t(esz), esz, host, bytes);
}
int fx_scene_write_field(fx_sim* s, uint32_t scene, int32_t id, const float* host, size_t bytes) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  const Grid* gp = scene_grid(s, scene);
  if (!gp) return FX_ERR_INVALID_ARG;
  if (!s->multi) return fx_write_field(s, id, host, bytes);
  const Grid g = *gp;
  int esz = 4;
  uint8_t* f = static_cast<uint8_t*>(field_ptr(s, id, &esz));
  if (!f) return fail(s, FX_ERR_INVALID_ARG, "unknown field id");
  if (bytes != size_t(g.W) * g.H * g.D * 4) return fail(s, FX_ERR_INVALID_ARG, "buffer size mismatch");
  HIP_OK(s, hipSetDevice(s->device));
  snapshot_stale(s);
  const size_t nb = brick_count(g);
  std::vector<float> tmp(nb * 512, 0.f);
  scatter_brickmajor(g, host, tmp.data());
  HIP_OK(s, hipStreamSynchronize(s->stream));
  uint8_t* dst = f + size_t(g.base) * 512 * size_t(esz);
  if (esz == 4) {
    HIP_OK(s, hipMemcpy(dst, tmp.data(), tmp.size() * 4, hipMemcpyHostToDevice));
  } else {
    const int prec = s->desc.precision;
    std::vector<uint16_t> h(tmp.size());
    for (size_t i = 0; i < tmp.size(); ++i) h[i] = store16(prec, tmp[i]);
    HIP_OK(s, hipMemcpy(dst, h.data(), h.size() * 2, hipMemcpyHostToDevice));
  }
  // bricks holding non-zero data join the next activation's candidates
  std::vector<uint32_t> nz;
  for (size_t b = 0; b < nb; ++b) {
    const float* q = &tmp[b << 9];
    for (int k = 0; k < 512; ++k)
      if (q[k] != 0.f) { nz.push_back(g.base + uint32_t(b)); break; }
  }
  if (!nz.empty()) {
    HIP_OK(s, hipMemcpy(s->d_cand, nz.data(), nz.size() * 4, hipMemcpyHostToDevice));
    launch_mark(s->stream, s->d_cand, uint32_t(nz.size()), s->touched);
    HIP_OK(s, hipStreamSynchronize(s->stream));
    std::vector<uint8_t> m(s->nb, 0);
    for (uint32_t b : s->pending) m[b] = 1;
    for (uint32_t b : nz) if (!m[b]) s->pending.push_back(b);
  }
  return FX_OK;
}
int fx_scene_read_flags(fx_sim* s, uint32_t scene, int32_t which, uint32_t* host, size_t count) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  const Grid* g = scene_grid(s, scene);
  if (!g) return FX_ERR_INVALID_ARG;
  if (count != brick_count(*g)) return fail(s, FX_ERR_INVALID_ARG, "count must equal the scene's brick count");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  const uint32_t* f = which == FX_FLAGS_INJECTED ? s->flagsInj : s->flagsRender;
  HIP_OK(s, hipMemcpy(host, f + g->base, count * 4, hipMemcpyDeviceToHost));
  return FX_OK;
}
int fx_scene_read_active_list(fx_sim* s, uint32_t scene, uint32_t* host, size_t capacity, uint32_t* count) {
  if (!s || !count) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  const Grid* g = scene_grid(s, scene);
  if (!g) return FX_ERR_INVALID_ARG;
  if (!s->multi) return fx_read_active_list(s, host, capacity, count);
  HIP_OK(s, hipSetDevice(s->device));
  std::vector<uint32_t> list;
  uint32_t k0 = 0, n = 0;
  int rc = scene_list_range(s, *g, list, k0, n);
  if (rc != FX_OK) return rc;
  *count = n;
  for (uint32_t i = 0; host && i < n && i < capacity; ++i) host[i] = list[k0 + i] - g->base;
  return FX_OK;
}
int fx_scene_read_page_table(fx_sim* s, uint32_t scene, int32_t* host, size_t count) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  const Grid* g = scene_grid(s, scene);
  if (!g) return FX_ERR_INVALID_ARG;
  if (!s->multi) return fx_read_page_table(s, host, count);
  if (count != brick_count(*g)) return fail(s, FX_ERR_INVALID_ARG, "count must equal the scene's brick count");
  HIP_OK(s, hipSetDevice(s->device));
  std::vector<uint32_t> list;
  uint32_t k0 = 0, n = 0;
  int rc = scene_list_range(s, *g, list, k0, n);
  if (rc != FX_OK) return rc;
  HIP_OK(s, hipMemcpy(host, s->lv[0].L.page + g->base, count * 4, hipMemcpyDeviceToHost));
  // positions within the scene's own list
  for (size_t b = 0; b < count; ++b)
    if (host[b] >= 0) host[b] -= int32_t(k0);
  return FX_OK;
}
int fx_scene_active_brick_count(fx_sim* s, uint32_t scene, uint32_t* out) {
  if (!s || !out) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  const Grid* g = scene_grid(s, scene);
  if (!g) return FX_ERR_INVALID_ARG;
  if (!s->multi) return fx_active_brick_count(s, out);
  HIP_OK(s, hipSetDevice(s->device));
  std::vector<uint32_t> list;
  uint32_t k0 = 0, n = 0;
  int rc = scene_list_range(s, *g, list, k0, n);
  if (rc != FX_OK) return rc;
  *out = n;
  return FX_OK;
}

// g: the grid of the scene rendered.
static int build_ray_const(fx_sim* s, const Grid& g, const fx_camera* c, const fx_render_params* p, uint32_t w, uint32_t h,
                           RayConst& R) {
  if (!c || !p || w == 0 || h == 0) return fail(s, FX_ERR_INVALID_ARG, "invalid raymarch arguments");
  // createVolumeRaymarchPass.ts:83-88, 196-262; volumeWorldBounds.ts:14-31
  double maxH = double(std::max(g.W, g.D));
  float hx = float(2.05 * 13.5 * (g.W / maxH)), hy = float(2.05 * 13.5 * (g.H / maxH)), hz = float(2.05 * 13.5 * (g.D / maxH));
  float cy = float(double(hy) - 0.25);
  std::memset(&R, 0, sizeof(R));
  for (int k = 0; k < 3; ++k) { R.cam_pos[k] = c->position[k]; R.right[k] = c->right[k]; R.up[k] = c->up[k]; R.fwd[k] = c->forward[k]; }
  R.aspect = c->aspect; R.tan_half = c->tan_half_fov_y;
  const float center[3] = {0.f, cy, 0.f}, half[3] = {hx, hy, hz};
  for (int k = 0; k < 3; ++k) { R.box_min[k] = center[k] - half[k]; R.box_max[k] = center[k] + half[k]; R.half_ext[k] = half[k]; }
  R.mode = float(p->display_mode);
  R.step_count = uint32_t(float(p->step_count));
  R.light_count = float(p->light_count);
  R.scat_f = p->scattering_forward; R.scat_b = p->scattering_back;
  for (int i = 0; i < 3; ++i) {
    if (i < p->light_count) {
      const fx_light& l = p->lights[i];
      for (int k = 0; k < 3; ++k) { R.light_dir[i][k] = l.direction[k]; R.light_pos[i][k] = l.position[k]; R.light_color[i][k] = l.color[k]; }
      R.light_kind[i] = l.type == 1 ? 1.f : 0.f;
      R.light_intensity[i] = l.intensity;
    }
  }
  R.width = w; R.height = h;
  return FX_OK;
}

}  // extern "C"
namespace {
// The raymarch of the scene with grid g (one scene: the simulation's grid) from
// the render snapshot, on the render stream.
int raymarch_scene(fx_sim* s, const Grid& g, const fx_camera* c, const fx_render_params* p, uint32_t w, uint32_t h, void* dev) {
  RayConst R;
  int rc = build_ray_const(s, g, c, p, w, h, R);
  if (rc != FX_OK) return rc;
  HIP_OK(s, hipSetDevice(s->device));
  if (!s->scalars_marked) {  // simulation work since the last step: snapshot everything now
    rc = pack_render(s, true);
    if (rc != FX_OK) return rc;
  }
  const int lane = s->multi && s->render_overlap ? int(g.scene % uint32_t(s->rlanes)) : 0;
  hipStream_t rs = !s->render_overlap ? s->stream : lane ? s->rl_stream[lane] : s->rstream;
  if (s->render_overlap) HIP_OK(s, hipStreamWaitEvent(rs, s->ev_scalars, 0));
  {
    StageTimer timer(s, FX_TIME_RAYMARCH, rs);
    launch_raymarch(rs, g, R, s->d_render, s->d_bits, s->d_rqueue + lane, s->ray_blocks, static_cast<uint16_t*>(dev));
  }
  if (s->render_overlap && lane) {
    HIP_OK(s, hipEventRecord(s->rl_ev[lane], rs));
    s->rl_pending[lane] = true;
  } else if (s->render_overlap) {
    HIP_OK(s, hipEventRecord(s->ev_render, s->rstream));
    s->render_pending = true;
  }
  s->last_render_stream = rs;
  s->render_wanted = true;  // the next step packs its snapshot
  HIP_OK(s, hipGetLastError());
  return FX_OK;
}
}  // namespace
extern "C" {

int fx_raymarch_device(fx_sim* s, const fx_camera* c, const fx_render_params* p, uint32_t w, uint32_t h, void* dev) {
  if (!s || !dev) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  return raymarch_scene(s, s->g, c, p, w, h, dev);
}
int fx_scene_raymarch_device(fx_sim* s, uint32_t scene, const fx_camera* c, const fx_render_params* p, uint32_t w, uint32_t h,
                             void* dev) {
  if (!s || !dev) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  const Grid* g = scene_grid(s, scene);
  if (!g) return FX_ERR_INVALID_ARG;
  return raymarch_scene(s, *g, c, p, w, h, dev);
}

int fx_bloom_device(fx_sim* s, const fx_bloom_params* p, uint32_t w, uint32_t h, const void* src, void* dst) {
  if (!s || !p || !src || !dst) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  if (w == 0 || h == 0 || uint64_t(w) * h > (1ull << 28)) return fail(s, FX_ERR_INVALID_ARG, "image size out of range");
  HIP_OK(s, hipSetDevice(s->device));
  hipStream_t rs = s->render_overlap ? s->rstream : s->stream;
  const uint32_t hw = std::max(1u, w / 2), hh = std::max(1u, h / 2);  // createBloomPass.ts:84-89
  if (w != s->bloom_w || h != s->bloom_h) {
    if (s->d_bloom) {
      HIP_OK(s, hipStreamSynchronize(rs));
      HIP_OK(s, hipFree(s->d_bloom));
      s->d_bloom = nullptr;
    }
    HIP_OK(s, hipMalloc(&s->d_bloom, size_t(hw) * hh * 8 * 3));
    s->bloom_w = w;
    s->bloom_h = h;
  }
  BloomArgs a{p->enabled ? 1 : 0, p->threshold, p->strength, p->radius, w, h, hw, hh};
  char* b = static_cast<char*>(s->d_bloom);
  const size_t img = size_t(hw) * hh * 8;
  launch_bloom(rs, a, src, b, b + img, b + 2 * img, dst);
  HIP_OK(s, hipGetLastError());
  return FX_OK;
}

int fx_bloom(fx_sim* s, const fx_bloom_params* p, uint32_t w, uint32_t h, const uint16_t* host_src, uint8_t* host_dst) {
  if (!s || !p || !host_src || !host_dst) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  if (w == 0 || h == 0 || uint64_t(w) * h > (1ull << 28)) return fail(s, FX_ERR_INVALID_ARG, "image size out of range");
  HIP_OK(s, hipSetDevice(s->device));
  hipStream_t rs = s->render_overlap ? s->rstream : s->stream;
  const size_t in = size_t(w) * h * 8, out = size_t(w) * h * 4;
  if (in + out > s->bloom_io_bytes) {
    HIP_OK(s, hipStreamSynchronize(rs));
    if (s->d_bloom_io) HIP_OK(s, hipFree(s->d_bloom_io));
    s->d_bloom_io = nullptr;
    HIP_OK(s, hipMalloc(&s->d_bloom_io, in + out));
    s->bloom_io_bytes = in + out;
  }
  char* io = static_cast<char*>(s->d_bloom_io);
  HIP_OK(s, hipMemcpyAsync(io, host_src, in, hipMemcpyHostToDevice, rs));
  int rc = fx_bloom_device(s, p, w, h, io, io + in);
  if (rc != FX_OK) return rc;
  HIP_OK(s, hipMemcpyAsync(host_dst, io + in, out, hipMemcpyDeviceToHost, rs));
  HIP_OK(s, hipStreamSynchronize(rs));
  return FX_OK;
}

}  // extern "C"
namespace {
// fx_raymarch of the scene with grid g into host memory (synchronous).
int raymarch_host(fx_sim* s, const Grid& g, const fx_camera* c, const fx_render_params* p, uint32_t w, uint32_t h, uint16_t* host) {
  HIP_OK(s, hipSetDevice(s->device));
  void* dev = nullptr;
  HIP_OK(s, hipMalloc(&dev, size_t(w) * h * 8));
  int rc = raymarch_scene(s, g, c, p, w, h, dev);
  if (rc == FX_OK) {
    // the stream the raymarch was enqueued on (the simulation stream when rendering is not overlapped)
    hipStream_t rs = s->last_render_stream;
    hipError_t e = hipMemcpyAsync(host, dev, size_t(w) * h * 8, hipMemcpyDeviceToHost, rs);
    if (e == hipSuccess) e = hipStreamSynchronize(rs);
    if (e != hipSuccess) rc = fail(s, FX_ERR_DEVICE, std::string("raymarch: ") + hipGetErrorString(e));
  }
  (void)hipFree(dev);
  return rc;
}
}  // namespace
extern "C" {
int fx_raymarch(fx_sim* s, const fx_camera* c, const fx_render_params* p, uint32_t w, uint32_t h, uint16_t* host) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  ONE_SCENE(s);
  return raymarch_host(s, s->g, c, p, w, h, host);
}
int fx_scene_raymarch(fx_sim* s, uint32_t scene, const fx_camera* c, const fx_render_params* p, uint32_t w, uint32_t h,
                      uint16_t* host) {
  if (!s || !host) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  const Grid* g = scene_grid(s, scene);
  if (!g) return FX_ERR_INVALID_ARG;
  return raymarch_host(s, *g, c, p, w, h, host);
}

}  // extern "C"

struct fx_batch {
  std::vector<fx_sim*> scenes;
};

extern "C" {

int fx_batch_create(const fx_create_desc* descs, uint32_t n, fx_batch** out) {
  if (!descs || !out || n == 0) return fail(nullptr, FX_ERR_INVALID_ARG, "invalid batch arguments");
  *out = nullptr;
  fx_batch* b = new (std::nothrow) fx_batch;
  if (!b) return fail(nullptr, FX_ERR_OUT_OF_MEMORY, "batch allocation failed");
  for (uint32_t i = 0; i < n; ++i) {
    fx_sim* sim = nullptr;
    const int rc = fx_create(&descs[i], &sim);
    if (rc != FX_OK) {
      for (fx_sim* q : b->scenes) fx_destroy(q);
      delete b;
      return rc;  // the message stays in fx_last_error(NULL)
    }
    b->scenes.push_back(sim);
  }
  *out = b;
  return FX_OK;
}
void fx_batch_destroy(fx_batch* b) {
  if (!b) return;
  for (fx_sim* q : b->scenes) fx_destroy(q);
  delete b;
}
uint32_t fx_batch_size(const fx_batch* b) { return b ? uint32_t(b->scenes.size()) : 0u; }
fx_sim* fx_batch_scene(fx_batch* b, uint32_t i) { return (b && i < b->scenes.size()) ? b->scenes[i] : nullptr; }
int fx_batch_step(fx_batch* b, double elapsed_s, double step_s) {
  if (!b) return fail(nullptr, FX_ERR_INVALID_ARG, "null batch");
  for (fx_sim* q : b->scenes) {
    const int rc = fx_step(q, elapsed_s, step_s);
    if (rc != FX_OK) return rc;
  }
  return FX_OK;
}
int fx_batch_synchronize(fx_batch* b) {
  if (!b) return fail(nullptr, FX_ERR_INVALID_ARG, "null batch");
  for (fx_sim* q : b->scenes) {
    const int rc = fx_synchronize(q);
    if (rc != FX_OK) return rc;
  }
  return FX_OK;
}

int fx_set_vcycle_plan(fx_sim* s, uint32_t down_split, uint32_t up_split) {
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  s->split_down = down_split;
  s->split_up = up_split;
  return FX_OK;
}
int fx_get_vcycle_plan(fx_sim* s, uint32_t* down_split, uint32_t* up_split) {
  if (!s || !down_split || !up_split) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  *down_split = s->split_down;
  *up_split = s->split_up;
  return FX_OK;
}
int fx_autotune_vcycle(fx_sim* s, uint32_t repeats, uint32_t* down_split, uint32_t* up_split) {
  if (!s || !down_split || !up_split) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  if (s->desc.solver != FX_SOLVER_VCYCLE) return fail(s, FX_ERR_UNSUPPORTED, "the autotuner plans the V-cycle solver");
  if (s->timing) return fail(s, FX_ERR_UNSUPPORTED, "stage timing is enabled");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  HIP_OK(s, hipStreamSynchronize(s->rstream));
  for (int r = 1; r < s->rlanes; ++r) HIP_OK(s, hipStreamSynchronize(s->rl_stream[r]));
  repeats = std::max(1u, repeats);
  update_coarse_plan(s);  // the last step's active count has arrived
  hipEvent_t a = nullptr, b = nullptr;
  HIP_OK(s, hipEventCreate(&a));
  HIP_OK(s, hipEventCreate(&b));
  // Mean time of one graph-captured solve under the current plan.
  auto time_plan = [&](float* ms) -> int {
    HIP_OK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    const int last = solve_vcycle(s);
    hipGraph_t g = nullptr;
    HIP_OK(s, hipStreamEndCapture(s->stream, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiateWithFlags(&ge, g, 0);
    if (e == hipSuccess) e = hipGraphLaunch(ge, s->stream);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(a, s->stream);
    for (uint32_t r = 0; r < repeats && e == hipSuccess; ++r) e = hipGraphLaunch(ge, s->stream);
    if (e == hipSuccess) e = hipEventRecord(b, s->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    if (ge) (void)hipGraphExecDestroy(ge);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(s, FX_ERR_DEVICE, std::string("autotune: ") + hipGetErrorString(e));
    *ms /= float(repeats);
    s->lastPressure = last;
    return FX_OK;
  };
  int rc = FX_OK;
  const int L = int(s->lv.size());
  {
    // bring the clocks up before the first measurement
    float ms = 0.f;
    for (int w = 0; w < 3 && rc == FX_OK; ++w) rc = time_plan(&ms);
  }
  // Per leg: 0 fused LDS kernel, 1 split with one sweep per launch, 2 split with
  // two sweeps per launch (only where the leg has two or more sweeps). The split
  // legs with the boundary sweep fused (plan bit 16 + l) are not timed: on a 15%
  // active 256^3 plume they were slower at every level (0.985-1.18 against
  // 0.941 ms per step) and would nearly double the tuning time; callers can still
  // pin them with fx_set_vcycle_plan.
  auto set_leg = [&](uint32_t& mask, int l, uint32_t opt) {
    mask &= ~((1u << l) | (1u << (kPlanPairShift + l)) | (1u << (kPlanFuseShift + l)));
    if (opt >= 1) mask |= 1u << l;
    if (opt == 2) mask |= 1u << (kPlanPairShift + l);
  };
  const fx_vcycle& vc = s->desc.vcycle;
  std::vector<uint32_t> downs = {0u, 1u}, ups = {0u, 1u};
  if (vc.pre_sweeps >= 2) downs.push_back(2u);
  if (vc.post_sweeps >= 2) ups.push_back(2u);
  const uint32_t np = uint32_t(downs.size() * ups.size());
  for (int l = 0; l + 1 < L && rc == FX_OK; ++l) {
    s->split_up &= ~(3u << (kPlanTileShift + 2 * l));  // legs are compared on 8^3 bricks
    // each plan's best of two interleaved rounds
    std::vector<float> t(np, 0.f);
    for (int round = 0; round < 2 && rc == FX_OK; ++round)
      for (uint32_t c = 0; c < np && rc == FX_OK; ++c) {
        set_leg(s->split_down, l, downs[c % downs.size()]);
        set_leg(s->split_up, l, ups[c / downs.size()]);
        float ms = 0.f;
        rc = time_plan(&ms);
        if (rc == FX_OK && (round == 0 || ms < t[c])) t[c] = ms;
      }
    uint32_t best_c = 0;
    for (uint32_t c = 1; c < np; ++c)
      if (t[c] < t[best_c]) best_c = c;
    set_leg(s->split_down, l, downs[best_c % downs.size()]);
    set_leg(s->split_up, l, ups[best_c / downs.size()]);
    // a split leg at this level: its sweeps' compute tile (8^3 bricks, 4^3 octants
    // or 16^3 brick groups; the pair bits chosen above give the halo)
    if ((((s->split_down | s->split_up) >> l) & 1u) && rc == FX_OK) {
      const uint32_t tmask = 3u << (kPlanTileShift + 2 * l);
      float tt[3] = {t[best_c], 0.f, 0.f};
      for (int round = 0; round < 2 && rc == FX_OK; ++round)
        for (uint32_t tc = 1; tc < 3 && rc == FX_OK; ++tc) {
          s->split_up = (s->split_up & ~tmask) | (tc << (kPlanTileShift + 2 * l));
          float ms = 0.f;
          rc = time_plan(&ms);
          if (rc == FX_OK && (round == 0 || ms < tt[tc])) tt[tc] = ms;
        }
      const uint32_t best_t = tt[1] < tt[0] ? (tt[2] < tt[1] ? 2u : 1u) : (tt[2] < tt[0] ? 2u : 0u);
      s->split_up = (s->split_up & ~tmask) | (best_t << (kPlanTileShift + 2 * l));
    }
  }
  if (s->coarse_split && rc == FX_OK) {
    // the coarsest level beyond the one-workgroup smoother: 1, 2 or 4 sweeps per launch
    float t[3] = {0.f, 0.f, 0.f};
    for (int round = 0; round < 2 && rc == FX_OK; ++round)
      for (uint32_t opt = 0; opt < 3 && rc == FX_OK; ++opt) {
        s->split_down = (s->split_down & ~(3u << kPlanCoarseShift)) | (opt << kPlanCoarseShift);
        float ms = 0.f;
        rc = time_plan(&ms);
        if (rc == FX_OK && (round == 0 || ms < t[opt])) t[opt] = ms;
      }
    const uint32_t best = t[1] < t[0] ? (t[2] < t[1] ? 2u : 1u) : (t[2] < t[0] ? 2u : 0u);
    s->split_down = (s->split_down & ~(3u << kPlanCoarseShift)) | (best << kPlanCoarseShift);
  }
  // leave the pressure buffers as the chosen plan computes them
  float ms = 0.f;
  if (rc == FX_OK) rc = time_plan(&ms);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *down_split = s->split_down;
  *up_split = s->split_up;
  return rc;
}

int fx_time_fine_sweep(fx_sim* s, uint32_t repeats, double* avg_ms, uint64_t* voxels) {
  if (!s || !avg_ms || !voxels) return fail(s, FX_ERR_INVALID_ARG, "null argument");
  if (s->desc.solver != FX_SOLVER_VCYCLE) return fail(s, FX_ERR_UNSUPPORTED, "the fine sweep belongs to the V-cycle solver");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  repeats = std::max(1u, repeats);
  Level& V = s->lv[0];
  const float* src = V.p[s->lastPressure];
  float* dst = V.p[1 - s->lastPressure];
  hipEvent_t a = nullptr, b = nullptr;
  HIP_OK(s, hipEventCreate(&a));
  HIP_OK(s, hipEventCreate(&b));
  launch_sweep(s->stream, grid_for(s, V.nb), level_list(s, 0), level_count(s, 0), V.g, bits_of(s, 0), s->d_const, V.div, src, dst,
               true);
  hipError_t e = hipEventRecord(a, s->stream);
  for (uint32_t r = 0; r < repeats; ++r)
    launch_sweep(s->stream, grid_for(s, V.nb), level_list(s, 0), level_count(s, 0), V.g, bits_of(s, 0), s->d_const, V.div, src, dst,
                 true);
  if (e == hipSuccess) e = hipEventRecord(b, s->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  uint32_t n = 0;
  if (e == hipSuccess) e = hipMemcpy(&n, level_count(s, 0), 4, hipMemcpyDeviceToHost);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess) return fail(s, FX_ERR_DEVICE, std::string("fine sweep clock: ") + hipGetErrorString(e));
  *avg_ms = double(ms) / double(repeats);
  *voxels = uint64_t(n) * 512u;
  return FX_OK;
}

int fx_set_stage_timing(fx_sim* s, int32_t enable) {
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  HIP_OK(s, hipStreamSynchronize(s->rstream));
  for (int r = 1; r < s->rlanes; ++r) HIP_OK(s, hipStreamSynchronize(s->rl_stream[r]));
  drain_timing(s);
  std::memset(s->acc_ms, 0, sizeof(s->acc_ms));
  std::memset(s->acc_calls, 0, sizeof(s->acc_calls));
  s->timing = enable != 0;
  return FX_OK;
}

int fx_read_stage_times(fx_sim* s, double* total_ms, uint32_t* calls, size_t n) {
  if (!s) return fail(nullptr, FX_ERR_INVALID_ARG, "null simulation");
  HIP_OK(s, hipSetDevice(s->device));
  HIP_OK(s, hipStreamSynchronize(s->stream));
  HIP_OK(s, hipStreamSynchronize(s->rstream));
  for (int r = 1; r < s->rlanes; ++r) HIP_OK(s, hipStreamSynchronize(s->rl_stream[r]));
  drain_timing(s);
  for (size_t i = 0; i < std::min<size_t>(n, FX_TIME_COUNT); ++i) {
    if (total_ms) total_ms[i] = s->acc_ms[i];
    if (calls) calls[i] = s->acc_calls[i];
  }
  return FX_OK;
}

const char* fx_last_error(fx_sim* s) { return s ? s->err.c_str() : g_create_error.c_str(); }

}  // extern "C"

// ---- hooks for the differentiable step (fenix_diff.cpp) ---------------------
namespace fx {

SimView sim_view(fx_sim* s) {
  return SimView{s->stream, s->g, fine_list(s), fine_count(s), grid_for(s, s->nb), s->nb, storage_count(s->g)};
}
int sim_fail(fx_sim* s, int code, const char* msg) { return fail(s, code, msg); }
int sim_device(fx_sim* s) { return s->device; }
hipStream_t sim_render_stream(fx_sim* s) { return s->render_overlap ? s->rstream : s->stream; }

void sim_set_blocks_per_cu(fx_sim* s, int per_cu) {
  if (!s->blocks_env) s->blocks = s->cus * std::max(1, per_cu);
}
int sim_prepare_projection(fx_sim* s) {
  return upload_step(s, std::vector<fx_sim_params>(1), nullptr, 0);
}
int sim_project(fx_sim* s, Vec3 us, Vec3 u) {
  const SimView v = sim_view(s);
  launch_divergence(v.stream, v.blocks, v.list, v.count, v.g, us, s->lv[0].div);
  const int cur = solve_vcycle(s);
  launch_projection(v.stream, v.blocks, v.list, v.count, v.g, us, s->lv[0].p[cur], u);
  s->lastPressure = cur;
  HIP_OK(s, hipGetLastError());
  return FX_OK;
}
void bricks_from_rowmajor(const Grid& g, const float* rm, float* bm) { scatter_brickmajor(g, rm, bm); }
void bricks_to_rowmajor(const Grid& g, const float* bm, float* rm) { gather_rowmajor(g, bm, rm); }

}  // namespace fx
