// fenix_diff.cpp — the differentiable smoke step of include/fenix.h (fx_diff_*):
// rollouts of the forward model, the loss, and its adjoint with checkpointed
// recomputation. The pressure projection runs through the V-cycle of an
// internal dense simulation (fx_sim), which is also the projection's adjoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "fenix_internal.h"

using namespace fx;

namespace {
thread_local std::string g_diff_create_error;

// Per-step intermediates of the forward model that its adjoint reads.
struct Tape {
  float* ds = nullptr;                        // density after the source
  float* us[3] = {nullptr, nullptr, nullptr};  // velocity after the source
  float* ua[3] = {nullptr, nullptr, nullptr};  // after velocity advection
  float* ub[3] = {nullptr, nullptr, nullptr};  // after buoyancy
};
Vec3 v3(float* const* p) { return Vec3{p[0], p[1], p[2]}; }
}  // namespace

struct fx_diff {
  fx_diff_desc desc{};
  fx_sim* sim = nullptr;  // dense V-cycle simulation: stream, brick list, projection
  SimView v{};
  size_t n = 0;           // floats per brick-major field
  size_t cfl = 0;         // control floats
  std::vector<void*> owned;

  float* d = nullptr;  // state of the last rollout
  float* u[3] = {};
  float* rd = nullptr;  // state of a recomputed segment
  float* ru[3] = {};
  float* target = nullptr;
  float* ctrl = nullptr;
  float* cbar = nullptr;
  std::vector<Tape> tape;          // checkpoint_every slots
  std::vector<float*> ckpt;        // 4 fields per checkpoint: d, u_x, u_y, u_z
  float* uc[3] = {};               // velocity after confinement
  float* w[3] = {};                // vorticity (forward and adjoint)
  float* wmag = nullptr;
  float* dA = nullptr;             // adjoint of the state after the current step
  float* uA[3] = {};
  float* dB = nullptr;             // adjoint of the state after the source (scatter target)
  float* uB[3] = {};
  float* aA[3] = {};               // adjoint of the velocity inside the step
  float* gbar[3] = {};
  float* wbar[3] = {};
  float* stage = nullptr;          // staged scatter tiles of the advection adjoints
  double* loss_acc = nullptr;      // kNormLanes partial sums
  std::string err;
};

namespace {

int fail(fx_diff* f, int code, const std::string& msg) {
  if (f) f->err = msg; else g_diff_create_error = msg;
  return code;
}
#define DIFF_HIP(f, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(f, FX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

int alloc(fx_diff* f, float** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(float));
  if (e != hipSuccess) return fail(f, FX_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  f->owned.push_back(q);
  *p = static_cast<float*>(q);
  return FX_OK;
}

// One forward step from state (d, u) in place; intermediates into T.
void forward_step(fx_diff* f, float* d, float* const* u, const Tape& T) {
  const SimView& v = f->v;
  const fx_diff_desc& D = f->desc;
  launch_df_source(v.stream, v.blocks, v.list, v.count, v.g, f->ctrl, D.control_n, D.slab, D.dt, d, v3(u), T.ds, v3(T.us));
  launch_df_advect(v.stream, v.blocks, v.list, v.count, v.g, v3(T.us), v3(T.ua), T.ds, d, D.dt);
  launch_df_forces(v.stream, v.blocks, v.list, v.count, v.g, d, v3(T.ua), v3(T.ub), v3(f->w), f->wmag, v3(f->uc), D.dt * D.buoyancy,
                   D.dt * D.confinement);
}

// Rollout of desc.steps steps from the zero state into (f->d, f->u); with
// checkpoints, the state before every checkpoint_every-th step is kept.
int rollout(fx_diff* f, bool checkpoints) {
  const SimView& v = f->v;
  launch_clear_f32(v.stream, f->d, f->n);
  for (int c = 0; c < 3; ++c) launch_clear_f32(v.stream, f->u[c], f->n);
  const uint32_t ce = f->desc.checkpoint_every;
  for (uint32_t k = 0; k < f->desc.steps; ++k) {
    if (checkpoints && k % ce == 0) {
      float** ck = &f->ckpt[size_t(k / ce) * 4];
      DIFF_HIP(f, hipMemcpyAsync(ck[0], f->d, f->n * 4, hipMemcpyDeviceToDevice, v.stream));
      for (int c = 0; c < 3; ++c) DIFF_HIP(f, hipMemcpyAsync(ck[1 + c], f->u[c], f->n * 4, hipMemcpyDeviceToDevice, v.stream));
    }
    forward_step(f, f->d, f->u, f->tape[0]);
    int rc = sim_project(f->sim, v3(f->uc), v3(f->u));
    if (rc != FX_OK) return fail(f, rc, fx_last_error(f->sim));
  }
  DIFF_HIP(f, hipGetLastError());
  return FX_OK;
}

int loss_of(fx_diff* f, double* loss, bool seed) {
  const SimView& v = f->v;
  launch_clear_u32(v.stream, reinterpret_cast<uint32_t*>(f->loss_acc), size_t(kNormLanes) * 2, 0u);
  launch_df_loss(v.stream, v.blocks, v.list, v.count, v.g, f->d, f->target, seed ? f->dA : nullptr, f->loss_acc);
  double h[kNormLanes];
  DIFF_HIP(f, hipMemcpyAsync(h, f->loss_acc, sizeof(h), hipMemcpyDeviceToHost, v.stream));
  DIFF_HIP(f, hipStreamSynchronize(v.stream));
  double s = 0.0;
  for (int k = 0; k < kNormLanes; ++k) s += h[k];
  if (loss) *loss = s;
  return FX_OK;
}

int upload_control(fx_diff* f, const float* control) {
  DIFF_HIP(f, hipMemcpyAsync(f->ctrl, control, f->cfl * 4, hipMemcpyHostToDevice, f->v.stream));
  return FX_OK;
}

// Adjoint of one step: (dA, uA) = adjoint of the state after step k on entry,
// of the state before it on exit; cbar accumulates the control gradient.
int adjoint_step(fx_diff* f, const Tape& T) {
  const SimView& v = f->v;
  const fx_diff_desc& D = f->desc;
  // projection (self-adjoint): aA = P(uA) with the boundary adjoint dropped
  launch_da_zero_boundary(v.stream, v.blocks, v.list, v.count, v.g, v3(f->uA));
  int rc = sim_project(f->sim, v3(f->uA), v3(f->aA));
  if (rc != FX_OK) return fail(f, rc, fx_last_error(f->sim));
  // confinement and buoyancy
  launch_da_forces(v.stream, v.blocks, v.list, v.count, v.g, v3(T.ub), v3(f->w), f->wmag, v3(f->aA), v3(f->gbar), v3(f->wbar), f->dA,
                   D.dt * D.buoyancy, D.dt * D.confinement);
  // advection of density and velocity
  launch_clear_f32(v.stream, f->dB, f->n);
  for (int c = 0; c < 3; ++c) launch_clear_f32(v.stream, f->uB[c], f->n);
  launch_da_advect(v.stream, v.blocks, v.list, v.count, v.g, v.nb, T.ds, v3(T.us), v3(T.ua), f->dA, f->dB, v3(f->aA), v3(f->uB), D.dt,
                   f->stage);
  // source
  launch_da_source(v.stream, v.blocks, v.list, v.count, v.g, D.control_n, D.slab, D.dt, f->dB, v3(f->uB), f->cbar);
  std::swap(f->dA, f->dB);
  for (int c = 0; c < 3; ++c) std::swap(f->uA[c], f->uB[c]);
  DIFF_HIP(f, hipGetLastError());
  return FX_OK;
}

}  // namespace

extern "C" {

void fx_diff_default_desc(fx_diff_desc* d) {
  if (!d) return;
  std::memset(d, 0, sizeof(*d));
  d->width = d->height = d->depth = 64;
  d->steps = 64;
  d->checkpoint_every = 8;
  d->control_n = 32;
  d->slab = 4;
  d->dt = 1.0f;
  d->buoyancy = 0.02f;
  d->confinement = 0.02f;
  d->vcycle.levels = 4;
  d->vcycle.pre_sweeps = 2;
  d->vcycle.post_sweeps = 2;
  d->vcycle.cycles = 4;
  d->vcycle.coarse_sweeps = 16;
  d->vcycle.omega = 2.0f / 3.0f;
  d->device = 0;
}

void fx_diff_destroy(fx_diff* f) {
  if (!f) return;
  if (f->sim) {
    fx_synchronize(f->sim);
    hipSetDevice(f->desc.device);
  }
  for (void* p : f->owned) hipFree(p);
  if (f->sim) fx_destroy(f->sim);
  delete f;
}

int fx_diff_create(const fx_diff_desc* desc, fx_diff** out) {
  if (!desc || !out) return fail(nullptr, FX_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  const fx_diff_desc& D = *desc;
  if (D.width < 8 || D.height < 8 || D.depth < 8) return fail(nullptr, FX_ERR_INVALID_ARG, "grid must be at least 8 in every axis");
  if (D.steps == 0 || D.checkpoint_every == 0) return fail(nullptr, FX_ERR_INVALID_ARG, "steps and checkpoint_every must be positive");
  if (D.control_n < 2 || D.control_n > 1024) return fail(nullptr, FX_ERR_INVALID_ARG, "control_n must be in [2, 1024]");
  if (D.slab < 1 || D.slab + 2 > D.height) return fail(nullptr, FX_ERR_INVALID_ARG, "slab must be in [1, height - 2]");
  if (!(D.dt > 0.f)) return fail(nullptr, FX_ERR_INVALID_ARG, "dt must be positive");
  if (D.vcycle.pre_sweeps != D.vcycle.post_sweeps)
    return fail(nullptr, FX_ERR_INVALID_ARG, "the adjoint projection needs a symmetric V-cycle (pre_sweeps == post_sweeps)");
  fx_create_desc cd;
  fx_default_create_desc(&cd);
  cd.width = D.width; cd.height = D.height; cd.depth = D.depth;
  cd.domain_mode = FX_DOMAIN_DENSE;
  cd.solver = FX_SOLVER_VCYCLE;
  cd.precision = FX_PRECISION_FP32;
  cd.vcycle = D.vcycle;
  cd.debug_fields = 0;
  cd.device = D.device;
  fx_diff* f = new (std::nothrow) fx_diff();
  if (!f) return fail(nullptr, FX_ERR_OUT_OF_MEMORY, "host allocation failed");
  f->desc = D;
  int rc = fx_create(&cd, &f->sim);
  if (rc != FX_OK) {
    std::string m = fx_last_error(nullptr);
    delete f;
    return fail(nullptr, rc, "internal simulation: " + m);
  }
  sim_set_blocks_per_cu(f->sim, 4);  // nothing renders beside the differentiable step
  f->v = sim_view(f->sim);
  f->n = f->v.storage;
  f->cfl = size_t(4) * D.control_n * D.control_n * D.control_n;
  const uint32_t nck = (D.steps + D.checkpoint_every - 1) / D.checkpoint_every;
  const uint32_t nslots = std::min(D.checkpoint_every, D.steps);
  auto A = [&](float** p, size_t cnt) { return rc == FX_OK ? (rc = alloc(f, p, cnt)) : rc; };
  A(&f->d, f->n); A(&f->rd, f->n); A(&f->target, f->n); A(&f->wmag, f->n); A(&f->dA, f->n); A(&f->dB, f->n);
  A(&f->ctrl, f->cfl); A(&f->cbar, f->cfl); A(&f->stage, da_stage_floats(f->v.nb));
  for (int c = 0; c < 3; ++c) {
    A(&f->u[c], f->n); A(&f->ru[c], f->n); A(&f->uc[c], f->n); A(&f->w[c], f->n);
    A(&f->uA[c], f->n); A(&f->uB[c], f->n); A(&f->aA[c], f->n); A(&f->gbar[c], f->n); A(&f->wbar[c], f->n);
  }
  f->tape.resize(nslots);
  for (Tape& T : f->tape) {
    A(&T.ds, f->n);
    for (int c = 0; c < 3; ++c) { A(&T.us[c], f->n); A(&T.ua[c], f->n); A(&T.ub[c], f->n); }
  }
  f->ckpt.assign(size_t(nck) * 4, nullptr);
  for (float*& p : f->ckpt) A(&p, f->n);
  if (rc == FX_OK) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(double) * kNormLanes) != hipSuccess) {
      rc = fail(f, FX_ERR_OUT_OF_MEMORY, "hipMalloc: loss accumulator");
    } else {
      f->owned.push_back(q);
      f->loss_acc = static_cast<double*>(q);
    }
  }
  if (rc == FX_OK && (rc = sim_prepare_projection(f->sim)) != FX_OK) fail(f, rc, fx_last_error(f->sim));
  if (rc == FX_OK) {
    // Kernel plan of the projection's V-cycle for this dense grid (every plan
    // gives bit-identical results; timing does not depend on the field values).
    uint32_t down = 0, up = 0;
    if ((rc = fx_autotune_vcycle(f->sim, 3, &down, &up)) != FX_OK) fail(f, rc, fx_last_error(f->sim));
  }
  if (rc == FX_OK) {
    launch_clear_f32(f->v.stream, f->target, f->n);
    if (hipStreamSynchronize(f->v.stream) != hipSuccess) rc = fail(f, FX_ERR_DEVICE, "initialisation failed");
  }
  if (rc != FX_OK) {
    g_diff_create_error = f->err;
    fx_diff_destroy(f);
    return rc;
  }
  *out = f;
  return FX_OK;
}

int fx_diff_set_target(fx_diff* f, const float* host, size_t bytes) {
  if (!f || !host) return fail(f, FX_ERR_INVALID_ARG, "null argument");
  const Grid& g = f->v.g;
  if (bytes != size_t(g.W) * g.H * g.D * 4) return fail(f, FX_ERR_INVALID_ARG, "buffer size mismatch");
  std::vector<float> bm(f->n, 0.f);
  bricks_from_rowmajor(g, host, bm.data());
  DIFF_HIP(f, hipSetDevice(f->desc.device));
  DIFF_HIP(f, hipMemcpyAsync(f->target, bm.data(), f->n * 4, hipMemcpyHostToDevice, f->v.stream));
  DIFF_HIP(f, hipStreamSynchronize(f->v.stream));
  return FX_OK;
}

int fx_diff_forward(fx_diff* f, const float* control, double* loss) {
  if (!f || !control) return fail(f, FX_ERR_INVALID_ARG, "null argument");
  DIFF_HIP(f, hipSetDevice(f->desc.device));
  int rc = upload_control(f, control);
  if (rc == FX_OK) rc = rollout(f, false);
  if (rc == FX_OK) rc = loss_of(f, loss, false);
  return rc;
}

int fx_diff_gradient(fx_diff* f, const float* control, double* loss, float* grad) {
  if (!f || !control || !grad) return fail(f, FX_ERR_INVALID_ARG, "null argument");
  DIFF_HIP(f, hipSetDevice(f->desc.device));
  const SimView& v = f->v;
  int rc = upload_control(f, control);
  if (rc == FX_OK) rc = rollout(f, true);
  if (rc == FX_OK) rc = loss_of(f, loss, true);  // seeds dA = d_T - target
  if (rc != FX_OK) return rc;
  for (int c = 0; c < 3; ++c) launch_clear_f32(v.stream, f->uA[c], f->n);
  launch_clear_f32(v.stream, f->cbar, f->cfl);
  const uint32_t ce = f->desc.checkpoint_every, steps = f->desc.steps;
  const uint32_t nseg = (steps + ce - 1) / ce;
  for (uint32_t sgi = nseg; sgi-- > 0;) {
    const uint32_t k0 = sgi * ce, k1 = std::min(k0 + ce, steps);
    float** ck = &f->ckpt[size_t(sgi) * 4];
    DIFF_HIP(f, hipMemcpyAsync(f->rd, ck[0], f->n * 4, hipMemcpyDeviceToDevice, v.stream));
    for (int c = 0; c < 3; ++c) DIFF_HIP(f, hipMemcpyAsync(f->ru[c], ck[1 + c], f->n * 4, hipMemcpyDeviceToDevice, v.stream));
    for (uint32_t k = k0; k < k1; ++k) {
      forward_step(f, f->rd, f->ru, f->tape[k - k0]);
      // the last step's projection feeds nothing the adjoint reads
      if (k + 1 < k1) {
        rc = sim_project(f->sim, v3(f->uc), v3(f->ru));
        if (rc != FX_OK) return fail(f, rc, fx_last_error(f->sim));
      }
    }
    for (uint32_t k = k1; k-- > k0;) {
      rc = adjoint_step(f, f->tape[k - k0]);
      if (rc != FX_OK) return rc;
    }
  }
  DIFF_HIP(f, hipMemcpyAsync(grad, f->cbar, f->cfl * 4, hipMemcpyDeviceToHost, v.stream));
  DIFF_HIP(f, hipStreamSynchronize(v.stream));
  return FX_OK;
}

int fx_diff_read_field(fx_diff* f, int32_t field, float* host, size_t bytes) {
  if (!f || !host) return fail(f, FX_ERR_INVALID_ARG, "null argument");
  if (field < 0 || field > 3) return fail(f, FX_ERR_INVALID_ARG, "field must be 0 (density) or 1..3 (velocity)");
  const Grid& g = f->v.g;
  if (bytes != size_t(g.W) * g.H * g.D * 4) return fail(f, FX_ERR_INVALID_ARG, "buffer size mismatch");
  DIFF_HIP(f, hipSetDevice(f->desc.device));
  std::vector<float> bm(f->n);
  const float* src = field == 0 ? f->d : f->u[field - 1];
  DIFF_HIP(f, hipStreamSynchronize(f->v.stream));
  DIFF_HIP(f, hipMemcpy(bm.data(), src, f->n * 4, hipMemcpyDeviceToHost));
  bricks_to_rowmajor(g, bm.data(), host);
  return FX_OK;
}

const char* fx_diff_last_error(fx_diff* f) { return f ? f->err.c_str() : g_diff_create_error.c_str(); }

}  // extern "C"
