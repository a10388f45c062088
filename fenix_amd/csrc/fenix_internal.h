// fenix_internal.h — shared declarations between the HIP kernels and the host
// side of libfenix. Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fenix.h"

#define FX_MAX_LEVELS_DEV 8
// Residual norms: each cycle's FP64 sum of squares is spread over this many
// accumulators (per-block atomics on one address serialise); the host adds them.
constexpr int kNormLanes = 32;
#define FX_MAX_ZERO_BUFS 24

namespace fx {

// Dimensions of one grid level and its 8^3 brick tiling. Storage is
// brick-major: brick b = bx + BX*(by + BY*bz) owns elements [b*512, b*512+512),
// voxel (x,y,z) sits at b*512 + (z&7)*64 + (y&7)*8 + (x&7).
// One level of one scene. Several scenes (a brick-queue batch) share each level's
// storage: scene s owns bricks [base, base + BX*BY*BZ) of it, and a level's
// launch walks the bricks of every scene. There the grid passed to a kernel is a
// selector (tab != nullptr): the kernel reads the grid of brick b's scene,
// tab[owner[b]], which carries that scene's base and index. `span` is the brick
// count of the level's storage (all scenes), the bound of brick arithmetic.
struct Grid {
  uint32_t W, H, D;
  uint32_t BX, BY, BZ;
  uint32_t base;          // first brick of the scene in the level's storage
  uint32_t scene;         // scene index (its step constants and sources)
  uint32_t span;          // bricks in the level's storage
  const Grid* tab;        // selector: device grid of every scene of the level
  const uint32_t* owner;  // selector: scene of every brick of the level
};

inline Grid make_grid(uint32_t W, uint32_t H, uint32_t D) {
  const uint32_t BX = (W + 7) / 8, BY = (H + 7) / 8, BZ = (D + 7) / 8;
  return Grid{W, H, D, BX, BY, BZ, 0u, 0u, BX * BY * BZ, nullptr, nullptr};
}
inline size_t brick_count(const Grid& g) { return size_t(g.BX) * g.BY * g.BZ; }
inline size_t storage_count(const Grid& g) { return brick_count(g) * 512; }

// Per-step constants read by every kernel from device memory (uploaded once per
// step; scalar loads).
struct StepConst {
  fx_sim_params p;        // the reference's 80-byte uniform
  uint32_t nsrc;          // emitters iterated (>= 1, the reference's 128-byte minimum buffer)
  float ignition_lo, ignition_hi, burn_rate;
  float act_density, act_speed;
  float omega;            // V-cycle weighted-Jacobi factor
  uint32_t inj_count;     // sparse: bricks visited by injection (head of the candidate list)
  uint32_t cand_count;    // sparse: injection bricks + bricks written by the host
  uint32_t step_key;      // steps completed: the step term of the stochastic-rounding hash
  uint32_t sr;            // 16-bit stores round stochastically (1) or to nearest even (0)
  uint32_t src_off;       // first float of the scene's emitters in the source buffer
};

// Brick list of one level: ascending brick indices, device count, page table.
struct BrickList {
  uint32_t* list;    // capacity = brick count
  uint32_t* count;   // device scalar
  int32_t* page;     // brick -> list position or -1
  uint32_t* retire;  // bricks to zero (sparse)
  uint32_t* retire_count;
};

// Activity of the voxels of one pressure level in a sparse domain: level-l
// voxel J lies inside fine brick (J << l) >> 3 and is active iff that brick is.
// flags == nullptr: every voxel of every listed brick is active (dense).
struct LvMask {
  const uint32_t* flags;  // fine-level brick flags
  Grid fine;              // fine level (a selector when several scenes share it)
  uint32_t shift;         // level index
};

// Voxel activity bits of one pressure level, built once per solve: 8 words per
// listed brick (one per z-slice, bit x + 8y), indexed by list position. A voxel's
// bit is set iff it lies in the domain, off the boundary faces, and is active
// under the level's LvMask. page == nullptr: list position = brick index (dense).
struct LvBits {
  const uint64_t* bits;
  const int32_t* page;
  // Fine level of a sparse V-cycle: per list position, bit d set when the face
  // neighbour brick d (+x, -x, +y, -y, +z, -z) of the brick is listed; an
  // unlisted brick holds zeros, so the fine sweep reads its faces as 0 without
  // loading them. nullptr: every neighbour is loaded.
  const uint32_t* nbr;
};
struct LevelMaskArgs {
  int nlevels;
  const uint32_t* list[FX_MAX_LEVELS_DEV];
  const uint32_t* count[FX_MAX_LEVELS_DEV];
  Grid g[FX_MAX_LEVELS_DEV];
  LvMask m[FX_MAX_LEVELS_DEV];
  uint64_t* bits[FX_MAX_LEVELS_DEV];
  uint32_t* zero;  // zero_words words cleared by workgroup 0 (the solve's FP64 norm accumulators)
  size_t zero_words;
  uint32_t* nbr0;        // LvBits::nbr of the fine level (nullptr: not built)
  const int32_t* page0;  // the fine level's page table
};

// Launch helpers implemented in fenix_kernels.hip. All are asynchronous on `s`.
// `list`/`count` select the bricks visited; `grid_blocks` is the persistent grid.
struct Fields;  // host-side view, see fenix_api.cpp

// One scalar set. Density, temperature and fuel use the simulation's storage
// precision (FX_PRECISION_*: float, bf16 bits, half or 16-bit fixed point);
// reaction is FP32.
struct Scalars {
  void* d;
  void* t;
  void* f;
  float* r;
  int prec;
};
inline size_t precision_bytes(int prec) { return prec == FX_PRECISION_FP32 ? 4 : 2; }
struct Vec3 {
  float* x;
  float* y;
  float* z;
};

// clear (optional): clear_n words zeroed by the same launch.
void launch_inject(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                   const StepConst* c, const float* src, Scalars set0, Vec3 u, uint32_t* touched, int parts,
                   uint32_t* clear = nullptr, uint32_t clear_n = 0);
void launch_clear_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t value);
void launch_mark(hipStream_t s, const uint32_t* list, uint32_t n, uint32_t* touched);
void launch_clear_f32(hipStream_t s, float* p, size_t n);
struct LevelFlags;
void launch_activate(hipStream_t s, int blocks, const uint32_t* listA, const uint32_t* countA,
                     const uint32_t* listB, const uint32_t* countB, Grid g, const StepConst* c, int rule,
                     Scalars set, Vec3 vel, uint32_t* flags, const LevelFlags& lf);
// Coarse-level brick flags written by the activation kernels in a sparse
// domain: marking fine brick f also marks level-l brick (f >> l) per axis.
struct LevelFlags {
  int n;                                 // coarse levels (0 when dense)
  uint32_t* f[FX_MAX_LEVELS_DEV];        // index 1..n
  Grid g[FX_MAX_LEVELS_DEV];             // index 1..n (selectors when several scenes share the levels)
};

// Brick-list compaction of every level (see k_compact_* in fenix_kernels.hip).
struct CompactArgs {
  int nlevels;
  uint32_t nb[FX_MAX_LEVELS_DEV];          // bricks of each level's storage
  uint32_t* flags[FX_MAX_LEVELS_DEV];      // level 0: activation output; l > 0: written
  BrickList L[FX_MAX_LEVELS_DEV];
  uint32_t block_start[FX_MAX_LEVELS_DEV + 1];  // first block of each level (1024 bricks per block)
  uint32_t* touched;
  uint32_t* counts;   // 2 per block: active, retiring
  uint32_t* offsets;  // 2 per block
};
void launch_compact(hipStream_t s, const CompactArgs& A);
void launch_level_masks(hipStream_t s, int blocks, const LevelMaskArgs& a);
// Fused V-cycle legs (LDS temporal blocking); return false for unsupported sweep counts.
// vdown: p_out == nullptr skips the write of the (unchanged, pre == 0) pressure.
bool launch_vdown(hipStream_t s, int blocks, int pre, const uint32_t* list, const uint32_t* count, Grid f, LvBits bf, Grid c,
                  LvMask mc, const StepConst* cst, const float* rhs, const float* p_in, float* p_out, float* crhs, double* norm2);
bool launch_vup(hipStream_t s, int blocks, int post, const uint32_t* list, const uint32_t* count, Grid c, Grid f, LvBits bf,
                const StepConst* cst, const float* e, const float* rhs, const float* p_in, float* p_out);
// Residual of a level into r (0 on inactive voxels), FP64 norm when norm2 != nullptr.
void launch_residual_lean(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvBits b,
                          const float* rhs, const float* p, float* r, double* norm2);
// Full-weighting restriction of r into the coarse right-hand side over the coarse level's bricks.
void launch_restrict_lean(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid f, Grid c, LvBits cb,
                          const float* r, float* crhs);
// p += prolongation of the coarse correction e, in place, on the level's active voxels.
void launch_prolong_add(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid c, Grid f, LvBits b,
                        const float* e, float* p);
// One weighted-Jacobi sweep of a V-cycle level; src == nullptr starts from p = 0.
// ev_start/ev_stop (fine level with src only): events that take the kernel's own
// start and end timestamps.
void launch_sweep(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvBits b, const StepConst* c,
                  const float* rhs, const float* src, float* dst, bool fine, hipEvent_t ev_start = nullptr,
                  hipEvent_t ev_stop = nullptr);
// Two sweeps in one launch (LDS temporal blocking), bit-identical to two
// launch_sweep calls; src == nullptr starts from p = 0.
void launch_sweep_pair(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvBits b, const StepConst* c,
                       const float* rhs, const float* src, float* dst);
// Four sweeps in one launch (LDS tiles of radius 4), bit-identical to four
// launch_sweep calls.
void launch_sweep_quad(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvBits b, const StepConst* c,
                       const float* rhs, const float* src, float* dst);
// ns (1 or 2) sweeps in one launch on compute tiles of 4^3 (a brick's octants,
// one wave each) or 16^3 (2x2x2 brick groups) with halo ns, bit-identical to ns
// launch_sweep calls.
void launch_sweep_tile(hipStream_t s, int blocks, int tile, int ns, const uint32_t* list, const uint32_t* count, Grid g, LvBits b,
                       const StepConst* c, const float* rhs, const float* src, float* dst);
// Bits 8.. of a V-cycle plan mask: the split leg of level l runs its sweeps two
// per launch (launch_sweep_pair).
constexpr int kPlanPairShift = 8;
// Bits 20 + 2l .. 21 + 2l of the up mask: the compute tile of level l's split-leg
// sweeps (0: the 8^3 brick, 1: 4^3 octants, 2: 16^3 brick groups; launch_sweep_tile);
// the pair bit gives the halo (sweeps per launch).
constexpr int kPlanTileShift = 20;
// Bits 24-25 of the down mask: sweeps per launch of the coarsest level when its
// active voxels exceed the one-workgroup smoother (0: one, 1: two, 2: four).
constexpr int kPlanCoarseShift = 24;
// Bits 16.. of a V-cycle plan mask: the split leg of level l fuses its sweep next to
// the level change with the restriction (down: k_vdown<1> after pre - 1 sweeps) or
// with the prolongation (up: k_vup<1> before post - 1 sweeps).
constexpr int kPlanFuseShift = 16;
// Coarsest-level slot table (kCoarseTableBytes at tab, count at tab_n), built
// once per solve; zeroes the inactive voxels of the level's listed bricks in p0/p1.
constexpr uint32_t kCoarseCap = 8192;  // active voxels the one-workgroup LDS smoother holds
constexpr size_t kCoarseTableBytes = size_t(kCoarseCap) * 16;
// scenes > 1 (g a selector): one table per scene, kCoarseTableBytes apart, counts tab_n[scene].
void launch_coarse_setup(hipStream_t s, const uint32_t* list, const uint32_t* count, Grid g, LvBits b, void* tab, uint32_t* tab_n,
                         float* p0, float* p1, uint32_t scenes = 1);
// tab == nullptr (or an overflowed table): global-memory sweeps in one workgroup.
void launch_smooth_block(hipStream_t s, const uint32_t* list, const uint32_t* count, Grid g, LvMask m, const StepConst* c,
                         const float* rhs, float* p0, float* p1, int sweeps, int weighted, int zero_start, const void* tab,
                         const uint32_t* tab_n, uint32_t scenes = 1);
void launch_advect_velocity(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                            const StepConst* c, Vec3 u, Vec3 us);
void launch_advect_scalars(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                           const StepConst* c, int mode, Scalars set0, Vec3 us, Scalars dst, const uint32_t* flags);
void launch_combustion(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                       const StepConst* c, Scalars src, Scalars dst, const uint32_t* flags);
void launch_buoyancy(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                     const StepConst* c, Scalars set0, Vec3 us, const uint32_t* flags);
void launch_vorticity(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                      Vec3 us, Vec3 w, float* wmag);
void launch_confinement(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                        const StepConst* c, Vec3 w, const float* wmag, Vec3 us, float* cmag, Scalars set0);
void launch_divergence(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                       Vec3 us, float* div);
// Listed fine bricks of each scene of a brick-queue batch into out[scene] (tab: the
// level's grid table).
void launch_scene_counts(hipStream_t s, const uint32_t* list, const uint32_t* count, const Grid* tab, uint32_t scenes,
                         uint32_t* out);
void launch_zero_bricks(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count,
                        float* const* bufs, int nbufs);
// Retiring bricks of every level, zeroed in one launch: level l's buffers are
// b[first[l] .. first[l+1]) and its brick list is list[l] / count[l].
struct RetireArgs {
  int nlevels;
  const uint32_t* list[FX_MAX_LEVELS_DEV];
  const uint32_t* count[FX_MAX_LEVELS_DEV];
  int first[FX_MAX_LEVELS_DEV + 1];
  void* b[FX_MAX_ZERO_BUFS + 3 * FX_MAX_LEVELS_DEV];
  uint8_t esz[FX_MAX_ZERO_BUFS + 3 * FX_MAX_LEVELS_DEV];  // element bytes (4, 2 or 16)
};
void launch_retire(hipStream_t s, int blocks, const RetireArgs& R);
void launch_jacobi(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvMask m,
                   const StepConst* c, const float* rhs, const float* src, float* dst, int weighted);
void launch_residual(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvMask m,
                     const float* rhs, const float* p, float* r, double* norm2);
void launch_restrict(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid fine,
                     Grid coarse, LvMask mc, const float* src, float* dst, int full_weight);
void launch_prolong(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid coarse,
                    Grid fine, LvMask mf, const float* src, float* dst);
void launch_projection(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g,
                       Vec3 us, const float* p, Vec3 u);
void launch_debug(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 u,
                  float* umag);

struct RayConst {
  float cam_pos[3], right[3], up[3], fwd[3];
  float aspect, tan_half;
  float box_min[3], box_max[3], half_ext[3];
  float mode;
  uint32_t step_count;
  float light_count;
  float scat_f, scat_b;
  float light_dir[3][3];
  float light_kind[3];
  float light_pos[3][3];
  float light_intensity[3];
  float light_color[3][3];
  uint32_t width, height;
};
// Render snapshot (16 bytes per voxel, storage order of the fields) of the listed
// bricks, zero on the retire list (rlist may be nullptr), and the brick flags
// packed 32 per word into bits.
void launch_render_pack(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Scalars set0, const uint32_t* rlist,
                        const uint32_t* rcount, void* rb, const uint32_t* flags, uint32_t nb, uint32_t* bits);
// Raymarch of the snapshot rb with packed flags bits; queue: one device word
// (tile counter); blocks: resident workgroups taking tiles.
void launch_raymarch(hipStream_t s, Grid g, const RayConst& rc, const void* rb, const uint32_t* bits, uint32_t* queue, int blocks,
                     uint16_t* out);

// ---- differentiable smoke step (fenix_diff.cpp) ----
void launch_df_source(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const float* ctrl, uint32_t cn,
                      uint32_t slab, float dt, const float* d, Vec3 u, float* ds, Vec3 us);
// u_a = semi-Lagrangian(u_s), d' = MacCormack(d_s, u_a)
void launch_df_advect(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 us, Vec3 ua,
                      const float* ds, float* dn, float dt);
// u_b = u_a + buoyancy(d'); w, |w| = curl u_b; u_c = u_b + confinement
void launch_df_forces(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const float* dn, Vec3 ua, Vec3 ub,
                      Vec3 w, float* wmag, Vec3 uc, float beta_dt, float eps_dt);
// out[kNormLanes] += 0.5 sum (d - target)^2 (FP64); db = d - target when non-null
void launch_df_loss(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const float* d, const float* target,
                    float* db, double* out);
void launch_da_zero_boundary(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 a);
// abar: adjoint of u_c in, of u_a out; dnb += buoyancy adjoint. w, wmag, gbar, wbar: scratch.
void launch_da_forces(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 ub, Vec3 w, float* wmag,
                      Vec3 abar, Vec3 gbar, Vec3 wbar, float* dnb, float beta_dt, float eps_dt);
// dsb, usb (cleared by the caller) += adjoints of d_s, u_s; abar gains the MacCormack velocity terms first.
// The dense identity list over nb bricks; stage: da_stage_floats(nb) floats of scratch.
void launch_da_advect(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, size_t nb, const float* ds,
                      Vec3 us, Vec3 ua, const float* dnb, float* dsb, Vec3 abar, Vec3 usb, float dt, float* stage);
size_t da_stage_floats(size_t nb);
void launch_da_source(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, uint32_t cn, uint32_t slab,
                      float dt, const float* dsb, Vec3 usb, float* cbar);

// Simulation internals used by the differentiable step: a dense simulation's
// stream, grid and brick list, and its V-cycle projection u = us - grad P(div us).
struct SimView {
  hipStream_t stream;
  Grid g;
  const uint32_t* list;
  const uint32_t* count;
  int blocks;
  size_t nb, storage;  // bricks, floats per brick-major field
};
SimView sim_view(fx_sim* s);
// Persistent workgroups per CU of the brick kernels (a simulation that never
// renders fills the CUs with 4); FENIX_SIM_BLOCKS_PER_CU takes precedence.
void sim_set_blocks_per_cu(fx_sim* s, int per_cu);
int sim_prepare_projection(fx_sim* s);  // uploads the step constants the V-cycle reads
int sim_project(fx_sim* s, Vec3 us, Vec3 u);
void bricks_from_rowmajor(const Grid& g, const float* rm, float* bm);
void bricks_to_rowmajor(const Grid& g, const float* bm, float* rm);

// Bloom and composite after the raymarch (createBloomPass.ts): src is the w x h
// rgba16f image, bright/blur_h/blur_v hold hw x hh rgba16f each, out is w x h
// RGBA8 (R in the low byte).
struct BloomArgs {
  int enabled;
  float threshold, strength, radius;
  uint32_t w, h, hw, hh;
};
void launch_bloom(hipStream_t s, const BloomArgs& a, const void* src, void* bright, void* blur_h, void* blur_v, void* out);

// Simulation internals used by the entry points of fenix_static.cpp.
int sim_fail(fx_sim* s, int code, const char* msg);  // records the message for fx_last_error; returns code
int sim_device(fx_sim* s);
hipStream_t sim_render_stream(fx_sim* s);            // where the raymarch and bloom run
// Address of the static scratch array on `device` (current), resolved once per device (fenix_static.cpp).
int static_scratch_setup(int device, void** out);

}  // namespace fx
