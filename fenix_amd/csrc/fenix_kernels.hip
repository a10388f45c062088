// fenix_kernels.hip — gfx950 kernels of the per-step combustion-volume pipeline.
//
// Every kernel walks a device-resident brick list with a persistent grid-stride
// loop (one 512-thread workgroup = one 8^3 brick = eight wave64 z-slices of 8x8
// voxels; each wave reads/writes 256 contiguous bytes per field). The same
// kernels serve the dense domain (list = every brick) and the sparse domain
// (list = active bricks; every buffer holds zeros outside the active bricks, so
// stencils and trilinear gathers read memory unmasked).
//
// Arithmetic follows the reference WGSL term by term (paths relative to the
// reference root, SH = src/engine/simulation/shaders/passes); the library is
// built with -ffp-contract=off so no FMA contraction changes the rounding.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>

#include "fenix_internal.h"

namespace fx {

// --------------------------------------------------------------- helpers --
struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ F3 mul3(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(F3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ F3 norm3(F3 a) { float l = len3(a); return F3{a.x / l, a.y / l, a.z / l}; }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
  return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float clamp01(float v) { return clampf(v, 0.f, 1.f); }
// WGSL mix(e1,e2,e3) = e1*(1-e3) + e2*e3
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.f - t) + b * t; }
__device__ __forceinline__ F3 mix3(F3 a, F3 b, float t) { return F3{mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; }
__device__ __forceinline__ float smoothf(float e0, float e1, float x) {
  float t = clampf((x - e0) / (e1 - e0), 0.f, 1.f);
  return t * t * (3.f - 2.f * t);
}
// n / d without the IEEE division sequence (~10 instructions): q0 = n * r with
// r = RN(1 / d), then one Markstein correction with the exact remainder
// fma(-q0, d, n), 3 instructions. Only used where the result is the correctly
// rounded quotient for every input that reaches it: smoothk's numerators in
// [0, d] for each literal edge pair, and (y + 0.5) / H for H <= 32768, both
// checked exhaustively against the IEEE quotient on the CPU.
__device__ __forceinline__ float div_mk(float n, float d, float r) {
  const float q0 = n * r;
  return fmaf(fmaf(-q0, d, n), r, q0);
}
// smoothf for literal edges: clamping x - e0 to [0, d] before dividing gives the
// same t as clamping the quotient to [0, 1] (NaN clamps to 0 either way).
__device__ __forceinline__ float smoothk(float e0, float e1, float x) {
  const float d = e1 - e0;
  const float t = div_mk(clampf(x - e0, 0.f, d), d, 1.f / d);
  return t * t * (3.f - 2.f * t);
}
// (y + 0.5) / H, the normalised height of a voxel.
__device__ __forceinline__ float height01(uint32_t y, uint32_t H) {
  const float n = float(y) + 0.5f, d = float(H);
  if (__builtin_expect(H <= 32768u, 1)) return div_mk(n, d, 1.f / d);
  return n / d;
}
__device__ __forceinline__ float stepf(float edge, float x) { return edge <= x ? 1.f : 0.f; }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
__device__ __forceinline__ uint32_t decu(uint32_t v) { return v == 0u ? 0u : v - 1u; }
__device__ __forceinline__ uint32_t incu(uint32_t v, uint32_t m) { return min(v + 1u, m); }

// The grid and owner tables of a brick-queue batch, read through the constant
// address space: the kernels never write them, and only loads the compiler knows
// to be unclobbered become scalar loads of a wave-uniform address. Through a
// generic pointer the table reads of a wave-uniform brick were vector loads, and
// every grid field after them lived in vector registers.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* const_space(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}
__device__ __forceinline__ Grid table_grid(const Grid* tab, uint32_t k) {
  const __attribute__((address_space(4))) Grid* t = const_space(tab) + k;
  Grid g;
  g.W = t->W; g.H = t->H; g.D = t->D;
  g.BX = t->BX; g.BY = t->BY; g.BZ = t->BZ;
  g.base = t->base; g.scene = t->scene; g.span = t->span;
  g.tab = t->tab; g.owner = t->owner;
  return g;
}

// Grid of brick b's scene. MS (several scenes share the level's storage): G is a
// selector and the scene's grid comes from its table; otherwise G is the grid of
// the only scene, whose base is 0.
template <bool MS>
__device__ __forceinline__ Grid grid_at(const Grid& G, uint32_t b) {
  if constexpr (MS) {
    return table_grid(G.tab, const_space(G.owner)[b]);
  } else {
    Grid g = G;
    g.base = 0u;
    g.scene = 0u;
    return g;
  }
}
// Grid of scene s on the level of selector G.
template <bool MS>
__device__ __forceinline__ Grid grid_scene(const Grid& G, uint32_t s) {
  if constexpr (MS) {
    return table_grid(G.tab, s);
  } else {
    Grid g = G;
    g.base = 0u;
    g.scene = 0u;
    return g;
  }
}
// grid_at for a loop over listed bricks. A brick-queue batch lays out the scenes'
// bricks one scene after another and every brick list is ascending, so the
// bricks one wave visits change scene rarely. The cache keeps the wave's last
// scene index and that scene's brick range (three scalar registers): the grid
// table entry of that scene is loaded together with the brick id (the table is
// small and stays in the scalar cache), and the per-brick owner table is read only
// for a brick outside the range. Without the cache every brick waited on the owner
// load (an L2 access) and then on the table load before its own loads could
// start. Single scenes fold the selector as grid_at does. `b` must be
// wave-uniform.
template <bool MS>
struct GridCache {
  uint32_t o = 0u, lo = 0u, n = 0u;
  __device__ __forceinline__ Grid at(const Grid& G, uint32_t b) {
    if constexpr (MS) {
      Grid g = table_grid(G.tab, o);
      if (b - lo >= n) {
        o = const_space(G.owner)[b];
        g = table_grid(G.tab, o);
        lo = g.base;
        n = g.BX * g.BY * g.BZ;
      }
      return g;
    } else {
      return grid_at<false>(G, b);
    }
  }
  // Grid of the last brick's scene on another level (selector S): its table entry
  // is addressed by the cached scene index, not by a loaded grid's.
  __device__ __forceinline__ Grid scene(const Grid& S) const {
    if constexpr (MS) {
      return table_grid(S.tab, o);
    } else {
      return grid_scene<false>(S, 0u);
    }
  }
};
// A value that is the same in every lane of the wave (brick ids of kernels that
// give each wave one brick), so that the table loads it feeds are scalar.
__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t vidx(const Grid& g, uint32_t x, uint32_t y, uint32_t z) {
  return ((g.base + ((z >> 3) * g.BY + (y >> 3)) * g.BX + (x >> 3)) << 9) | ((z & 7u) << 6) | ((y & 7u) << 3) | (x & 7u);
}
__device__ __forceinline__ uint32_t cidx(const Grid& g, uint32_t x, uint32_t y, uint32_t z) {
  return vidx(g, min(x, g.W - 1), min(y, g.H - 1), min(z, g.D - 1));
}
// Brick-major indices of the six face neighbours of the in-domain voxel (x, y, z)
// at index i, clamped to the domain like cidx(incu(x, W - 1), y, z) and
// cidx(decu(x), y, z) (a neighbour past a face is the voxel itself): a step of 1,
// 8 or 64 inside the brick, or into the next brick when the local coordinate is 7
// (0 going down). Per-axis steps instead of a full cidx per neighbour.
struct Nbr6 {
  uint32_t xp, xm, yp, ym, zp, zm;
};
__device__ __forceinline__ Nbr6 nbr6(const Grid& g, uint32_t x, uint32_t y, uint32_t z, uint32_t i) {
  const uint32_t sy = g.BX * 512u - 56u, sz = g.BX * g.BY * 512u - 448u;
  Nbr6 n;
  n.xp = x + 1u < g.W ? i + ((x & 7u) < 7u ? 1u : 505u) : i;
  n.xm = x > 0u ? i - ((x & 7u) > 0u ? 1u : 505u) : i;
  n.yp = y + 1u < g.H ? i + ((y & 7u) < 7u ? 8u : sy) : i;
  n.ym = y > 0u ? i - ((y & 7u) > 0u ? 8u : sy) : i;
  n.zp = z + 1u < g.D ? i + ((z & 7u) < 7u ? 64u : sz) : i;
  n.zm = z > 0u ? i - ((z & 7u) > 0u ? 64u : sz) : i;
  return n;
}
__device__ __forceinline__ bool on_boundary(const Grid& g, uint32_t x, uint32_t y, uint32_t z) {
  return x == 0 || y == 0 || z == 0 || x == g.W - 1 || y == g.H - 1 || z == g.D - 1;
}
// isActiveCoord (src/engine/simulation/shaders/common/active-brick-mask.wgsl.ts:13-18)
// clamps the voxel's brick coordinates to the grid; for an in-domain voxel of
// brick b that brick is b itself, so the kernels read flags[b], one wave-uniform
// (scalar) load per brick.

// Voxel owned by this thread inside brick b.
struct Vox {
  uint32_t x, y, z, i;
  bool in;
};
__device__ __forceinline__ Vox vox_of(const Grid& g, uint32_t b) {
  uint32_t t = threadIdx.x;
  const uint32_t lb = b - g.base;
  uint32_t bx = lb % g.BX, by = (lb / g.BX) % g.BY, bz = lb / (g.BX * g.BY);
  Vox v;
  v.x = bx * 8 + (t & 7u);
  v.y = by * 8 + ((t >> 3) & 7u);
  v.z = bz * 8 + (t >> 6);
  v.i = (b << 9) | t;
  v.in = v.x < g.W && v.y < g.H && v.z < g.D;
  return v;
}

// Scene-local brick coordinates of brick b.
__device__ __forceinline__ void brick_coords(const Grid& g, uint32_t b, int& bx, int& by, int& bz) {
  const uint32_t lb = b - g.base;
  bx = int(lb % g.BX); by = int((lb / g.BX) % g.BY); bz = int(lb / (g.BX * g.BY));
}
// Trilinear sampler, src/engine/simulation/shaders/common/sampling-scalar.wgsl.ts:11-34.
struct Tri {
  uint32_t i[8];
  float fx, fy, fz;
};
// Corners of the clamped trilinear sample at p. The upper corner along each axis
// is min(c + 1, n - 1); in brick-major order it sits 1 / 8 / 64 elements on,
// or in the next brick when the local coordinate is 7, so the eight indices are
// one brick-major index plus per-axis steps.
__device__ __forceinline__ Tri tri(const Grid& g, F3 p) {
  float mx = float(g.W - 1), my = float(g.H - 1), mz = float(g.D - 1);
  float cx = clampf(p.x, 0.f, mx), cy = clampf(p.y, 0.f, my), cz = clampf(p.z, 0.f, mz);
  uint32_t bx = uint32_t(floorf(cx)), by = uint32_t(floorf(cy)), bz = uint32_t(floorf(cz));
  const uint32_t i0 = vidx(g, bx, by, bz);
  const uint32_t sx = bx + 1u < g.W ? ((bx & 7u) < 7u ? 1u : 505u) : 0u;
  const uint32_t sy = by + 1u < g.H ? ((by & 7u) < 7u ? 8u : g.BX * 512u - 56u) : 0u;
  const uint32_t sz = bz + 1u < g.D ? ((bz & 7u) < 7u ? 64u : g.BX * g.BY * 512u - 448u) : 0u;
  Tri t;
  t.fx = fractf(cx); t.fy = fractf(cy); t.fz = fractf(cz);
  t.i[0] = i0; t.i[1] = i0 + sx; t.i[2] = i0 + sy; t.i[3] = i0 + sx + sy;
  t.i[4] = i0 + sz; t.i[5] = i0 + sx + sz; t.i[6] = i0 + sy + sz; t.i[7] = i0 + sx + sy + sz;
  return t;
}
__device__ __forceinline__ float tri_eval(const Tri& t, const float* __restrict__ f) {
  float x00 = mixf(f[t.i[0]], f[t.i[1]], t.fx);
  float x10 = mixf(f[t.i[2]], f[t.i[3]], t.fx);
  float x01 = mixf(f[t.i[4]], f[t.i[5]], t.fx);
  float x11 = mixf(f[t.i[6]], f[t.i[7]], t.fx);
  return mixf(mixf(x00, x10, t.fy), mixf(x01, x11, t.fy), t.fz);
}
// Storage of density, temperature and fuel: FP32, BF16 (round to nearest even)
// or FP16. Every load widens to FP32; all arithmetic is FP32.
// ld4/st4: elements i .. i+3 (i a multiple of 4) as one 16- or 8-byte access.
// Stochastic rounding of 16-bit stores (StepConst::sr): a value between two
// representable neighbours is stored as the upper one with probability equal to
// its distance from the lower one in ulps, so the stored value is unbiased and
// per-step increments below half an ulp survive on average. The decision compares
// the dropped low bits with bits of a 32-bit hash of (brick-major voxel index
// within the scene, step, pass, field) (MurmurHash3's fmix32 finalizer), so it is reproducible.
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t x) {
  x ^= x >> 16; x *= 0x85ebca6bu; x ^= x >> 13; x *= 0xc2b2ae35u; x ^= x >> 16;
  return x;
}
// Per-site hash term: site = 4 * pass + field (pass 0 scalar injection, 1 igniter,
// 2 scalar advection, 3 combustion; field 0 density, 1 temperature, 2 fuel).
__device__ __forceinline__ uint32_t sr_site(const StepConst* C, uint32_t site) { return fmix32(C->step_key * 16u + site); }
__device__ __forceinline__ uint32_t sr_hash(uint32_t i, uint32_t site_term) { return fmix32((i * 0x9e3779b1u) ^ site_term); }
__device__ __forceinline__ uint16_t bf16_sr(float v, uint32_t h) {
  const uint32_t u = __float_as_uint(v);
  return uint16_t((u >> 16) + ((h & 0xffffu) < (u & 0xffffu) ? 1u : 0u));
}
__device__ __forceinline__ uint16_t f16_sr(float v, uint32_t h) {
  const uint32_t u = __float_as_uint(v), sgn = (u >> 16) & 0x8000u, a = u & 0x7fffffffu;
  if (a >= 0x477ff000u) return __half_as_ushort(__float2half(v));  // beyond the largest half: nearest
  const uint32_t e = a >> 23;
  if (e >= 113u) {  // normal half: 13 dropped bits
    const uint32_t hb = ((e - 112u) << 10) | ((a >> 13) & 0x3ffu);
    return uint16_t(sgn | (hb + ((h & 0x1fffu) < (a & 0x1fffu) ? 1u : 0u)));
  }
  const uint32_t d = 126u - e;  // subnormal half (unit 2^-24): 14 or more dropped bits
  if (d >= 32u) return uint16_t(sgn);
  const uint32_t m = (a & 0x7fffffu) | 0x800000u;
  return uint16_t(sgn | ((m >> d) + ((h >> (32u - d)) < (m & ((1u << d) - 1u)) ? 1u : 0u)));
}

struct StF32 {
  using T = float;
  static __device__ __forceinline__ T enc_sr(float v, uint32_t, uint32_t, uint32_t) { return v; }
  static __device__ __forceinline__ float dec(T v) { return v; }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, float v, uint32_t, uint32_t, uint32_t) { p[i] = v; }
  static __device__ __forceinline__ float ld(const T* p, uint32_t i) { return p[i]; }
  static __device__ __forceinline__ void st(T* p, uint32_t i, float v) { p[i] = v; }
  static __device__ __forceinline__ float4 ld4(const T* p, uint32_t i) { return *reinterpret_cast<const float4*>(p + i); }
  static __device__ __forceinline__ void st4(T* p, uint32_t i, float4 v) { *reinterpret_cast<float4*>(p + i) = v; }
};
struct StBF16 {
  using T = uint16_t;
  static __device__ __forceinline__ uint16_t enc(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite values)
    return uint16_t(u >> 16);
  }
  static __device__ __forceinline__ float dec(T v) { return __uint_as_float(uint32_t(v) << 16); }
  static __device__ __forceinline__ float ld(const T* p, uint32_t i) { return dec(p[i]); }
  static __device__ __forceinline__ void st(T* p, uint32_t i, float v) { p[i] = enc(v); }
  // sr != 0: stochastic rounding with the hash term of the store site; hi is the
  // voxel's brick-major index within its scene (i less the scene's first element)
  static __device__ __forceinline__ T enc_sr(float v, uint32_t sr, uint32_t site, uint32_t hi) {
    return sr ? bf16_sr(v, sr_hash(hi, site)) : enc(v);
  }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, float v, uint32_t sr, uint32_t site, uint32_t hi) {
    p[i] = enc_sr(v, sr, site, hi);
  }
  static __device__ __forceinline__ float4 ld4(const T* p, uint32_t i) {
    const uint2 w = *reinterpret_cast<const uint2*>(p + i);
    return make_float4(__uint_as_float(w.x << 16), __uint_as_float(w.x & 0xffff0000u), __uint_as_float(w.y << 16),
                       __uint_as_float(w.y & 0xffff0000u));
  }
  static __device__ __forceinline__ void st4(T* p, uint32_t i, float4 v) {
    *reinterpret_cast<uint2*>(p + i) = make_uint2(uint32_t(enc(v.x)) | (uint32_t(enc(v.y)) << 16),
                                                  uint32_t(enc(v.z)) | (uint32_t(enc(v.w)) << 16));
  }
};
struct StF16 {
  using T = __half;
  static __device__ __forceinline__ float dec(T v) { return __half2float(v); }
  static __device__ __forceinline__ float ld(const T* p, uint32_t i) { return dec(p[i]); }
  static __device__ __forceinline__ void st(T* p, uint32_t i, float v) { p[i] = __float2half(v); }
  static __device__ __forceinline__ T enc_sr(float v, uint32_t sr, uint32_t site, uint32_t hi) {
    return sr ? __ushort_as_half(f16_sr(v, sr_hash(hi, site))) : __float2half(v);
  }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, float v, uint32_t sr, uint32_t site, uint32_t hi) {
    p[i] = enc_sr(v, sr, site, hi);
  }
  static __device__ __forceinline__ float4 ld4(const T* p, uint32_t i) {
    const uint2 w = *reinterpret_cast<const uint2*>(p + i);
    return make_float4(__half2float(__ushort_as_half(uint16_t(w.x))), __half2float(__ushort_as_half(uint16_t(w.x >> 16))),
                       __half2float(__ushort_as_half(uint16_t(w.y))), __half2float(__ushort_as_half(uint16_t(w.y >> 16))));
  }
  static __device__ __forceinline__ void st4(T* p, uint32_t i, float4 v) {
    const uint32_t a = __half_as_ushort(__float2half(v.x)), b = __half_as_ushort(__float2half(v.y));
    const uint32_t c = __half_as_ushort(__float2half(v.z)), d = __half_as_ushort(__float2half(v.w));
    *reinterpret_cast<uint2*>(p + i) = make_uint2(a | (b << 16), c | (d << 16));
  }
};
// 16-bit unsigned fixed point, value = u * 2^-16 (FX_PRECISION_U16): the three
// scalars are clamped to [0, 1] by every pass, so a uniform step of 2^-16 keeps
// 16 bits where a float format spends them on exponents. Encoding and decoding
// are exact scalings by 2^16; stores clamp to [0, 1 - 2^-16] and round to
// nearest even, or (sr) up with probability equal to the dropped fraction,
// compared against 24 hash bits.
__device__ __forceinline__ uint16_t u16_rne(float v) {
  return uint16_t(fminf(rintf(clamp01(v) * 65536.f), 65535.f));
}
__device__ __forceinline__ uint16_t u16_sr(float v, uint32_t h) {
  const float x = clamp01(v) * 65536.f, fl = floorf(x);
  const float up = float(h >> 8) < (x - fl) * 16777216.f ? 1.f : 0.f;
  return uint16_t(fminf(fl + up, 65535.f));
}
struct StU16 {
  using T = uint16_t;
  static __device__ __forceinline__ float dec(uint32_t u) { return float(u) * 1.52587890625e-05f; }  // 2^-16
  static __device__ __forceinline__ float ld(const T* p, uint32_t i) { return dec(p[i]); }
  static __device__ __forceinline__ void st(T* p, uint32_t i, float v) { p[i] = u16_rne(v); }
  static __device__ __forceinline__ T enc_sr(float v, uint32_t sr, uint32_t site, uint32_t hi) {
    return sr ? u16_sr(v, sr_hash(hi, site)) : u16_rne(v);
  }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, float v, uint32_t sr, uint32_t site, uint32_t hi) {
    p[i] = enc_sr(v, sr, site, hi);
  }
  static __device__ __forceinline__ float4 ld4(const T* p, uint32_t i) {
    const uint2 w = *reinterpret_cast<const uint2*>(p + i);
    return make_float4(dec(w.x & 0xffffu), dec(w.x >> 16), dec(w.y & 0xffffu), dec(w.y >> 16));
  }
  static __device__ __forceinline__ void st4(T* p, uint32_t i, float4 v) {
    *reinterpret_cast<uint2*>(p + i) = make_uint2(uint32_t(u16_rne(v.x)) | (uint32_t(u16_rne(v.y)) << 16),
                                                  uint32_t(u16_rne(v.z)) | (uint32_t(u16_rne(v.w)) << 16));
  }
};
// Unsigned 16-bit float with 4 exponent and 12 mantissa bits (FX_PRECISION_UF16):
// the three scalars lie in [0, 1], so the sign bit and the binades above 2 of a
// half are spent on mantissa instead. Code c >= 4096 is the f32 with bits
// (c + 0x70000) << 11, i.e. (1 + m/4096) * 2^(E - 15) for E = c >> 12 in 1..15;
// c < 4096 is c * 2^-26 (steps of 2^-26 below 2^-14). Every value of [0, 1] is
// held to within 2^-13 relative (a quarter of a half's step); 1.0 is exact.
// Decoding is one shift and one exact multiply by 2^112: c << 11 is the f32 with
// exponent field E and mantissa m (a subnormal f32 m * 2^-138 when E = 0), which
// needs the kernels' IEEE f32 denormal mode (denormals neither flushed on input
// nor on output, the HIP default).
__device__ __forceinline__ float uf16_dec(uint32_t c) { return __uint_as_float(c << 11) * 0x1p112f; }
__device__ __forceinline__ uint16_t uf16_rne(float v) {
  const float a = clamp01(v);
  uint32_t b = __float_as_uint(a);
  if (b < 0x38800000u) return uint16_t(rintf(a * 67108864.f));  // below 2^-14: multiples of 2^-26
  b += 0x3ffu + ((b >> 11) & 1u);
  return uint16_t((b >> 11) - 0x70000u);
}
// Stochastic: up with probability equal to the dropped fraction, from the low 11
// hash bits (normal codes) or the top 24 (small values).
__device__ __forceinline__ uint16_t uf16_sr(float v, uint32_t h) {
  const float a = clamp01(v);
  uint32_t b = __float_as_uint(a);
  if (b < 0x38800000u) {
    const float x = a * 67108864.f, fl = floorf(x);
    return uint16_t(uint32_t(fl) + (float(h >> 8) < (x - fl) * 16777216.f ? 1u : 0u));
  }
  b += (h & 0x7ffu) < (b & 0x7ffu) ? 0x800u : 0u;
  return uint16_t((b >> 11) - 0x70000u);
}
struct StUF16 {
  using T = uint16_t;
  static __device__ __forceinline__ float dec(T v) { return uf16_dec(v); }
  static __device__ __forceinline__ float ld(const T* p, uint32_t i) { return dec(p[i]); }
  static __device__ __forceinline__ void st(T* p, uint32_t i, float v) { p[i] = uf16_rne(v); }
  static __device__ __forceinline__ T enc_sr(float v, uint32_t sr, uint32_t site, uint32_t hi) {
    return sr ? uf16_sr(v, sr_hash(hi, site)) : uf16_rne(v);
  }
  static __device__ __forceinline__ void st_sr(T* p, uint32_t i, float v, uint32_t sr, uint32_t site, uint32_t hi) {
    p[i] = enc_sr(v, sr, site, hi);
  }
  static __device__ __forceinline__ float4 ld4(const T* p, uint32_t i) {
    const uint2 w = *reinterpret_cast<const uint2*>(p + i);
    return make_float4(uf16_dec(w.x & 0xffffu), uf16_dec(w.x >> 16), uf16_dec(w.y & 0xffffu), uf16_dec(w.y >> 16));
  }
  static __device__ __forceinline__ void st4(T* p, uint32_t i, float4 v) {
    *reinterpret_cast<uint2*>(p + i) = make_uint2(uint32_t(uf16_rne(v.x)) | (uint32_t(uf16_rne(v.y)) << 16),
                                                  uint32_t(uf16_rne(v.z)) | (uint32_t(uf16_rne(v.w)) << 16));
  }
};
template <class ST>
struct SS {
  typename ST::T* d;
  typename ST::T* t;
  typename ST::T* f;
  float* r;
};
template <class ST>
__host__ __device__ inline SS<ST> typed(const Scalars& s) {
  return SS<ST>{static_cast<typename ST::T*>(s.d), static_cast<typename ST::T*>(s.t), static_cast<typename ST::T*>(s.f), s.r};
}
template <class ST>
__device__ __forceinline__ float tri_evalP(const Tri& t, const typename ST::T* __restrict__ f) {
  float x00 = mixf(ST::ld(f, t.i[0]), ST::ld(f, t.i[1]), t.fx);
  float x10 = mixf(ST::ld(f, t.i[2]), ST::ld(f, t.i[3]), t.fx);
  float x01 = mixf(ST::ld(f, t.i[4]), ST::ld(f, t.i[5]), t.fx);
  float x11 = mixf(ST::ld(f, t.i[6]), ST::ld(f, t.i[7]), t.fx);
  return mixf(mixf(x00, x10, t.fy), mixf(x01, x11, t.fy), t.fz);
}

__device__ __forceinline__ F3 sample_u(const Grid& g, const Vec3& u, F3 p) {
  Tri t = tri(g, p);
  return F3{tri_eval(t, u.x), tri_eval(t, u.y), tri_eval(t, u.z)};
}

// XCD-aware brick loop. Workgroups are dispatched to the 8 XCDs round-robin by
// block id, and each XCD has its own L2: a plain grid-stride loop sends list
// neighbours (x-neighbour bricks) to different XCDs, so every stencil halo is
// fetched from memory once per XCD. Here XCD x (blocks b with b % 8 == x) walks
// one contiguous eighth of the ascending brick list, i.e. a z-slab, whose
// neighbours mostly sit in the same L2. Grids that are not a multiple of 8 blocks
// use the plain loop.
struct BrickRange {
  uint32_t k, end, step;
};
__device__ __forceinline__ BrickRange brick_range(uint32_t n) {
  if (gridDim.x < 8u || (gridDim.x & 7u) != 0u) return BrickRange{blockIdx.x, n, gridDim.x};
  const uint32_t x = blockIdx.x & 7u, per = gridDim.x >> 3, seg = (n + 7u) >> 3;
  const uint32_t lo = min(x * seg, n), hi = min(lo + seg, n);
  return BrickRange{lo + (blockIdx.x >> 3), hi, per};
}
// brick_range for kernels that process `lanes` bricks side by side in one
// workgroup (blockDim / lanes threads per brick): slot s of workgroup w walks the
// list as virtual workgroup w * lanes + s of its XCD's segment.
__device__ __forceinline__ BrickRange brick_range_slots(uint32_t n, uint32_t lanes, uint32_t slot) {
  if (gridDim.x < 8u || (gridDim.x & 7u) != 0u) return BrickRange{blockIdx.x * lanes + slot, n, gridDim.x * lanes};
  const uint32_t x = blockIdx.x & 7u, per = (gridDim.x >> 3) * lanes, seg = (n + 7u) >> 3;
  const uint32_t lo = min(x * seg, n), hi = min(lo + seg, n);
  return BrickRange{lo + (blockIdx.x >> 3) * lanes + slot, hi, per};
}
// Float4 brick rows with the domain clamps of the reference's stencils
// (incu/decu + cidx: a neighbour past a domain face is the voxel itself). A
// thread owns the 4 voxels x = x .. x+3 of one brick row (128 threads per
// brick); RowCtx holds the row's index, the indices of its +-y and +-z
// neighbour rows (the row itself where the neighbour is past a face) and of the
// x neighbour across the brick face. Every index lies in the allocation.
struct RowCtx {
  uint32_t i0, iyp, iym, izp, izm, iedge;
  uint32_t x, y, z, h;
  uint32_t nin;  // elements inside the domain (x + e < W, y < H, z < D)
};
__device__ __forceinline__ RowCtx row_ctx(const Grid& g, uint32_t b, uint32_t t) {
  RowCtx R;
  const uint32_t h = t & 1u, ly = (t >> 1) & 7u, lz = t >> 4, x0 = h * 4u;
  const uint32_t BX = g.BX, BXY = g.BX * g.BY, lb = b - g.base;
  R.h = h;
  R.x = (lb % g.BX) * 8 + x0;
  R.y = ((lb / g.BX) % g.BY) * 8 + ly;
  R.z = (lb / BXY) * 8 + lz;
  R.i0 = (b << 9) | (lz << 6) | (ly << 3) | x0;
  R.iyp = R.y + 1u < g.H ? (ly < 7u ? R.i0 + 8u : ((b + BX) << 9) | (lz << 6) | x0) : R.i0;
  R.iym = R.y > 0u ? (ly > 0u ? R.i0 - 8u : ((b - BX) << 9) | (lz << 6) | (7u << 3) | x0) : R.i0;
  R.izp = R.z + 1u < g.D ? (lz < 7u ? R.i0 + 64u : ((b + BXY) << 9) | (ly << 3) | x0) : R.i0;
  R.izm = R.z > 0u ? (lz > 0u ? R.i0 - 64u : ((b - BXY) << 9) | (7u << 6) | (ly << 3) | x0) : R.i0;
  if (h) R.iedge = R.x + 4u < g.W ? ((b + 1u) << 9) | (lz << 6) | (ly << 3) : R.i0 + 3u;
  else R.iedge = R.x > 0u ? ((b - 1u) << 9) | (lz << 6) | (ly << 3) | 7u : R.i0;
  R.nin = (R.y < g.H && R.z < g.D) ? min(g.W - min(R.x, g.W), 4u) : 0u;
  return R;
}
__device__ __forceinline__ void unpack4(float4 v, float* a) { a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w; }
// The row's own values and its clamped x neighbours (one shuffle with the
// partner lane, which holds the other half of the row; call with both lanes
// active).
template <class ST>
__device__ __forceinline__ void row_x(const Grid& g, const RowCtx& R, const typename ST::T* __restrict__ f, float* c, float* xp,
                                      float* xm) {
  unpack4(ST::ld4(f, R.i0), c);
  const float edge = ST::ld(f, R.iedge);
  const float other = __shfl_xor(R.h ? c[0] : c[3], 1, 64);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t x = R.x + uint32_t(e);
    const float next = e < 3 ? c[e + 1] : (R.h ? edge : other);
    const float prev = e > 0 ? c[e - 1] : (R.h ? other : edge);
    xp[e] = x + 1u < g.W ? next : c[e];
    xm[e] = x > 0u ? prev : c[e];
  }
}
template <class ST>
__device__ __forceinline__ void row_y(const RowCtx& R, const typename ST::T* __restrict__ f, float* yp, float* ym) {
  unpack4(ST::ld4(f, R.iyp), yp);
  unpack4(ST::ld4(f, R.iym), ym);
}
template <class ST>
__device__ __forceinline__ void row_z(const RowCtx& R, const typename ST::T* __restrict__ f, float* zp, float* zm) {
  unpack4(ST::ld4(f, R.izp), zp);
  unpack4(ST::ld4(f, R.izm), zm);
}
// Store the row's in-domain elements.
template <class ST>
__device__ __forceinline__ void row_store(const RowCtx& R, typename ST::T* __restrict__ f, const float* v) {
  if (R.nin == 4u) {
    ST::st4(f, R.i0, make_float4(v[0], v[1], v[2], v[3]));
  } else {
    for (uint32_t e = 0; e < R.nin; ++e) ST::st(f, R.i0 + e, v[e]);
  }
}
// Simulation kernels share the CUs with the raymarch, which runs overlapped on
// its own stream. s_setprio raises the wave's instruction-issue priority (0..3,
// default 0): where both kinds of waves are resident the simulation's issue
// first. The raymarch keeps priority 0 and finishes within the step anyway
// (measured on the 256^3 plume: 1.135 -> 1.074 ms per step).
#define FX_SIM_PRIORITY() __builtin_amdgcn_s_setprio(2)
// Row kernels: 4 bricks side by side per 512-thread workgroup, thread fx_t
// (0..127) of its brick owns one half brick row (row_ctx).
#define FX_ROW_LOOP(list, count)                                           \
  FX_SIM_PRIORITY();                                                       \
  const uint32_t fx_t = threadIdx.x & 127u;                                \
  const BrickRange fx_r = brick_range_slots(*(count), 4u, threadIdx.x >> 7); \
  for (uint32_t fx_k = fx_r.k; fx_k < fx_r.end; fx_k += fx_r.step)
#define FX_BRICK_LOOP(list, count)                                       \
  FX_SIM_PRIORITY();                                                     \
  const BrickRange fx_r = brick_range(*(count));                         \
  for (uint32_t fx_k = fx_r.k; fx_k < fx_r.end; fx_k += fx_r.step)

// ------------------------------------------------------------ emitters --
// EmitterSource (SH/inject-source-common.wgsl.ts:3-12), 32 floats per source.
struct Src {
  const float* s;
  __device__ float px() const { return s[0]; } __device__ float py() const { return s[1]; }
  __device__ float pz() const { return s[2]; } __device__ float radius() const { return s[3]; }
  __device__ float start() const { return s[4]; } __device__ float duration() const { return s[5]; }
  __device__ float yd() const { return s[8]; } __device__ float yt() const { return s[9]; } __device__ float yf() const { return s[10]; }
  __device__ float speed() const { return s[12]; } __device__ float mode() const { return s[13]; }
  __device__ float falloff() const { return s[14]; } __device__ float vw() const { return s[15]; }
  __device__ float dx_() const { return s[16]; } __device__ float dy_() const { return s[17]; } __device__ float dz_() const { return s[18]; }
  __device__ float nscale() const { return s[20]; } __device__ float nmix() const { return s[21]; }
  __device__ float intensity() const { return s[22]; }
  __device__ float kind() const { return s[28]; } __device__ float seed() const { return s[29]; }
};

// SH/inject-source-common.wgsl.ts:102-121
// Sine and cosine for the smooth procedural fields (thermal noise, buoyancy
// shear): two-part Cody-Waite reduction by 2*pi, then the hardware
// v_sin_f32 / v_cos_f32 (argument in revolutions). The arguments stay below a
// few thousand radians, where this keeps the error near 1e-6. hash31 below
// scales its sine by 43758 and takes the fraction, so a one-ulp difference in
// the sine moves the hash by ~3e-3: it takes the double-precision sine rounded
// to float (the correctly rounded float sine, barring double-rounding ties of
// probability ~2^-29), which the oracle computes the same way.
__device__ __forceinline__ float rev2pi(float x) {
  const float k = rintf(x * 0.15915494309189535f);
  float r = fmaf(-k, 6.28318548202514648f, x);
  r = fmaf(-k, -1.7484556000744487e-07f, r);
  return r * 0.15915494309189535f;
}
__device__ __forceinline__ float sin_s(float x) { return __builtin_amdgcn_sinf(rev2pi(x)); }
__device__ __forceinline__ float cos_s(float x) { return __builtin_amdgcn_cosf(rev2pi(x)); }
__device__ __forceinline__ float hash31(F3 p) {
  const float s = float(sin(double(dot3(p, f3(127.1f, 311.7f, 74.7f)))));
  return fractf(s * 43758.5453123f);
}
__device__ __noinline__ float value_noise3(F3 p) {
  F3 c = f3(floorf(p.x), floorf(p.y), floorf(p.z));
  F3 f = f3(fractf(p.x), fractf(p.y), fractf(p.z));
  F3 u = f3(f.x * f.x * (3.f - 2.f * f.x), f.y * f.y * (3.f - 2.f * f.y), f.z * f.z * (3.f - 2.f * f.z));
  float h000 = hash31(c), h100 = hash31(c + f3(1, 0, 0)), h010 = hash31(c + f3(0, 1, 0)), h110 = hash31(c + f3(1, 1, 0));
  float h001 = hash31(c + f3(0, 0, 1)), h101 = hash31(c + f3(1, 0, 1)), h011 = hash31(c + f3(0, 1, 1)), h111 = hash31(c + f3(1, 1, 1));
  return mixf(mixf(mixf(h000, h100, u.x), mixf(h010, h110, u.x), u.y), mixf(mixf(h001, h101, u.x), mixf(h011, h111, u.x), u.y), u.z);
}
__device__ __forceinline__ float emitter_noise(F3 pos, float scale, float seed) {
  return value_noise3(pos * fmaxf(scale, 0.1f) + f3(seed * 0.137f, seed * 0.271f, -seed * 0.193f));
}
__device__ __forceinline__ float active_gate(const fx_sim_params& P, Src s) {
  float age = P.time - s.start();
  return stepf(0.f, age) * stepf(age, s.duration());
}
__device__ __forceinline__ float source_duration(Src s) { return fmaxf(s.duration(), 0.0001f); }
__device__ __forceinline__ float sphere_falloff(float dist, float radius, float falloff) {
  float t = clampf(1.f - dist / fmaxf(radius, 0.0001f), 0.f, 1.f);
  return mixf(stepf(0.f, t - 0.0001f), t * t, falloff);
}
__device__ __forceinline__ float source_feather(const fx_sim_params& P, float r) {
  float vw = P.dx / fmaxf(P.world_size, 0.0001f);
  return fmaxf(r * 0.24f, vw * 3.f);
}
__device__ __forceinline__ float outer_limit(const fx_sim_params& P, float r) { return r + source_feather(P, r); }
__device__ __noinline__ float eruptive_falloff(const fx_sim_params& P, F3 pos, float dist, Src s, float falloff) {
  float targetRadius = s.radius();
  float growthSeconds = clampf(targetRadius * 5.5f, 0.1f, 0.75f);
  float age = fmaxf(P.time - s.start(), 0.f);
  float growth = smoothf(0.f, 1.f, clampf(age / growthSeconds, 0.f, 1.f));
  float feather = source_feather(P, targetRadius);
  float seed = s.seed();
  float coarse = emitter_noise(pos + f3(age * 0.17f, -age * 0.11f, age * 0.07f), 6.f, seed);
  float fine = emitter_noise(pos + f3(-age * 0.33f, age * 0.21f, -age * 0.15f), 18.f, seed + 41.f);
  float lobe = clampf(coarse * 0.72f + fine * 0.28f, 0.f, 1.f);
  float reach = targetRadius * growth * mixf(0.62f, 1.f, smoothk(0.2f, 0.92f, lobe));
  float front = 1.f - smoothf(reach - feather, reach + feather, dist);
  float coreRadius = fmaxf(targetRadius * (0.18f + growth * 0.38f), feather);
  float core = sphere_falloff(dist, coreRadius, falloff) * (1.f - growth * 0.35f);
  float body = sphere_falloff(dist, fmaxf(reach + feather, feather), falloff) * front;
  return clampf(fmaxf(core, body) * mixf(0.72f, 1.2f, lobe), 0.f, 1.f);
}
__device__ __noinline__ F3 turbulent_dir(F3 pos, float seed) {  // SH/inject-velocity.wgsl.ts:39-44
  float nx = emitter_noise(pos, 4.f, seed);
  float ny = emitter_noise(pos, 4.f, seed + 17.3f);
  float nz = emitter_noise(pos, 4.f, seed + 31.7f);
  return norm3(f3(nx, ny, nz) - f3(0.5f, 0.5f, 0.5f));
}

// SH/inject-scalar.wgsl.ts:32-70, SH/inject-velocity.wgsl.ts:46-96 and
// SH/inject-igniter.wgsl.ts:32-66 fused per voxel in dispatch order (each pass
// touches only its own voxel, so fusing keeps the sequential result).
template <class ST, bool MS>
__global__ __launch_bounds__(512) void k_inject(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                const StepConst* __restrict__ C0, const float* __restrict__ srcs0,
                                                SS<ST> S, Vec3 u, uint32_t* touched, int parts,
                                                uint32_t* __restrict__ clear, uint32_t clear_n) {
  // the step's clear of the injected-activation flags (ORCH:271), which the next
  // launch (the activation) fills: one launch fewer than a separate fill
  if (clear)
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < clear_n; q += gridDim.x * blockDim.x) clear[q] = 0u;
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid g = gk.at(G, b);
    const StepConst* __restrict__ C = C0 + g.scene;
    const fx_sim_params& P = C->p;
    const uint32_t ns = C->nsrc;
    const uint32_t sr = C->sr;
    const uint32_t s0 = sr_site(C, 0), s1 = sr_site(C, 1), s2 = sr_site(C, 2), s5 = sr_site(C, 5);
    const float* __restrict__ srcs = srcs0 + C->src_off;
    if (touched && threadIdx.x == 0) touched[b] = 1u;
    Vox v = vox_of(g, b);
    if (!v.in) continue;
    const uint32_t i = v.i, hi = i - (g.base << 9);
    F3 pos = f3((float(v.x) + 0.5f) / float(g.W), (float(v.y) + 0.5f) / float(g.H), (float(v.z) + 0.5f) / float(g.D));
    // the stored (decoded) temperature and fuel of the scalar part, which the
    // igniter part reads (the values a reload returns)
    float tt = 0.f, ff = 0.f;
    if (parts & 1) {
      float den = ST::ld(S.d, i), tem = ST::ld(S.t, i), fu = ST::ld(S.f, i);
      for (uint32_t k = 0; k < ns; ++k) {
        Src s{srcs + size_t(k) * FX_FLOATS_PER_EMITTER};
        bool burst = s.kind() == 3.f;
        if (!(s.kind() == 0.f || burst)) continue;
        if (active_gate(P, s) < 0.0001f) continue;
        F3 off = pos - f3(s.px(), s.py(), s.pz());
        float dist = len3(off);
        float radius = s.radius();
        if (dist > (burst ? outer_limit(P, radius) : radius)) continue;
        float weight = burst ? eruptive_falloff(P, pos, dist, s, 0.5f) : sphere_falloff(dist, radius, 0.5f);
        float nm = clampf(s.nmix(), 0.f, 1.f);
        float modulation = nm == 0.f ? 1.f : mixf(1.f, emitter_noise(pos, s.nscale(), s.seed()), nm);
        float release = burst ? 1.f / source_duration(s) : 1.f;
        float contribution = weight * modulation * release * P.delta_time;
        den = clamp01(den + s.yd() * contribution);
        tem = clamp01(tem + s.yt() * contribution);
        fu = clamp01(fu + s.yf() * contribution);
      }
      const typename ST::T et = ST::enc_sr(tem, sr, s1, hi), ef = ST::enc_sr(fu, sr, s2, hi);
      ST::st_sr(S.d, i, den, sr, s0, hi); S.t[i] = et; S.f[i] = ef;
      tt = ST::dec(et); ff = ST::dec(ef);
    }
    if (parts & 2) {
      F3 vel = f3(u.x[i], u.y[i], u.z[i]);
      for (uint32_t k = 0; k < ns; ++k) {
        Src s{srcs + size_t(k) * FX_FLOATS_PER_EMITTER};
        bool burst = s.kind() == 3.f;
        if (!(s.kind() == 1.f || burst)) continue;
        if (active_gate(P, s) < 0.0001f) continue;
        F3 off = pos - f3(s.px(), s.py(), s.pz());
        float dist = len3(off);
        float radius = s.radius();
        if (dist > (burst ? outer_limit(P, radius) : radius)) continue;
        float weight = burst ? eruptive_falloff(P, pos, dist, s, s.falloff()) : sphere_falloff(dist, radius, s.falloff());
        float speed = s.speed();
        float mode = s.mode();
        F3 dir = f3(0, 0, 0);
        F3 up = f3(0.f, 1.f, 0.f);
        if (burst) {
          F3 radial = dist > 0.0001f ? off / fmaxf(dist, 0.0001f) : up;
          F3 turb = turbulent_dir(pos + f3(P.time * 0.19f, -P.time * 0.13f, P.time * 0.07f), s.seed() + 83.f);
          F3 bv = radial * speed + f3(0.f, s.mode(), 0.f) + turb * s.vw();
          F3 impulse = bv * weight * P.delta_time / source_duration(s);
          vel = vel + impulse;
          continue;
        } else if (mode == 0.f) {
          dir = dist > 0.0001f ? off / fmaxf(dist, 0.0001f) : up;
        } else if (mode == 1.f) {
          F3 dd = f3(s.dx_(), s.dy_(), s.dz_());
          dir = dot3(dd, dd) > 0.0001f ? norm3(dd) : up;
        } else {
          dir = turbulent_dir(pos, s.seed());
        }
        F3 force = dir * speed * weight * P.delta_time;
        float tight = clampf(s.vw(), 0.f, 1.f) * weight;
        vel = mix3(vel + force, dir * speed, tight);
      }
      float lim = 80.f / fmaxf(P.dx, 0.001f);
      u.x[i] = clampf(vel.x, -lim, lim); u.y[i] = clampf(vel.y, -lim, lim); u.z[i] = clampf(vel.z, -lim, lim);
    }
    if (parts & 4) {
      float rea = S.r[i], tem = (parts & 1) ? tt : ST::ld(S.t, i);
      const float fuel = (parts & 1) ? ff : ST::ld(S.f, i);
      for (uint32_t k = 0; k < ns; ++k) {
        Src s{srcs + size_t(k) * FX_FLOATS_PER_EMITTER};
        bool burst = s.kind() == 3.f;
        if (!(s.kind() == 2.f || burst)) continue;
        if (active_gate(P, s) < 0.0001f) continue;
        float dist = len3(pos - f3(s.px(), s.py(), s.pz()));
        float radius = s.radius();
        if (dist > (burst ? outer_limit(P, radius) : radius)) continue;
        float weight = burst ? eruptive_falloff(P, pos, dist, s, 0.6f) : sphere_falloff(dist, radius, 0.6f);
        float intensity = s.intensity();
        float avail = burst ? 1.f : smoothk(0.02f, 0.22f, fuel);
        float release = burst ? 1.f / source_duration(s) : 1.6f;
        float seedv = intensity * weight * avail * P.delta_time * release;
        rea = clamp01(rea + seedv);
        tem = clamp01(tem + intensity * weight * P.delta_time * 4.f);
      }
      S.r[i] = rea; ST::st_sr(S.t, i, tem, sr, s5, hi);
    }
  }
}

// ----------------------------------------------------------- clearing --
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t value) {
  FX_SIM_PRIORITY();
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = value;
}
__global__ void k_mark(const uint32_t* __restrict__ list, uint32_t n, uint32_t* touched) {
  FX_SIM_PRIORITY();
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) touched[list[k]] = 1u;
}
__global__ void k_fill_f32(float* p, size_t n) {
  FX_SIM_PRIORITY();
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0.f;
}

// ---------------------------------------------------------- activation --
// REFERENCE: SH/compute-active-bricks.wgsl.ts:24-79 (one thread per brick).
// THRESHOLD: density > a or |u| > b anywhere in the brick, dilation 1 (one
// workgroup per brick). Candidates = listA (+ listB).
// Marks fine brick (x, y, z) of g's scene and the coarse bricks that hold it.
template <bool MS>
__device__ __forceinline__ void mark(const Grid& g, uint32_t* flags, const LevelFlags& lf, uint32_t x, uint32_t y, uint32_t z) {
  flags[g.base + x + g.BX * (y + g.BY * z)] = 1u;
  for (int l = 1; l <= lf.n; ++l) {
    const Grid c = grid_scene<MS>(lf.g[l], g.scene);
    uint32_t cx = min(x >> l, c.BX - 1), cy = min(y >> l, c.BY - 1), cz = min(z >> l, c.BZ - 1);
    lf.f[l][c.base + cx + c.BX * (cy + c.BY * cz)] = 1u;
  }
}
__device__ __forceinline__ void dilate(const Grid& g, uint32_t b, int r, uint32_t* flags, const LevelFlags& lf) {
  int bx = int(b % g.BX), by = int((b / g.BX) % g.BY), bz = int(b / (g.BX * g.BY));
  for (int z = -r; z <= r; ++z)
    for (int y = -r; y <= r; ++y)
      for (int x = -r; x <= r; ++x) {
        int nx = min(max(bx + x, 0), int(g.BX) - 1);
        int ny = min(max(by + y, 0), int(g.BY) - 1);
        int nz = min(max(bz + z, 0), int(g.BZ) - 1);
        mark<false>(g, flags, lf, uint32_t(nx), uint32_t(ny), uint32_t(nz));
      }
}
template <class ST>
__global__ __launch_bounds__(256) void k_activate_reference(const uint32_t* __restrict__ listA, const uint32_t* __restrict__ countA,
                                                            const uint32_t* __restrict__ listB, const uint32_t* __restrict__ countB,
                                                            Grid g, SS<ST> S, uint32_t* flags, LevelFlags lf) {
  FX_SIM_PRIORITY();
  const uint32_t na = *countA, nb = listB ? *countB : 0u;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < na + nb; k += gridDim.x * blockDim.x) {
    const uint32_t b = k < na ? listA[k] : listB[k - na];
    uint32_t bx = b % g.BX, by = (b / g.BX) % g.BY, bz = b / (g.BX * g.BY);
    uint32_t base[3] = {bx * 8, by * 8, bz * 8};
    bool active = false;
    float peak = 0.f;
    for (uint32_t z = 0; z < 8; z += 2)
      for (uint32_t y = 0; y < 8; y += 2)
        for (uint32_t x = 0; x < 8; x += 2) {
          uint32_t cx = base[0] + x, cy = base[1] + y, cz = base[2] + z;
          if (cx < g.W && cy < g.H && cz < g.D) {
            uint32_t i = vidx(g, cx, cy, cz);
            float e = ST::ld(S.d, i) + ST::ld(S.t, i) * 0.7f + S.r[i] * 0.9f;
            peak = fmaxf(peak, e);
            if (e > 0.012f) active = true;
          }
        }
    uint32_t ci = vidx(g, min(base[0] + 4, g.W - 1), min(base[1] + 4, g.H - 1), min(base[2] + 4, g.D - 1));
    float ce = ST::ld(S.d, ci) + ST::ld(S.t, ci) * 0.7f + S.r[ci] * 0.9f;
    peak = fmaxf(peak, ce);
    if (ce > 0.012f) active = true;
    if (!active) continue;
    dilate(g, b, (peak > 0.05f || ce > 0.035f) ? 1 : 0, flags, lf);
  }
}
template <class ST, bool MS>
__global__ __launch_bounds__(512) void k_activate_threshold(const uint32_t* __restrict__ listA, const uint32_t* __restrict__ countA,
                                                            const uint32_t* __restrict__ listB, const uint32_t* __restrict__ countB,
                                                            Grid G, const StepConst* __restrict__ C0, const typename ST::T* __restrict__ den,
                                                            Vec3 vel, uint32_t* flags, LevelFlags lf) {
  FX_SIM_PRIORITY();
  // One wave per brick (8 bricks per block): lane (x, y) of the brick reads its
  // column of 8 voxels, ballots reduce; flags are only ever set to 1, so the
  // 27 neighbour marks run on 27 lanes at once.
  const uint32_t na = *countA, nb = listB ? *countB : 0u;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  GridCache<MS> gk;
  for (uint32_t k = blockIdx.x * 8 + wave; k < na + nb; k += gridDim.x * 8) {
    const uint32_t b = wave_uniform(k < na ? listA[k] : listB[k - na]);
    const Grid g = gk.at(G, b);
    const StepConst* __restrict__ C = C0 + g.scene;
    const float dth = C->act_density, sth = C->act_speed;
    const uint32_t lb = b - g.base;
    const uint32_t bx = lb % g.BX, by = (lb / g.BX) % g.BY, bz = lb / (g.BX * g.BY);
    const uint32_t x = bx * 8 + (lane & 7u), y = by * 8 + (lane >> 3);
    bool dens = false, fast = false;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      const uint32_t i = (b << 9) | (j << 6) | lane;
      if (x < g.W && y < g.H && bz * 8 + j < g.D) {
        dens |= ST::ld(den, i) > dth;
        fast |= len3(f3(vel.x[i], vel.y[i], vel.z[i])) > sth;
      }
    }
    // Density-active bricks dilate by one brick; velocity-only bricks stay
    // active without dilating (velocity lives inside the active set only).
    if (__ballot(dens) != 0ull) {
      if (lane < 27) {
        const int nx = min(max(int(bx) + int(lane % 3) - 1, 0), int(g.BX) - 1);
        const int ny = min(max(int(by) + int((lane / 3) % 3) - 1, 0), int(g.BY) - 1);
        const int nz = min(max(int(bz) + int(lane / 9) - 1, 0), int(g.BZ) - 1);
        mark<MS>(g, flags, lf, uint32_t(nx), uint32_t(ny), uint32_t(nz));
      }
    } else if (__ballot(fast) != 0ull && lane == 0) {
      mark<MS>(g, flags, lf, bx, by, bz);
    }
  }
}

// ---------------------------------------------------------- compaction --
// Three launches for every level at once. Each block owns 1024 consecutive
// bricks of one level (256 threads x 4 bricks, 16-byte loads):
//   count: level flags (level 0 given; level l > 0: OR over the 2^l x 2^l x 2^l
//          fine bricks a level-l brick covers), per-block counts of active and
//          retiring bricks (retiring = was listed or touched, now inactive);
//   scan:  one block turns the per-block counts into offsets and list lengths;
//   write: ascending active list, page table (brick -> position or -1), retire
//          list; level 0 also resets `touched` to the new flags.
__device__ __forceinline__ uint32_t level_flag(const CompactArgs& A, int l, uint32_t b) {
  return A.flags[l][b] != 0u ? 1u : 0u;
}
__device__ __forceinline__ void block_of(const CompactArgs& A, uint32_t blk, int* level, uint32_t* base) {
  int l = 0;
  while (l + 1 < A.nlevels && blk >= A.block_start[l + 1]) ++l;
  *level = l;
  *base = (blk - A.block_start[l]) * 1024u;
}
// Exclusive scan of one value per thread over a 256-thread block; returns the total.
__device__ __forceinline__ uint32_t scan256(uint32_t v, uint32_t* excl, uint32_t* lds) {
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= uint32_t(off)) x += y;
  }
  if (lane == 63) lds[wv] = x;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    uint32_t w = lds[k];
    if (k < wv) before += w;
    total += w;
  }
  *excl = before + x - v;
  __syncthreads();
  return total;
}
__global__ __launch_bounds__(256) void k_compact_count(CompactArgs A) {
  FX_SIM_PRIORITY();
  __shared__ uint32_t lds[8];
  int l;
  uint32_t base;
  block_of(A, blockIdx.x, &l, &base);
  const uint32_t nb = A.nb[l];
  const BrickList L = A.L[l];
  uint32_t on_n = 0, ret_n = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    uint32_t b = base + threadIdx.x * 4 + k;
    if (b >= nb) break;
    uint32_t on = level_flag(A, l, b);
    bool was = L.page[b] >= 0 || (l == 0 && A.touched[b] != 0u);
    on_n += on;
    ret_n += (was && !on) ? 1u : 0u;
  }
  uint32_t e;
  uint32_t ta = scan256(on_n, &e, lds);
  uint32_t tr = scan256(ret_n, &e, lds + 4);
  if (threadIdx.x == 0) { A.counts[2 * blockIdx.x] = ta; A.counts[2 * blockIdx.x + 1] = tr; }
}
// Exclusive offsets of every block's active and retiring counts, per level, and
// the level totals: one 1024-thread workgroup, a workgroup-wide scan per chunk of
// 1024 blocks with a running carry (one chunk up to 1024^2 bricks per level).
__global__ __launch_bounds__(1024) void k_compact_scan(CompactArgs A) {
  FX_SIM_PRIORITY();
  __shared__ uint32_t wa[16], wr[16];
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  for (int l = 0; l < A.nlevels; ++l) {
    uint32_t carry_a = 0, carry_r = 0;
    for (uint32_t c0 = A.block_start[l]; c0 < A.block_start[l + 1]; c0 += 1024u) {
      const uint32_t j = c0 + t;
      const bool in = j < A.block_start[l + 1];
      const uint32_t ca = in ? A.counts[2 * j] : 0u, cr = in ? A.counts[2 * j + 1] : 0u;
      uint32_t xa = ca, xr = cr;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t ya = __shfl_up(xa, off, 64), yr = __shfl_up(xr, off, 64);
        if (lane >= uint32_t(off)) { xa += ya; xr += yr; }
      }
      if (lane == 63u) { wa[wv] = xa; wr[wv] = xr; }
      __syncthreads();
      uint32_t ba = 0, br = 0, ta = 0, tr = 0;
      for (uint32_t k = 0; k < 16; ++k) {
        if (k < wv) { ba += wa[k]; br += wr[k]; }
        ta += wa[k]; tr += wr[k];
      }
      if (in) { A.offsets[2 * j] = carry_a + ba + xa - ca; A.offsets[2 * j + 1] = carry_r + br + xr - cr; }
      carry_a += ta; carry_r += tr;
      __syncthreads();  // wa / wr are rewritten by the next chunk
    }
    if (t == 0) { *A.L[l].count = carry_a; *A.L[l].retire_count = carry_r; }
  }
}
__global__ __launch_bounds__(256) void k_compact_write(CompactArgs A) {
  FX_SIM_PRIORITY();
  __shared__ uint32_t lds[8];
  int l;
  uint32_t base;
  block_of(A, blockIdx.x, &l, &base);
  const uint32_t nb = A.nb[l];
  const BrickList L = A.L[l];
  uint32_t on[4], ret[4];
  uint32_t on_n = 0, ret_n = 0;
  for (uint32_t k = 0; k < 4; ++k) {
    uint32_t b = base + threadIdx.x * 4 + k;
    on[k] = ret[k] = 0;
    if (b < nb) {
      on[k] = level_flag(A, l, b);
      bool was = L.page[b] >= 0 || (l == 0 && A.touched[b] != 0u);
      ret[k] = (was && !on[k]) ? 1u : 0u;
    }
    on_n += on[k];
    ret_n += ret[k];
  }
  uint32_t ea, er;
  scan256(on_n, &ea, lds);
  scan256(ret_n, &er, lds + 4);
  ea += A.offsets[2 * blockIdx.x];
  er += A.offsets[2 * blockIdx.x + 1];
  for (uint32_t k = 0; k < 4; ++k) {
    uint32_t b = base + threadIdx.x * 4 + k;
    if (b >= nb) break;
    if (on[k]) { L.list[ea] = b; L.page[b] = int32_t(ea); ++ea; } else { L.page[b] = -1; }
    if (ret[k]) L.retire[er++] = b;
    if (l == 0) A.touched[b] = on[k];
    else A.flags[l][b] = 0u;  // coarse flags start cleared for the next activation
  }
}

// Zero the retiring bricks of every level in one launch: level l's buffers are
// Z.b[Z.first[l] .. Z.first[l+1]) and its list is retire[l] / count[l].
__global__ __launch_bounds__(512) void k_retire(RetireArgs R) {
  FX_SIM_PRIORITY();
  for (int l = 0; l < R.nlevels; ++l) {
    const uint32_t n = *R.count[l];
    for (uint32_t k = blockIdx.x; k < n; k += gridDim.x) {
      const uint32_t i = (R.list[l][k] << 9) | threadIdx.x;
      for (int q = R.first[l]; q < R.first[l + 1]; ++q) {
        if (R.esz[q] == 4) static_cast<float*>(R.b[q])[i] = 0.f;
        else if (R.esz[q] == 16) static_cast<float4*>(R.b[q])[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        else static_cast<uint16_t*>(R.b[q])[i] = 0;
      }
    }
  }
}

// Zero `nbufs` buffers in every brick of the list.
struct ZeroArgs { float* b[FX_MAX_ZERO_BUFS]; int n; };
__global__ __launch_bounds__(512) void k_zero_bricks(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, ZeroArgs Z) {
  FX_BRICK_LOOP(list, count) {
    const uint32_t i = (list[fx_k] << 9) | threadIdx.x;
    for (int k = 0; k < Z.n; ++k) Z.b[k][i] = 0.f;
  }
}

// ----------------------------------------------------------- advection --
// SH/advect-velocity-semilagrangian.wgsl.ts. The velocity of the brick +- 1 voxel
// is staged in LDS (one 10^3 tile per component, loaded a brick ahead), and the
// eight trilinear corners of each backtrace are read there: a 10^3 tile is 2
// bricks of coalesced row loads, where 24 scattered gathers per voxel pulled far
// more cache lines from L2. A backtrace whose corners leave the tile (|u dt| of a
// voxel or more along an axis; the plume's |u dt| stays below 0.2) gathers from
// memory instead. Corner coordinates, fractions and blend order are those of
// tri / tri_eval, so results are unchanged. A 12^3 tile (halo 2) measured 41.9 us
// on the 256^3 plume, the 10^3 tile 27.5 us.
constexpr int kAdvH = 1, kAdvS = 8 + 2 * kAdvH, kAdvN = kAdvS * kAdvS * kAdvS, kAdvIT = (kAdvN + 511) / 512;
// At most 64 VGPRs, so that three workgroups fit on a CU beside a raymarch workgroup.
template <bool MS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void k_advect_velocity(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                         const StepConst* __restrict__ C0, Vec3 u, Vec3 us) {
  FX_SIM_PRIORITY();
  __shared__ float tux[kAdvN], tuy[kAdvN], tuz[kAdvN];
  GridCache<MS> gk;
  const BrickRange br = brick_range(*count);
  const int tid = int(threadIdx.x);
  float rx[kAdvIT], ry[kAdvIT], rz[kAdvIT];
  uint32_t bn = 0;
  Grid gn = G;
  // the tile loads of brick list[k] into registers (0 outside the domain)
  auto fetch = [&](uint32_t k) {
    bn = list[k];
    gn = gk.at(G, bn);
    int bx, by, bz;
    brick_coords(gn, bn, bx, by, bz);
    const int ox = bx * 8 - kAdvH, oy = by * 8 - kAdvH, oz = bz * 8 - kAdvH;
#pragma unroll
    for (int j = 0; j < kAdvIT; ++j) {
      const int q = tid + j * 512;
      const int x = ox + q % kAdvS, y = oy + (q / kAdvS) % kAdvS, z = oz + q / (kAdvS * kAdvS);
      const bool in = q < kAdvN && x >= 0 && y >= 0 && z >= 0 && x < int(gn.W) && y < int(gn.H) && z < int(gn.D);
      const uint32_t id = in ? vidx(gn, uint32_t(x), uint32_t(y), uint32_t(z)) : 0u;
      rx[j] = in ? u.x[id] : 0.f;
      ry[j] = in ? u.y[id] : 0.f;
      rz[j] = in ? u.z[id] : 0.f;
    }
  };
  if (br.k < br.end) fetch(br.k);
  for (uint32_t k = br.k; k < br.end; k += br.step) {
    const uint32_t b = bn;
    const Grid g = gn;
    __syncthreads();  // the previous brick's tiles are no longer read
#pragma unroll
    for (int j = 0; j < kAdvIT; ++j) {
      const int q = tid + j * 512;
      if (q < kAdvN) { tux[q] = rx[j]; tuy[q] = ry[j]; tuz[q] = rz[j]; }
    }
    __syncthreads();
    if (k + br.step < br.end) fetch(k + br.step);
    const float dt = C0[g.scene].p.delta_time;
    const Vox v = vox_of(g, b);
    if (!v.in) continue;
    if (on_boundary(g, v.x, v.y, v.z)) {
      us.x[v.i] = 0.f; us.y[v.i] = 0.f; us.z[v.i] = 0.f;
      continue;
    }
    int bx, by, bz;
    brick_coords(g, b, bx, by, bz);
    const int ox = bx * 8 - kAdvH, oy = by * 8 - kAdvH, oz = bz * 8 - kAdvH;
    const int self = (int(v.x) - ox) + kAdvS * ((int(v.y) - oy) + kAdvS * (int(v.z) - oz));
    F3 coord = f3(float(v.x) + 0.5f, float(v.y) + 0.5f, float(v.z) + 0.5f);
    F3 half = f3(0.5f, 0.5f, 0.5f);
    // sampleVelocity(coord - 0.5) lands on the voxel with all fractions 0; the
    // trilinear blend a*(1-0) + b*0 returns a exactly for finite fields, so the
    // voxel's own value is read instead of 8 corners per component.
    F3 back = coord - f3(tux[self], tuy[self], tuz[self]) * dt;
    const F3 p = back - half;
    // tri's corners: the clamped sample's floor and the next voxel (itself at the last one)
    const float cx = clampf(p.x, 0.f, float(g.W - 1)), cy = clampf(p.y, 0.f, float(g.H - 1)), cz = clampf(p.z, 0.f, float(g.D - 1));
    const int x0 = int(floorf(cx)), y0 = int(floorf(cy)), z0 = int(floorf(cz));
    const int x1 = x0 + 1 < int(g.W) ? x0 + 1 : x0, y1 = y0 + 1 < int(g.H) ? y0 + 1 : y0, z1 = z0 + 1 < int(g.D) ? z0 + 1 : z0;
    const int ax = x0 - ox, ay = y0 - oy, az = z0 - oz, bxx = x1 - ox, byy = y1 - oy, bzz = z1 - oz;
    F3 smp;
    if (ax >= 0 && ay >= 0 && az >= 0 && bxx < kAdvS && byy < kAdvS && bzz < kAdvS) {
      const float fx = fractf(cx), fy = fractf(cy), fz = fractf(cz);
      const int i0 = ax + kAdvS * (ay + kAdvS * az), i1 = bxx + kAdvS * (ay + kAdvS * az);
      const int i2 = ax + kAdvS * (byy + kAdvS * az), i3 = bxx + kAdvS * (byy + kAdvS * az);
      const int i4 = ax + kAdvS * (ay + kAdvS * bzz), i5 = bxx + kAdvS * (ay + kAdvS * bzz);
      const int i6 = ax + kAdvS * (byy + kAdvS * bzz), i7 = bxx + kAdvS * (byy + kAdvS * bzz);
      auto blend = [&](const float* f) {
        const float x00 = mixf(f[i0], f[i1], fx), x10 = mixf(f[i2], f[i3], fx);
        const float x01 = mixf(f[i4], f[i5], fx), x11 = mixf(f[i6], f[i7], fx);
        return mixf(mixf(x00, x10, fy), mixf(x01, x11, fy), fz);
      };
      smp = F3{blend(tux), blend(tuy), blend(tuz)};
    } else {
      smp = sample_u(g, u, p);
    }
    F3 adv = smp * (1.f - dt * 0.025f);
    us.x[v.i] = adv.x; us.y[v.i] = adv.y; us.z[v.i] = adv.z;
  }
}

// SH/advect-scalar-maccormack.wgsl.ts:52-57 (and semilagrangian :69-74)
__device__ __forceinline__ float thermal_noise(F3 pos, float time) {
  F3 p = pos + f3(time * 1.7f, -time * 3.1f, time * 1.1f);
  float a = 0.5f + 0.5f * sin_s(dot3(p, f3(0.097f, 0.061f, -0.083f)));
  float b = 0.5f + 0.5f * sin_s(dot3(p, f3(-0.043f, 0.137f, 0.071f)) + sin_s(p.y * 0.079f) * 1.6f);
  return clampf(a * 0.58f + b * 0.42f, 0.f, 1.f);
}
// SH/advect-scalar-maccormack.wgsl.ts:75-100. Keeps smoothf's IEEE division: with
// smoothk the scalar advection, which sits at 128 VGPRs and waits on its gathers,
// measured 2-4 us slower.
template <class ST>
__device__ __forceinline__ void decay_write(const Grid& g, uint32_t i, uint32_t y, float den, float tem, float fu, float rea,
                                            float tn, float dt, SS<ST> D, uint32_t sr, const uint32_t* site) {
  float ny = (float(y) + 0.5f) / float(g.H);
  float hot = smoothf(0.22f, 0.78f, tem + fu * 0.2f + rea * 0.26f);
  float coolingPocket = smoothf(0.44f, 0.9f, 1.f - tn) * smoothf(0.08f, 0.74f, den) * (1.f - hot * 0.72f);
  float cooling = dt * (0.016f + ny * 0.052f) * mixf(1.28f, 0.58f, hot) * mixf(0.86f, 1.9f, coolingPocket);
  float smokeLoss = dt * (0.008f + ny * 0.026f) * mixf(0.75f, 1.38f, tn);
  float coldBreakup = dt * den * (1.f - smoothf(0.2f, 0.86f, hot)) * smoothf(0.18f, 0.72f, den) * smoothf(0.34f, 0.92f, tn) * 0.34f;
  float clearing = dt * den * hot * mixf(0.1f, 0.72f, tn);
  const uint32_t hi = i - (g.base << 9);
  ST::st_sr(D.d, i, clamp01(fmaxf(0.f, den * (1.f - smokeLoss) - clearing - coldBreakup)), sr, site[0], hi);
  ST::st_sr(D.t, i, clamp01(fmaxf(0.f, tem - cooling)), sr, site[1], hi);
  ST::st_sr(D.f, i, clamp01(fmaxf(0.f, fu - dt * mixf(0.012f, 0.034f, tn))), sr, site[2], hi);
  D.r[i] = clamp01(fmaxf(0.f, rea - dt * mixf(0.46f, 0.82f, tn)));
}
// SH/advect-scalar-maccormack.wgsl.ts:33-101 / SH/advect-scalar-semilagrangian.wgsl.ts:26-44.
// The four advect<Name> calls share every position; sampleVelocity(coord-0.5)
// and the `original` sample hit the voxel exactly (fraction 0). One workgroup per
// brick stages the brick +- 1 voxel in LDS, loaded into registers a brick ahead:
// set 0's scalars as stored (decoded) and the advected velocity, each as one
// float4 per voxel, so that every trilinear corner is one 16-byte LDS read. A
// sample whose corners leave the tile (a backtrace of a voxel or more along an
// axis) gathers from memory instead. Corner coordinates, fractions and blend
// order are those of tri / tri_eval.
constexpr int kScH = 1, kScS = 8 + 2 * kScH, kScN = kScS * kScS * kScS, kScIT = (kScN + 511) / 512;
struct TileCorners {
  int i[8];
  float fx, fy, fz;
  bool in;  // every corner lies in the tile
};
// tri(g, p)'s corners as indices of the tile whose voxel 0 sits at (ox, oy, oz).
__device__ __forceinline__ TileCorners tile_corners(const Grid& g, F3 p, int ox, int oy, int oz) {
  TileCorners c;
  const float cx = clampf(p.x, 0.f, float(g.W - 1)), cy = clampf(p.y, 0.f, float(g.H - 1)), cz = clampf(p.z, 0.f, float(g.D - 1));
  const int x0 = int(floorf(cx)), y0 = int(floorf(cy)), z0 = int(floorf(cz));
  const int x1 = x0 + 1 < int(g.W) ? x0 + 1 : x0, y1 = y0 + 1 < int(g.H) ? y0 + 1 : y0, z1 = z0 + 1 < int(g.D) ? z0 + 1 : z0;
  const int ax = x0 - ox, ay = y0 - oy, az = z0 - oz, bx = x1 - ox, by = y1 - oy, bz = z1 - oz;
  c.in = ax >= 0 && ay >= 0 && az >= 0 && bx < kScS && by < kScS && bz < kScS;
  c.fx = fractf(cx); c.fy = fractf(cy); c.fz = fractf(cz);
  c.i[0] = ax + kScS * (ay + kScS * az); c.i[1] = bx + kScS * (ay + kScS * az);
  c.i[2] = ax + kScS * (by + kScS * az); c.i[3] = bx + kScS * (by + kScS * az);
  c.i[4] = ax + kScS * (ay + kScS * bz); c.i[5] = bx + kScS * (ay + kScS * bz);
  c.i[6] = ax + kScS * (by + kScS * bz); c.i[7] = bx + kScS * (by + kScS * bz);
  return c;
}
// Each component blended in tri_eval's order.
__device__ __forceinline__ float4 blend4(const TileCorners& t, const float4* f) {
  const float4 c0 = f[t.i[0]], c1 = f[t.i[1]], c2 = f[t.i[2]], c3 = f[t.i[3]];
  const float4 c4 = f[t.i[4]], c5 = f[t.i[5]], c6 = f[t.i[6]], c7 = f[t.i[7]];
  auto blend = [&](float a0, float a1, float a2, float a3, float a4, float a5, float a6, float a7) {
    const float x00 = mixf(a0, a1, t.fx), x10 = mixf(a2, a3, t.fx), x01 = mixf(a4, a5, t.fx), x11 = mixf(a6, a7, t.fx);
    return mixf(mixf(x00, x10, t.fy), mixf(x01, x11, t.fy), t.fz);
  };
  return make_float4(blend(c0.x, c1.x, c2.x, c3.x, c4.x, c5.x, c6.x, c7.x), blend(c0.y, c1.y, c2.y, c3.y, c4.y, c5.y, c6.y, c7.y),
                     blend(c0.z, c1.z, c2.z, c3.z, c4.z, c5.z, c6.z, c7.z), blend(c0.w, c1.w, c2.w, c3.w, c4.w, c5.w, c6.w, c7.w));
}
// The scalars of set 0 at p from memory (corners outside the tile).
template <class ST>
__device__ __forceinline__ float4 sample_scalars(const Grid& g, const SS<ST>& S, F3 p) {
  const Tri t = tri(g, p);
  return make_float4(tri_evalP<ST>(t, S.d), tri_evalP<ST>(t, S.t), tri_evalP<ST>(t, S.f), tri_eval(t, S.r));
}
// At most 128 VGPRs: two workgroups per CU.
template <class ST, bool MS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_advect_scalars(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                        const StepConst* __restrict__ C0, int mode, SS<ST> S, Vec3 us, SS<ST> D,
                                                        const uint32_t* __restrict__ flags) {
  FX_SIM_PRIORITY();
  __shared__ float4 ts[kScN], tv[kScN];
  const int tid = int(threadIdx.x);
  GridCache<MS> gk;
  const BrickRange br = brick_range(*count);
  float4 rs[kScIT];
  float rvx[kScIT], rvy[kScIT], rvz[kScIT];
  uint32_t bn = 0;
  Grid gn = G;
  // the tile loads of brick list[k] into registers (0 outside the domain)
  auto fetch = [&](uint32_t k) {
    bn = list[k];
    gn = gk.at(G, bn);
    int bx, by, bz;
    brick_coords(gn, bn, bx, by, bz);
    const int ox = bx * 8 - kScH, oy = by * 8 - kScH, oz = bz * 8 - kScH;
#pragma unroll
    for (int j = 0; j < kScIT; ++j) {
      const int q = tid + j * 512;
      const int x = ox + q % kScS, y = oy + (q / kScS) % kScS, z = oz + q / (kScS * kScS);
      const bool in = q < kScN && x >= 0 && y >= 0 && z >= 0 && x < int(gn.W) && y < int(gn.H) && z < int(gn.D);
      const uint32_t id = in ? vidx(gn, uint32_t(x), uint32_t(y), uint32_t(z)) : 0u;
      rs[j] = in ? make_float4(ST::ld(S.d, id), ST::ld(S.t, id), ST::ld(S.f, id), S.r[id]) : make_float4(0.f, 0.f, 0.f, 0.f);
      rvx[j] = in ? us.x[id] : 0.f;
      rvy[j] = in ? us.y[id] : 0.f;
      rvz[j] = in ? us.z[id] : 0.f;
    }
  };
  if (br.k < br.end) fetch(br.k);
  for (uint32_t k = br.k; k < br.end; k += br.step) {
    const uint32_t b = bn;
    const Grid g = gn;
    __syncthreads();  // the previous brick's tiles are no longer read
#pragma unroll
    for (int j = 0; j < kScIT; ++j) {
      const int q = tid + j * 512;
      if (q < kScN) { ts[q] = rs[j]; tv[q] = make_float4(rvx[j], rvy[j], rvz[j], 0.f); }
    }
    __syncthreads();
    if (k + br.step < br.end) fetch(k + br.step);
    int bx, by, bz;
    brick_coords(g, b, bx, by, bz);
    const int ox = bx * 8 - kScH, oy = by * 8 - kScH, oz = bz * 8 - kScH;
    const StepConst* __restrict__ C = C0 + g.scene;
    const float dt = C->p.delta_time, time = C->p.time;
    const uint32_t sr = C->sr;
    const uint32_t site[3] = {sr_site(C, 8), sr_site(C, 9), sr_site(C, 10)};
    const Vox v = vox_of(g, b);
    if (!v.in) continue;
    const uint32_t i = v.i;
    // isActiveCoord of an in-domain voxel of brick b is brick b's flag
    if (flags[b] == 0u) { ST::st(D.d, i, 0.f); ST::st(D.t, i, 0.f); ST::st(D.f, i, 0.f); D.r[i] = 0.f; continue; }
    const int self = (int(v.x) - ox) + kScS * ((int(v.y) - oy) + kScS * (int(v.z) - oz));
    F3 coord = f3(float(v.x) + 0.5f, float(v.y) + 0.5f, float(v.z) + 0.5f);
    F3 half = f3(0.5f, 0.5f, 0.5f);
    const float4 uv = tv[self];
    F3 velocity = f3(uv.x, uv.y, uv.z);
    F3 back = coord - velocity * dt;
    F3 sp = back - half;
    const TileCorners tp = tile_corners(g, sp, ox, oy, oz);
    if (mode == FX_ADVECT_SEMI_LAGRANGIAN) {
      const float4 s4 = tp.in ? blend4(tp, ts) : sample_scalars<ST>(g, S, sp);
      decay_write<ST>(g, i, v.y, s4.x, s4.y, s4.z, s4.w, thermal_noise(sp, time), dt, D, sr, site);
      continue;
    }
    F3 fwd;
    float4 pr;
    if (tp.in) {
      const float4 f4 = blend4(tp, tv);
      fwd = f3(f4.x, f4.y, f4.z);
      pr = blend4(tp, ts);
    } else {
      fwd = sample_u(g, us, sp);
      pr = sample_scalars<ST>(g, S, sp);
    }
    F3 rec = back + fwd * dt;
    const F3 rp = rec - half;
    const TileCorners tr = tile_corners(g, rp, ox, oy, oz);
    const float4 rc = tr.in ? blend4(tr, ts) : sample_scalars<ST>(g, S, rp);
    const float4 og = ts[self];
    auto mac = [&](float predicted, float recovered, float original) {
      float corrected = predicted + 0.5f * (original - recovered);
      return clampf(corrected, fminf(original, predicted), fmaxf(original, predicted));
    };
    float out[4];
    out[0] = mac(pr.x, rc.x, og.x);
    out[1] = mac(pr.y, rc.y, og.y);
    out[2] = mac(pr.z, rc.z, og.z);
    out[3] = mac(pr.w, rc.w, og.w);
    decay_write<ST>(g, i, v.y, out[0], out[1], out[2], out[3], thermal_noise(coord, time), dt, D, sr, site);
  }
}

// ---------------------------------------------------------- combustion --
// SH/combustion-update.wgsl.ts:27-93 with read-old semantics: advected set ->
// render set; voxels outside active bricks are copied.
__device__ __forceinline__ float thermal_cluster(F3 c, float time) {
  F3 p = c + f3(time * 2.2f, time * 5.2f, -time * 1.7f);
  float body = 0.5f + 0.5f * sin_s(dot3(p, f3(0.071f, 0.093f, -0.064f)) + sin_s(p.y * 0.057f) * 1.8f);
  float pocket = 0.5f + 0.5f * sin_s(dot3(p, f3(0.143f, -0.081f, 0.119f)) + sin_s(p.x * 0.11f + p.z * 0.07f) * 1.2f);
  float rising = 0.5f + 0.5f * sin_s(p.x * 0.18f + p.y * 0.13f - p.z * 0.16f + time * 1.4f);
  return clampf(body * 0.46f + pocket * 0.34f + rising * 0.2f, 0.f, 1.f);
}
template <class ST, bool MS>
__global__ __launch_bounds__(512) void k_combustion(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                    const StepConst* __restrict__ C0, SS<ST> S, SS<ST> D,
                                                    const uint32_t* __restrict__ flags) {
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid g = gk.at(G, b);
    const StepConst* __restrict__ C = C0 + g.scene;
    const float dt = C->p.delta_time, time = C->p.time;
    const float ilo = C->ignition_lo, ihi = C->ignition_hi, rate = C->burn_rate;
    const uint32_t sr = C->sr, sd = sr_site(C, 12), st = sr_site(C, 13), sf = sr_site(C, 14);
    Vox v = vox_of(g, b);
    if (!v.in) continue;
    const uint32_t i = v.i, x = v.x, y = v.y, z = v.z;
    if (flags[b] == 0u) {  // isActiveCoord: brick b's flag for its in-domain voxels
      // copied: the stored values are representable, so storing them is exact
      ST::st(D.d, i, ST::ld(S.d, i)); ST::st(D.t, i, ST::ld(S.t, i)); ST::st(D.f, i, ST::ld(S.f, i)); D.r[i] = S.r[i];
      continue;
    }
    auto hotgas = [&](uint32_t k) {
      return clampf(ST::ld(S.t, k) * 1.55f + ST::ld(S.f, k) * 0.26f + S.r[k] * 0.44f, 0.f, 1.f);
    };
    const Nbr6 nb = nbr6(g, x, y, z, i);
    float ny = height01(y, g.H);
    float cluster = thermal_cluster(f3(float(x), float(y), float(z)), time);
    float nh = (hotgas(nb.xp) + hotgas(nb.xm) + hotgas(nb.yp) + hotgas(nb.ym) + hotgas(nb.zp) + hotgas(nb.zm)) * 0.16666667f;
    float localBefore = hotgas(i);
    float den = ST::ld(S.d, i), tem = ST::ld(S.t, i), fu = ST::ld(S.f, i), rea = S.r[i];
    float radiative = fmaxf(nh - localBefore * 0.55f, 0.f) * smoothk(0.04f, 0.76f, den) * (0.45f + (1.f - cluster) * 0.55f) * dt * 2.6f;
    float hotPocket = smoothk(0.58f, 0.92f, cluster + rea * 0.22f);
    float coolingPocket = smoothk(0.42f, 0.82f, 1.f - cluster) * smoothk(0.08f, 0.68f, den) * (1.f - smoothk(0.54f, 0.94f, rea));
    float ignition = smoothf(ilo, ihi, tem);
    float burn = fminf(fu, fu * ignition * dt * mixf(4.2f, 8.6f, hotPocket) * rate);
    float hotSoot = smoothk(0.24f, 0.88f, tem + rea * 0.2f) * smoothk(0.36f, 0.9f, 1.f - cluster + den * 0.18f);
    float smokeYield = burn * mixf(0.16f, 0.58f, coolingPocket + (1.f - ignition) * 0.5f + hotSoot * 0.45f);
    float hotCooling = mixf(1.3f, 0.62f, hotPocket);
    float heatLoss = dt * (0.018f + ny * 0.052f) * hotCooling * mixf(1.f, 2.25f, coolingPocket);
    float densityLoss = dt * den * (0.02f + ny * 0.07f);
    float clearing = den * smoothk(0.18f, 0.82f, tem + rea * 0.22f + fu * 0.1f) * dt * mixf(0.55f, 2.45f, hotPocket);
    float coolingSmoke = den * tem * (coolingPocket + hotSoot * 0.32f) * dt * 0.16f;
    float reactionDecay = rea * (1.f - dt * mixf(0.68f, 0.22f, hotPocket));
    const uint32_t hi = i - (g.base << 9);
    ST::st_sr(D.f, i, clamp01(fu - burn), sr, sf, hi);
    ST::st_sr(D.t, i, clamp01(tem + burn * mixf(2.1f, 3.25f, hotPocket) + radiative - heatLoss), sr, st, hi);
    ST::st_sr(D.d, i, clamp01(fmaxf(0.f, den + smokeYield + coolingSmoke - densityLoss - clearing * 0.72f)), sr, sd, hi);
    D.r[i] = clamp01(fmaxf(reactionDecay, burn * mixf(4.f, 7.5f, hotPocket) + radiative * 1.8f));
  }
}

// ------------------------------------------------------------ buoyancy --
// SH/apply-buoyancy.wgsl.ts:27-110 (render set, velocityScratch in place).
template <class ST, bool MS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void k_buoyancy(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                  const StepConst* __restrict__ C0, SS<ST> S, Vec3 us, const uint32_t* __restrict__ flags) {
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid g = gk.at(G, b);
    const fx_sim_params& P = C0[g.scene].p;
    Vox v = vox_of(g, b);
    if (!v.in) continue;
    const uint32_t i = v.i, x = v.x, y = v.y, z = v.z;
    if (flags[b] == 0u) continue;  // isActiveCoord: brick b's flag for its in-domain voxels
    auto hot = [&](uint32_t k) { return clampf(ST::ld(S.t, k) * 1.f + S.r[k] * 0.85f, 0.f, 1.f); };
    auto rd = [&](uint32_t k) { return ST::ld(S.d, k); };
    const Nbr6 nb = nbr6(g, x, y, z, i);
    const float dt = P.delta_time;
    const float den = ST::ld(S.d, i);
    float response = P.buoyancy * (1.f + smoothf(6.f, 24.f, fabsf(P.buoyancy)) * 0.38f);
    float hotGas = hot(i);
    float lift = hotGas * response;
    F3 gdir = f3(P.gravity[0], P.gravity[1], P.gravity[2]);
    F3 gravityForce = gdir * den * P.gravity[3];
    float entrain = den * smoothk(0.04f, 0.3f, hotGas);
    float ny = height01(y, g.H);
    float shear = sin_s(float(y) * 0.073f + P.time * 0.9f) * 0.35f + sin_s(float(x) * 0.041f - float(z) * 0.052f + P.time * 0.43f) * 0.22f;
    float windAmount = P.wind[3] * (0.18f + den * 0.55f + ny * 0.45f);
    F3 wdir = f3(P.wind[0], P.wind[1], P.wind[2]);
    F3 wind = wdir * windAmount + f3(-P.wind[2], 0.f, P.wind[0]) * windAmount * shear;
    F3 tg = f3(hot(nb.xp) - hot(nb.xm), hot(nb.yp) - hot(nb.ym), hot(nb.zp) - hot(nb.zm)) * 0.5f;
    F3 dg = f3(rd(nb.xp) - rd(nb.xm), rd(nb.yp) - rd(nb.ym), rd(nb.zp) - rd(nb.zm)) * 0.5f;
    float dgl = fmaxf(len3(dg), 0.0001f);
    float coldSmoke = den * (1.f - smoothk(0.12f, 0.72f, hotGas));
    float surfaceBreakup = smoothk(0.008f, 0.06f, dgl) * coldSmoke;
    F3 shred = norm3(f3(sin_s(float(y) * 0.117f + P.time * 1.9f), sin_s(float(x) * 0.071f - float(z) * 0.083f + P.time * 1.35f) * 0.35f,
                        cos_s(float(x) * 0.061f + float(y) * 0.049f - P.time * 1.6f)));
    float collapse = den * (1.f - smoothk(0.16f, 0.64f, hotGas));
    float smokeDrag = den * smoothk(0.03f, 0.45f, hotGas);
    F3 vel = f3(us.x[i], us.y[i], us.z[i]) * (1.f - dt * 0.08f);
    vel = vel + wind * dt;
    vel = vel + gravityForce * dt;
    vel = vel + tg * collapse * dt * (3.2f + fabsf(P.buoyancy) * 0.08f);
    vel = vel + (-dg / dgl) * surfaceBreakup * dt * 2.8f;
    vel = vel + shred * coldSmoke * dt * (0.72f + surfaceBreakup * 2.4f);
    vel = vel + wdir * coldSmoke * P.wind[3] * dt * 0.9f;
    vel = vel + f3(0.f, smokeDrag * response * 0.55f, 0.f) * dt;
    float lim = 80.f / fmaxf(P.dx, 0.001f);
    vel.y = clampf(vel.y + (lift + entrain * response * 0.24f) * dt, -lim, lim);
    if (x == 0 || x == g.W - 1 || z == 0 || z == g.D - 1) { vel.x = 0.f; vel.z = 0.f; }
    if (y == 0) vel.y = fmaxf(vel.y, 0.f);
    us.x[i] = vel.x; us.y[i] = vel.y; us.z[i] = vel.z;
  }
}

// ------------------------------------------------ vorticity/confinement --
// SH/compute-vorticity.wgsl.ts:23-50, on float4 brick rows: x neighbours of
// u_y and u_z, y neighbours of u_x and u_z, z neighbours of u_x and u_y.
template <bool MS>
__global__ __launch_bounds__(512) void k_vorticity(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G, Vec3 us,
                                                   Vec3 w, float* wmag) {
  GridCache<MS> gk;
  FX_ROW_LOOP(list, count) {
    const uint32_t b = wave_uniform(list[fx_k]);
    const Grid g = gk.at(G, b);
    const RowCtx R = row_ctx(g, b, fx_t);
    float c[4], yxp[4], yxm[4], zxp[4], zxm[4], xyp[4], xym[4], zyp[4], zym[4], xzp[4], xzm[4], yzp[4], yzm[4];
    row_x<StF32>(g, R, us.y, c, yxp, yxm);
    row_x<StF32>(g, R, us.z, c, zxp, zxm);
    row_y<StF32>(R, us.x, xyp, xym);
    row_y<StF32>(R, us.z, zyp, zym);
    row_z<StF32>(R, us.x, xzp, xzm);
    row_z<StF32>(R, us.y, yzp, yzm);
    float ox[4], oy[4], oz[4], om[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (on_boundary(g, R.x + uint32_t(e), R.y, R.z)) { ox[e] = oy[e] = oz[e] = om[e] = 0.f; continue; }
      F3 curl = f3((zyp[e] - zym[e]) - (yzp[e] - yzm[e]), (xzp[e] - xzm[e]) - (zxp[e] - zxm[e]), (yxp[e] - yxm[e]) - (xyp[e] - xym[e])) * 0.5f;
      ox[e] = curl.x; oy[e] = curl.y; oz[e] = curl.z;
      om[e] = len3(curl);
    }
    row_store<StF32>(R, w.x, ox);
    row_store<StF32>(R, w.y, oy);
    row_store<StF32>(R, w.z, oz);
    row_store<StF32>(R, wmag, om);
  }
}
// SH/apply-vorticity-confinement.wgsl.ts:30-66 (render set)
template <class ST, bool MS>
__global__ __launch_bounds__(512) void k_confinement(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                     const StepConst* __restrict__ C0, Vec3 w, const float* __restrict__ wmag, Vec3 us,
                                                     float* cmag, SS<ST> S) {
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid g = gk.at(G, b);
    const fx_sim_params& P = C0[g.scene].p;
    Vox v = vox_of(g, b);
    if (!v.in) continue;
    const uint32_t i = v.i, x = v.x, y = v.y, z = v.z;
    if (on_boundary(g, x, y, z)) { cmag[i] = 0.f; continue; }
    const Nbr6 nb = nbr6(g, x, y, z, i);
    F3 grad = f3(wmag[nb.xp] - wmag[nb.xm], wmag[nb.yp] - wmag[nb.ym], wmag[nb.zp] - wmag[nb.zm]) * 0.5f;
    F3 dir = grad / fmaxf(len3(grad), 0.0001f);
    float hotGas = clampf(ST::ld(S.t, i) + S.r[i] * 0.85f, 0.f, 1.f);
    float mask = clampf(P.vorticity_mask[0] + smoothk(0.02f, 0.38f, wmag[i]) * P.vorticity_mask[1] +
                            smoothk(0.04f, 0.7f, hotGas) * P.vorticity_mask[2] + smoothk(0.04f, 0.65f, ST::ld(S.d, i)) * P.vorticity_mask[3],
                        0.f, 4.f);
    F3 force = cross3(dir, f3(w.x[i], w.y[i], w.z[i])) * P.vorticity_strength * mask * 1.45f;
    F3 vel = f3(us.x[i], us.y[i], us.z[i]) + force * P.delta_time;
    float lim = 80.f / fmaxf(P.dx, 0.001f);
    us.x[i] = clampf(vel.x, -lim, lim); us.y[i] = clampf(vel.y, -lim, lim); us.z[i] = clampf(vel.z, -lim, lim);
    cmag[i] = len3(force);
  }
}

// ------------------------------------------------ divergence / pressure --
// SH/compute-divergence.wgsl.ts:22-42, on float4 brick rows.
template <bool MS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void k_divergence(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G, Vec3 us,
                                                    float* div) {
  GridCache<MS> gk;
  FX_ROW_LOOP(list, count) {
    const uint32_t b = wave_uniform(list[fx_k]);
    const Grid g = gk.at(G, b);
    const RowCtx R = row_ctx(g, b, fx_t);
    float c[4], xp[4], xm[4], yp[4], ym[4], zp[4], zm[4], out[4];
    row_x<StF32>(g, R, us.x, c, xp, xm);
    row_y<StF32>(R, us.y, yp, ym);
    row_z<StF32>(R, us.z, zp, zm);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (on_boundary(g, R.x + uint32_t(e), R.y, R.z)) { out[e] = 0.f; continue; }
      const float dx = xp[e] - xm[e], dy = yp[e] - ym[e], dz = zp[e] - zm[e];
      out[e] = (dx + dy + dz) * 0.5f;
    }
    row_store<StF32>(R, div, out);
  }
}
// fg: the fine grid of the voxel's scene (grid_scene of m.fine).
__device__ __forceinline__ bool lv_active(const LvMask& m, const Grid& fg, uint32_t x, uint32_t y, uint32_t z) {
  if (!m.flags) return true;
  uint32_t fx = min((x << m.shift) >> 3, fg.BX - 1), fy = min((y << m.shift) >> 3, fg.BY - 1),
           fz = min((z << m.shift) >> 3, fg.BZ - 1);
  return m.flags[fg.base + fx + fg.BX * (fy + fg.BY * fz)] != 0u;
}
// One scene only (the reference schedule's kernels).
__device__ __forceinline__ bool lv_active(const LvMask& m, uint32_t x, uint32_t y, uint32_t z) {
  return lv_active(m, grid_scene<false>(m.fine, 0u), x, y, z);
}
__device__ __forceinline__ float nsum(const Grid& g, const float* __restrict__ s, uint32_t x, uint32_t y, uint32_t z) {
  return s[cidx(g, incu(x, g.W - 1), y, z)] + s[cidx(g, decu(x), y, z)] + s[cidx(g, x, incu(y, g.H - 1), z)] +
         s[cidx(g, x, decu(y), z)] + s[cidx(g, x, y, incu(z, g.D - 1))] + s[cidx(g, x, y, decu(z))];
}
// SH/pressure-jacobi.wgsl.ts:15-36 (weighted: mix(p, jacobi, omega))
__global__ __launch_bounds__(512) void k_jacobi(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                LvMask m, const StepConst* __restrict__ C, const float* __restrict__ rhs,
                                                const float* __restrict__ src, float* __restrict__ dst, int weighted) {
  const float om = C->omega;
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    if (on_boundary(g, v.x, v.y, v.z) || !lv_active(m, v.x, v.y, v.z)) { dst[v.i] = 0.f; continue; }
    // src == nullptr: the sweep starts from p = 0 (sum of six zeros, centre 0).
    float pj = ((src ? nsum(g, src, v.x, v.y, v.z) : 0.f) - rhs[v.i]) / 6.f;
    dst[v.i] = weighted ? mixf(src ? src[v.i] : 0.f, pj, om) : pj;
  }
}
// All `sweeps` sweeps of one small level in a single workgroup (1024 threads,
// __syncthreads between sweeps): p0 -> p1 -> p0 ... The first sweep starts from
// zero when zero_start is set. Used for the coarsest level, whose dispatches
// would otherwise be launch-bound.
//
// Staged path: k_coarse_setup (once per solve) compacts the active voxels of the
// listed bricks into slots (per-brick 512-bit masks, an exclusive scan of the
// per-brick counts) and writes one 16-byte table entry per slot: the voxel's
// global index and its six neighbour slots (16 bits each; kSmoothCap names the
// always-zero slot: inactive, boundary, or in an unlisted brick, as in the
// sparse invariant). It also zeroes the inactive voxels of both buffers, which
// the sweeps then never write. Each k_smooth_block thread owns up to kSmoothPer
// slots; every sweep is LDS-only.
constexpr uint32_t kSmoothCap = kCoarseCap;  // active voxels staged in LDS
constexpr uint32_t kSmoothBricks = 512;    // listed bricks whose voxel masks fit in LDS
constexpr uint32_t kSmoothTable = 4096;    // level bricks addressable through the LDS slot table
constexpr uint32_t kSmoothOverflow = 0xffffffffu;  // table header: the level takes the global path
constexpr int kSmoothPer = int(kSmoothCap / 1024);

// Bit position of the r-th (0-based) set bit of w; w has more than r set bits.
__device__ __forceinline__ uint32_t select_bit(uint64_t w, uint32_t r) {
  uint32_t pos = 0;
  for (uint32_t width = 32; width >= 1; width >>= 1) {
    const uint32_t c = uint32_t(__popcll(w & ((1ull << width) - 1ull)));
    if (r >= c) { r -= c; w >>= width; pos += width; }
  }
  return pos;
}
// Active voxels of listed brick k that precede local voxel t.
__device__ __forceinline__ uint32_t rank_in_brick(const uint64_t* am, uint32_t k, uint32_t t) {
  uint32_t r = 0;
  for (uint32_t j = 0; j < (t >> 6); ++j) r += uint32_t(__popcll(am[k * 8 + j]));
  return r + uint32_t(__popcll(am[k * 8 + (t >> 6)] & ((1ull << (t & 63u)) - 1ull)));
}

// First list position at or after brick b in an ascending brick list of n entries.
__device__ __forceinline__ uint32_t list_lower_bound(const uint32_t* __restrict__ list, uint32_t n, uint32_t b) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (list[mid] < b) lo = mid + 1; else hi = mid;
  }
  return lo;
}
// The listed bricks of g's scene: positions [k0, k0 + n) of the level's ascending list.
template <bool MS>
__device__ __forceinline__ void scene_list(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, const Grid& g,
                                           uint32_t& k0, uint32_t& n) {
  const uint32_t all = *count;
  if constexpr (MS) {
    k0 = list_lower_bound(list, all, g.base);
    n = list_lower_bound(list, all, g.base + g.BX * g.BY * g.BZ) - k0;
  } else {
    k0 = 0u;
    n = all;
  }
}

// Several scenes (MS): workgroup row blockIdx.y builds the table of scene
// blockIdx.y (tab + kSmoothCap * scene, tab_n + scene) from its bricks of the list.
template <bool MS>
__global__ __launch_bounds__(1024) void k_coarse_setup(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                       LvBits lb, uint4* __restrict__ tab, uint32_t* __restrict__ tab_n, float* p0,
                                                       float* p1) {
  FX_SIM_PRIORITY();
  __shared__ uint64_t amask[kSmoothBricks * 8];
  __shared__ uint32_t base[kSmoothBricks];
  __shared__ int16_t slot[kSmoothTable];
  __shared__ uint32_t wtot[16];
  const Grid g = grid_scene<MS>(G, blockIdx.y);
  tab += size_t(kSmoothCap) * blockIdx.y;
  tab_n += blockIdx.y;
  uint32_t k0, n;
  scene_list<MS>(list, count, g, k0, n);
  list += k0;
  const uint32_t nbl = g.BX * g.BY * g.BZ;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const bool lead = blockIdx.x == 0 && tid == 0;
  if (n == 0 || n > kSmoothBricks || nbl > kSmoothTable) {
    if (lead) *tab_n = n == 0 ? 0u : kSmoothOverflow;
    return;
  }
  for (uint32_t q = tid; q < nbl; q += 1024) slot[q] = -1;
  __syncthreads();
  for (uint32_t k = tid; k < n; k += 1024) slot[list[k] - g.base] = int16_t(k);
  for (uint32_t wd = tid; wd < n * 8; wd += 1024) amask[wd] = lb.bits[size_t(k0) * 8 + wd];
  __syncthreads();
  // exclusive scan of the per-brick counts (n <= 512 <= blockDim)
  uint32_t cnt = 0;
  if (tid < n)
    for (uint32_t j = 0; j < 8; ++j) cnt += uint32_t(__popcll(amask[tid * 8 + j]));
  uint32_t incl = cnt;
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  if (lane == 63) wtot[wave] = incl;
  __syncthreads();
  uint32_t pre = 0, nact = 0;
  for (uint32_t w = 0; w < wave; ++w) pre += wtot[w];
  for (uint32_t w = 0; w < 16; ++w) nact += wtot[w];
  if (tid < n) base[tid] = pre + incl - cnt;
  if (nact > kSmoothCap) {
    if (lead) *tab_n = kSmoothOverflow;
    return;
  }
  __syncthreads();
  const uint32_t first = blockIdx.x * 1024u + tid, stride = gridDim.x * 1024u;
  for (uint32_t q = first; q < nact; q += stride) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (base[mid] <= q) lo = mid; else hi = mid - 1;
    }
    uint32_t r = q - base[lo], wd = 0;
    for (; wd < 7; ++wd) {
      const uint32_t c = uint32_t(__popcll(amask[lo * 8 + wd]));
      if (r < c) break;
      r -= c;
    }
    const uint32_t t = wd * 64 + select_bit(amask[lo * 8 + wd], r);
    const uint32_t b = list[lo], lb0 = b - g.base;
    const uint32_t x = (lb0 % g.BX) * 8 + (t & 7u), y = ((lb0 / g.BX) % g.BY) * 8 + ((t >> 3) & 7u),
                   z = (lb0 / (g.BX * g.BY)) * 8 + (t >> 6);
    const uint32_t nx[6] = {x + 1, x - 1, x, x, x, x}, ny[6] = {y, y, y + 1, y - 1, y, y}, nz[6] = {z, z, z, z, z + 1, z - 1};
    uint32_t nb[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int sl = slot[(nx[k] >> 3) + g.BX * ((ny[k] >> 3) + g.BY * (nz[k] >> 3))];
      nb[k] = kSmoothCap;
      if (sl < 0) continue;
      const uint32_t tn = ((nz[k] & 7u) << 6) | ((ny[k] & 7u) << 3) | (nx[k] & 7u);
      if ((amask[uint32_t(sl) * 8 + (tn >> 6)] >> (tn & 63u)) & 1ull) nb[k] = base[sl] + rank_in_brick(amask, uint32_t(sl), tn);
    }
    tab[q] = make_uint4((b << 9) | t, nb[0] | (nb[1] << 16), nb[2] | (nb[3] << 16), nb[4] | (nb[5] << 16));
  }
  // the inactive in-domain voxels of the listed bricks hold 0 in both buffers
  for (uint32_t q = first; q < n * 512u; q += stride) {
    const uint32_t k = q >> 9, t = q & 511u;
    if ((amask[k * 8 + (t >> 6)] >> (t & 63u)) & 1ull) continue;
    const uint32_t b = list[k], lb0 = b - g.base;
    const uint32_t x = (lb0 % g.BX) * 8 + (t & 7u), y = ((lb0 / g.BX) % g.BY) * 8 + ((t >> 3) & 7u), z = (lb0 / (g.BX * g.BY)) * 8 + (t >> 6);
    if (x < g.W && y < g.H && z < g.D) { p0[(b << 9) | t] = 0.f; p1[(b << 9) | t] = 0.f; }
  }
  if (lead) *tab_n = nact;
}

// Several scenes (MS): workgroup blockIdx.y smooths scene blockIdx.y.
template <bool MS>
__global__ __launch_bounds__(1024) void k_smooth_block(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                       LvMask m, const StepConst* __restrict__ C0, const float* __restrict__ rhs,
                                                       float* p0, float* p1, int sweeps, int weighted, int zero_start,
                                                       const uint4* __restrict__ tab, const uint32_t* __restrict__ tab_n) {
  FX_SIM_PRIORITY();
  __shared__ float sp[2][kSmoothCap + 1];  // slot kSmoothCap holds 0: the missing neighbour
  const Grid g = grid_scene<MS>(G, blockIdx.y);
  const float om = C0[g.scene].omega;
  if (tab) {
    tab += size_t(kSmoothCap) * blockIdx.y;
    tab_n += blockIdx.y;
  }
  const uint32_t tid = threadIdx.x;
  const uint32_t nact = tab ? *tab_n : kSmoothOverflow;
  if (nact != kSmoothOverflow) {
    uint32_t gid[kSmoothPer], nbr[kSmoothPer][6];
    float rv[kSmoothPer], pv[kSmoothPer];
#pragma unroll
    for (int j = 0; j < kSmoothPer; ++j) {
      const uint32_t q = tid + uint32_t(j) * 1024u;
      const uint4 e = q < nact ? tab[q] : make_uint4(0xffffffffu, 0u, 0u, 0u);
      gid[j] = e.x;
      nbr[j][0] = e.y & 0xffffu; nbr[j][1] = e.y >> 16;
      nbr[j][2] = e.z & 0xffffu; nbr[j][3] = e.z >> 16;
      nbr[j][4] = e.w & 0xffffu; nbr[j][5] = e.w >> 16;
    }
    // the right-hand side and the start value of every owned slot stay in
    // registers; LDS carries the values neighbours read
#pragma unroll
    for (int j = 0; j < kSmoothPer; ++j) {
      const bool own = gid[j] != 0xffffffffu;
      rv[j] = own ? rhs[gid[j]] : 0.f;
      pv[j] = (own && !zero_start) ? p0[gid[j]] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < kSmoothPer; ++j)
      if (gid[j] != 0xffffffffu) sp[0][tid + uint32_t(j) * 1024u] = pv[j];
    if (tid == 0) { sp[0][kSmoothCap] = 0.f; sp[1][kSmoothCap] = 0.f; }
    __syncthreads();
    int cur = 0;
    for (int sw = 0; sw < sweeps; ++sw) {
      const float* A = sp[cur];
      float* B = sp[1 - cur];
#pragma unroll
      for (int j = 0; j < kSmoothPer; ++j) {
        if (gid[j] == 0xffffffffu) continue;
        const float sum = A[nbr[j][0]] + A[nbr[j][1]] + A[nbr[j][2]] + A[nbr[j][3]] + A[nbr[j][4]] + A[nbr[j][5]];
        const float pj = (sum - rv[j]) / 6.f;
        pv[j] = weighted ? mixf(pv[j], pj, om) : pj;
        B[tid + uint32_t(j) * 1024u] = pv[j];
      }
      __syncthreads();
      cur = 1 - cur;
    }
    float* dst = (sweeps & 1) ? p1 : p0;
#pragma unroll
    for (int j = 0; j < kSmoothPer; ++j)
      if (gid[j] != 0xffffffffu) dst[gid[j]] = pv[j];
    return;
  }
  uint32_t k0, nl;
  scene_list<MS>(list, count, g, k0, nl);
  const uint32_t total = nl * 512u;
  const Grid fg = grid_scene<MS>(m.fine, g.scene);
  for (int sw = 0; sw < sweeps; ++sw) {
    const float* src = (sw & 1) ? p1 : p0;
    float* dst = (sw & 1) ? p0 : p1;
    const bool zero = zero_start && sw == 0;
    for (uint32_t q = threadIdx.x; q < total; q += blockDim.x) {
      const uint32_t b = list[k0 + (q >> 9)], t = q & 511u, lb0 = b - g.base;
      uint32_t bx = lb0 % g.BX, by = (lb0 / g.BX) % g.BY, bz = lb0 / (g.BX * g.BY);
      uint32_t x = bx * 8 + (t & 7u), y = by * 8 + ((t >> 3) & 7u), z = bz * 8 + (t >> 6);
      if (x >= g.W || y >= g.H || z >= g.D) continue;
      const uint32_t i = (b << 9) | t;
      if (on_boundary(g, x, y, z) || !lv_active(m, fg, x, y, z)) { dst[i] = 0.f; continue; }
      float pj = ((zero ? 0.f : nsum(g, src, x, y, z)) - rhs[i]) / 6.f;
      dst[i] = weighted ? mixf(zero ? 0.f : src[i], pj, om) : pj;
    }
    __syncthreads();
  }
}
// SH/pressure-residual.wgsl.ts:16-38 (+ FP64 sum of squares when norm2 != null)
__global__ __launch_bounds__(512) void k_residual(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                  LvMask m, const float* __restrict__ rhs, const float* __restrict__ p,
                                                  float* __restrict__ r, double* norm2) {
  __shared__ double part[8];
  double acc = 0.0;
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    if (on_boundary(g, v.x, v.y, v.z) || !lv_active(m, v.x, v.y, v.z)) { r[v.i] = 0.f; continue; }
    float lap = nsum(g, p, v.x, v.y, v.z) - p[v.i] * 6.f;
    float res = rhs[v.i] - lap;
    r[v.i] = res;
    acc += double(res) * double(res);
  }
  if (!norm2) return;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < 8; ++k) s += part[k];
    atomicAdd(norm2 + (blockIdx.x % kNormLanes), s);
  }
}
// SH/pressure-restriction.wgsl.ts:27-50 (8-child mean) or (1,3,3,1)/8 x 4 (V-cycle).
__global__ __launch_bounds__(512) void k_restrict(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid f, Grid c,
                                                  LvMask mc, const float* __restrict__ src, float* __restrict__ dst, int full) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(c, list[fx_k]);
    if (!v.in) continue;
    if (on_boundary(c, v.x, v.y, v.z) || !lv_active(mc, v.x, v.y, v.z)) { dst[v.i] = 0.f; continue; }
    float sum = 0.f;
    if (!full) {
      for (uint32_t k = 0; k < 2; ++k)
        for (uint32_t j = 0; j < 2; ++j)
          for (uint32_t a = 0; a < 2; ++a) sum += src[cidx(f, 2 * v.x + a, 2 * v.y + j, 2 * v.z + k)];
      dst[v.i] = sum * 0.125f;
    } else {
      const float wt[4] = {0.125f, 0.375f, 0.375f, 0.125f};
      for (int k = 0; k < 4; ++k) {
        uint32_t fz = uint32_t(min(max(int(2 * v.z) - 1 + k, 0), int(f.D) - 1));
        for (int j = 0; j < 4; ++j) {
          uint32_t fy = uint32_t(min(max(int(2 * v.y) - 1 + j, 0), int(f.H) - 1));
          for (int a = 0; a < 4; ++a) {
            uint32_t fx = uint32_t(min(max(int(2 * v.x) - 1 + a, 0), int(f.W) - 1));
            sum += (wt[k] * wt[j] * wt[a]) * src[vidx(f, fx, fy, fz)];
          }
        }
      }
      dst[v.i] = sum * 4.f;
    }
  }
}
// SH/pressure-prolongation.wgsl.ts:29-44
__global__ __launch_bounds__(512) void k_prolong(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid c, Grid f,
                                                 LvMask mf, const float* __restrict__ src, float* __restrict__ dst) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(f, list[fx_k]);
    if (!v.in) continue;
    if (on_boundary(f, v.x, v.y, v.z) || !lv_active(mf, v.x, v.y, v.z)) { dst[v.i] = 0.f; continue; }
    F3 cp = f3((float(v.x) + 0.5f) * 0.5f - 0.5f, (float(v.y) + 0.5f) * 0.5f - 0.5f, (float(v.z) + 0.5f) * 0.5f - 0.5f);
    Tri t = tri(c, cp);
    dst[v.i] += tri_eval(t, src);
  }
}
// ------------------------------------------------ level activity bits --
// One ballot per (listed brick, z-slice) of every level: bit x + 8y is set iff
// the voxel is in the domain, off the boundary faces and active under the
// level's mask. Built once per solve; every V-cycle kernel reads these bits.
template <bool MS>
__global__ __launch_bounds__(512) void k_level_masks(LevelMaskArgs a) {
  FX_SIM_PRIORITY();
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (blockIdx.x == 0 && a.zero)
    for (size_t q = threadIdx.x; q < a.zero_words; q += blockDim.x) a.zero[q] = 0u;
  uint32_t start[FX_MAX_LEVELS_DEV + 1];
  start[0] = 0;
  for (int l = 0; l < a.nlevels; ++l) start[l + 1] = start[l] + *a.count[l];
  const uint32_t total = start[a.nlevels];
  // one wave per brick, over the bricks of all levels: lane (x, y) = (lane & 7,
  // lane >> 3) evaluates its column of the brick, one ballot per z-slice, and
  // lane z stores word z (8 bricks per workgroup and round, so that few rounds of
  // list and flag loads are in flight one after another)
  const uint32_t waves = gridDim.x * 8u;
  for (uint32_t e = blockIdx.x * 8u + wave; e < total; e += waves) {
    int l = 0;
    while (e >= start[l + 1]) ++l;
    const uint32_t k = e - start[l];
    const uint32_t b = wave_uniform(a.list[l][k]);
    const Grid g = grid_at<MS>(a.g[l], b);
    const Grid fg = grid_scene<MS>(a.m[l].fine, g.scene);
    const uint32_t lb = b - g.base;
    const uint32_t bx = lb % g.BX, by = (lb / g.BX) % g.BY, bz = lb / (g.BX * g.BY);
    const uint32_t x = bx * 8 + (lane & 7u), y = by * 8 + (lane >> 3);
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t zz = 0; zz < 8; ++zz) {
      const uint32_t z = bz * 8 + zz;
      const bool act = x < g.W && y < g.H && z < g.D && !on_boundary(g, x, y, z) && lv_active(a.m[l], fg, x, y, z);
      const uint64_t w = __ballot(act);
      mine = lane == zz ? w : mine;
    }
    if (lane < 8u) a.bits[l][size_t(k) * 8 + lane] = mine;
    if (l == 0 && a.nbr0) {
      // lane d < 6: is face neighbour d (+x, -x, +y, -y, +z, -z) of the brick listed?
      const int d = int(lane);
      const int nx = int(bx) + (d == 0) - (d == 1), ny = int(by) + (d == 2) - (d == 3), nz = int(bz) + (d == 4) - (d == 5);
      bool on = false;
      if (d < 6 && nx >= 0 && ny >= 0 && nz >= 0 && nx < int(g.BX) && ny < int(g.BY) && nz < int(g.BZ))
        on = a.page0[g.base + uint32_t(nx) + g.BX * (uint32_t(ny) + g.BY * uint32_t(nz))] >= 0;
      const uint64_t nw = __ballot(on);
      if (lane == 0) a.nbr0[k] = uint32_t(nw) & 0x3fu;
    }
  }
}

// Activity words of the 3x3x3 bricks around brick (bx, by, bz): nb[c * 8 + z],
// c = (dx+1) + 3 (dy+1) + 9 (dz+1); bricks outside the domain or unlisted are 0.
__device__ __forceinline__ uint64_t nbhd_word(const Grid& g, const LvBits& lb, int bx, int by, int bz, int q) {
  const int c = q >> 3, w = q & 7;
  const int nx = bx + c % 3 - 1, ny = by + (c / 3) % 3 - 1, nz = bz + c / 9 - 1;
  uint64_t v = 0;
  if (nx >= 0 && ny >= 0 && nz >= 0 && nx < int(g.BX) && ny < int(g.BY) && nz < int(g.BZ)) {
    const uint32_t nbk = g.base + uint32_t(nx) + g.BX * (uint32_t(ny) + g.BY * uint32_t(nz));
    const int pos = lb.page ? lb.page[nbk] : int(nbk);
    if (pos >= 0) v = lb.bits[size_t(pos) * 8 + w];
  }
  return v;
}
// The neighbourhood words in two phases, like the tiles (blockDim >= 216):
// nbhd_fetch issues the page and bit loads of word threadIdx.x into a register
// before the barrier that frees the LDS array, so that they are in flight
// together with the tile loads; nbhd_store writes the word after it.
__device__ __forceinline__ uint64_t nbhd_fetch(const Grid& g, const LvBits& lb, int bx, int by, int bz) {
  return threadIdx.x < 216u ? nbhd_word(g, lb, bx, by, bz, int(threadIdx.x)) : 0ull;
}
__device__ __forceinline__ void nbhd_store(uint64_t w, uint64_t* nb) {
  if (threadIdx.x < 216u) nb[threadIdx.x] = w;
}

// ------------------------------------------------ fused V-cycle kernels --
// Temporal blocking in LDS: one 512-thread workgroup per brick of the level.
// Tiles share one origin (brick origin - R0) and side S = 8 + 2*R0; stage s of
// a chain computes on the brick +- (R0 - s). Halo voxels are recomputed with the
// same arithmetic their own brick uses, so results equal the unfused sweeps.
// Inactive voxels (and everything outside the domain) hold 0 in every tile.
template <int R0>
struct Tile {
  static_assert(R0 <= 8, "tiles reach at most one brick beyond their own");
  static constexpr int S = 8 + 2 * R0;
  static constexpr int N = S * S * S;
};
// Activity of tile voxel (tx, ty, tz) from the neighbourhood words.
template <int R0>
__device__ __forceinline__ bool tile_on(const uint64_t* nb, int tx, int ty, int tz) {
  const int lx = tx - R0 + 8, ly = ty - R0 + 8, lz = tz - R0 + 8;
  return (nb[((lx >> 3) + 3 * ((ly >> 3) + 3 * (lz >> 3))) * 8 + (lz & 7)] >> ((lx & 7) | ((ly & 7) << 3))) & 1ull;
}
// Tile loads in two phases so that every global load of a tile (and the
// neighbourhood words) is in flight at once: tile_fetch reads f on brick +- R0
// into registers (0 outside the domain); tile_store writes them to LDS after the
// neighbourhood words have landed, keeping active voxels only (inactive voxels
// hold 0 in memory as well, by the sparse invariant).
template <int R0>
struct TileRegs {
  static constexpr int IT = (Tile<R0>::N + 511) / 512;
  float v[IT];
};
template <int R0>
__device__ __forceinline__ void tile_fetch(const Grid& g, const float* __restrict__ f, int ox, int oy, int oz, TileRegs<R0>& r) {
  constexpr int S = Tile<R0>::S, N = Tile<R0>::N;
#pragma unroll
  for (int j = 0; j < TileRegs<R0>::IT; ++j) {
    const int q = int(threadIdx.x) + j * 512;
    const int x = ox + q % S, y = oy + (q / S) % S, z = oz + q / (S * S);
    r.v[j] = (q < N && x >= 0 && y >= 0 && z >= 0 && x < int(g.W) && y < int(g.H) && z < int(g.D))
                 ? f[vidx(g, uint32_t(x), uint32_t(y), uint32_t(z))]
                 : 0.f;
  }
}
template <int R0>
__device__ __forceinline__ void tile_store(const uint64_t* nb, const TileRegs<R0>& r, float* t) {
  constexpr int S = Tile<R0>::S, N = Tile<R0>::N;
#pragma unroll
  for (int j = 0; j < TileRegs<R0>::IT; ++j) {
    const int q = int(threadIdx.x) + j * 512;
    if (q < N) t[q] = tile_on<R0>(nb, q % S, (q / S) % S, q / (S * S)) ? r.v[j] : 0.f;
  }
}
template <int R0>
__device__ __forceinline__ void tile_zero(float* t) {
  for (int q = threadIdx.x; q < Tile<R0>::N; q += 512) t[q] = 0.f;
}
// One weighted-Jacobi sweep A -> B on brick +- R (tiles of radius R0).
template <int R0, int R>
__device__ __forceinline__ void tile_sweep(const uint64_t* nb, const float* A, const float* rhs, float* B, float om) {
  constexpr int S = Tile<R0>::S, side = 8 + 2 * R, n = side * side * side, off = R0 - R;
  for (int q = threadIdx.x; q < n; q += 512) {
    const int tx = off + q % side, ty = off + (q / side) % side, tz = off + q / (side * side);
    const int i = tx + S * (ty + S * tz);
    if (!tile_on<R0>(nb, tx, ty, tz)) { B[i] = 0.f; continue; }
    const float sum = A[i + 1] + A[i - 1] + A[i + S] + A[i - S] + A[i + S * S] + A[i - S * S];
    const float pj = (sum - rhs[i]) / 6.f;
    B[i] = mixf(A[i], pj, om);
  }
}

// Pre-smoothing (PRE sweeps), residual, FP64 norm and (1,3,3,1)/8 x 4 restriction.
// p_in == nullptr: start from p = 0. Writes the smoothed p of the brick to p_out
// (skipped when p_out == nullptr) and the coarse right-hand side of the 4^3
// coarse voxels the brick covers.
template <int PRE, bool MS>
__global__ __launch_bounds__(512) void k_vdown(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid F,
                                               LvBits bf, Grid CG, LvMask mc, const StepConst* __restrict__ C0,
                                               const float* __restrict__ rhs, const float* __restrict__ p_in,
                                               float* __restrict__ p_out, float* __restrict__ crhs, double* norm2) {
  constexpr int R0 = PRE + 2, S = Tile<R0>::S, N = Tile<R0>::N;
  __shared__ float ta[N], tb[N], tr[N];
  __shared__ uint64_t nb[216];
  __shared__ double part[8];
  double acc = 0.0;
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid f = gk.at(F, b);
    const Grid c = gk.scene(CG);
    const float om = C0->omega;  // the scenes of a batch share the V-cycle settings
    int bx, by, bz;
    brick_coords(f, b, bx, by, bz);
    const int ox = bx * 8 - R0, oy = by * 8 - R0, oz = bz * 8 - R0;
    // activity of the coarse voxel this thread restricts (threads 0..63)
    bool cact = false;
    if (threadIdx.x < 64) {
      const int t = threadIdx.x;
      const uint32_t cx = uint32_t(bx * 4 + (t & 3)), cy = uint32_t(by * 4 + ((t >> 2) & 3)), cz = uint32_t(bz * 4 + (t >> 4));
      cact = cx < c.W && cy < c.H && cz < c.D && !on_boundary(c, cx, cy, cz) &&
             lv_active(mc, gk.scene(mc.fine), cx, cy, cz);
    }
    TileRegs<R0> rr, rp;
    tile_fetch<R0>(f, rhs, ox, oy, oz, rr);
    if (p_in) tile_fetch<R0>(f, p_in, ox, oy, oz, rp);
    const uint64_t nw = nbhd_fetch(f, bf, bx, by, bz);
    __syncthreads();  // previous brick's tiles are no longer read
    nbhd_store(nw, nb);
    __syncthreads();
    tile_store<R0>(nb, rr, tr);
    if (p_in) tile_store<R0>(nb, rp, ta); else tile_zero<R0>(ta);
    __syncthreads();
    float* A = ta;
    float* B = tb;
    if constexpr (PRE >= 1) {
      tile_sweep<R0, R0 - 1>(nb, A, tr, B, om);
      __syncthreads();
      float* t = A; A = B; B = t;
    }
    if constexpr (PRE >= 2) {
      tile_sweep<R0, R0 - 2>(nb, A, tr, B, om);
      __syncthreads();
      float* t = A; A = B; B = t;
    }
    // residual on brick +- 1 into B
    {
      constexpr int side = 10, n = 1000, off = R0 - 1;
      for (int q = threadIdx.x; q < n; q += 512) {
        const int tx = off + q % side, ty = off + (q / side) % side, tz = off + q / (side * side);
        const int i = tx + S * (ty + S * tz);
        float res = 0.f;
        if (tile_on<R0>(nb, tx, ty, tz)) {
          const float sum = A[i + 1] + A[i - 1] + A[i + S] + A[i - S] + A[i + S * S] + A[i - S * S];
          const float lap = sum - A[i] * 6.f;
          res = tr[i] - lap;
        }
        B[i] = res;
        const bool inside = tx >= R0 && ty >= R0 && tz >= R0 && tx < R0 + 8 && ty < R0 + 8 && tz < R0 + 8;
        if (inside) acc += double(res) * double(res);
      }
    }
    // smoothed p of the brick
    if (p_out) {
      const int t = threadIdx.x;
      const int tx = R0 + (t & 7), ty = R0 + ((t >> 3) & 7), tz = R0 + (t >> 6);
      if (ox + tx < int(f.W) && oy + ty < int(f.H) && oz + tz < int(f.D)) p_out[(b << 9) | t] = A[tx + S * (ty + S * tz)];
    }
    __syncthreads();
    // (1,3,3,1)/8 x 4 restriction of the brick's 4^3 coarse voxels (the tap order
    // of the unfused restriction, taps clamped to the domain)
    if (threadIdx.x < 64) {
      const int t = threadIdx.x;
      const int cx = bx * 4 + (t & 3), cy = by * 4 + ((t >> 2) & 3), cz = bz * 4 + (t >> 4);
      if (cx < int(c.W) && cy < int(c.H) && cz < int(c.D)) {
        float out = 0.f;
        if (cact) {
          const float wt[4] = {0.125f, 0.375f, 0.375f, 0.125f};
          float sum = 0.f;
          for (int k = 0; k < 4; ++k) {
            const int fz = min(max(2 * cz - 1 + k, 0), int(f.D) - 1) - oz;
            for (int j = 0; j < 4; ++j) {
              const int fy = min(max(2 * cy - 1 + j, 0), int(f.H) - 1) - oy;
              for (int a = 0; a < 4; ++a) {
                const int fx = min(max(2 * cx - 1 + a, 0), int(f.W) - 1) - ox;
                sum += (wt[k] * wt[j] * wt[a]) * B[fx + S * (fy + S * fz)];
              }
            }
          }
          out = sum * 4.f;
        }
        crhs[vidx(c, uint32_t(cx), uint32_t(cy), uint32_t(cz))] = out;
      }
    }
  }
  if (!norm2) return;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sacc = 0.0;
    for (int k = 0; k < 8; ++k) sacc += part[k];
    atomicAdd(norm2 + (blockIdx.x % kNormLanes), sacc);
  }
}

// Prolongation of the coarse correction e (pressure-prolongation.wgsl.ts:29-44
// sampler) added to p_in on brick +- POST, then POST sweeps; writes p_out.
template <int POST, bool MS>
__global__ __launch_bounds__(512) void k_vup(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid CG, Grid F,
                                             LvBits bf, const StepConst* __restrict__ C0, const float* __restrict__ e,
                                             const float* __restrict__ rhs, const float* __restrict__ p_in,
                                             float* __restrict__ p_out) {
  constexpr int R0 = POST > 0 ? POST : 0, S = Tile<R0>::S, N = Tile<R0>::N;
  constexpr int CR = (POST + 3) / 2, CS = 4 + 2 * CR, CN = CS * CS * CS;
  __shared__ float ta[N], tb[N], tr[N], tc[CN];
  __shared__ uint64_t nb[216];
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid f = gk.at(F, b);
    const Grid c = gk.scene(CG);
    const float om = C0->omega;  // the scenes of a batch share the V-cycle settings
    const float mcx = float(c.W - 1), mcy = float(c.H - 1), mcz = float(c.D - 1);
    int bx, by, bz;
    brick_coords(f, b, bx, by, bz);
    const int ox = bx * 8 - R0, oy = by * 8 - R0, oz = bz * 8 - R0;
    const int cox = bx * 4 - CR, coy = by * 4 - CR, coz = bz * 4 - CR;
    constexpr int CIT = (CN + 511) / 512;
    float cv[CIT];
#pragma unroll
    for (int j = 0; j < CIT; ++j) {
      const int q = int(threadIdx.x) + j * 512;
      const int x = min(max(cox + q % CS, 0), int(c.W) - 1), y = min(max(coy + (q / CS) % CS, 0), int(c.H) - 1),
                z = min(max(coz + q / (CS * CS), 0), int(c.D) - 1);
      cv[j] = q < CN ? e[vidx(c, uint32_t(x), uint32_t(y), uint32_t(z))] : 0.f;
    }
    TileRegs<R0> rp, rr;
    tile_fetch<R0>(f, p_in, ox, oy, oz, rp);
    if (POST > 0) tile_fetch<R0>(f, rhs, ox, oy, oz, rr);
    const uint64_t nw = nbhd_fetch(f, bf, bx, by, bz);
    __syncthreads();
    nbhd_store(nw, nb);
#pragma unroll
    for (int j = 0; j < CIT; ++j) {
      const int q = int(threadIdx.x) + j * 512;
      if (q < CN) tc[q] = cv[j];
    }
    __syncthreads();
    tile_store<R0>(nb, rp, ta);
    if (POST > 0) tile_store<R0>(nb, rr, tr);
    __syncthreads();
    // p + P(e) on brick +- R0 into tb
    for (int q = threadIdx.x; q < N; q += 512) {
      const int tx = q % S, ty = (q / S) % S, tz = q / (S * S);
      if (!tile_on<R0>(nb, tx, ty, tz)) { tb[q] = 0.f; continue; }
      const int x = ox + tx, y = oy + ty, z = oz + tz;
      const float px = clampf((float(x) + 0.5f) * 0.5f - 0.5f, 0.f, mcx), py = clampf((float(y) + 0.5f) * 0.5f - 0.5f, 0.f, mcy),
                  pz = clampf((float(z) + 0.5f) * 0.5f - 0.5f, 0.f, mcz);
      const int x0 = int(floorf(px)), y0 = int(floorf(py)), z0 = int(floorf(pz));
      const int x1 = min(x0 + 1, int(c.W) - 1), y1 = min(y0 + 1, int(c.H) - 1), z1 = min(z0 + 1, int(c.D) - 1);
      const float fx = fractf(px), fy = fractf(py), fz = fractf(pz);
      auto at = [&](int a, int bb, int cc) { return tc[(a - cox) + CS * ((bb - coy) + CS * (cc - coz))]; };
      const float x00 = mixf(at(x0, y0, z0), at(x1, y0, z0), fx);
      const float x10 = mixf(at(x0, y1, z0), at(x1, y1, z0), fx);
      const float x01 = mixf(at(x0, y0, z1), at(x1, y0, z1), fx);
      const float x11 = mixf(at(x0, y1, z1), at(x1, y1, z1), fx);
      tb[q] = ta[q] + mixf(mixf(x00, x10, fy), mixf(x01, x11, fy), fz);
    }
    __syncthreads();
    float* A = tb;
    float* B = ta;
    if constexpr (POST >= 1) {
      tile_sweep<R0, R0 - 1>(nb, A, tr, B, om);
      __syncthreads();
      float* t = A; A = B; B = t;
    }
    if constexpr (POST >= 2) {
      tile_sweep<R0, R0 - 2>(nb, A, tr, B, om);
      __syncthreads();
      float* t = A; A = B; B = t;
    }
    if constexpr (POST >= 3) {
      tile_sweep<R0, R0 - 3>(nb, A, tr, B, om);
      __syncthreads();
      float* t = A; A = B; B = t;
    }
    const int t = threadIdx.x;
    const int tx = R0 + (t & 7), ty = R0 + ((t >> 3) & 7), tz = R0 + (t >> 6);
    if (ox + tx < int(f.W) && oy + ty < int(f.H) && oz + tz < int(f.D)) p_out[(b << 9) | t] = A[tx + S * (ty + S * tz)];
  }
}

// p += P(e) on the active voxels of a level: the prolongation sampler of
// pressure-prolongation.wgsl.ts:29-44 with the same arithmetic as the fused
// up-leg, in place, on float4 brick rows (the layout of k_sweep). Voxel
// x = 4m + q samples coarse columns x0 = floor(px), x1 = x0 + 1 with
// px = 2m + q/2 - 1/4: columns (2m-1, 2m), (2m, 2m+1), (2m, 2m+1), (2m+1, 2m+2)
// for q = 0..3. The thread loads columns 2m-1 .. 2m+2 (clamped to the level) at
// the 4 (y, z) corners once. Where px is clamped to the level, the weight fx is
// 0 and the clamped column holds the value the clamped x0 reads, so every blend
// equals a one-voxel-per-thread prolongation's.
template <bool MS>
__global__ __launch_bounds__(512) void k_prolong_add(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid CG,
                                                     Grid F, LvBits lb, const float* __restrict__ e, float* __restrict__ p) {
  FX_SIM_PRIORITY();
  const uint32_t slot = threadIdx.x >> 7, t = threadIdx.x & 127u;
  const uint32_t h = t & 1u, ly = (t >> 1) & 7u, lz = t >> 4;
  const BrickRange br = brick_range_slots(*count, 4u, slot);
  GridCache<MS> gk;
  for (uint32_t k = br.k; k < br.end; k += br.step) {
    const uint32_t b = wave_uniform(list[k]);
    const uint32_t act = uint32_t(lb.bits[size_t(k) * 8 + lz] >> ((ly << 3) | (h * 4u))) & 0xFu;
    if (!act) continue;
    const Grid f = gk.at(F, b);
    const Grid c = gk.scene(CG);
    const float mcx = float(c.W - 1), mcy = float(c.H - 1), mcz = float(c.D - 1);
    const uint32_t lbk = b - f.base;
    const uint32_t xb = (lbk % f.BX) * 8 + h * 4u, y = ((lbk / f.BX) % f.BY) * 8 + ly, z = (lbk / (f.BX * f.BY)) * 8 + lz;
    const float py = clampf((float(y) + 0.5f) * 0.5f - 0.5f, 0.f, mcy), pz = clampf((float(z) + 0.5f) * 0.5f - 0.5f, 0.f, mcz);
    const uint32_t y0 = uint32_t(floorf(py)), z0 = uint32_t(floorf(pz));
    const uint32_t y1 = min(y0 + 1u, c.H - 1u), z1 = min(z0 + 1u, c.D - 1u);
    const float fy = fractf(py), fz = fractf(pz);
    const int base = int(xb / 2u) - 1;  // 2m - 1
    float col[4][4];                    // [column j][corner y0z0, y1z0, y0z1, y1z1]
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t xc = uint32_t(min(max(base + j, 0), int(c.W) - 1));
      col[j][0] = e[vidx(c, xc, y0, z0)];
      col[j][1] = e[vidx(c, xc, y1, z0)];
      col[j][2] = e[vidx(c, xc, y0, z1)];
      col[j][3] = e[vidx(c, xc, y1, z1)];
    }
    const uint32_t i0 = (b << 9) | (lz << 6) | (ly << 3) | (h * 4u);
    float4 pv = *reinterpret_cast<const float4*>(p + i0);
    float pr[4] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j0 = q == 0 ? 0 : q == 3 ? 2 : 1, j1 = j0 + 1;
      const float fx = fractf(clampf((float(xb + uint32_t(q)) + 0.5f) * 0.5f - 0.5f, 0.f, mcx));
      const float x00 = mixf(col[j0][0], col[j1][0], fx);
      const float x10 = mixf(col[j0][1], col[j1][1], fx);
      const float x01 = mixf(col[j0][2], col[j1][2], fx);
      const float x11 = mixf(col[j0][3], col[j1][3], fx);
      const float add = pr[q] + mixf(mixf(x00, x10, fy), mixf(x01, x11, fy), fz);
      pr[q] = ((act >> q) & 1u) ? add : pr[q];
    }
    *reinterpret_cast<float4*>(p + i0) = make_float4(pr[0], pr[1], pr[2], pr[3]);
  }
}

// Rows of a brick as float4s (4 voxels per thread, 128 threads per brick): the
// layout of k_sweep. RowNbrs holds a row's own values and its six neighbour
// sets in the order xp, xm, yp, ym, zp, zm; brick ids are clamped to the
// allocation so that rows can be read before activity says whether they are used.
struct RowNbrs {
  float c[4], xp[4], xm[4], yp[4], ym[4], zp[4], zm[4];
};
// nb6: the listed face neighbours of the brick (LvBits::nbr bits); a row in an
// unlisted neighbour reads the brick's own row (in cache) and is taken as 0.
__device__ __forceinline__ void row_load(const float* __restrict__ f, uint32_t b, uint32_t h, uint32_t ly, uint32_t lz,
                                         uint32_t BX, uint32_t BXY, uint32_t last, RowNbrs& R, uint32_t nb6 = 0x3fu) {
  const uint32_t x0 = h * 4u, i0 = (b << 9) | (lz << 6) | (ly << 3) | x0;
  const bool typ = ly < 7u || (nb6 & 4u), tym = ly > 0u || (nb6 & 8u), tzp = lz < 7u || (nb6 & 16u),
             tzm = lz > 0u || (nb6 & 32u), te = h ? (nb6 & 1u) : (nb6 & 2u);
  const uint32_t iyp = !typ ? i0 : ly < 7u ? i0 + 8u : ((min(b + BX, last) << 9) | (lz << 6) | x0);
  const uint32_t iym = !tym ? i0 : ly > 0u ? i0 - 8u : (((b >= BX ? b - BX : 0u) << 9) | (lz << 6) | (7u << 3) | x0);
  const uint32_t izp = !tzp ? i0 : lz < 7u ? i0 + 64u : ((min(b + BXY, last) << 9) | (ly << 3) | x0);
  const uint32_t izm = !tzm ? i0 : lz > 0u ? i0 - 64u : (((b >= BXY ? b - BXY : 0u) << 9) | (7u << 6) | (ly << 3) | x0);
  const uint32_t iedge = !te ? i0 : h ? ((min(b + 1u, last) << 9) | (lz << 6) | (ly << 3))
                                      : (((b > 0u ? b - 1u : 0u) << 9) | (lz << 6) | (ly << 3) | 7u);
  const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 c = *reinterpret_cast<const float4*>(f + i0);
  float4 yp = *reinterpret_cast<const float4*>(f + iyp);
  float4 ym = *reinterpret_cast<const float4*>(f + iym);
  float4 zp = *reinterpret_cast<const float4*>(f + izp);
  float4 zm = *reinterpret_cast<const float4*>(f + izm);
  float edge = f[iedge];
  if (!typ) yp = z4;
  if (!tym) ym = z4;
  if (!tzp) zp = z4;
  if (!tzm) zm = z4;
  if (!te) edge = 0.f;
  const float other = __shfl_xor(h ? c.x : c.w, 1, 64);
  R.c[0] = c.x; R.c[1] = c.y; R.c[2] = c.z; R.c[3] = c.w;
  R.xp[0] = c.y; R.xp[1] = c.z; R.xp[2] = c.w; R.xp[3] = h ? edge : other;
  R.xm[0] = h ? other : edge; R.xm[1] = c.x; R.xm[2] = c.y; R.xm[3] = c.z;
  R.yp[0] = yp.x; R.yp[1] = yp.y; R.yp[2] = yp.z; R.yp[3] = yp.w;
  R.ym[0] = ym.x; R.ym[1] = ym.y; R.ym[2] = ym.z; R.ym[3] = ym.w;
  R.zp[0] = zp.x; R.zp[1] = zp.y; R.zp[2] = zp.z; R.zp[3] = zp.w;
  R.zm[0] = zm.x; R.zm[1] = zm.y; R.zm[2] = zm.z; R.zm[3] = zm.w;
}

// Residual r = rhs - (sum of 6 neighbours - 6 p) of a V-cycle level
// (pressure-residual.wgsl.ts:16-38) on float4 brick rows, with the FP64 sum of
// squares over the active voxels when norm2 != nullptr. Each voxel's arithmetic
// is that of a one-voxel-per-thread residual; only the FP64 summation order of
// the norm differs.
template <bool MS>
__global__ __launch_bounds__(512) void k_residual_lean(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                       LvBits lb, const float* __restrict__ rhs, const float* __restrict__ p,
                                                       float* __restrict__ r, double* norm2) {
  FX_SIM_PRIORITY();
  __shared__ double part[8];
  const uint32_t slot = threadIdx.x >> 7, t = threadIdx.x & 127u;
  const uint32_t h = t & 1u, ly = (t >> 1) & 7u, lz = t >> 4;
  const uint32_t last = G.span - 1u;  // brick ids of unused neighbour rows stay in the allocation
  const uint32_t rofs = (lz << 6) | (ly << 3) | (h * 4u);
  double acc = 0.0;
  const BrickRange br = brick_range_slots(*count, 4u, slot);
  GridCache<MS> gk;
  for (uint32_t k = br.k; k < br.end; k += br.step) {
    const uint32_t b = wave_uniform(list[k]);
    const Grid g = gk.at(G, b);
    const uint32_t BX = g.BX, BXY = g.BX * g.BY;
    const uint32_t i0 = (b << 9) | rofs;
    const uint32_t act = uint32_t(lb.bits[size_t(k) * 8 + lz] >> ((ly << 3) | (h * 4u))) & 0xFu;
    const float4 rh = *reinterpret_cast<const float4*>(rhs + i0);
    RowNbrs R;
    row_load(p, b, h, ly, lz, BX, BXY, last, R, lb.nbr ? lb.nbr[k] : 0x3fu);
    const float rs[4] = {rh.x, rh.y, rh.z, rh.w};
    float out[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      out[e] = 0.f;
      if ((act >> e) & 1u) {
        const float sum = R.xp[e] + R.xm[e] + R.yp[e] + R.ym[e] + R.zp[e] + R.zm[e];
        const float lap = sum - R.c[e] * 6.f;
        out[e] = rs[e] - lap;
        acc += double(out[e]) * double(out[e]);
      }
    }
    *reinterpret_cast<float4*>(r + i0) = make_float4(out[0], out[1], out[2], out[3]);
  }
  if (!norm2) return;
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sacc = 0.0;
    for (int k = 0; k < 8; ++k) sacc += part[k];
    atomicAdd(norm2 + (blockIdx.x % kNormLanes), sacc);
  }
}
// (1,3,3,1)/8 x 4 restriction of r into the coarse right-hand side, one
// workgroup per listed coarse brick: the 18^3 fine region its taps reach is
// staged in LDS (every load in flight at once), then each thread sums the 64
// taps of its coarse voxel in the oracle's order (taps clamped to the domain).
template <bool MS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8))) void k_restrict_lean(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid F,
                                                       Grid CG, LvBits cb, const float* __restrict__ r, float* __restrict__ crhs) {
  constexpr int S = 18, N = S * S * S, IT = (N + 511) / 512;
  __shared__ float tf[N];
  const uint32_t t = threadIdx.x;
  const float wt[4] = {0.125f, 0.375f, 0.375f, 0.125f};
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid c = gk.at(CG, b);
    const Grid f = gk.scene(F);
    const uint32_t lb = b - c.base;
    const int cbx = int(lb % c.BX), cby = int((lb / c.BX) % c.BY), cbz = int(lb / (c.BX * c.BY));
    // fine origin of the region: 2 * (coarse brick origin) - 1
    const int ox = cbx * 16 - 1, oy = cby * 16 - 1, oz = cbz * 16 - 1;
    float v[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int q = int(t) + j * 512;
      const int x = min(max(ox + q % S, 0), int(f.W) - 1), y = min(max(oy + (q / S) % S, 0), int(f.H) - 1),
                z = min(max(oz + q / (S * S), 0), int(f.D) - 1);
      v[j] = q < N ? r[vidx(f, uint32_t(x), uint32_t(y), uint32_t(z))] : 0.f;
    }
    const bool act = (cb.bits[size_t(fx_k) * 8 + (t >> 6)] >> (t & 63u)) & 1ull;
    __syncthreads();  // the previous brick's taps have been read
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int q = int(t) + j * 512;
      if (q < N) tf[q] = v[j];
    }
    __syncthreads();
    const uint32_t i = (b << 9) | t;
    if (!act) { crhs[i] = 0.f; continue; }
    const int cx = cbx * 8 + int(t & 7u), cy = cby * 8 + int((t >> 3) & 7u), cz = cbz * 8 + int(t >> 6);
    float sum = 0.f;
    for (int k = 0; k < 4; ++k) {
      const int fz = min(max(2 * cz - 1 + k, 0), int(f.D) - 1) - oz;
      for (int j = 0; j < 4; ++j) {
        const int fy = min(max(2 * cy - 1 + j, 0), int(f.H) - 1) - oy;
        for (int a = 0; a < 4; ++a) {
          const int fx = min(max(2 * cx - 1 + a, 0), int(f.W) - 1) - ox;
          sum += (wt[k] * wt[j] * wt[a]) * tf[fx + S * (fy + S * fz)];
        }
      }
    }
    crhs[i] = sum * 4.f;
  }
}

// One weighted-Jacobi sweep of a V-cycle level (pressure-jacobi.wgsl.ts:15-36,
// mix(p, jacobi, omega)) on float4 brick rows (row_load): each thread owns 4
// consecutive voxels of a brick row (x = 0..3 or 4..7), 128 threads per brick, 4
// bricks side by side in a workgroup. The row and its +-y and +-z neighbour rows
// are float4 loads; the in-row x neighbours come from the registers and the
// partner lane (the other half of the row, one shuffle); only the x neighbour
// across the brick face is a scalar load. Every voxel sums its neighbours in the
// same order as a one-voxel-per-thread sweep (xp, xm, yp, ym, zp, zm).
// Neighbours are addressed through home-page brick arithmetic (active voxels are
// interior, so every neighbour brick exists; unlisted ones hold 0). ZERO: the
// sweep starts from p = 0. FINE: the finest level, where every in-domain voxel
// of a listed brick off the boundary faces is active; the test uses coordinates
// instead of the activity words.
template <bool ZERO, bool FINE, bool MS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_sweep(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                               LvBits lb, const StepConst* __restrict__ C0, const float* __restrict__ rhs,
                                               const float* __restrict__ src, float* __restrict__ dst) {
  FX_SIM_PRIORITY();
  // one wave per brick, 8 bricks per workgroup: lane t owns x = 4h .. 4h+3 of row
  // y in the two z-layers z0 = 2 zq and z1 = z0 + 1 (each layer's z neighbour on
  // that side is the other layer, in registers); a wave's loads of one layer pair
  // stay contiguous (1 KB per instruction)
  const uint32_t slot = threadIdx.x >> 6, t = threadIdx.x & 63u;
  const uint32_t h = t & 1u, ly = (t >> 1) & 7u, zq = t >> 4, x0 = h * 4u, z0 = 2u * zq;
  const uint32_t last = G.span - 1u;  // brick ids of unused neighbour rows stay in the allocation
  const BrickRange br = brick_range_slots(*count, 4u, slot);
  const uint32_t rofs = (z0 << 6) | (ly << 3) | x0;
  GridCache<MS> gk;
  for (uint32_t k = br.k; k < br.end; k += br.step) {
    const uint32_t b = wave_uniform(list[k]);
    const Grid g = gk.at(G, b);
    const float om = C0->omega;  // the scenes of a batch share the V-cycle settings
    const uint32_t BX = g.BX, BXY = g.BX * g.BY;
    const uint32_t i0 = (b << 9) | rofs, i1 = i0 + 64u;
    uint32_t act = 0;  // bits 0-3: layer z0, bits 4-7: layer z1
    if (FINE) {
      // in the domain and off its faces: 1 <= c <= n - 2 on every axis (separable)
      const uint32_t lbk = b - g.base;
      const uint32_t x = (lbk % g.BX) * 8 + x0, y = ((lbk / g.BX) % g.BY) * 8 + ly, z = (lbk / BXY) * 8 + z0;
      uint32_t xs = 0;
#pragma unroll
      for (uint32_t e = 0; e < 4; ++e) xs |= (x + e >= 1u && x + e + 2u <= g.W) ? (1u << e) : 0u;
      const bool yok = y >= 1u && y + 2u <= g.H;
      const bool z0ok = z >= 1u && z + 2u <= g.D, z1ok = z + 3u <= g.D;
      act = yok ? ((z0ok ? xs : 0u) | (z1ok ? xs << 4 : 0u)) : 0u;
    } else {
      act = (uint32_t(lb.bits[size_t(k) * 8 + z0] >> ((ly << 3) | x0)) & 0xFu) |
            ((uint32_t(lb.bits[size_t(k) * 8 + z0 + 1] >> ((ly << 3) | x0)) & 0xFu) << 4);
    }
    float r[8], c[8], yp[8], ym[8], zm0[4], zp1[4];
    float e0 = 0.f, e1 = 0.f;
    unpack4(*reinterpret_cast<const float4*>(rhs + i0), r);
    unpack4(*reinterpret_cast<const float4*>(rhs + i1), r + 4);
    if (!ZERO) {
      const uint32_t iyp0 = ly < 7u ? i0 + 8u : ((min(b + BX, last) << 9) | (z0 << 6) | x0);
      const uint32_t iym0 = ly > 0u ? i0 - 8u : (((b >= BX ? b - BX : 0u) << 9) | (z0 << 6) | (7u << 3) | x0);
      const uint32_t izm0 = z0 > 0u ? i0 - 64u : (((b >= BXY ? b - BXY : 0u) << 9) | (7u << 6) | (ly << 3) | x0);
      const uint32_t izp1 = z0 + 1u < 7u ? i1 + 64u : ((min(b + BXY, last) << 9) | (ly << 3) | x0);
      const uint32_t ie0 = h ? ((min(b + 1u, last) << 9) | (z0 << 6) | (ly << 3)) : (((b > 0u ? b - 1u : 0u) << 9) | (z0 << 6) | (ly << 3) | 7u);
      // Rows in an unlisted face neighbour are 0 (the sparse zero invariant). Their
      // lanes load the brick's own row instead, which is in cache, and take 0: the
      // cold zero pages of inactive neighbours stay out of the traffic, and every
      // load stays one unconditional instruction (a branch around the loads made
      // the sweep 50% slower).
      const uint32_t nb6 = (FINE && lb.nbr) ? lb.nbr[k] : 0x3fu;
      const bool typ = ly < 7u || (nb6 & 4u), tym = ly > 0u || (nb6 & 8u), tzm = z0 > 0u || (nb6 & 32u),
                 tzp = z0 + 1u < 7u || (nb6 & 16u), te = h ? (nb6 & 1u) : (nb6 & 2u);
      const uint32_t jyp = typ ? iyp0 : i0, jym = tym ? iym0 : i0, jzm = tzm ? izm0 : i0, jzp = tzp ? izp1 : i1,
                     je = te ? ie0 : i0;
      unpack4(*reinterpret_cast<const float4*>(src + i0), c);
      unpack4(*reinterpret_cast<const float4*>(src + i1), c + 4);
      unpack4(*reinterpret_cast<const float4*>(src + jyp), yp);
      unpack4(*reinterpret_cast<const float4*>(src + jyp + 64u), yp + 4);
      unpack4(*reinterpret_cast<const float4*>(src + jym), ym);
      unpack4(*reinterpret_cast<const float4*>(src + jym + 64u), ym + 4);
      unpack4(*reinterpret_cast<const float4*>(src + jzm), zm0);
      unpack4(*reinterpret_cast<const float4*>(src + jzp), zp1);
      e0 = src[je];
      e1 = src[je + 64u];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!typ) { yp[e] = 0.f; yp[4 + e] = 0.f; }
        if (!tym) { ym[e] = 0.f; ym[4 + e] = 0.f; }
        if (!tzm) zm0[e] = 0.f;
        if (!tzp) zp1[e] = 0.f;
      }
      if (!te) { e0 = 0.f; e1 = 0.f; }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) c[e] = yp[e] = ym[e] = 0.f;
#pragma unroll
      for (int e = 0; e < 4; ++e) zm0[e] = zp1[e] = 0.f;
    }
    // the partner lane holds the other half of each row
    const float o0 = __shfl_xor(h ? c[0] : c[3], 1, 64), o1 = __shfl_xor(h ? c[4] : c[7], 1, 64);
    float out[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = j & 3, L = j >> 2;
      const float edge = L ? e1 : e0, other = L ? o1 : o0;
      const float xpv = e < 3 ? c[j + 1] : (h ? edge : other);
      const float xmv = e > 0 ? c[j - 1] : (h ? other : edge);
      const float zpv = L ? zp1[e] : c[4 + e];
      const float zmv = L ? c[e] : zm0[e];
      out[j] = 0.f;
      if ((act >> j) & 1u) {
        const float sum = ZERO ? 0.f : xpv + xmv + yp[j] + ym[j] + zpv + zmv;
        const float pj = (sum - r[j]) / 6.f;
        out[j] = mixf(ZERO ? 0.f : c[j], pj, om);
      }
    }
    *reinterpret_cast<float4*>(dst + i0) = make_float4(out[0], out[1], out[2], out[3]);
    *reinterpret_cast<float4*>(dst + i1) = make_float4(out[4], out[5], out[6], out[7]);
  }
}

// Sweeps A -> B on brick +- R, then B -> A on brick +- (R - 1), ... down to the brick.
template <int R0, int R>
__device__ __forceinline__ void tile_sweeps(const uint64_t* nb, float* A, float* B, const float* rhs, float om) {
  tile_sweep<R0, R>(nb, A, rhs, B, om);
  __syncthreads();
  if constexpr (R > 0) tile_sweeps<R0, R - 1>(nb, B, A, rhs, om);
}
// NS weighted-Jacobi sweeps of a V-cycle level in one launch, temporally blocked
// in LDS: tiles of radius NS around the brick, sweep j on the brick +- (NS - j).
// The next brick's tile loads and neighbourhood words are issued before the
// current brick's sweeps, so their latency overlaps the LDS work. Every voxel sees
// the same arithmetic as in NS k_sweep launches (same neighbour order, inactive
// voxels 0), so results are bit-identical. ZERO: the first sweep starts from p = 0.
template <int NS, bool ZERO, bool MS>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(NS == 2 ? 8 : 4))) void k_sweepn(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                LvBits lb, const StepConst* __restrict__ C0, const float* __restrict__ rhs,
                                                const float* __restrict__ src, float* __restrict__ dst) {
  FX_SIM_PRIORITY();
  constexpr int R0 = NS, S = Tile<R0>::S;
  __shared__ float ta[Tile<R0>::N], tb[Tile<R0>::N], tr[Tile<R0>::N];
  __shared__ uint64_t nb[216];
  const int tid = int(threadIdx.x);
  const BrickRange br = brick_range(*count);
  TileRegs<R0> rr, rp;
  uint64_t nw = 0;
  uint32_t b = 0;
  GridCache<MS> gk;
  Grid gn = G;
  auto fetch = [&](uint32_t k) {
    b = list[k];
    gn = gk.at(G, b);
    const Grid g = gn;
    int bx, by, bz;
    brick_coords(g, b, bx, by, bz);
    const int ox = bx * 8 - R0, oy = by * 8 - R0, oz = bz * 8 - R0;
#pragma unroll
    for (int j = 0; j < TileRegs<R0>::IT; ++j) {
      const int q = tid + j * 512;
      const int x = ox + q % S, y = oy + (q / S) % S, z = oz + q / (S * S);
      const bool in = q < Tile<R0>::N && x >= 0 && y >= 0 && z >= 0 && x < int(g.W) && y < int(g.H) && z < int(g.D);
      const uint32_t id = in ? vidx(g, uint32_t(x), uint32_t(y), uint32_t(z)) : 0u;
      rr.v[j] = in ? rhs[id] : 0.f;
      rp.v[j] = (!ZERO && in) ? src[id] : 0.f;
    }
    nw = tid < 216 ? nbhd_word(g, lb, bx, by, bz, tid) : 0ull;
  };
  if (br.k < br.end) fetch(br.k);
  for (uint32_t k = br.k; k < br.end; k += br.step) {
    const uint32_t cb = b;
    const Grid g = gn;
    __syncthreads();  // the previous brick's tiles are no longer read
    if (tid < 216) nb[tid] = nw;
    __syncthreads();
    tile_store<R0>(nb, rr, tr);
    if (!ZERO) tile_store<R0>(nb, rp, ta); else tile_zero<R0>(ta);
    __syncthreads();
    if (k + br.step < br.end) fetch(k + br.step);
    const float om = C0->omega;  // the scenes of a batch share the V-cycle settings
    tile_sweeps<R0, R0 - 1>(nb, ta, tb, tr, om);
    const float* res = (NS & 1) ? tb : ta;
    int bx, by, bz;
    brick_coords(g, cb, bx, by, bz);
    const int x = bx * 8 + (tid & 7), y = by * 8 + ((tid >> 3) & 7), z = bz * 8 + (tid >> 6);
    if (x < int(g.W) && y < int(g.H) && z < int(g.D))
      dst[(cb << 9) | uint32_t(tid)] = res[(R0 + (tid & 7)) + S * ((R0 + ((tid >> 3) & 7)) + S * (R0 + (tid >> 6)))];
  }
}

// ------------------------------------------------ sweep compute tiles --
// Split-leg sweeps on compute tiles other than the 8^3 brick, the tile
// dimension of the V-cycle autotuner: NS (1 or 2) weighted-Jacobi sweeps per
// launch on a cube of E voxels per axis, staged in LDS with a halo of NS.
// - E = 4: the eight 4^3 octants of a brick, one wave each with its own tiles:
//   eight times the waves of a brick-per-wave sweep on levels with few bricks;
// - E = 16: a 2x2x2 group of bricks per workgroup, taken by the group's first
//   listed brick (lowest index): one halo for eight bricks.
// Every voxel's arithmetic and neighbour order are those of k_sweep (xp, xm, yp,
// ym, zp, zm; inactive voxels 0; ZERO starts from p = 0), so results are
// bit-identical to it and to every other plan.

// Activity of the voxel at coordinates (lx, ly, lz) in [0, 8K)^3 of a K^3 brick
// neighbourhood whose words nb hold brick (i, j, k) at ((i + K (j + K k)) * 8 + z).
template <int K>
__device__ __forceinline__ bool nb_on(const uint64_t* nb, int lx, int ly, int lz) {
  return (nb[((lx >> 3) + K * ((ly >> 3) + K * (lz >> 3))) * 8 + (lz & 7)] >> ((lx & 7) | ((ly & 7) << 3))) & 1ull;
}
// Word q of the K^3 brick neighbourhood whose first brick is (bx0, by0, bz0);
// bricks outside the level or unlisted give 0.
template <int K>
__device__ __forceinline__ uint64_t nbk_word(const Grid& g, const LvBits& lb, int bx0, int by0, int bz0, int q) {
  const int c = q >> 3, w = q & 7;
  const int nx = bx0 + c % K, ny = by0 + (c / K) % K, nz = bz0 + c / (K * K);
  uint64_t v = 0;
  if (q < K * K * K * 8 && nx >= 0 && ny >= 0 && nz >= 0 && nx < int(g.BX) && ny < int(g.BY) && nz < int(g.BZ)) {
    const uint32_t nbk = g.base + uint32_t(nx) + g.BX * (uint32_t(ny) + g.BY * uint32_t(nz));
    const int pos = lb.page ? lb.page[nbk] : int(nbk);
    if (pos >= 0) v = lb.bits[size_t(pos) * 8 + w];
  }
  return v;
}
// One sweep A -> B over the cube [lo, lo + side)^3 of an S^3 tile whose voxel
// (0, 0, 0) sits at neighbourhood coordinates (ox, oy, oz); `lane`/`nt` split the
// cube over the calling threads.
template <int K, int S>
__device__ __forceinline__ void tile_sweep_k(const uint64_t* nb, int ox, int oy, int oz, const float* A, const float* R, float* B,
                                             int lo, int side, int lane, int nt, float om) {
  const int n = side * side * side;
  for (int q = lane; q < n; q += nt) {
    const int tx = lo + q % side, ty = lo + (q / side) % side, tz = lo + q / (side * side);
    const int i = tx + S * (ty + S * tz);
    if (!nb_on<K>(nb, ox + tx, oy + ty, oz + tz)) { B[i] = 0.f; continue; }
    const float sum = A[i + 1] + A[i - 1] + A[i + S] + A[i - S] + A[i + S * S] + A[i - S * S];
    const float pj = (sum - R[i]) / 6.f;
    B[i] = mixf(A[i], pj, om);
  }
}
template <int NS, bool ZERO, bool MS>
__global__ __launch_bounds__(512) void k_sweep_t4(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                  LvBits lb, const StepConst* __restrict__ C0, const float* __restrict__ rhs,
                                                  const float* __restrict__ src, float* __restrict__ dst) {
  FX_SIM_PRIORITY();
  constexpr int S = 4 + 2 * NS, N = S * S * S, IT = (N + 63) / 64;
  __shared__ float tr[8][N], ta[8][N], tb[8][NS > 1 ? N : 1];
  __shared__ uint64_t nb[216];
  const int lane = int(threadIdx.x & 63u), w = int(threadIdx.x >> 6);
  // the wave's octant: origin (sx, sy, sz) in the brick; its tile starts NS before
  const int sx = (w & 1) * 4, sy = ((w >> 1) & 1) * 4, sz = (w >> 2) * 4;
  const int off_x = sx - NS + 8, off_y = sy - NS + 8, off_z = sz - NS + 8;  // in the 3^3 neighbourhood
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid g = gk.at(G, b);
    int bx, by, bz;
    brick_coords(g, b, bx, by, bz);
    const int ox = bx * 8 + sx - NS, oy = by * 8 + sy - NS, oz = bz * 8 + sz - NS;
    float vr[IT], vp[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int q = lane + j * 64;
      const int x = ox + q % S, y = oy + (q / S) % S, z = oz + q / (S * S);
      const bool in = q < N && x >= 0 && y >= 0 && z >= 0 && x < int(g.W) && y < int(g.H) && z < int(g.D);
      const uint32_t id = in ? vidx(g, uint32_t(x), uint32_t(y), uint32_t(z)) : 0u;
      vr[j] = in ? rhs[id] : 0.f;
      vp[j] = (!ZERO && in) ? src[id] : 0.f;
    }
    const uint64_t nw = nbhd_fetch(g, lb, bx, by, bz);
    __syncthreads();  // the previous brick's words and tiles are no longer read
    nbhd_store(nw, nb);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int q = lane + j * 64;
      if (q < N) {
        const bool on = nb_on<3>(nb, off_x + q % S, off_y + (q / S) % S, off_z + q / (S * S));
        tr[w][q] = on ? vr[j] : 0.f;
        ta[w][q] = on ? vp[j] : 0.f;
      }
    }
    __syncthreads();
    const float om = C0->omega;  // the scenes of a batch share the V-cycle settings
    const float* A = ta[w];
    if constexpr (NS == 2) {
      tile_sweep_k<3, S>(nb, off_x, off_y, off_z, ta[w], tr[w], tb[w], 1, 6, lane, 64, om);  // the octant +- 1
      __syncthreads();
      A = tb[w];
    }
    // the octant itself, one voxel per lane
    const int tx = NS + (lane & 3), ty = NS + ((lane >> 2) & 3), tz = NS + (lane >> 4);
    const int i = tx + S * (ty + S * tz);
    float out = 0.f;
    if (nb_on<3>(nb, off_x + tx, off_y + ty, off_z + tz)) {
      const float sum = A[i + 1] + A[i - 1] + A[i + S] + A[i - S] + A[i + S * S] + A[i - S * S];
      const float pj = (sum - tr[w][i]) / 6.f;
      out = mixf(A[i], pj, om);
    }
    const int lx = sx + (lane & 3), ly = sy + ((lane >> 2) & 3), lz = sz + (lane >> 4);
    if (bx * 8 + lx < int(g.W) && by * 8 + ly < int(g.H) && bz * 8 + lz < int(g.D))
      dst[(b << 9) | uint32_t((lz << 6) | (ly << 3) | lx)] = out;
  }
}
template <int NS, bool ZERO, bool MS>
__global__ __launch_bounds__(512) void k_sweep_t16(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G,
                                                   LvBits lb, const StepConst* __restrict__ C0, const float* __restrict__ rhs,
                                                   const float* __restrict__ src, float* __restrict__ dst) {
  FX_SIM_PRIORITY();
  constexpr int S = 16 + 2 * NS, N = S * S * S, IT = (N + 511) / 512;
  __shared__ float tr[N], ta[N], tb[NS > 1 ? N : 1];
  __shared__ uint64_t nb[512];
  const int tid = int(threadIdx.x);
  constexpr int off = 8 - NS;  // tile voxel 0 in the 4^3 neighbourhood that starts one brick before the group
  GridCache<MS> gk;
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Grid g = gk.at(G, b);
    int bx, by, bz;
    brick_coords(g, b, bx, by, bz);
    const int ax = bx & ~1, ay = by & ~1, az = bz & ~1;
    // the group's listed bricks (bit m: brick (ax + (m & 1), ay + (m >> 1 & 1), az + (m >> 2)));
    // the lowest-index one takes the group
    uint32_t listed = 0;
    bool lead = true;
    for (int m = 0; m < 8; ++m) {
      const int nx = ax + (m & 1), ny = ay + ((m >> 1) & 1), nz = az + (m >> 2);
      if (nx >= int(g.BX) || ny >= int(g.BY) || nz >= int(g.BZ)) continue;
      const uint32_t nbk = g.base + uint32_t(nx) + g.BX * (uint32_t(ny) + g.BY * uint32_t(nz));
      if (lb.page && lb.page[nbk] < 0) continue;
      listed |= 1u << m;
      if (nbk < b) lead = false;
    }
    if (!lead) continue;
    const int ox = ax * 8 - NS, oy = ay * 8 - NS, oz = az * 8 - NS;
    float vr[IT], vp[IT];
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int q = tid + j * 512;
      const int x = ox + q % S, y = oy + (q / S) % S, z = oz + q / (S * S);
      const bool in = q < N && x >= 0 && y >= 0 && z >= 0 && x < int(g.W) && y < int(g.H) && z < int(g.D);
      const uint32_t id = in ? vidx(g, uint32_t(x), uint32_t(y), uint32_t(z)) : 0u;
      vr[j] = in ? rhs[id] : 0.f;
      vp[j] = (!ZERO && in) ? src[id] : 0.f;
    }
    const uint64_t nw = nbk_word<4>(g, lb, ax - 1, ay - 1, az - 1, tid);
    __syncthreads();  // the previous group's words and tiles are no longer read
    nb[tid] = nw;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IT; ++j) {
      const int q = tid + j * 512;
      if (q < N) {
        const bool on = nb_on<4>(nb, off + q % S, off + (q / S) % S, off + q / (S * S));
        tr[q] = on ? vr[j] : 0.f;
        ta[q] = on ? vp[j] : 0.f;
      }
    }
    __syncthreads();
    const float om = C0->omega;  // the scenes of a batch share the V-cycle settings
    const float* A = ta;
    if constexpr (NS == 2) {
      tile_sweep_k<4, S>(nb, off, off, off, ta, tr, tb, 1, 18, tid, 512, om);  // the group +- 1
      __syncthreads();
      A = tb;
    }
    // the group's listed bricks, 8 voxels per thread
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int q = tid + j * 512;
      const int cx = q & 15, cy = (q >> 4) & 15, cz = q >> 8;
      const int m = (cx >> 3) | ((cy >> 3) << 1) | ((cz >> 3) << 2);
      if (!((listed >> m) & 1u)) continue;
      const int tx = NS + cx, ty = NS + cy, tz = NS + cz;
      const int i = tx + S * (ty + S * tz);
      float out = 0.f;
      if (nb_on<4>(nb, off + tx, off + ty, off + tz)) {
        const float sum = A[i + 1] + A[i - 1] + A[i + S] + A[i - S] + A[i + S * S] + A[i - S * S];
        const float pj = (sum - tr[i]) / 6.f;
        out = mixf(A[i], pj, om);
      }
      if (ax * 8 + cx < int(g.W) && ay * 8 + cy < int(g.H) && az * 8 + cz < int(g.D)) {
        const uint32_t gb = g.base + uint32_t(ax + (cx >> 3)) + g.BX * (uint32_t(ay + (cy >> 3)) + g.BY * uint32_t(az + (cz >> 3)));
        dst[(gb << 9) | uint32_t(((cz & 7) << 6) | ((cy & 7) << 3) | (cx & 7))] = out;
      }
    }
  }
}

// SH/project-velocity.wgsl.ts:21-43, on float4 brick rows.
template <bool MS>
__global__ __launch_bounds__(512) void k_projection(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid G, Vec3 us,
                                                    const float* __restrict__ p, Vec3 u) {
  GridCache<MS> gk;
  FX_ROW_LOOP(list, count) {
    const uint32_t b = wave_uniform(list[fx_k]);
    const Grid g = gk.at(G, b);
    const RowCtx R = row_ctx(g, b, fx_t);
    float c[4], xp[4], xm[4], yp[4], ym[4], zp[4], zm[4], ux[4], uy[4], uz[4];
    row_x<StF32>(g, R, p, c, xp, xm);
    row_y<StF32>(R, p, yp, ym);
    row_z<StF32>(R, p, zp, zm);
    unpack4(StF32::ld4(us.x, R.i0), ux); unpack4(StF32::ld4(us.y, R.i0), uy); unpack4(StF32::ld4(us.z, R.i0), uz);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (on_boundary(g, R.x + uint32_t(e), R.y, R.z)) { ux[e] = uy[e] = uz[e] = 0.f; continue; }
      const float gx = (xp[e] - xm[e]) * 0.5f, gy = (yp[e] - ym[e]) * 0.5f, gz = (zp[e] - zm[e]) * 0.5f;
      ux[e] = ux[e] - gx; uy[e] = uy[e] - gy; uz[e] = uz[e] - gz;
    }
    row_store<StF32>(R, u.x, ux);
    row_store<StF32>(R, u.y, uy);
    row_store<StF32>(R, u.z, uz);
  }
}
// SH/compute-debug-fields.wgsl.ts:18-25
__global__ __launch_bounds__(512) void k_debug(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g, Vec3 u,
                                               float* umag) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    umag[v.i] = len3(f3(u.x[v.i], u.y[v.i], u.z[v.i]));
  }
}

// ------------------------------------------------------------ raymarch --
// src/engine/render/passes/createVolumeRaymarchPass.ts:361-724, one thread per
// pixel, row 0 = top; discarded pixels are written as 0 (cleared target).
__device__ __forceinline__ F3 fire_palette(float t) {
  F3 ember = f3(0.18f, 0.018f, 0.f), red = f3(1.f, 0.075f, 0.012f), orange = f3(1.f, 0.42f, 0.035f), white = f3(1.f, 0.86f, 0.56f);
  float rm = smoothk(0.01f, 0.11f, t), om = smoothk(0.09f, 0.38f, t), wm = smoothk(0.58f, 0.92f, t);
  return mix3(mix3(mix3(ember, red, rm), orange, om), white, wm);
}
__device__ __forceinline__ F3 smoke_albedo(float den, float tem, float hot) {
  float warm = clampf(tem * 0.38f, 0.f, 1.f);
  float soot = smoothk(0.16f, 0.9f, den) * (1.f - smoothk(0.2f, 0.74f, hot));
  return mix3(mix3(f3(0.26f, 0.275f, 0.295f), f3(0.04f, 0.041f, 0.045f), soot * 0.86f), f3(0.36f, 0.27f, 0.18f), warm * (0.35f + hot * 0.45f));
}
// Shading transcendentals on the hardware base-2 units (v_exp_f32 / v_log_f32,
// about 1 ulp): the image is FP16 and its parity tolerance is 2e-3, so the
// correctly rounded library forms buy nothing here and cost ~100 instructions.
__device__ __forceinline__ float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
// x^y for x >= 0, y > 0 (0^y = 0).
__device__ __forceinline__ float pow_hw(float x, float y) {
  return x > 0.f ? __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)) : 0.f;
}
__device__ __forceinline__ float phase_hg(float c, float gg) {
  float g2 = gg * gg;
  float base = fmaxf(1.f + g2 - 2.f * gg * c, 0.001f);
  float denom = base * sqrtf(base);  // base^1.5
  return (1.f - g2) / (12.566370614359172f * denom);
}

// Brick flags packed 32 per word (the raymarch keeps them in LDS).
// One ray (createVolumeRaymarchPass.ts fsMain) for pixel (px, py), reading the
// render snapshot RB (density, temperature, fuel, reaction per voxel) and the
// packed brick flags (sbits in LDS, else gbits in memory). g is one scene's grid;
// its flags start at bit (g.base & 31) of gbits[0] (the launcher offsets gbits by
// g.base / 32 words).
__device__ __forceinline__ void march_pixel(const Grid& g, const RayConst& R, const float4* __restrict__ RB,
                                            const uint32_t* __restrict__ gbits, const uint32_t* sbits, uint32_t px, uint32_t py,
                                            uint16_t* out) {
  // isActiveSample (RP:378-381)
  auto active_brick = [&](uint32_t x, uint32_t y, uint32_t z) {
    uint32_t bx = min(x >> 3, g.BX - 1), by = min(y >> 3, g.BY - 1), bz = min(z >> 3, g.BZ - 1);
    uint32_t b = (g.base & 31u) + bx + g.BX * (by + g.BY * bz);
    return (((sbits ? sbits[b >> 5] : gbits[b >> 5]) >> (b & 31u)) & 1u) != 0u;
  };
  if (px >= R.width || py >= R.height) return;
  __half* o = reinterpret_cast<__half*>(out) + (size_t(py) * R.width + px) * 4;
  const F3 dims = f3(float(g.W) - 1.f, float(g.H) - 1.f, float(g.D) - 1.f);
  const F3 camPos = f3(R.cam_pos[0], R.cam_pos[1], R.cam_pos[2]);
  float uvx = (float(px) + 0.5f) / float(R.width), uvy = 1.f - (float(py) + 0.5f) / float(R.height);
  float ndcx = uvx * 2.f - 1.f, ndcy = uvy * 2.f - 1.f;
  F3 rayDir = norm3(f3(R.fwd[0], R.fwd[1], R.fwd[2]) + f3(R.right[0], R.right[1], R.right[2]) * ndcx * R.aspect * R.tan_half +
                    f3(R.up[0], R.up[1], R.up[2]) * ndcy * R.tan_half);
  F3 bmin = f3(R.box_min[0], R.box_min[1], R.box_min[2]), bmax = f3(R.box_max[0], R.box_max[1], R.box_max[2]);
  F3 inv = f3(1.f / rayDir.x, 1.f / rayDir.y, 1.f / rayDir.z);
  F3 t0 = mul3(bmin - camPos, inv), t1 = mul3(bmax - camPos, inv);
  float bx = fmaxf(fminf(t0.x, t1.x), fmaxf(fminf(t0.y, t1.y), fminf(t0.z, t1.z)));
  float by = fminf(fmaxf(t0.x, t1.x), fminf(fmaxf(t0.y, t1.y), fmaxf(t0.z, t1.z)));
  const __half zero = __float2half(0.f);
  if (bx > by) { o[0] = zero; o[1] = zero; o[2] = zero; o[3] = zero; return; }
  bx = fmaxf(bx, 0.f);
  F3 invExt = f3(1.f / (R.half_ext[0] * 2.f), 1.f / (R.half_ext[1] * 2.f), 1.f / (R.half_ext[2] * 2.f));
  F3 vrd = mul3(mul3(rayDir, invExt), dims);
  const uint32_t stepCount = R.step_count;
  float voxDelta = len3(vrd) * (by - bx) / fmaxf(float(stepCount), 1.f);
  float worldDelta = (by - bx) / fmaxf(float(stepCount), 1.f);
  uint32_t nl = uint32_t(clampf(R.light_count, 0.f, 3.f));
  auto lightDir = [&](int k, F3 wp) {
    if (R.light_kind[k] > 0.5f) return norm3(f3(R.light_pos[k][0], R.light_pos[k][1], R.light_pos[k][2]) - wp);
    F3 raw = f3(R.light_dir[k][0], R.light_dir[k][1], R.light_dir[k][2]);
    return norm3(dot3(raw, raw) > 0.0001f ? raw : f3(-0.34f, 0.88f, 0.31f));
  };
  auto lightAtt = [&](int k, F3 wp) {
    if (R.light_kind[k] <= 0.5f) return 1.f;
    float dist = len3(f3(R.light_pos[k][0], R.light_pos[k][1], R.light_pos[k][2]) - wp);
    return 1.f / (1.f + dist * 0.09f + dist * dist * 0.02f);
  };
  F3 primary = nl > 0 ? lightDir(0, camPos) : f3(0.f, 1.f, 0.f);
  // light 0 when directional: direction, power, phase terms and rim factor are
  // fixed per ray (other lights are evaluated per sample)
  const bool dir0 = nl > 0 && R.light_kind[0] <= 0.5f;
  F3 dirL0 = f3(0.f, 1.f, 0.f);
  float powL0 = 0.f, fsL0 = 0.f, bsL0 = 0.f, rimL0 = 0.f;
  if (dir0) {
    dirL0 = lightDir(0, camPos);
    float raw = fmaxf(R.light_intensity[0], 0.f);
    powL0 = raw / (1.f + raw * 0.18f);
    float lvc = clampf(dot3(dirL0, -rayDir), -1.f, 1.f);
    fsL0 = phase_hg(lvc, R.scat_f) * 5.2f * powL0;
    bsL0 = phase_hg(lvc, R.scat_b) * 1.55f * powL0;
    rimL0 = pow_hw(clampf(dot3(dirL0, -rayDir) * 0.5f + 0.5f, 0.f, 1.f), 1.2f);
  }
  F3 lightStep = mul3(primary, dims) * 0.036f;
  F3 vp = mul3(mul3(camPos + rayDir * bx - bmin, invExt), dims);
  float travelled = 0.f;
  float rayLen = len3(vrd) * (by - bx);
  F3 acc = f3(0, 0, 0);
  float alpha = 0.f;
  F3 nrd = norm3(vrd);
  const float vrdLen = len3(vrd);
  const uint32_t mxu = uint32_t(dims.x), myu = uint32_t(dims.y), mzu = uint32_t(dims.z);
  // per-ray reciprocals: the samples multiply where the shader divides (within
  // an ulp or two; the image is FP16)
  const F3 invDims = f3(1.f / dims.x, 1.f / dims.y, 1.f / dims.z);
  const F3 invVrd = f3(1.f / vrd.x, 1.f / vrd.y, 1.f / vrd.z);
  for (uint32_t k = 0; k < stepCount && travelled <= rayLen; ++k) {
    if (alpha > 0.985f) break;
    float vstep = voxDelta;
    F3 sp = vp;
    const float cxf = clampf(sp.x, 0.f, dims.x), cyf = clampf(sp.y, 0.f, dims.y), czf = clampf(sp.z, 0.f, dims.z);
    const uint32_t ax = uint32_t(floorf(cxf)), ay = uint32_t(floorf(cyf)), az = uint32_t(floorf(czf));
    if (active_brick(ax, ay, az)) {
      Tri t = tri(g, sp);
      float4 cr[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) cr[k] = RB[t.i[k]];
      // the tri_eval order per channel
      auto tri4 = [&](float c0, float c1, float c2, float c3, float c4, float c5, float c6, float c7) {
        const float x00 = mixf(c0, c1, t.fx), x10 = mixf(c2, c3, t.fx), x01 = mixf(c4, c5, t.fx), x11 = mixf(c6, c7, t.fx);
        return mixf(mixf(x00, x10, t.fy), mixf(x01, x11, t.fy), t.fz);
      };
      float den = fmaxf(tri4(cr[0].x, cr[1].x, cr[2].x, cr[3].x, cr[4].x, cr[5].x, cr[6].x, cr[7].x) - 0.004f, 0.f);
      float tem = tri4(cr[0].y, cr[1].y, cr[2].y, cr[3].y, cr[4].y, cr[5].y, cr[6].y, cr[7].y);
      float fu = tri4(cr[0].z, cr[1].z, cr[2].z, cr[3].z, cr[4].z, cr[5].z, cr[6].z, cr[7].z);
      float rea = tri4(cr[0].w, cr[1].w, cr[2].w, cr[3].w, cr[4].w, cr[5].w, cr[6].w, cr[7].w);
      float hot = clampf(tem * 1.45f + fu * 0.22f + rea * 0.38f, 0.f, 1.f);
      if (den > 0.001f || hot > 0.018f) {
        if (den < 0.012f && hot < 0.04f) {
          vstep = voxDelta * 1.8f;
        } else {
          uint32_t cx = min(uint32_t(fmaxf(rintf(sp.x), 0.f)), mxu), cy = min(uint32_t(fmaxf(rintf(sp.y), 0.f)), myu),
                   cz = min(uint32_t(fmaxf(rintf(sp.z), 0.f)), mzu);
          F3 dg = f3(RB[vidx(g, min(cx + 1, mxu), cy, cz)].x - RB[vidx(g, cx == 0 ? 0 : cx - 1, cy, cz)].x,
                     RB[vidx(g, cx, min(cy + 1, myu), cz)].x - RB[vidx(g, cx, cy == 0 ? 0 : cy - 1, cz)].x,
                     RB[vidx(g, cx, cy, min(cz + 1, mzu))].x - RB[vidx(g, cx, cy, cz == 0 ? 0 : cz - 1)].x);
          float gl = fmaxf(len3(dg), 0.0001f);
          F3 normal = gl > 0.00011f ? (-dg) * (1.f / gl) : f3(0.f, 1.f, 0.f);
          float detail = clampf(gl * 5.2f + hot * 0.28f + rea * 0.18f, 0.f, 1.f);
          float dd = fmaxf(pow_hw(den, 0.96f) * mixf(0.82f, 1.42f, detail), 0.f);
          F3 uvw = mul3(sp, invDims);
          F3 wp = f3(mixf(bmin.x, bmax.x, uvw.x), mixf(bmin.y, bmax.y, uvw.y), mixf(bmin.z, bmax.z, uvw.z));
          float topFade = 1.f - smoothk(0.9f, 0.995f, uvw.y);
          F3 total = f3(0, 0, 0);
          float primT = 1.f;
          if (nl > 0) {
            auto nearest = [&](F3 p) {
              F3 c = f3(clampf(p.x, 0.f, dims.x), clampf(p.y, 0.f, dims.y), clampf(p.z, 0.f, dims.z));
              return RB[vidx(g, min(uint32_t(floorf(c.x)), mxu), min(uint32_t(floorf(c.y)), myu), min(uint32_t(floorf(c.z)), mzu))].x;
            };
            float d1 = fmaxf(nearest(sp + lightStep * 0.9f) - 0.004f, 0.f);
            float d2 = fmaxf(nearest(sp + lightStep * 2.f) - 0.004f, 0.f);
            float d3 = fmaxf(nearest(sp + lightStep * 3.6f) - 0.004f, 0.f);
            float od = dd * 0.34f + d1 * 1.15f + d2 * 0.72f + d3 * 0.4f - hot * 0.06f;
            primT = exp_hw(-fmaxf(od, 0.f) * 2.15f);
          }
          const float rimView = pow_hw(1.f - fabsf(dot3(normal, -rayDir)), 2.4f);
#pragma unroll
          for (uint32_t li = 0; li < 3; ++li) {
            if (li >= nl) continue;
            F3 dl;
            float power, fs, bs, rimLight;
            if (li == 0 && dir0) {  // per-ray constants
              dl = dirL0; power = powL0; fs = fsL0; bs = bsL0; rimLight = rimL0;
            } else {
              dl = lightDir(int(li), wp);
              float raw = fmaxf(R.light_intensity[li], 0.f) * lightAtt(int(li), wp);
              power = raw / (1.f + raw * 0.18f);
              float lvc = clampf(dot3(dl, -rayDir), -1.f, 1.f);
              fs = phase_hg(lvc, R.scat_f) * 5.2f * power;
              bs = phase_hg(lvc, R.scat_b) * 1.55f * power;
              rimLight = pow_hw(clampf(dot3(dl, -rayDir) * 0.5f + 0.5f, 0.f, 1.f), 1.2f);
            }
            F3 color = f3(R.light_color[li][0], R.light_color[li][1], R.light_color[li][2]);
            float nf = clampf(dot3(normal, dl) * 0.5f + 0.5f, 0.f, 1.f);
            float surface = pow_hw(nf, 1.35f) * mixf(1.f, 0.56f, smoothk(0.16f, 0.9f, dd));
            float rim = rimView * rimLight;
            float directional = (surface * 0.42f + rim * 0.22f) * power;
            float scatter = (fs + bs) * mixf(0.36f, 0.74f, hot);
            float trans = li > 0 ? 1.f : primT;
            total = total + color * trans * (directional + scatter);
          }
          F3 tl = f3(1.f - exp_hw(-total.x * 0.72f), 1.f - exp_hw(-total.y * 0.72f), 1.f - exp_hw(-total.z * 0.72f));
          F3 smoke = smoke_albedo(dd, tem, hot);
          float absorb = mixf(0.92f, 0.34f, smoothk(0.26f, 1.05f, dd));
          float heat = clampf(tem * 1.25f + fu * 0.18f + rea * 0.34f, 0.f, 1.f);
          float flame = clampf(tem * 0.9f + fu * 0.12f + rea * 0.18f, 0.f, 1.f);
          F3 heatScatter = fire_palette(flame) * heat * heat * smoothk(0.012f, 0.38f, dd) * 0.18f;
          F3 lit = mul3(smoke, f3(0.045f, 0.048f, 0.052f)) + mul3(smoke, tl) * absorb + tl * smoothk(0.018f, 0.22f, dd) * 0.18f + heatScatter;
          float flameHeat = clampf(tem * 0.9f + fu * 0.14f + rea * 0.16f, 0.f, 1.f);
          F3 emissive = fire_palette(flameHeat) * hot * (0.48f + fu * 0.16f + rea * 0.24f);
          float contribution = clampf(dd * 1.35f + hot * 1.25f, 0.f, 1.f);
          float stepDistance = worldDelta * mixf(1.02f, 0.76f, contribution);
          vstep = voxDelta * mixf(1.02f, 0.76f, contribution);
          F3 comp = lit + emissive;
          float opacity = (1.f - exp_hw(-(dd * 4.75f + hot * 0.42f) * stepDistance)) * topFade;
          if (R.mode > 0.5f && R.mode < 1.5f) {
            comp = mix3(f3(0.01f, 0.01f, 0.012f), f3(0.88f, 0.9f, 0.94f), clampf(dd * 1.4f, 0.f, 1.f));
            opacity = 1.f - exp_hw(-dd * stepDistance * 8.6f);
          } else if (R.mode >= 1.5f) {
            comp = mix3(f3(0.04f, 0.012f, 0.f), f3(1.f, 0.72f, 0.22f), clampf(fu * 1.3f, 0.f, 1.f));
            opacity = 1.f - exp_hw(-fu * stepDistance * 7.f);
          }
          acc = acc + comp * ((1.f - alpha) * opacity);
          alpha = alpha + (1.f - alpha) * opacity;
        }
      } else {
        vstep = voxDelta * 1.35f;
      }
    } else {
      // voxelDistToNextBrick, RP:383-396
      F3 bc = f3(floorf(sp.x / 8.f), floorf(sp.y / 8.f), floorf(sp.z / 8.f));
      F3 bnd = f3(vrd.x > 0.f ? (bc.x + 1.f) * 8.f : bc.x * 8.f, vrd.y > 0.f ? (bc.y + 1.f) * 8.f : bc.y * 8.f,
                  vrd.z > 0.f ? (bc.z + 1.f) * 8.f : bc.z * 8.f);
      float tx = fabsf(vrd.x) > 0.00001f ? (bnd.x - sp.x) * invVrd.x : 1e9f;
      float ty = fabsf(vrd.y) > 0.00001f ? (bnd.y - sp.y) * invVrd.y : 1e9f;
      float tz = fabsf(vrd.z) > 0.00001f ? (bnd.z - sp.z) * invVrd.z : 1e9f;
      vstep = fmaxf(fminf(tx, fminf(ty, tz)) * vrdLen + 0.5f, voxDelta);
    }
    vp = vp + nrd * vstep;
    travelled += vstep;
  }
  if (alpha < 0.01f) { o[0] = zero; o[1] = zero; o[2] = zero; o[3] = zero; return; }
  o[0] = __float2half(acc.x); o[1] = __float2half(acc.y); o[2] = __float2half(acc.z); o[3] = __float2half(alpha);
}


// Render snapshot: RB = (density, temperature, fuel, reaction) of the listed
// bricks (FP32 whatever the storage), 0 on the bricks in the retire list. The
// raymarch reads only the snapshot and the packed flags, so the next step may
// rewrite the simulation fields while a render is still running.
template <class ST>
__global__ __launch_bounds__(512) void k_pack_render(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, SS<ST> S,
                                                     const uint32_t* __restrict__ rlist, const uint32_t* __restrict__ rcount,
                                                     float4* __restrict__ RB, const uint32_t* __restrict__ flags, uint32_t nb,
                                                     uint32_t* __restrict__ bits) {
  // the brick flags packed 32 per word for the raymarch's empty-brick test: each
  // wave packs 64 flags (two words) with one coalesced load and one ballot
  {
    const uint32_t lane = threadIdx.x & 63u, wpb = blockDim.x >> 6, nw = (nb + 31u) / 32u;
    for (uint32_t v = blockIdx.x * wpb + (threadIdx.x >> 6); v * 64u < nb; v += gridDim.x * wpb) {
      const uint32_t f = v * 64u + lane;
      const uint64_t m = __ballot(f < nb && flags[f] != 0u);
      if (lane < 2u && v * 2u + lane < nw) bits[v * 2u + lane] = uint32_t(m >> (32u * lane));
    }
  }
  FX_BRICK_LOOP(list, count) {
    const uint32_t i = (list[fx_k] << 9) | threadIdx.x;
    RB[i] = make_float4(ST::ld(S.d, i), ST::ld(S.t, i), ST::ld(S.f, i), S.r[i]);
  }
  if (!rlist) return;
  const uint32_t nr = *rcount;
  for (uint32_t k = blockIdx.x; k < nr; k += gridDim.x) RB[(rlist[k] << 9) | threadIdx.x] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Work-queue raymarch: a bounded number of resident workgroups (so that the
// simulation's own kernels keep CUs while both run) take 16x16-pixel tiles from
// `queue` (zeroed before launch) until none are left; each wave64 marches a
// compact 8x8 pixel tile (coherent rays), 2x2 of them per workgroup tile.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_raymarch(Grid g, RayConst R, const float4* __restrict__ RB, const uint32_t* __restrict__ gbits,
                                                  uint32_t nwords, int use_lds, uint16_t* out, uint32_t* __restrict__ queue) {
  extern __shared__ uint32_t sbits[];
  __shared__ uint32_t tile;
  if (use_lds)
    for (uint32_t q = threadIdx.x; q < nwords; q += 256) sbits[q] = gbits[q];
  const uint32_t tiles_x = (R.width + 15) / 16, ntiles = tiles_x * ((R.height + 15) / 16);
  for (;;) {
    __syncthreads();  // flags staged; the previous tile index has been read
    if (threadIdx.x == 0) tile = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t tl = tile;
    if (tl >= ntiles) break;
    const uint32_t px = (tl % tiles_x) * 16 + ((threadIdx.x >> 6) & 1u) * 8 + (threadIdx.x & 7u),
                   py = (tl / tiles_x) * 16 + (threadIdx.x >> 7) * 8 + ((threadIdx.x >> 3) & 7u);
    march_pixel(g, R, RB, gbits, use_lds ? sbits : nullptr, px, py, out);
  }
}

// --------------------------------------------------------------- bloom --
// src/engine/render/passes/createBloomPass.ts: threshold into a half-resolution
// image (:119), 9-tap separable blur (:121-125, blurShader :237-262), composite
// with the source into the 8-bit canvas (:128, compositeShader :276-293), or the
// plain blit when bloom is off (:114-116, blitShader :264-274). Each render pass
// draws a full-screen triangle: the fragment of target pixel (i, j) reads its
// sources at uv = ((i + 0.5) / W, (j + 0.5) / H), row 0 at the top, through a
// linear, clamp-to-edge sampler. Intermediate targets are rgba16float.
struct BloomConst {
  float threshold, strength, radius;
  uint32_t w, h, hw, hh;
};
__device__ __forceinline__ float4 ld_h4(const __half* img, uint32_t i) {
  const __half* p = img + size_t(i) * 4;
  return make_float4(__half2float(p[0]), __half2float(p[1]), __half2float(p[2]), __half2float(p[3]));
}
__device__ __forceinline__ void st_h4(__half* img, uint32_t i, float4 v) {
  __half* p = img + size_t(i) * 4;
  p[0] = __float2half(v.x); p[1] = __float2half(v.y); p[2] = __float2half(v.z); p[3] = __float2half(v.w);
}
// Linear, clamp-to-edge sample of a w x h rgba16f image at uv.
__device__ __forceinline__ float4 sample_lin(const __half* img, uint32_t w, uint32_t h, float u, float v) {
  const float tx = u * float(w) - 0.5f, ty = v * float(h) - 0.5f;
  const float fx0 = floorf(tx), fy0 = floorf(ty);
  const float ax = tx - fx0, ay = ty - fy0;
  const int x0 = int(fx0), y0 = int(fy0);
  const int xa = min(max(x0, 0), int(w) - 1), xb = min(max(x0 + 1, 0), int(w) - 1);
  const int ya = min(max(y0, 0), int(h) - 1), yb = min(max(y0 + 1, 0), int(h) - 1);
  const float4 c00 = ld_h4(img, uint32_t(ya) * w + uint32_t(xa)), c10 = ld_h4(img, uint32_t(ya) * w + uint32_t(xb));
  const float4 c01 = ld_h4(img, uint32_t(yb) * w + uint32_t(xa)), c11 = ld_h4(img, uint32_t(yb) * w + uint32_t(xb));
  auto lerp2 = [&](float a, float b, float c, float d) { return mixf(mixf(a, b, ax), mixf(c, d, ax), ay); };
  return make_float4(lerp2(c00.x, c10.x, c01.x, c11.x), lerp2(c00.y, c10.y, c01.y, c11.y), lerp2(c00.z, c10.z, c01.z, c11.z),
                     lerp2(c00.w, c10.w, c01.w, c11.w));
}
// WebGPU float -> unorm8: clamp to [0, 1], round to nearest.
__device__ __forceinline__ uint32_t unorm8(float v) { return uint32_t(rintf(clampf(v, 0.f, 1.f) * 255.f)); }
__device__ __forceinline__ uint32_t pack_rgba8(float r, float g, float b, float a) {
  return unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (unorm8(a) << 24);
}
__global__ __launch_bounds__(256) void k_bloom_threshold(BloomConst P, const __half* __restrict__ src, __half* __restrict__ bright) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= P.hw * P.hh) return;
  const uint32_t x = i % P.hw, y = i / P.hw;
  const float4 c = sample_lin(src, P.w, P.h, (float(x) + 0.5f) / float(P.hw), (float(y) + 0.5f) / float(P.hh));
  const float lum = c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
  const float knee = P.threshold * 0.5f;
  const float rq = clampf(lum - P.threshold + knee, 0.f, 2.f * knee);
  const float wgt = (rq * rq) / (4.f * knee * knee + 0.00001f);
  const float mask = fmaxf(wgt, stepf(P.threshold, lum));
  st_h4(bright, i, make_float4(c.x * mask * c.w, c.y * mask * c.w, c.z * mask * c.w, c.w * mask));
}
__global__ __launch_bounds__(256) void k_bloom_blur(BloomConst P, int horizontal, const __half* __restrict__ src, __half* __restrict__ dst) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= P.hw * P.hh) return;
  const uint32_t x = i % P.hw, y = i / P.hw;
  const float u = (float(x) + 0.5f) / float(P.hw), v = (float(y) + 0.5f) / float(P.hh);
  const float scale = 1.f + P.radius * 2.f;
  const float su = horizontal ? (1.f / float(P.hw)) * scale : 0.f, sv = horizontal ? 0.f : (1.f / float(P.hh)) * scale;
  const float wt[5] = {0.2270270270f, 0.1945945946f, 0.1216216216f, 0.0540540541f, 0.0162162162f};
  float4 col = sample_lin(src, P.hw, P.hh, u, v);
  col = make_float4(col.x * wt[0], col.y * wt[0], col.z * wt[0], col.w * wt[0]);
#pragma unroll
  for (int k = 1; k <= 4; ++k) {
    const float4 a = sample_lin(src, P.hw, P.hh, u + su * float(k), v + sv * float(k));
    col = make_float4(col.x + a.x * wt[k], col.y + a.y * wt[k], col.z + a.z * wt[k], col.w + a.w * wt[k]);
    const float4 b = sample_lin(src, P.hw, P.hh, u - su * float(k), v - sv * float(k));
    col = make_float4(col.x + b.x * wt[k], col.y + b.y * wt[k], col.z + b.z * wt[k], col.w + b.w * wt[k]);
  }
  st_h4(dst, i, col);
}
// glow == nullptr: the blit (bloom disabled).
__global__ __launch_bounds__(256) void k_bloom_composite(BloomConst P, const __half* __restrict__ src, const __half* __restrict__ glow,
                                                         uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= P.w * P.h) return;
  const uint32_t x = i % P.w, y = i / P.w;
  const float u = (float(x) + 0.5f) / float(P.w), v = (float(y) + 0.5f) / float(P.h);
  const float4 base = sample_lin(src, P.w, P.h, u, v);
  float3 hdr;
  float alpha;
  if (glow) {
    const float4 g = sample_lin(glow, P.hw, P.hh, u, v);
    hdr = make_float3(fmaxf(base.x + fmaxf(g.x, 0.f) * P.strength, 0.f), fmaxf(base.y + fmaxf(g.y, 0.f) * P.strength, 0.f),
                      fmaxf(base.z + fmaxf(g.z, 0.f) * P.strength, 0.f));
    const float ga = clampf(g.w * P.strength, 0.f, 1.f);
    alpha = clampf(base.w + ga * (1.f - base.w), 0.f, 1.f);
  } else {
    hdr = make_float3(fmaxf(base.x, 0.f), fmaxf(base.y, 0.f), fmaxf(base.z, 0.f));
    alpha = clampf(base.w, 0.f, 1.f);
  }
  out[i] = pack_rgba8(pow_hw(hdr.x, 0.92f), pow_hw(hdr.y, 0.92f), pow_hw(hdr.z, 0.92f), alpha);
}

// ------------------------------------------------ differentiable smoke --
// A differentiable smoke step for inflow-control optimisation (no counterpart in
// the reference, which has no gradients). State: density d and velocity u on a
// dense grid. Per step:
//   source:      d_s = d + dt S_rho, u_s = u + dt S_u on the floor slab (interior
//                voxels with 1 <= y <= slab); S is the trilinear upsampling of a
//                cn^3 x 4 control, row-major [channel][z][y][x], channels
//                (u_x, u_y, u_z, rho), cell-centred over the slab box;
//   advection:   u_a = u_s(x - u_s(x) dt) (semi-Lagrangian, 0 on the boundary);
//                d' = MacCormack(d_s, u_a) with the min/max(original, predicted)
//                limiter (the scalar advection's formula without the decay terms);
//   buoyancy:    u_b = u_a + dt beta d' e_y (interior);
//   confinement: u_c = u_b + dt eps (N x w), w = curl u_b, N = grad|w| / max(|grad|w||, 1e-4)
//                (the confinement kernel's direction, without its masks and clamp);
//   projection:  u' = u_c - grad P(div u_c), P the V-cycle (0 on the boundary).
// The adjoint kernels reverse each stage; transposed gathers are float atomics.
__device__ __forceinline__ bool interior(const Grid& g, uint32_t x, uint32_t y, uint32_t z) {
  return x > 0 && y > 0 && z > 0 && x + 1 < g.W && y + 1 < g.H && z + 1 < g.D;
}
__device__ __forceinline__ bool in_slab(const Grid& g, uint32_t slab, uint32_t x, uint32_t y, uint32_t z) {
  return y >= 1 && y <= slab && interior(g, x, y, z);
}
// The 8 control cells (per channel) and weights of slab voxel (x, y, z).
__device__ __forceinline__ void control_taps(const Grid& g, uint32_t cn, uint32_t slab, uint32_t x, uint32_t y, uint32_t z,
                                             uint32_t idx[8], float w[8]) {
  const float m = float(cn - 1), c = float(cn);
  const float px = clampf((float(x) + 0.5f) / float(g.W) * c - 0.5f, 0.f, m);
  const float py = clampf((float(y - 1) + 0.5f) / float(slab) * c - 0.5f, 0.f, m);
  const float pz = clampf((float(z) + 0.5f) / float(g.D) * c - 0.5f, 0.f, m);
  const uint32_t x0 = uint32_t(floorf(px)), y0 = uint32_t(floorf(py)), z0 = uint32_t(floorf(pz));
  const uint32_t x1 = min(x0 + 1u, cn - 1u), y1 = min(y0 + 1u, cn - 1u), z1 = min(z0 + 1u, cn - 1u);
  const float fx = px - float(x0), fy = py - float(y0), fz = pz - float(z0);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    idx[k] = ((k & 1) ? x1 : x0) + cn * (((k & 2) ? y1 : y0) + cn * ((k & 4) ? z1 : z0));
    w[k] = ((k & 1) ? fx : 1.f - fx) * ((k & 2) ? fy : 1.f - fy) * ((k & 4) ? fz : 1.f - fz);
  }
}
__global__ __launch_bounds__(512) void k_df_source(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                   const float* __restrict__ ctrl, uint32_t cn, uint32_t slab, float dt,
                                                   const float* __restrict__ d, Vec3 u, float* __restrict__ ds, Vec3 us) {
  const size_t plane = size_t(cn) * cn * cn;
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    if (in_slab(g, slab, v.x, v.y, v.z)) {
      uint32_t idx[8];
      float w[8];
      control_taps(g, cn, slab, v.x, v.y, v.z, idx, w);
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) s[ch] += w[k] * ctrl[ch * plane + idx[k]];
    }
    us.x[v.i] = u.x[v.i] + dt * s[0];
    us.y[v.i] = u.y[v.i] + dt * s[1];
    us.z[v.i] = u.z[v.i] + dt * s[2];
    ds[v.i] = d[v.i] + dt * s[3];
  }
}
__global__ __launch_bounds__(512) void k_df_advect_u(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     Vec3 u, Vec3 ua, float dt) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    if (!interior(g, v.x, v.y, v.z)) { ua.x[v.i] = 0.f; ua.y[v.i] = 0.f; ua.z[v.i] = 0.f; continue; }
    const Tri t = tri(g, f3(float(v.x) - u.x[v.i] * dt, float(v.y) - u.y[v.i] * dt, float(v.z) - u.z[v.i] * dt));
    ua.x[v.i] = tri_eval(t, u.x);
    ua.y[v.i] = tri_eval(t, u.y);
    ua.z[v.i] = tri_eval(t, u.z);
  }
}
__global__ __launch_bounds__(512) void k_df_advect_d(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     const float* __restrict__ d, Vec3 ua, float* __restrict__ dn, float dt) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    const F3 sp = f3(float(v.x) - ua.x[v.i] * dt, float(v.y) - ua.y[v.i] * dt, float(v.z) - ua.z[v.i] * dt);
    const Tri tp = tri(g, sp);
    const F3 fwd = f3(tri_eval(tp, ua.x), tri_eval(tp, ua.y), tri_eval(tp, ua.z));
    const Tri tr = tri(g, sp + fwd * dt);
    const float pred = tri_eval(tp, d), rec = tri_eval(tr, d), orig = d[v.i];
    const float corr = pred + 0.5f * (orig - rec);
    dn[v.i] = clampf(corr, fminf(orig, pred), fmaxf(orig, pred));
  }
}
__global__ __launch_bounds__(512) void k_df_buoyancy(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     const float* __restrict__ d, Vec3 ua, Vec3 ub, float beta_dt) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    ub.x[v.i] = ua.x[v.i];
    ub.y[v.i] = ua.y[v.i] + (interior(g, v.x, v.y, v.z) ? beta_dt * d[v.i] : 0.f);
    ub.z[v.i] = ua.z[v.i];
  }
}
constexpr float kConfineFloor = 1e-4f;  // SH/apply-vorticity-confinement.wgsl.ts: max(length(grad), 0.0001)
// grad|w| at an interior voxel (central differences; |w| is 0 on the boundary)
__device__ __forceinline__ F3 grad_mag(const Grid& g, const float* wmag, uint32_t x, uint32_t y, uint32_t z) {
  return f3(wmag[vidx(g, x + 1, y, z)] - wmag[vidx(g, x - 1, y, z)], wmag[vidx(g, x, y + 1, z)] - wmag[vidx(g, x, y - 1, z)],
            wmag[vidx(g, x, y, z + 1)] - wmag[vidx(g, x, y, z - 1)]) * 0.5f;
}
__global__ __launch_bounds__(512) void k_df_confine(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                    Vec3 w, const float* __restrict__ wmag, Vec3 ub, Vec3 uc, float eps_dt) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    F3 f = f3(0.f, 0.f, 0.f);
    if (interior(g, v.x, v.y, v.z)) {
      const F3 gm = grad_mag(g, wmag, v.x, v.y, v.z);
      const F3 N = gm / fmaxf(len3(gm), kConfineFloor);
      f = cross3(N, f3(w.x[v.i], w.y[v.i], w.z[v.i])) * eps_dt;
    }
    uc.x[v.i] = ub.x[v.i] + f.x;
    uc.y[v.i] = ub.y[v.i] + f.y;
    uc.z[v.i] = ub.z[v.i] + f.z;
  }
}
// Loss 0.5 sum (d - target)^2 over the grid: FP64 per-workgroup sums into
// out[blockIdx % kNormLanes]; db (optional) receives the adjoint seed d - target.
__global__ __launch_bounds__(512) void k_df_loss(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                 const float* __restrict__ d, const float* __restrict__ target, float* db,
                                                 double* out) {
  __shared__ double part[8];
  double acc = 0.0;
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    const float r = d[v.i] - target[v.i];
    if (db) db[v.i] = r;
    acc += 0.5 * double(r) * double(r);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < 8; ++k) t += part[k];
    atomicAdd(out + (blockIdx.x % kNormLanes), t);
  }
}

// ---- adjoints ----
// Derivative of the trilinear sample t of f with respect to its position p (0
// along an axis where tri() clamped p).
__device__ __forceinline__ F3 tri_grad(const Grid& g, F3 p, const Tri& t, const float* __restrict__ f) {
  const float c0 = f[t.i[0]], c1 = f[t.i[1]], c2 = f[t.i[2]], c3 = f[t.i[3]];
  const float c4 = f[t.i[4]], c5 = f[t.i[5]], c6 = f[t.i[6]], c7 = f[t.i[7]];
  const float ax = 1.f - t.fx, ay = 1.f - t.fy, az = 1.f - t.fz;
  F3 r;
  r.x = ((c1 - c0) * ay + (c3 - c2) * t.fy) * az + ((c5 - c4) * ay + (c7 - c6) * t.fy) * t.fz;
  r.y = ((c2 - c0) * ax + (c3 - c1) * t.fx) * az + ((c6 - c4) * ax + (c7 - c5) * t.fx) * t.fz;
  r.z = ((c4 - c0) * ax + (c5 - c1) * t.fx) * ay + ((c6 - c2) * ax + (c7 - c3) * t.fx) * t.fy;
  if (!(p.x > 0.f && p.x < float(g.W - 1))) r.x = 0.f;
  if (!(p.y > 0.f && p.y < float(g.H - 1))) r.y = 0.f;
  if (!(p.z > 0.f && p.z < float(g.D - 1))) r.z = 0.f;
  return r;
}
// Zero a velocity adjoint on the boundary faces (the forward velocity there is
// the constant 0 written by the projection, so its adjoint is discarded).
__global__ __launch_bounds__(512) void k_da_zero_boundary(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                          Vec3 a) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (v.in && !interior(g, v.x, v.y, v.z)) { a.x[v.i] = 0.f; a.y[v.i] = 0.f; a.z[v.i] = 0.f; }
  }
}
// Confinement adjoint, part 1. With f = eps dt (N x w) and abar = adjoint of u_c:
// Nbar = w x fbar, wbar = fbar x N, gbar = J_N(g)^T Nbar where N = g / max(|g|, floor).
__global__ __launch_bounds__(512) void k_da_confine1(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     Vec3 w, const float* __restrict__ wmag, Vec3 abar, float eps_dt, Vec3 gbar,
                                                     Vec3 wbar) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in) continue;
    const uint32_t i = v.i;
    F3 gb = f3(0.f, 0.f, 0.f), wb = f3(0.f, 0.f, 0.f);
    if (interior(g, v.x, v.y, v.z)) {
      const F3 fb = f3(abar.x[i], abar.y[i], abar.z[i]) * eps_dt;
      const F3 om = f3(w.x[i], w.y[i], w.z[i]);
      const F3 gm = grad_mag(g, wmag, v.x, v.y, v.z);
      const float n = len3(gm);
      const F3 Nb = cross3(om, fb);
      if (n > kConfineFloor) {
        const F3 N = gm / n;
        wb = cross3(fb, N);
        gb = (Nb - N * dot3(N, Nb)) / n;
      } else {
        wb = cross3(fb, gm / kConfineFloor);
        gb = Nb / kConfineFloor;
      }
    }
    gbar.x[i] = gb.x; gbar.y[i] = gb.y; gbar.z[i] = gb.z;
    wbar.x[i] = wb.x; wbar.y[i] = wb.y; wbar.z[i] = wb.z;
  }
}
// Part 2: |w|bar = (grad)^T gbar = 0.5 sum_a (gbar_a[x - e_a] - gbar_a[x + e_a])
// (gbar is 0 off the interior), then wbar += |w|bar w / |w|.
__global__ __launch_bounds__(512) void k_da_confine2(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     Vec3 w, const float* __restrict__ wmag, Vec3 gbar, Vec3 wbar) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in || !interior(g, v.x, v.y, v.z)) continue;
    const uint32_t x = v.x, y = v.y, z = v.z, i = v.i;
    const float m = wmag[i];
    if (!(m > 0.f)) continue;
    const float mb = 0.5f * ((gbar.x[vidx(g, x - 1, y, z)] - gbar.x[vidx(g, x + 1, y, z)]) +
                             (gbar.y[vidx(g, x, y - 1, z)] - gbar.y[vidx(g, x, y + 1, z)]) +
                             (gbar.z[vidx(g, x, y, z - 1)] - gbar.z[vidx(g, x, y, z + 1)]));
    const float s = mb / m;
    wbar.x[i] += s * w.x[i];
    wbar.y[i] += s * w.y[i];
    wbar.z[i] += s * w.z[i];
  }
}
// Part 3: abar += curl^T wbar. The central-difference curl is its own transpose
// on fields that vanish off the interior, so this is curl(wbar) at interior voxels.
__global__ __launch_bounds__(512) void k_da_confine3(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     Vec3 wbar, Vec3 abar) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in || !interior(g, v.x, v.y, v.z)) continue;
    const uint32_t i = v.i, x = v.x, y = v.y, z = v.z;
    auto rv = [&](uint32_t a, uint32_t b2, uint32_t c2) {
      const uint32_t k = vidx(g, a, b2, c2);
      return f3(wbar.x[k], wbar.y[k], wbar.z[k]);
    };
    const F3 x0 = rv(x - 1, y, z), x1 = rv(x + 1, y, z), y0 = rv(x, y - 1, z), y1 = rv(x, y + 1, z);
    const F3 z0 = rv(x, y, z - 1), z1 = rv(x, y, z + 1);
    const F3 c = f3((y1.z - y0.z) - (z1.y - z0.y), (z1.x - z0.x) - (x1.z - x0.z), (x1.y - x0.y) - (y1.x - y0.x)) * 0.5f;
    abar.x[i] += c.x;
    abar.y[i] += c.y;
    abar.z[i] += c.z;
  }
}
// Buoyancy adjoint: dbar' += beta dt abar_y (interior).
__global__ __launch_bounds__(512) void k_da_buoyancy(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     const float* __restrict__ aby, float* dbar, float beta_dt) {
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (v.in && interior(g, v.x, v.y, v.z)) dbar[v.i] += beta_dt * aby[v.i];
  }
}
// Transposed gathers staged in LDS, summed without atomics. The adjoint of a
// semi-Lagrangian sample is a scatter to the 8 corners of the sample; with
// |u dt| < kScatHalo every corner of a brick's voxels lies in the brick grown by
// kScatHalo voxels. Phase 1 (the adjoint kernels below): each target field gets
// a (8 + 2 kScatHalo)^3 LDS tile around the brick, the voxels add into it with
// LDS atomics, and the tile is stored whole to a staging buffer (plain
// coalesced stores); corners outside the tile go to the output with a global
// atomic. Phase 2 (k_da_gather): every voxel adds the entries of the (at most 8)
// tiles that cover it, in a fixed order. Global float atomics on runs of
// neighbouring addresses serialise in L2: a per-entry atomic flush of the tiles
// cost ~10x the rest of the kernel, and so does a halo of 1 when the flow moves
// faster than a voxel per step.
constexpr int kScatHalo = 2;
static_assert(kScatHalo <= 4, "a voxel is covered by at most one neighbour tile per axis");
constexpr int kScatT = 8 + 2 * kScatHalo;
constexpr int kScatN = kScatT * kScatT * kScatT;
struct TileOrigin { int x, y, z; };
__device__ __forceinline__ TileOrigin tile_origin(const Grid& g, uint32_t b) {
  const uint32_t bx = b % g.BX, by = (b / g.BX) % g.BY, bz = b / (g.BX * g.BY);
  return TileOrigin{int(bx * 8) - kScatHalo, int(by * 8) - kScatHalo, int(bz * 8) - kScatHalo};
}
// Corner coordinates and fractions of the clamped trilinear sample at p (as tri()).
struct TriC { int x0, y0, z0, dx, dy, dz; float fx, fy, fz; };
__device__ __forceinline__ TriC tri_coords(const Grid& g, F3 p) {
  const float cx = clampf(p.x, 0.f, float(g.W - 1)), cy = clampf(p.y, 0.f, float(g.H - 1)), cz = clampf(p.z, 0.f, float(g.D - 1));
  TriC c;
  c.x0 = int(floorf(cx)); c.y0 = int(floorf(cy)); c.z0 = int(floorf(cz));
  c.dx = c.x0 + 1 < int(g.W) ? 1 : 0; c.dy = c.y0 + 1 < int(g.H) ? 1 : 0; c.dz = c.z0 + 1 < int(g.D) ? 1 : 0;
  c.fx = fractf(cx); c.fy = fractf(cy); c.fz = fractf(cz);
  return c;
}
// The tiles are passed as generic pointers; the explicit LDS address space keeps
// the staged adds on ds_add_f32 instead of flat atomics.
typedef __attribute__((address_space(3))) float lds_float;
__device__ __forceinline__ void stage_add(float* tile, const TileOrigin& o, float* glob, const Grid& g, int x, int y, int z, float v) {
  const int lx = x - o.x, ly = y - o.y, lz = z - o.z;
  if (unsigned(lx) < unsigned(kScatT) && unsigned(ly) < unsigned(kScatT) && unsigned(lz) < unsigned(kScatT))
    __hip_atomic_fetch_add((lds_float*)tile + ((lz * kScatT + ly) * kScatT + lx), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    atomicAdd(&glob[vidx(g, uint32_t(x), uint32_t(y), uint32_t(z))], v);
}
__device__ __forceinline__ void stage_scatter(float* tile, const TileOrigin& o, float* glob, const Grid& g, const TriC& c, float a) {
  if (a == 0.f) return;
  const float wx[2] = {1.f - c.fx, c.fx}, wy[2] = {1.f - c.fy, c.fy}, wz[2] = {1.f - c.fz, c.fz};
#pragma unroll
  for (int k = 0; k < 8; ++k)
    stage_add(tile, o, glob, g, c.x0 + ((k & 1) ? c.dx : 0), c.y0 + ((k & 2) ? c.dy : 0), c.z0 + ((k & 4) ? c.dz : 0),
              wx[k & 1] * wy[(k >> 1) & 1] * wz[k >> 2] * a);
}
// Scatter of NF fields sharing one trilinear cell c, with the 8 corner values of
// neighbouring lanes merged first. The 64 lanes of a wave are the 8x8 voxels of a
// brick z-slice (lane = x + 8 y). Where the cell of lane - 1 is the cell of this
// lane shifted by -1 in x, its +x corners are this lane's -x corners: they are
// added here through a shuffle and the left lane drops them; then the same along
// y with lane - 8. In smooth flow this leaves about a third of the LDS atomics.
// Every lane of the wave must call this (zero values where nothing scatters);
// the addition order differs from a plain scatter only as atomics already do.
template <int NF>
__device__ __forceinline__ void stage_scatter_merged(float* const* tiles, const TileOrigin& o, float* const* glob, const Grid& g,
                                                     const TriC& c, const float* a) {
  const int lane = int(threadIdx.x & 63u), lx = lane & 7, ly = lane >> 3;
  const float wx[2] = {1.f - c.fx, c.fx}, wy[2] = {1.f - c.fy, c.fy}, wz[2] = {1.f - c.fz, c.fz};
  float v[NF][8];
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int k = 0; k < 8; ++k) v[f][k] = wx[k & 1] * wy[(k >> 1) & 1] * wz[k >> 2] * a[f];
  // x: lane - 1
  {
    const int px = __shfl_up(c.x0, 1, 64), py = __shfl_up(c.y0, 1, 64), pz = __shfl_up(c.z0, 1, 64);
    const int pdx = __shfl_up(c.dx, 1, 64), pdy = __shfl_up(c.dy, 1, 64), pdz = __shfl_up(c.dz, 1, 64);
    const bool from_prev = lx > 0 && pdx == 1 && px + 1 == c.x0 && py == c.y0 && pz == c.z0 && pdy == c.dy && pdz == c.dz;
    // shuffles run with every lane active (a lane reading an inactive lane gets 0)
    const int next_takes = __shfl_down(from_prev ? 1 : 0, 1, 64);
    const bool to_next = lx < 7 && next_takes != 0;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 1; k < 8; k += 2) {
        const float pv = __shfl_up(v[f][k], 1, 64);
        v[f][k - 1] += from_prev ? pv : 0.f;
      }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 1; k < 8; k += 2)
        if (to_next) v[f][k] = 0.f;
  }
  // y: lane - 8
  {
    const int px = __shfl_up(c.x0, 8, 64), py = __shfl_up(c.y0, 8, 64), pz = __shfl_up(c.z0, 8, 64);
    const int pdx = __shfl_up(c.dx, 8, 64), pdy = __shfl_up(c.dy, 8, 64), pdz = __shfl_up(c.dz, 8, 64);
    const bool from_prev = ly > 0 && pdy == 1 && py + 1 == c.y0 && px == c.x0 && pz == c.z0 && pdx == c.dx && pdz == c.dz;
    // shuffles run with every lane active (a lane reading an inactive lane gets 0)
    const int next_takes = __shfl_down(from_prev ? 1 : 0, 8, 64);
    const bool to_next = ly < 7 && next_takes != 0;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (!(k & 2)) continue;
        const float pv = __shfl_up(v[f][k], 8, 64);
        v[f][k - 2] += from_prev ? pv : 0.f;
      }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((k & 2) && to_next) v[f][k] = 0.f;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = c.x0 + ((k & 1) ? c.dx : 0), y = c.y0 + ((k & 2) ? c.dy : 0), z = c.z0 + ((k & 4) ? c.dz : 0);
#pragma unroll
    for (int f = 0; f < NF; ++f)
      if (v[f][k] != 0.f) stage_add(tiles[f], o, glob[f], g, x, y, z, v[f][k]);
  }
}
// Stores the tile of list entry k to stage (layout [list entry][kScatN]) and clears it.
__device__ __forceinline__ void stage_store(float* tile, float* __restrict__ stage, uint32_t k) {
  float* dst = stage + size_t(k) * kScatN;
  for (int t = int(threadIdx.x); t < kScatN; t += int(blockDim.x)) {
    dst[t] = tile[t];
    tile[t] = 0.f;
  }
}
// Phase 2: out[f][v] += sum of the staged tiles covering v, for nf fields
// (stage[f] holds one tile per list entry; page maps a brick to its list entry,
// or the list is the identity when page is null).
struct GatherArgs {
  const float* stage[4];
  float* out[4];
  int nf;
};
__global__ __launch_bounds__(512) void k_da_gather(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                   const int32_t* __restrict__ page, GatherArgs A) {
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Vox v = vox_of(g, b);
    if (!v.in) continue;
    const int lx = int(v.x & 7u), ly = int(v.y & 7u), lz = int(v.z & 7u);
    const int bx = int(b % g.BX), by = int((b / g.BX) % g.BY), bz = int(b / (g.BX * g.BY));
    // neighbour offsets whose tile covers this voxel (own brick first)
    auto side = [](int l) { return l < kScatHalo ? -1 : (l >= 8 - kScatHalo ? 1 : 0); };
    const int ox = side(lx), oy = side(ly), oz = side(lz);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 8; ++k) {
      const int sx = (k & 1) ? ox : 0, sy = (k & 2) ? oy : 0, sz = (k & 4) ? oz : 0;
      if (((k & 1) && !ox) || ((k & 2) && !oy) || ((k & 4) && !oz)) continue;
      const int nbx = bx + sx, nby = by + sy, nbz = bz + sz;
      if (nbx < 0 || nby < 0 || nbz < 0 || nbx >= int(g.BX) || nby >= int(g.BY) || nbz >= int(g.BZ)) continue;
      const uint32_t nb = uint32_t(nbx) + g.BX * (uint32_t(nby) + g.BY * uint32_t(nbz));
      const int32_t e = page ? page[nb] : int32_t(nb);
      if (e < 0) continue;
      const int tx = lx - 8 * sx + kScatHalo, ty = ly - 8 * sy + kScatHalo, tz = lz - 8 * sz + kScatHalo;
      const size_t off = size_t(e) * kScatN + size_t((tz * kScatT + ty) * kScatT + tx);
      for (int f = 0; f < A.nf; ++f) acc[f] += A.stage[f][off];
    }
    for (int f = 0; f < A.nf; ++f) A.out[f][v.i] += acc[f];
  }
}
// MacCormack adjoint. From dnb (adjoint of d') it scatter-adds into dsb (adjoint
// of d_s, cleared by the caller) and abar (adjoint of u_a). The limiter is a
// clamp to [min(orig, pred), max(orig, pred)]; its subgradient passes the adjoint
// to the corrected value inside the interval and to the active bound outside.
__global__ __launch_bounds__(512) void k_da_advect_d(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     const float* __restrict__ d, Vec3 ua, const float* __restrict__ dnb, float* dsb,
                                                     Vec3 abar, float dt, float* __restrict__ stage, size_t stage_field) {
  __shared__ float tile[4][kScatN];  // dsb, abar x, y, z
  for (int t = int(threadIdx.x); t < 4 * kScatN; t += int(blockDim.x)) (&tile[0][0])[t] = 0.f;
  __syncthreads();
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Vox v = vox_of(g, b);
    const TileOrigin o = tile_origin(g, b);
    const float ob = v.in ? dnb[v.i] : 0.f;
    // Every lane takes part in the merged scatters below, so the per-lane work
    // leaves its scatter values here (zero where nothing scatters): predb and the
    // forward-velocity adjoint fb at the backtrace cell cp, recb at the cell cr.
    float predb = 0.f, recb = 0.f;
    F3 fb = f3(0.f, 0.f, 0.f);
    F3 sp = f3(float(v.x), float(v.y), float(v.z)), sr = sp;
    if (ob != 0.f) {
      sp = f3(float(v.x) - ua.x[v.i] * dt, float(v.y) - ua.y[v.i] * dt, float(v.z) - ua.z[v.i] * dt);
      const Tri tp = tri(g, sp);
      const F3 fwd = f3(tri_eval(tp, ua.x), tri_eval(tp, ua.y), tri_eval(tp, ua.z));
      sr = sp + fwd * dt;
      const Tri tr = tri(g, sr);
      const float pred = tri_eval(tp, d), rec = tri_eval(tr, d), orig = d[v.i];
      const float corr = pred + 0.5f * (orig - rec);
      const float lo = fminf(orig, pred), hi = fmaxf(orig, pred);
      float cb = 0.f, origb = 0.f;
      if (corr < lo) {
        if (orig <= pred) origb = ob; else predb = ob;
      } else if (corr > hi) {
        if (orig >= pred) origb = ob; else predb = ob;
      } else {
        cb = ob;
      }
      predb += cb;
      origb += 0.5f * cb;
      recb = -0.5f * cb;
      stage_add(tile[0], o, dsb, g, int(v.x), int(v.y), int(v.z), origb);
      // sp = x - u_a(x) dt; sr = sp + u_a(sp) dt
      F3 spb = f3(0.f, 0.f, 0.f);
      if (predb != 0.f) spb = tri_grad(g, sp, tp, d) * predb;
      if (recb != 0.f) {
        const F3 srb = tri_grad(g, sr, tr, d) * recb;
        fb = srb * dt;
        spb = spb + srb + tri_grad(g, sp, tp, ua.x) * fb.x + tri_grad(g, sp, tp, ua.y) * fb.y + tri_grad(g, sp, tp, ua.z) * fb.z;
      }
      if (spb.x != 0.f) stage_add(tile[1], o, abar.x, g, int(v.x), int(v.y), int(v.z), -dt * spb.x);
      if (spb.y != 0.f) stage_add(tile[2], o, abar.y, g, int(v.x), int(v.y), int(v.z), -dt * spb.y);
      if (spb.z != 0.f) stage_add(tile[3], o, abar.z, g, int(v.x), int(v.y), int(v.z), -dt * spb.z);
    }
    {
      float* tiles[4] = {tile[0], tile[1], tile[2], tile[3]};
      float* globs[4] = {dsb, abar.x, abar.y, abar.z};
      const float ap[4] = {predb, fb.x, fb.y, fb.z};
      stage_scatter_merged<4>(tiles, o, globs, g, tri_coords(g, sp), ap);
      const float ar[1] = {recb};
      stage_scatter_merged<1>(tiles, o, globs, g, tri_coords(g, sr), ar);
    }
    __syncthreads();
    for (int f = 0; f < 4; ++f) stage_store(tile[f], stage + f * stage_field, fx_k);
    __syncthreads();
  }
}
// Velocity advection adjoint: from abar (adjoint of u_a) scatter-adds into usb
// (adjoint of u_s, cleared by the caller).
__global__ __launch_bounds__(512) void k_da_advect_u(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                     Vec3 u, Vec3 abar, Vec3 usb, float dt, float* __restrict__ stage,
                                                     size_t stage_field) {
  __shared__ float tile[3][kScatN];
  for (int t = int(threadIdx.x); t < 3 * kScatN; t += int(blockDim.x)) (&tile[0][0])[t] = 0.f;
  __syncthreads();
  FX_BRICK_LOOP(list, count) {
    const uint32_t b = list[fx_k];
    const Vox v = vox_of(g, b);
    const TileOrigin o = tile_origin(g, b);
    const bool live = v.in && interior(g, v.x, v.y, v.z);
    const F3 a = live ? f3(abar.x[v.i], abar.y[v.i], abar.z[v.i]) : f3(0.f, 0.f, 0.f);
    const bool any = a.x != 0.f || a.y != 0.f || a.z != 0.f;
    // every lane takes part in the merged scatter (zero values where nothing scatters)
    const F3 sp = any ? f3(float(v.x) - u.x[v.i] * dt, float(v.y) - u.y[v.i] * dt, float(v.z) - u.z[v.i] * dt)
                      : f3(float(v.x), float(v.y), float(v.z));
    const TriC c = tri_coords(g, sp);
    {
      float* tiles[3] = {tile[0], tile[1], tile[2]};
      float* globs[3] = {usb.x, usb.y, usb.z};
      const float av[3] = {a.x, a.y, a.z};
      stage_scatter_merged<3>(tiles, o, globs, g, c, av);
    }
    if (any) {
      const Tri t = tri(g, sp);
      const F3 spb = tri_grad(g, sp, t, u.x) * a.x + tri_grad(g, sp, t, u.y) * a.y + tri_grad(g, sp, t, u.z) * a.z;
      if (spb.x != 0.f) stage_add(tile[0], o, usb.x, g, int(v.x), int(v.y), int(v.z), -dt * spb.x);
      if (spb.y != 0.f) stage_add(tile[1], o, usb.y, g, int(v.x), int(v.y), int(v.z), -dt * spb.y);
      if (spb.z != 0.f) stage_add(tile[2], o, usb.z, g, int(v.x), int(v.y), int(v.z), -dt * spb.z);
    }
    __syncthreads();
    for (int f = 0; f < 3; ++f) stage_store(tile[f], stage + f * stage_field, fx_k);
    __syncthreads();
  }
}
// Source adjoint: cbar[ch] += dt * taps^T (u_s bar, d_s bar) over the slab.
__global__ __launch_bounds__(512) void k_da_source(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g,
                                                   uint32_t cn, uint32_t slab, float dt, const float* __restrict__ dsb, Vec3 usb,
                                                   float* cbar) {
  const size_t plane = size_t(cn) * cn * cn;
  FX_BRICK_LOOP(list, count) {
    Vox v = vox_of(g, list[fx_k]);
    if (!v.in || !in_slab(g, slab, v.x, v.y, v.z)) continue;
    uint32_t idx[8];
    float w[8];
    control_taps(g, cn, slab, v.x, v.y, v.z, idx, w);
    const float a[4] = {dt * usb.x[v.i], dt * usb.y[v.i], dt * usb.z[v.i], dt * dsb[v.i]};
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
      if (a[ch] == 0.f) continue;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (w[k] != 0.f) atomicAdd(&cbar[ch * plane + idx[k]], w[k] * a[ch]);
    }
  }
}

// ------------------------------------------------------------ launchers --
#define FX_LAUNCH(kernel, blocks, threads, stream, ...) \
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), 0, stream, __VA_ARGS__)

// Listed fine bricks of every scene of a brick-queue batch (the ascending list
// holds each scene's bricks as one run): out[k] = bricks in [base_k, base_k + bricks_k).
__global__ void k_scene_counts(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, const Grid* __restrict__ tab,
                               uint32_t scenes, uint32_t* __restrict__ out) {
  const uint32_t n = *count;
  for (uint32_t k = threadIdx.x; k < scenes; k += blockDim.x) {
    const Grid g = tab[k];
    out[k] = list_lower_bound(list, n, g.base + g.BX * g.BY * g.BZ) - list_lower_bound(list, n, g.base);
  }
}

// MS launches: kernel template KT (several scenes) when G is a selector, KF otherwise.
#define FX_LAUNCH_MS(G, KT, KF, blocks, threads, stream, ...)    \
  do {                                                          \
    if ((G).tab) FX_LAUNCH(KT, blocks, threads, stream, __VA_ARGS__); \
    else FX_LAUNCH(KF, blocks, threads, stream, __VA_ARGS__);   \
  } while (0)
void launch_inject(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const StepConst* c,
                   const float* src, Scalars set0, Vec3 u, uint32_t* touched, int parts, uint32_t* clear,
                   uint32_t clear_n) {
  switch (set0.prec) {
    case FX_PRECISION_BF16: FX_LAUNCH_MS(g, (k_inject<StBF16, true>), (k_inject<StBF16, false>), blocks, 512, s, list, count, g, c, src, typed<StBF16>(set0), u, touched, parts, clear, clear_n); break;
    case FX_PRECISION_FP16: FX_LAUNCH_MS(g, (k_inject<StF16, true>), (k_inject<StF16, false>), blocks, 512, s, list, count, g, c, src, typed<StF16>(set0), u, touched, parts, clear, clear_n); break;
    case FX_PRECISION_U16: FX_LAUNCH_MS(g, (k_inject<StU16, true>), (k_inject<StU16, false>), blocks, 512, s, list, count, g, c, src, typed<StU16>(set0), u, touched, parts, clear, clear_n); break;
    case FX_PRECISION_UF16: FX_LAUNCH_MS(g, (k_inject<StUF16, true>), (k_inject<StUF16, false>), blocks, 512, s, list, count, g, c, src, typed<StUF16>(set0), u, touched, parts, clear, clear_n); break;
    default: FX_LAUNCH_MS(g, (k_inject<StF32, true>), (k_inject<StF32, false>), blocks, 512, s, list, count, g, c, src, typed<StF32>(set0), u, touched, parts, clear, clear_n);
  }
}
void launch_clear_u32(hipStream_t s, uint32_t* p, size_t n, uint32_t value) {
  int blocks = int(std::min<size_t>((n + 255) / 256, 4096));
  if (blocks < 1) blocks = 1;
  FX_LAUNCH(k_fill_u32, blocks, 256, s, p, n, value);
}
void launch_mark(hipStream_t s, const uint32_t* list, uint32_t n, uint32_t* touched) {
  if (n == 0) return;
  FX_LAUNCH(k_mark, int(std::min<uint32_t>((n + 255) / 256, 1024)), 256, s, list, n, touched);
}
void launch_clear_f32(hipStream_t s, float* p, size_t n) {
  int blocks = int(std::min<size_t>((n + 255) / 256, 8192));
  if (blocks < 1) blocks = 1;
  FX_LAUNCH(k_fill_f32, blocks, 256, s, p, n);
}
void launch_activate(hipStream_t s, int blocks, const uint32_t* listA, const uint32_t* countA, const uint32_t* listB,
                     const uint32_t* countB, Grid g, const StepConst* c, int rule, Scalars set, Vec3 vel, uint32_t* flags,
                     const LevelFlags& lf) {
  if (rule == FX_ACTIVATION_REFERENCE) {  // one scene only (the host rejects it for several)
    int b = std::max(1, std::min(blocks, int((brick_count(g) + 255) / 256) + 1));
    switch (set.prec) {
      case FX_PRECISION_BF16: FX_LAUNCH(k_activate_reference<StBF16>, b, 256, s, listA, countA, listB, countB, g, typed<StBF16>(set), flags, lf); break;
      case FX_PRECISION_FP16: FX_LAUNCH(k_activate_reference<StF16>, b, 256, s, listA, countA, listB, countB, g, typed<StF16>(set), flags, lf); break;
      case FX_PRECISION_U16: FX_LAUNCH(k_activate_reference<StU16>, b, 256, s, listA, countA, listB, countB, g, typed<StU16>(set), flags, lf); break;
      case FX_PRECISION_UF16: FX_LAUNCH(k_activate_reference<StUF16>, b, 256, s, listA, countA, listB, countB, g, typed<StUF16>(set), flags, lf); break;
      default: FX_LAUNCH(k_activate_reference<StF32>, b, 256, s, listA, countA, listB, countB, g, typed<StF32>(set), flags, lf);
    }
  } else {
    switch (set.prec) {
      case FX_PRECISION_BF16: FX_LAUNCH_MS(g, (k_activate_threshold<StBF16, true>), (k_activate_threshold<StBF16, false>), blocks, 512, s, listA, countA, listB, countB, g, c, typed<StBF16>(set).d, vel, flags, lf); break;
      case FX_PRECISION_FP16: FX_LAUNCH_MS(g, (k_activate_threshold<StF16, true>), (k_activate_threshold<StF16, false>), blocks, 512, s, listA, countA, listB, countB, g, c, typed<StF16>(set).d, vel, flags, lf); break;
      case FX_PRECISION_U16: FX_LAUNCH_MS(g, (k_activate_threshold<StU16, true>), (k_activate_threshold<StU16, false>), blocks, 512, s, listA, countA, listB, countB, g, c, typed<StU16>(set).d, vel, flags, lf); break;
      case FX_PRECISION_UF16: FX_LAUNCH_MS(g, (k_activate_threshold<StUF16, true>), (k_activate_threshold<StUF16, false>), blocks, 512, s, listA, countA, listB, countB, g, c, typed<StUF16>(set).d, vel, flags, lf); break;
      default: FX_LAUNCH_MS(g, (k_activate_threshold<StF32, true>), (k_activate_threshold<StF32, false>), blocks, 512, s, listA, countA, listB, countB, g, c, typed<StF32>(set).d, vel, flags, lf);
    }
  }
}
void launch_compact(hipStream_t s, const CompactArgs& A) {
  const uint32_t blocks = A.block_start[A.nlevels];
  FX_LAUNCH(k_compact_count, blocks, 256, s, A);
  FX_LAUNCH(k_compact_scan, 1, 1024, s, A);
  FX_LAUNCH(k_compact_write, blocks, 256, s, A);
}
void launch_advect_velocity(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const StepConst* c,
                            Vec3 u, Vec3 us) {
  FX_LAUNCH_MS(g, (k_advect_velocity<true>), (k_advect_velocity<false>), blocks, 512, s, list, count, g, c, u, us);
}
void launch_advect_scalars(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const StepConst* c,
                           int mode, Scalars set0, Vec3 us, Scalars dst, const uint32_t* flags) {
  switch (dst.prec) {
    case FX_PRECISION_BF16: FX_LAUNCH_MS(g, (k_advect_scalars<StBF16, true>), (k_advect_scalars<StBF16, false>), blocks, 512, s, list, count, g, c, mode, typed<StBF16>(set0), us, typed<StBF16>(dst), flags); break;
    case FX_PRECISION_FP16: FX_LAUNCH_MS(g, (k_advect_scalars<StF16, true>), (k_advect_scalars<StF16, false>), blocks, 512, s, list, count, g, c, mode, typed<StF16>(set0), us, typed<StF16>(dst), flags); break;
    case FX_PRECISION_U16: FX_LAUNCH_MS(g, (k_advect_scalars<StU16, true>), (k_advect_scalars<StU16, false>), blocks, 512, s, list, count, g, c, mode, typed<StU16>(set0), us, typed<StU16>(dst), flags); break;
    case FX_PRECISION_UF16: FX_LAUNCH_MS(g, (k_advect_scalars<StUF16, true>), (k_advect_scalars<StUF16, false>), blocks, 512, s, list, count, g, c, mode, typed<StUF16>(set0), us, typed<StUF16>(dst), flags); break;
    default: FX_LAUNCH_MS(g, (k_advect_scalars<StF32, true>), (k_advect_scalars<StF32, false>), blocks, 512, s, list, count, g, c, mode, typed<StF32>(set0), us, typed<StF32>(dst), flags);
  }
}
void launch_combustion(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const StepConst* c,
                       Scalars src, Scalars dst, const uint32_t* flags) {
  switch (src.prec) {
    case FX_PRECISION_BF16: FX_LAUNCH_MS(g, (k_combustion<StBF16, true>), (k_combustion<StBF16, false>), blocks, 512, s, list, count, g, c, typed<StBF16>(src), typed<StBF16>(dst), flags); break;
    case FX_PRECISION_FP16: FX_LAUNCH_MS(g, (k_combustion<StF16, true>), (k_combustion<StF16, false>), blocks, 512, s, list, count, g, c, typed<StF16>(src), typed<StF16>(dst), flags); break;
    case FX_PRECISION_U16: FX_LAUNCH_MS(g, (k_combustion<StU16, true>), (k_combustion<StU16, false>), blocks, 512, s, list, count, g, c, typed<StU16>(src), typed<StU16>(dst), flags); break;
    case FX_PRECISION_UF16: FX_LAUNCH_MS(g, (k_combustion<StUF16, true>), (k_combustion<StUF16, false>), blocks, 512, s, list, count, g, c, typed<StUF16>(src), typed<StUF16>(dst), flags); break;
    default: FX_LAUNCH_MS(g, (k_combustion<StF32, true>), (k_combustion<StF32, false>), blocks, 512, s, list, count, g, c, typed<StF32>(src), typed<StF32>(dst), flags);
  }
}
void launch_buoyancy(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const StepConst* c,
                     Scalars set0, Vec3 us, const uint32_t* flags) {
  switch (set0.prec) {
    case FX_PRECISION_BF16: FX_LAUNCH_MS(g, (k_buoyancy<StBF16, true>), (k_buoyancy<StBF16, false>), blocks, 512, s, list, count, g, c, typed<StBF16>(set0), us, flags); break;
    case FX_PRECISION_FP16: FX_LAUNCH_MS(g, (k_buoyancy<StF16, true>), (k_buoyancy<StF16, false>), blocks, 512, s, list, count, g, c, typed<StF16>(set0), us, flags); break;
    case FX_PRECISION_U16: FX_LAUNCH_MS(g, (k_buoyancy<StU16, true>), (k_buoyancy<StU16, false>), blocks, 512, s, list, count, g, c, typed<StU16>(set0), us, flags); break;
    case FX_PRECISION_UF16: FX_LAUNCH_MS(g, (k_buoyancy<StUF16, true>), (k_buoyancy<StUF16, false>), blocks, 512, s, list, count, g, c, typed<StUF16>(set0), us, flags); break;
    default: FX_LAUNCH_MS(g, (k_buoyancy<StF32, true>), (k_buoyancy<StF32, false>), blocks, 512, s, list, count, g, c, typed<StF32>(set0), us, flags);
  }
}
void launch_vorticity(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 us, Vec3 w, float* wmag) {
  FX_LAUNCH_MS(g, (k_vorticity<true>), (k_vorticity<false>), blocks, 512, s, list, count, g, us, w, wmag);
}
void launch_confinement(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const StepConst* c, Vec3 w,
                        const float* wmag, Vec3 us, float* cmag, Scalars set0) {
  switch (set0.prec) {
    case FX_PRECISION_BF16: FX_LAUNCH_MS(g, (k_confinement<StBF16, true>), (k_confinement<StBF16, false>), blocks, 512, s, list, count, g, c, w, wmag, us, cmag, typed<StBF16>(set0)); break;
    case FX_PRECISION_FP16: FX_LAUNCH_MS(g, (k_confinement<StF16, true>), (k_confinement<StF16, false>), blocks, 512, s, list, count, g, c, w, wmag, us, cmag, typed<StF16>(set0)); break;
    case FX_PRECISION_U16: FX_LAUNCH_MS(g, (k_confinement<StU16, true>), (k_confinement<StU16, false>), blocks, 512, s, list, count, g, c, w, wmag, us, cmag, typed<StU16>(set0)); break;
    case FX_PRECISION_UF16: FX_LAUNCH_MS(g, (k_confinement<StUF16, true>), (k_confinement<StUF16, false>), blocks, 512, s, list, count, g, c, w, wmag, us, cmag, typed<StUF16>(set0)); break;
    default: FX_LAUNCH_MS(g, (k_confinement<StF32, true>), (k_confinement<StF32, false>), blocks, 512, s, list, count, g, c, w, wmag, us, cmag, typed<StF32>(set0));
  }
}
void launch_divergence(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 us, float* div) {
  FX_LAUNCH_MS(g, (k_divergence<true>), (k_divergence<false>), blocks, 512, s, list, count, g, us, div);
}
void launch_scene_counts(hipStream_t s, const uint32_t* list, const uint32_t* count, const Grid* tab, uint32_t scenes,
                         uint32_t* out) {
  FX_LAUNCH(k_scene_counts, 1, 256, s, list, count, tab, scenes, out);
}
void launch_retire(hipStream_t s, int blocks, const RetireArgs& R) { FX_LAUNCH(k_retire, blocks, 512, s, R); }
void launch_zero_bricks(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, float* const* bufs, int nbufs) {
  for (int base = 0; base < nbufs; base += FX_MAX_ZERO_BUFS) {
    ZeroArgs Z;
    Z.n = std::min(FX_MAX_ZERO_BUFS, nbufs - base);
    for (int k = 0; k < Z.n; ++k) Z.b[k] = bufs[base + k];
    FX_LAUNCH(k_zero_bricks, blocks, 512, s, list, count, Z);
  }
}
void launch_jacobi(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvMask m, const StepConst* c,
                   const float* rhs, const float* src, float* dst, int weighted) {
  FX_LAUNCH(k_jacobi, blocks, 512, s, list, count, g, m, c, rhs, src, dst, weighted);
}
void launch_level_masks(hipStream_t s, int blocks, const LevelMaskArgs& a) {
  FX_LAUNCH_MS(a.g[0], k_level_masks<true>, k_level_masks<false>, blocks, 512, s, a);
}
void launch_coarse_setup(hipStream_t s, const uint32_t* list, const uint32_t* count, Grid g, LvBits b, void* tab, uint32_t* tab_n,
                         float* p0, float* p1, uint32_t scenes) {
  // every workgroup of a scene's row rebuilds its brick scan (at most kSmoothBricks
  // bricks) and writes its share of the slots and of the zeroing
  const dim3 grid(16, std::max(1u, scenes));
  if (g.tab) hipLaunchKernelGGL(k_coarse_setup<true>, grid, dim3(1024), 0, s, list, count, g, b, static_cast<uint4*>(tab), tab_n, p0, p1);
  else hipLaunchKernelGGL(k_coarse_setup<false>, grid, dim3(1024), 0, s, list, count, g, b, static_cast<uint4*>(tab), tab_n, p0, p1);
}
void launch_smooth_block(hipStream_t s, const uint32_t* list, const uint32_t* count, Grid g, LvMask m, const StepConst* c,
                         const float* rhs, float* p0, float* p1, int sweeps, int weighted, int zero_start, const void* tab,
                         const uint32_t* tab_n, uint32_t scenes) {
  const dim3 grid(1, std::max(1u, scenes));
  const uint4* T = static_cast<const uint4*>(tab);
  if (g.tab) hipLaunchKernelGGL(k_smooth_block<true>, grid, dim3(1024), 0, s, list, count, g, m, c, rhs, p0, p1, sweeps, weighted, zero_start, T, tab_n);
  else hipLaunchKernelGGL(k_smooth_block<false>, grid, dim3(1024), 0, s, list, count, g, m, c, rhs, p0, p1, sweeps, weighted, zero_start, T, tab_n);
}
bool launch_vdown(hipStream_t s, int blocks, int pre, const uint32_t* list, const uint32_t* count, Grid f, LvBits bf, Grid c,
                  LvMask mc, const StepConst* cst, const float* rhs, const float* p_in, float* p_out, float* crhs, double* norm2) {
  switch (pre) {
    case 0: FX_LAUNCH_MS(f, (k_vdown<0, true>), (k_vdown<0, false>), blocks, 512, s, list, count, f, bf, c, mc, cst, rhs, p_in, p_out, crhs, norm2); return true;
    case 1: FX_LAUNCH_MS(f, (k_vdown<1, true>), (k_vdown<1, false>), blocks, 512, s, list, count, f, bf, c, mc, cst, rhs, p_in, p_out, crhs, norm2); return true;
    case 2: FX_LAUNCH_MS(f, (k_vdown<2, true>), (k_vdown<2, false>), blocks, 512, s, list, count, f, bf, c, mc, cst, rhs, p_in, p_out, crhs, norm2); return true;
    default: return false;  // larger halos exceed the LDS budget: unfused sweeps
  }
}
bool launch_vup(hipStream_t s, int blocks, int post, const uint32_t* list, const uint32_t* count, Grid c, Grid f, LvBits bf,
                const StepConst* cst, const float* e, const float* rhs, const float* p_in, float* p_out) {
  switch (post) {
    case 0: FX_LAUNCH_MS(f, (k_vup<0, true>), (k_vup<0, false>), blocks, 512, s, list, count, c, f, bf, cst, e, rhs, p_in, p_out); return true;
    case 1: FX_LAUNCH_MS(f, (k_vup<1, true>), (k_vup<1, false>), blocks, 512, s, list, count, c, f, bf, cst, e, rhs, p_in, p_out); return true;
    case 2: FX_LAUNCH_MS(f, (k_vup<2, true>), (k_vup<2, false>), blocks, 512, s, list, count, c, f, bf, cst, e, rhs, p_in, p_out); return true;
    case 3: FX_LAUNCH_MS(f, (k_vup<3, true>), (k_vup<3, false>), blocks, 512, s, list, count, c, f, bf, cst, e, rhs, p_in, p_out); return true;
    default: return false;
  }
}
void launch_residual_lean(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvBits b,
                          const float* rhs, const float* p, float* r, double* norm2) {
  FX_LAUNCH_MS(g, k_residual_lean<true>, k_residual_lean<false>, blocks, 512, s, list, count, g, b, rhs, p, r, norm2);
}
void launch_restrict_lean(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid f, Grid c, LvBits cb,
                          const float* r, float* crhs) {
  FX_LAUNCH_MS(c, k_restrict_lean<true>, k_restrict_lean<false>, blocks, 512, s, list, count, f, c, cb, r, crhs);
}
void launch_prolong_add(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid c, Grid f, LvBits b,
                        const float* e, float* p) {
  FX_LAUNCH_MS(f, k_prolong_add<true>, k_prolong_add<false>, blocks, 512, s, list, count, c, f, b, e, p);
}
void launch_sweep(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvBits b, const StepConst* c,
                  const float* rhs, const float* src, float* dst, bool fine, hipEvent_t ev_start, hipEvent_t ev_stop) {
  // 256-thread workgroups of one wave per brick, 5/3 as many as the 512-thread
  // persistent grid (five per CU fit beside a raymarch workgroup at 72 VGPRs)
  const int wg = blocks >= 8 ? ((blocks * 5 / 3) & ~7) : blocks;
  if (ev_start && ev_stop && src && fine && !g.tab) {
    // the events take the kernel's own start and end timestamps (no dispatch time)
    hipExtLaunchKernelGGL((k_sweep<false, true, false>), dim3(wg), dim3(256), 0, s, ev_start, ev_stop, 0, list, count, g, b, c, rhs,
                          src, dst);
  } else if (src && fine) {
    FX_LAUNCH_MS(g, (k_sweep<false, true, true>), (k_sweep<false, true, false>), wg, 256, s, list, count, g, b, c, rhs, src, dst);
  } else if (src) {
    FX_LAUNCH_MS(g, (k_sweep<false, false, true>), (k_sweep<false, false, false>), wg, 256, s, list, count, g, b, c, rhs, src, dst);
  } else if (fine) {
    FX_LAUNCH_MS(g, (k_sweep<true, true, true>), (k_sweep<true, true, false>), wg, 256, s, list, count, g, b, c, rhs, src, dst);
  } else {
    FX_LAUNCH_MS(g, (k_sweep<true, false, true>), (k_sweep<true, false, false>), wg, 256, s, list, count, g, b, c, rhs, src, dst);
  }
}
void launch_sweep_pair(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvBits b, const StepConst* c,
                       const float* rhs, const float* src, float* dst) {
  if (src) FX_LAUNCH_MS(g, (k_sweepn<2, false, true>), (k_sweepn<2, false, false>), blocks, 512, s, list, count, g, b, c, rhs, src, dst);
  else FX_LAUNCH_MS(g, (k_sweepn<2, true, true>), (k_sweepn<2, true, false>), blocks, 512, s, list, count, g, b, c, rhs, src, dst);
}
void launch_sweep_quad(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvBits b, const StepConst* c,
                       const float* rhs, const float* src, float* dst) {
  if (src) FX_LAUNCH_MS(g, (k_sweepn<4, false, true>), (k_sweepn<4, false, false>), blocks, 512, s, list, count, g, b, c, rhs, src, dst);
  else FX_LAUNCH_MS(g, (k_sweepn<4, true, true>), (k_sweepn<4, true, false>), blocks, 512, s, list, count, g, b, c, rhs, src, dst);
}
void launch_sweep_tile(hipStream_t s, int blocks, int tile, int ns, const uint32_t* list, const uint32_t* count, Grid g, LvBits b,
                       const StepConst* c, const float* rhs, const float* src, float* dst) {
#define FX_TILE_LAUNCH(K, NS)                                                                                                   \
  do {                                                                                                                          \
    if (src) FX_LAUNCH_MS(g, (K<NS, false, true>), (K<NS, false, false>), blocks, 512, s, list, count, g, b, c, rhs, src, dst); \
    else FX_LAUNCH_MS(g, (K<NS, true, true>), (K<NS, true, false>), blocks, 512, s, list, count, g, b, c, rhs, src, dst);      \
  } while (0)
  if (tile == 4) {
    if (ns == 2) FX_TILE_LAUNCH(k_sweep_t4, 2); else FX_TILE_LAUNCH(k_sweep_t4, 1);
  } else {
    if (ns == 2) FX_TILE_LAUNCH(k_sweep_t16, 2); else FX_TILE_LAUNCH(k_sweep_t16, 1);
  }
#undef FX_TILE_LAUNCH
}
void launch_residual(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, LvMask m, const float* rhs,
                     const float* p, float* r, double* norm2) {
  FX_LAUNCH(k_residual, blocks, 512, s, list, count, g, m, rhs, p, r, norm2);
}
void launch_restrict(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid fine, Grid coarse, LvMask mc,
                     const float* src, float* dst, int full_weight) {
  FX_LAUNCH(k_restrict, blocks, 512, s, list, count, fine, coarse, mc, src, dst, full_weight);
}
void launch_prolong(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid coarse, Grid fine, LvMask mf,
                    const float* src, float* dst) {
  FX_LAUNCH(k_prolong, blocks, 512, s, list, count, coarse, fine, mf, src, dst);
}
void launch_projection(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 us, const float* p,
                       Vec3 u) {
  FX_LAUNCH_MS(g, k_projection<true>, k_projection<false>, blocks, 512, s, list, count, g, us, p, u);
}
void launch_debug(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 u, float* umag) {
  FX_LAUNCH(k_debug, blocks, 512, s, list, count, g, u, umag);
}
void launch_render_pack(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Scalars set0, const uint32_t* rlist,
                        const uint32_t* rcount, void* rb, const uint32_t* flags, uint32_t nb, uint32_t* bits) {
  float4* RB = static_cast<float4*>(rb);
  switch (set0.prec) {
    case FX_PRECISION_BF16: FX_LAUNCH(k_pack_render<StBF16>, blocks, 512, s, list, count, typed<StBF16>(set0), rlist, rcount, RB, flags, nb, bits); break;
    case FX_PRECISION_FP16: FX_LAUNCH(k_pack_render<StF16>, blocks, 512, s, list, count, typed<StF16>(set0), rlist, rcount, RB, flags, nb, bits); break;
    case FX_PRECISION_U16: FX_LAUNCH(k_pack_render<StU16>, blocks, 512, s, list, count, typed<StU16>(set0), rlist, rcount, RB, flags, nb, bits); break;
    case FX_PRECISION_UF16: FX_LAUNCH(k_pack_render<StUF16>, blocks, 512, s, list, count, typed<StUF16>(set0), rlist, rcount, RB, flags, nb, bits); break;
    default: FX_LAUNCH(k_pack_render<StF32>, blocks, 512, s, list, count, typed<StF32>(set0), rlist, rcount, RB, flags, nb, bits);
  }
}
void launch_raymarch(hipStream_t s, Grid g, const RayConst& rc, const void* rb, const uint32_t* bits, uint32_t* queue, int blocks,
                     uint16_t* out) {
  // the flag words of g's scene (its bricks start at bit base & 31 of the first)
  const uint32_t nb = uint32_t(brick_count(g)), nwords = ((g.base & 31u) + nb + 31) / 32;
  bits += g.base >> 5;
  const int lds = nwords <= 8192 ? 1 : 0;  // up to 32 KB of flag bits in LDS
  launch_clear_u32(s, queue, 1, 0u);
  const uint32_t ntiles = ((rc.width + 15) / 16) * ((rc.height + 15) / 16);
  dim3 grid(std::max(1u, std::min(ntiles, uint32_t(std::max(blocks, 1)))));
  const size_t sh = lds ? nwords * 4 : 0;
  hipLaunchKernelGGL(k_raymarch, grid, dim3(256), sh, s, g, rc, static_cast<const float4*>(rb), bits, nwords, lds, out, queue);
}

// ---------------------------------------------- differentiable smoke launchers --
void launch_df_source(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const float* ctrl, uint32_t cn,
                      uint32_t slab, float dt, const float* d, Vec3 u, float* ds, Vec3 us) {
  FX_LAUNCH(k_df_source, blocks, 512, s, list, count, g, ctrl, cn, slab, dt, d, u, ds, us);
}
void launch_df_advect(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 us, Vec3 ua,
                      const float* ds, float* dn, float dt) {
  FX_LAUNCH(k_df_advect_u, blocks, 512, s, list, count, g, us, ua, dt);
  FX_LAUNCH(k_df_advect_d, blocks, 512, s, list, count, g, ds, ua, dn, dt);
}
void launch_df_forces(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const float* dn, Vec3 ua, Vec3 ub,
                      Vec3 w, float* wmag, Vec3 uc, float beta_dt, float eps_dt) {
  FX_LAUNCH(k_df_buoyancy, blocks, 512, s, list, count, g, dn, ua, ub, beta_dt);
  FX_LAUNCH(k_vorticity<false>, blocks, 512, s, list, count, g, ub, w, wmag);
  FX_LAUNCH(k_df_confine, blocks, 512, s, list, count, g, w, wmag, ub, uc, eps_dt);
}
void launch_df_loss(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, const float* d, const float* target,
                    float* db, double* out) {
  FX_LAUNCH(k_df_loss, blocks, 512, s, list, count, g, d, target, db, out);
}
void launch_da_zero_boundary(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 a) {
  FX_LAUNCH(k_da_zero_boundary, blocks, 512, s, list, count, g, a);
}
void launch_da_forces(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, Vec3 ub, Vec3 w, float* wmag,
                      Vec3 abar, Vec3 gbar, Vec3 wbar, float* dnb, float beta_dt, float eps_dt) {
  FX_LAUNCH(k_vorticity<false>, blocks, 512, s, list, count, g, ub, w, wmag);
  FX_LAUNCH(k_da_confine1, blocks, 512, s, list, count, g, w, wmag, abar, eps_dt, gbar, wbar);
  FX_LAUNCH(k_da_confine2, blocks, 512, s, list, count, g, w, wmag, gbar, wbar);
  FX_LAUNCH(k_da_confine3, blocks, 512, s, list, count, g, wbar, abar);
  FX_LAUNCH(k_da_buoyancy, blocks, 512, s, list, count, g, abar.y, dnb, beta_dt);
}
void launch_da_advect(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, size_t nb, const float* ds,
                      Vec3 us, Vec3 ua, const float* dnb, float* dsb, Vec3 abar, Vec3 usb, float dt, float* stage) {
  const size_t sf = nb * size_t(kScatN);
  FX_LAUNCH(k_da_advect_d, blocks, 512, s, list, count, g, ds, ua, dnb, dsb, abar, dt, stage, sf);
  GatherArgs A{{stage, stage + sf, stage + 2 * sf, stage + 3 * sf}, {dsb, abar.x, abar.y, abar.z}, 4};
  FX_LAUNCH(k_da_gather, blocks, 512, s, list, count, g, nullptr, A);
  FX_LAUNCH(k_da_advect_u, blocks, 512, s, list, count, g, us, abar, usb, dt, stage, sf);
  GatherArgs B{{stage, stage + sf, stage + 2 * sf, nullptr}, {usb.x, usb.y, usb.z, nullptr}, 3};
  FX_LAUNCH(k_da_gather, blocks, 512, s, list, count, g, nullptr, B);
}
size_t da_stage_floats(size_t nb) { return 4 * nb * size_t(kScatN); }
void launch_da_source(hipStream_t s, int blocks, const uint32_t* list, const uint32_t* count, Grid g, uint32_t cn, uint32_t slab,
                      float dt, const float* dsb, Vec3 usb, float* cbar) {
  FX_LAUNCH(k_da_source, blocks, 512, s, list, count, g, cn, slab, dt, dsb, usb, cbar);
}

// ----------------------------------------------------------- bloom launchers --
void launch_bloom(hipStream_t s, const BloomArgs& a, const void* src, void* bright, void* blur_h, void* blur_v, void* out) {
  BloomConst P{a.threshold, a.strength, a.radius, a.w, a.h, a.hw, a.hh};
  const __half* S = static_cast<const __half*>(src);
  const uint32_t nh = a.hw * a.hh, nf = a.w * a.h;
  if (a.enabled) {
    FX_LAUNCH(k_bloom_threshold, (nh + 255) / 256, 256, s, P, S, static_cast<__half*>(bright));
    FX_LAUNCH(k_bloom_blur, (nh + 255) / 256, 256, s, P, 1, static_cast<const __half*>(bright), static_cast<__half*>(blur_h));
    FX_LAUNCH(k_bloom_blur, (nh + 255) / 256, 256, s, P, 0, static_cast<const __half*>(blur_h), static_cast<__half*>(blur_v));
  }
  FX_LAUNCH(k_bloom_composite, (nf + 255) / 256, 256, s, P, S, a.enabled ? static_cast<const __half*>(blur_v) : nullptr,
            static_cast<uint32_t*>(out));
}

}  // namespace fx
