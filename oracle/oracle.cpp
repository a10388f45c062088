// oracle.cpp — CPU restatement of fenix's per-step combustion-volume pipeline.
//
// TEST INFRASTRUCTURE ONLY. This file is the oracle for libfenix: only tests,
// the smoke check and the CPU-baseline leg of the benchmark load it. The product
// (fenix_amd) never links or calls it.
//
// It restates the reference's WGSL passes (paths below are relative to the
// reference repository root; SH = src/engine/simulation/shaders/passes,
// SC = src/engine/simulation/shaders/common,
// ORCH = src/engine/simulation/runtime/createCombustionVolumeSimulation.ts,
// PSP = src/engine/simulation/runtime/passes/PressureSolvePass.ts,
// RP = src/engine/render/passes/createVolumeRaymarchPass.ts) in two builds:
//   T = float : WGSL op order, mix(a,b,t) = a*(1-t) + b*t, smoothstep per the
//               WGSL spec, no FMA contraction (-ffp-contract=off);
//   T = double: the same arithmetic in FP64 (drift reference).
// Fields are dense row-major (x + W*(y + H*z)), SC/indexing.wgsl.ts:2-4.
//
// Deliberate semantic choices (shared with libfenix and stated here):
//   * combustion reads the pre-pass values of every neighbour ("read-old"),
//     resolving the in-place race of SH/combustion-update.wgsl.ts:56-92;
//   * the render set is always set 0: scalar advection writes set 1 and
//     combustion writes set 0 back, which is the reference's set flip followed
//     by its in-place combustion (ORCH:274-302);
//   * FX_DOMAIN_SPARSE: one activation per step after injection; every field is
//     zero outside the active bricks; injection runs over the bricks that
//     overlap a gate-active source's bounding box; the pressure hierarchy uses
//     per-level brick masks coarsened from the fine one;
//   * FX_SOLVER_VCYCLE: weighted-Jacobi V-cycle, residual restricted with the
//     (1,3,3,1)/8 stencil (transpose of the trilinear prolongation / 8) and
//     scaled by 4 (h^2), coarse levels start from zero.
//
// Pinning: tests/test_oracle.py and tests/test_oracle_pins.py check each function
// against closed forms and invariants: Jacobi eigenvectors, exact shifts under
// integer advection, the trilinear sampler exact on affine fields at fractional
// positions, the MacCormack limiter on cases where it bites, central differences
// on polynomials, activation special cases, the prolongation exact on affine
// coarse fields and the restriction as its scaled transpose, the combustion
// 6-neighbour stencil and buoyancy gradient terms by differences of runs, the
// raymarch's slab opacity/colour closed forms, phase normalisation, palette and
// light helpers (orc_probe), zero-source identities.
// Parity unpinned (no execution of the WGSL is possible here): hash noise
// (inject-source-common.wgsl.ts:102-121) and the sin-based thermal noise beyond
// tolerance, WGSL mix/FMA choices, the raymarch's pow/exp and fp16 rounding.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../include/fenix.h"

namespace orc {

// ---------------------------------------------------------------- helpers --
template <class T> struct V3 { T x, y, z; };
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> mul(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T len3(V3<T> a) { return std::sqrt(dot3(a, a)); }
template <class T> inline V3<T> norm3(V3<T> a) { T l = len3(a); return {a.x / l, a.y / l, a.z / l}; }
template <class T> inline V3<T> cross3(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T clampT(T v, T lo, T hi) { return std::min(std::max(v, lo), hi); }
template <class T> inline T clamp01(T v) { return clampT(v, T(0), T(1)); }
// WGSL mix(e1,e2,e3) = e1*(1-e3) + e2*e3.
template <class T> inline T mixT(T a, T b, T t) { return a * (T(1) - t) + b * t; }
template <class T> inline V3<T> mix3(V3<T> a, V3<T> b, T t) { return {mixT(a.x, b.x, t), mixT(a.y, b.y, t), mixT(a.z, b.z, t)}; }
template <class T> inline T smooth(T e0, T e1, T x) {
  T t = clampT((x - e0) / (e1 - e0), T(0), T(1));
  return t * t * (T(3) - T(2) * t);
}
// WGSL step(edge, x) = 1 if edge <= x.
template <class T> inline T stepT(T edge, T x) { return edge <= x ? T(1) : T(0); }
template <class T> inline T fractT(T x) { return x - std::floor(x); }

// 16-bit storage of density, temperature and fuel (FX_PRECISION_BF16/FP16):
// values are rounded to nearest even on every store; FP32 arithmetic.
inline float round_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}
inline float round_f16(float f) {
  // IEEE binary16, round to nearest even, including subnormals (unit 2^-24).
  float a = std::fabs(f);
  if (!(a < 65520.f)) return std::copysign(INFINITY, f);
  if (a < 6.103515625e-05f) return std::copysign(std::nearbyint(a * 16777216.f) / 16777216.f, f);
  int e;
  std::frexp(a, &e);                       // a in [2^(e-1), 2^e)
  float ulp = std::ldexp(1.f, e - 11);     // 10 fraction bits
  return std::copysign(std::nearbyint(a / ulp) * ulp, f);
}

// Stochastic rounding (FX_PRECISION_BF16 / FP16): the stored value is the upper
// representable neighbour with probability (dropped fraction of an ulp); the
// decision compares the dropped low bits of the f32 value with bits of
// fmix32(i * 0x9e3779b1 ^ fmix32(step * 16 + site)), i the brick-major index of the
// voxel and site = 4 * pass + field (pass 0 scalar injection, 1 igniter,
// 2 scalar advection, 3 combustion). fmix32 is MurmurHash3's 32-bit finalizer.
inline uint32_t fmix32(uint32_t x) {
  x ^= x >> 16; x *= 0x85ebca6bu; x ^= x >> 13; x *= 0xc2b2ae35u; x ^= x >> 16;
  return x;
}
inline uint32_t sr_hash(uint32_t i, uint32_t step, uint32_t site) { return fmix32((i * 0x9e3779b1u) ^ fmix32(step * 16u + site)); }
inline float sr_bf16(float f, uint32_t h) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u = (u & 0xffff0000u) + ((h & 0xffffu) < (u & 0xffffu) ? 0x10000u : 0u);
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}
inline float sr_f16(float f, uint32_t h) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t a = u & 0x7fffffffu;
  if (a >= 0x477ff000u) return round_f16(f);
  const uint32_t e = a >> 23;
  uint32_t hb;  // half bits without the sign
  if (e >= 113u) {
    hb = (((e - 112u) << 10) | ((a >> 13) & 0x3ffu)) + ((h & 0x1fffu) < (a & 0x1fffu) ? 1u : 0u);
  } else {
    const uint32_t d = 126u - e;
    if (d >= 32u) return std::copysign(0.f, f);
    const uint32_t m = (a & 0x7fffffu) | 0x800000u;
    hb = (m >> d) + ((h >> (32u - d)) < (m & ((1u << d) - 1u)) ? 1u : 0u);
  }
  const uint32_t he = hb >> 10, hm = hb & 0x3ffu;  // decode the half
  const float mag = he == 0 ? std::ldexp(float(hm), -24) : std::ldexp(float(1024u + hm), int(he) - 25);
  return std::copysign(mag, f);
}

// 16-bit unsigned fixed point (FX_PRECISION_U16): value = u * 2^-16, u in
// [0, 65535]; the value is clamped to [0, 1] first. Nearest even, or stochastic:
// up when the top 24 hash bits, read as an integer, are below the dropped
// fraction scaled by 2^24 (exact in float: the fraction has at most 24 bits).
inline float round_u16(float f) {
  const float x = clamp01(f) * 65536.f;
  return std::fmin(std::nearbyint(x), 65535.f) / 65536.f;
}
inline float sr_u16(float f, uint32_t h) {
  const float x = clamp01(f) * 65536.f, fl = std::floor(x);
  const float up = float(h >> 8) < (x - fl) * 16777216.f ? 1.f : 0.f;
  return std::fmin(fl + up, 65535.f) / 65536.f;
}
// Unsigned 16-bit float (FX_PRECISION_UF16): 4 exponent bits, 12 mantissa bits,
// normal values (1 + m/4096) * 2^(E-15) for E = 1..15 and multiples of 2^-26 below
// 2^-14; the value is clamped to [0, 1] first. Rounding by the dropped 11 low bits
// of the f32 (normal values) or by the fraction of a 2^-26 step (small values).
inline float round_uf16(float f) {
  const float a = clamp01(f);
  if (a < 6.103515625e-05f) return std::nearbyint(a * 67108864.f) / 67108864.f;
  uint32_t b;
  std::memcpy(&b, &a, 4);
  b += 0x3ffu + ((b >> 11) & 1u);
  b &= ~0x7ffu;
  float r;
  std::memcpy(&r, &b, 4);
  return r;
}
inline float sr_uf16(float f, uint32_t h) {
  const float a = clamp01(f);
  if (a < 6.103515625e-05f) {
    const float x = a * 67108864.f, fl = std::floor(x);
    return (fl + (float(h >> 8) < (x - fl) * 16777216.f ? 1.f : 0.f)) / 67108864.f;
  }
  uint32_t b;
  std::memcpy(&b, &a, 4);
  b = (b & ~0x7ffu) + ((h & 0x7ffu) < (b & 0x7ffu) ? 0x800u : 0u);
  float r;
  std::memcpy(&r, &b, 4);
  return r;
}
inline float round_storage(int precision, float f) {
  return precision == FX_PRECISION_BF16 ? round_bf16(f) : precision == FX_PRECISION_U16 ? round_u16(f)
       : precision == FX_PRECISION_UF16 ? round_uf16(f) : round_f16(f);
}
inline float sr_storage(int precision, float f, uint32_t h) {
  return precision == FX_PRECISION_BF16 ? sr_bf16(f, h) : precision == FX_PRECISION_U16 ? sr_u16(f, h)
       : precision == FX_PRECISION_UF16 ? sr_uf16(f, h) : sr_f16(f, h);
}

struct Dims {
  uint32_t W, H, D;
  size_t N() const { return size_t(W) * H * D; }
  size_t idx(uint32_t x, uint32_t y, uint32_t z) const { return size_t(x) + size_t(W) * (size_t(y) + size_t(H) * z); }
  bool boundary(uint32_t x, uint32_t y, uint32_t z) const {
    return x == 0 || y == 0 || z == 0 || x == W - 1 || y == H - 1 || z == D - 1;
  }
  uint32_t cx(uint32_t x) const { return std::min(x, W - 1); }
  uint32_t cy(uint32_t y) const { return std::min(y, H - 1); }
  uint32_t cz(uint32_t z) const { return std::min(z, D - 1); }
  size_t cidx(uint32_t x, uint32_t y, uint32_t z) const { return idx(cx(x), cy(y), cz(z)); }
};
inline uint32_t dec(uint32_t v) { return v == 0 ? 0u : v - 1u; }
inline uint32_t inc(uint32_t v, uint32_t mx) { return std::min(v + 1u, mx); }
inline uint32_t halve(uint32_t n) { return std::max<uint32_t>(4u, n / 2u); }  // CVS/buffers.ts:10-16

// Trilinear sampler, SC/sampling-scalar.wgsl.ts:11-34 (clamp, floor, upper = min(base+1, max)).
template <class T> struct Tri {
  size_t i[8];
  T fx, fy, fz;
};
template <class T> inline Tri<T> tri(const Dims& d, V3<T> p) {
  T mx = T(d.W - 1), my = T(d.H - 1), mz = T(d.D - 1);
  T cx = clampT(p.x, T(0), mx), cy = clampT(p.y, T(0), my), cz = clampT(p.z, T(0), mz);
  uint32_t bx = uint32_t(std::floor(cx)), by = uint32_t(std::floor(cy)), bz = uint32_t(std::floor(cz));
  uint32_t ux = std::min(bx + 1u, d.W - 1), uy = std::min(by + 1u, d.H - 1), uz = std::min(bz + 1u, d.D - 1);
  Tri<T> t;
  t.fx = fractT(cx); t.fy = fractT(cy); t.fz = fractT(cz);
  t.i[0] = d.idx(bx, by, bz); t.i[1] = d.idx(ux, by, bz); t.i[2] = d.idx(bx, uy, bz); t.i[3] = d.idx(ux, uy, bz);
  t.i[4] = d.idx(bx, by, uz); t.i[5] = d.idx(ux, by, uz); t.i[6] = d.idx(bx, uy, uz); t.i[7] = d.idx(ux, uy, uz);
  return t;
}
template <class T> inline T tri_eval(const Tri<T>& t, const T* f) {
  T x00 = mixT(f[t.i[0]], f[t.i[1]], t.fx);
  T x10 = mixT(f[t.i[2]], f[t.i[3]], t.fx);
  T x01 = mixT(f[t.i[4]], f[t.i[5]], t.fx);
  T x11 = mixT(f[t.i[6]], f[t.i[7]], t.fx);
  return mixT(mixT(x00, x10, t.fy), mixT(x01, x11, t.fy), t.fz);
}

// -------------------------------------------------------- shading helpers --
// Raymarch shading functions (RP = src/engine/render/passes/createVolumeRaymarchPass.ts),
// free functions so that tests can evaluate them alone (orc_probe).
template <class T> inline V3<T> fire_palette(T t) {  // RP:477-486
  V3<T> ember{T(0.18), T(0.018), T(0)}, red{T(1), T(0.075), T(0.012)}, orange{T(1), T(0.42), T(0.035)}, white{T(1), T(0.86), T(0.56)};
  T rm = smooth(T(0.01), T(0.11), t), om = smooth(T(0.09), T(0.38), t), wm = smooth(T(0.58), T(0.92), t);
  return mix3(mix3(mix3(ember, red, rm), orange, om), white, wm);
}
template <class T> inline V3<T> smoke_albedo(T den, T tem, T hot) {  // RP:488-495
  T warm = clampT(tem * T(0.38), T(0), T(1));
  T soot = smooth(T(0.16), T(0.9), den) * (T(1) - smooth(T(0.2), T(0.74), hot));
  V3<T> cool{T(0.26), T(0.275), T(0.295)}, dsoot{T(0.04), T(0.041), T(0.045)}, warmS{T(0.36), T(0.27), T(0.18)};
  return mix3(mix3(cool, dsoot, soot * T(0.86)), warmS, warm * (T(0.35) + hot * T(0.45)));
}
template <class T> inline T phase_hg(T c, T g) {  // RP:497-501
  T g2 = g * g;
  T denom = std::pow(std::max(T(1) + g2 - T(2) * g * c, T(0.001)), T(1.5));
  return (T(1) - g2) / (T(12.566370614359172) * denom);
}
// kind > 0.5: point light at pos; else directional dir (default when |dir|^2 <= 1e-4). RP:503-510
template <class T> inline V3<T> light_dir(T kind, V3<T> dir, V3<T> pos, V3<T> wp) {
  if (kind > T(0.5)) return norm3(pos - wp);
  return norm3(dot3(dir, dir) > T(0.0001) ? dir : V3<T>{T(-0.34), T(0.88), T(0.31)});
}
template <class T> inline T light_att(T kind, V3<T> pos, V3<T> wp) {  // RP:512-520
  if (kind <= T(0.5)) return T(1);
  T dist = len3(pos - wp);
  return T(1) / (T(1) + dist * T(0.09) + dist * dist * T(0.02));
}
template <class T> inline T light_power(T intensity, T att) {  // RP:655-656, lightIntensity :522-524
  T raw = std::max(intensity, T(0)) * att;
  return raw / (T(1) + raw * T(0.18));
}

// ------------------------------------------------------------- emitters --
// EmitterSource layout, SH/inject-source-common.wgsl.ts:3-12 / packEmitterSources.ts:69-121.
struct SrcView {
  const float* s;
  float px() const { return s[0]; } float py() const { return s[1]; } float pz() const { return s[2]; }
  float radius() const { return s[3]; }
  float start() const { return s[4]; } float duration() const { return s[5]; }
  float yd() const { return s[8]; } float yt() const { return s[9]; } float yf() const { return s[10]; }
  float speed() const { return s[12]; } float mode() const { return s[13]; }
  float falloff() const { return s[14]; } float vw() const { return s[15]; }
  float dx_() const { return s[16]; } float dy_() const { return s[17]; } float dz_() const { return s[18]; }
  float nscale() const { return s[20]; } float nmix() const { return s[21]; } float intensity() const { return s[22]; }
  float kind() const { return s[28]; } float seed() const { return s[29]; }
};

template <class T> struct Params {
  T time, dt, prev, setIndex;
  T wind[4], gravity[4];
  T buoyancy, vort, dx, worldSize;
  T vmask[4];
  static Params from(const fx_sim_params& p) {
    Params q;
    q.time = p.time; q.dt = p.delta_time; q.prev = p.previous_time; q.setIndex = p.scalar_set_index;
    for (int i = 0; i < 4; ++i) { q.wind[i] = p.wind[i]; q.gravity[i] = p.gravity[i]; q.vmask[i] = p.vorticity_mask[i]; }
    q.buoyancy = p.buoyancy; q.vort = p.vorticity_strength; q.dx = p.dx; q.worldSize = p.world_size;
    return q;
  }
};

template <class T> struct Emit {
  const Params<T>& P;
  // SH/inject-source-common.wgsl.ts:102-121. The sine is the correctly rounded one of
  // the build's type: for float, the double sine rounded to float (WGSL leaves f32 sin
  // of large arguments to the implementation; the kernels compute it the same way).
  static T sinCR(T x) { return T(std::sin(double(x))); }
  static T hash31(V3<T> p) { return fractT(sinCR(dot3(p, V3<T>{T(127.1), T(311.7), T(74.7)})) * T(43758.5453123)); }
  static T valueNoise3(V3<T> p) {
    V3<T> c{std::floor(p.x), std::floor(p.y), std::floor(p.z)};
    V3<T> f{fractT(p.x), fractT(p.y), fractT(p.z)};
    V3<T> u{f.x * f.x * (T(3) - T(2) * f.x), f.y * f.y * (T(3) - T(2) * f.y), f.z * f.z * (T(3) - T(2) * f.z)};
    auto h = [&](T a, T b, T d) { return hash31(c + V3<T>{a, b, d}); };
    return mixT(mixT(mixT(h(0, 0, 0), h(1, 0, 0), u.x), mixT(h(0, 1, 0), h(1, 1, 0), u.x), u.y),
                mixT(mixT(h(0, 0, 1), h(1, 0, 1), u.x), mixT(h(0, 1, 1), h(1, 1, 1), u.x), u.y), u.z);
  }
  static T emitterNoise(V3<T> pos, T scale, T seed) {
    return valueNoise3(pos * std::max(scale, T(0.1)) + V3<T>{seed * T(0.137), seed * T(0.271), -seed * T(0.193)});
  }
  T activeGate(SrcView s) const {
    T age = P.time - T(s.start());
    return stepT(T(0), age) * stepT(age, T(s.duration()));
  }
  T sourceAge(SrcView s) const { return std::max(P.time - T(s.start()), T(0)); }
  static T sourceDuration(SrcView s) { return std::max(T(s.duration()), T(0.0001)); }
  static T sphereFalloff(T dist, T radius, T falloff) {
    T t = clampT(T(1) - dist / std::max(radius, T(0.0001)), T(0), T(1));
    return mixT(stepT(T(0), t - T(0.0001)), t * t, falloff);
  }
  T sourceFeather(T r) const {
    T voxelWidth = P.dx / std::max(P.worldSize, T(0.0001));
    return std::max(r * T(0.24), voxelWidth * T(3));
  }
  T sourceOuterLimit(T r) const { return r + sourceFeather(r); }
  T eruptive(V3<T> pos, T dist, SrcView s, T falloff) const {
    T targetRadius = s.radius();
    T growthSeconds = clampT(targetRadius * T(5.5), T(0.1), T(0.75));
    T age = sourceAge(s);
    T growth = smooth(T(0), T(1), clampT(age / growthSeconds, T(0), T(1)));
    T feather = sourceFeather(targetRadius);
    T seed = s.seed();
    T coarse = emitterNoise(pos + V3<T>{age * T(0.17), -age * T(0.11), age * T(0.07)}, T(6), seed);
    T fine = emitterNoise(pos + V3<T>{-age * T(0.33), age * T(0.21), -age * T(0.15)}, T(18), seed + T(41));
    T lobe = clampT(coarse * T(0.72) + fine * T(0.28), T(0), T(1));
    T reach = targetRadius * growth * mixT(T(0.62), T(1), smooth(T(0.2), T(0.92), lobe));
    T front = T(1) - smooth(reach - feather, reach + feather, dist);
    T coreRadius = std::max(targetRadius * (T(0.18) + growth * T(0.38)), feather);
    T core = sphereFalloff(dist, coreRadius, falloff) * (T(1) - growth * T(0.35));
    T body = sphereFalloff(dist, std::max(reach + feather, feather), falloff) * front;
    return clampT(std::max(core, body) * mixT(T(0.72), T(1.2), lobe), T(0), T(1));
  }
  // SH/inject-velocity.wgsl.ts:39-44
  static V3<T> turbulentDir(V3<T> pos, T seed) {
    T nx = emitterNoise(pos, T(4), seed);
    T ny = emitterNoise(pos, T(4), seed + T(17.3));
    T nz = emitterNoise(pos, T(4), seed + T(31.7));
    return norm3(V3<T>{nx, ny, nz} - V3<T>{T(0.5), T(0.5), T(0.5)});
  }
};

// ------------------------------------------------------------- the sim --
struct Base {
  virtual ~Base() {}
  virtual int update_sources(const float* packed, uint32_t n) = 0;
  virtual int set_runtime(const fx_runtime_params* p, uint32_t mask) = 0;
  virtual int set_quality(const fx_quality* q) = 0;
  virtual int set_mode(int m) = 0;
  virtual int step(double e, double s) = 0;
  virtual int reset() = 0;
  virtual int run_pass(int pass, const fx_sim_params* p) = 0;
  virtual int read_field(int id, double* out) = 0;
  virtual int write_field(int id, const double* in) = 0;
  virtual int read_flags(int which, uint32_t* out) = 0;
  virtual int read_level(int level, int which, double* out) = 0;
  virtual int level_dims(int level, uint32_t* dims) = 0;
  virtual int residuals(double* out, int cap) = 0;
  virtual int raymarch(const fx_camera* c, const fx_render_params* rp, uint32_t w, uint32_t h, float* out) = 0;
  virtual int active_count() = 0;
  // Test hooks: write one level buffer; apply one level operator (LevelOp).
  virtual int write_level(int level, int which, const double* in) = 0;
  virtual int request_debug() = 0;
  virtual int level_op(int op, int level, int src, int dst) = 0;
};
// Level operators exposed to the tests (orc_level_op).
enum LevelOp {
  OP_PROLONG = 0,        // pbuf(level, dst) += prolongation of pbuf(level + 1, src)
  OP_RESTRICT_MEAN = 1,  // div(level + 1) = 8-child mean of buf(level, src)
  OP_RESTRICT_FULL = 2,  // div(level + 1) = 4 * (1,3,3,1)/8 restriction of buf(level, src)
  OP_JACOBI = 3,         // buf(level, dst) = one plain Jacobi sweep of buf(level, src) with rhs div(level)
};

inline int clampInt(double v, int lo, int hi) {
  // Math.max(lo, Math.min(hi, Math.round(v))) — CVS/types.ts:26-28 (Math.round = floor(v + 0.5)).
  double r = std::floor(v + 0.5);
  return int(std::max(double(lo), std::min(double(hi), r)));
}
inline fx_quality clampQuality(const fx_quality& q) {
  fx_quality o;
  o.pressure_interval = clampInt(q.pressure_interval, 1, 16);
  o.vorticity_interval = clampInt(q.vorticity_interval, 1, 16);
  o.fine_pre_iterations = clampInt(q.fine_pre_iterations, 0, 32);
  o.fine_post_iterations = clampInt(q.fine_post_iterations, 0, 32);
  o.mid_pre_iterations = clampInt(q.mid_pre_iterations, 0, 32);
  o.mid_post_iterations = clampInt(q.mid_post_iterations, 0, 32);
  o.coarse_iterations = clampInt(q.coarse_iterations, 0, 64);
  return o;
}

// One pressure level. In a sparse domain a level-l voxel J lies inside exactly
// one fine brick ((J << l) >> 3 per axis, l <= 3), and it is active iff that fine
// brick is: every level solves on exactly the fine active region.
template <class T> struct Level {
  Dims d;
  uint32_t BX, BY, BZ;
  std::vector<T> div, pA, pB;
  std::vector<uint8_t> bact;            // per-brick: brick holds an active voxel (sparse); all 1 when dense
  int shift = 0;                        // level index l
  const std::vector<uint8_t>* fine = nullptr;  // fine brick activity (sparse only)
  uint32_t FBX = 0, FBY = 0, FBZ = 0;
  bool on(uint32_t x, uint32_t y, uint32_t z) const {
    if (!fine) return bact[(x >> 3) + BX * ((y >> 3) + BY * (z >> 3))] != 0;
    uint32_t fx = std::min((x << shift) >> 3, FBX - 1), fy = std::min((y << shift) >> 3, FBY - 1),
             fz = std::min((z << shift) >> 3, FBZ - 1);
    return (*fine)[fx + FBX * (fy + size_t(FBY) * fz)] != 0;
  }
};

template <class T> struct Sim : Base {
  fx_create_desc desc;
  Dims d;
  uint32_t BX, BY, BZ;
  size_t NB;
  std::vector<T> S[2][4];  // [0] render set, [1] advected set; d, T, f, r
  std::vector<T> u[3], us[3], w[3], wmag, cmag, umag;
  std::vector<uint32_t> flagsInj, flagsRender;
  std::vector<uint8_t> bact;  // sparse compute set (per brick); dense: all 1
  std::vector<Level<T>> lv;
  std::vector<float> src;     // >= 1 source (Q13: 128-byte minimum buffer)
  uint32_t nsrc = 0;
  fx_runtime_params rt;
  fx_quality q;
  int mode;
  bool initialized = false;
  double startS = 0, lastS = 0;
  uint64_t stepIndex = 0;
  int lastPressure = 0;  // 0 = 'a', 1 = 'b'
  bool debugRequested = false;  // ORCH:191, set by getDebugBuffers (:316)
  bool stochastic = false;      // 16-bit stores of the passes round stochastically
  int curSet = 0;        // the reference's currentScalarSet (params.scalarSetIndex only)
  std::vector<double> resid;

  bool sparse() const { return desc.domain_mode == FX_DOMAIN_SPARSE; }
  // Rounds a scalar array to the storage precision (FP32 build only; the FP64
  // build is the unrounded reference).
  void round_store(std::vector<T>& a) const {
    if (sizeof(T) != sizeof(float) || desc.precision == FX_PRECISION_FP32) return;
    for (auto& v : a) v = T(round_storage(desc.precision, float(v)));
  }
  // The stores of one pass: stochastic rounding with that pass's sites, or nearest even.
  void round_set(int set, bool with_fuel, int pass) {
    if (sizeof(T) != sizeof(float) || desc.precision == FX_PRECISION_FP32) return;
    if (!stochastic) {
      round_store(S[set][0]); round_store(S[set][1]);
      if (with_fuel) round_store(S[set][2]);
      return;
    }
    const int prec = desc.precision;
    const uint32_t step = uint32_t(stepIndex);
    for (int f = 0; f < (with_fuel ? 3 : 2); ++f) {
      auto& a = S[set][f];
      const uint32_t site = uint32_t(4 * pass + f);
#pragma omp parallel for schedule(static)
      for (int64_t z = 0; z < d.D; ++z)
        for (uint32_t y = 0; y < d.H; ++y)
          for (uint32_t x = 0; x < d.W; ++x) {
            const size_t i = d.idx(x, y, uint32_t(z));
            const uint32_t bm = uint32_t(brickOf(x, y, uint32_t(z)) * 512u + (uint32_t(z) & 7u) * 64u + (y & 7u) * 8u + (x & 7u));
            const uint32_t h = sr_hash(bm, step, site);
            a[i] = T(sr_storage(prec, float(a[i]), h));
          }
    }
  }
  size_t brickOf(uint32_t x, uint32_t y, uint32_t z) const { return (x >> 3) + BX * ((y >> 3) + size_t(BY) * (z >> 3)); }
  bool on(uint32_t x, uint32_t y, uint32_t z) const { return bact[brickOf(x, y, z)] != 0; }

  explicit Sim(const fx_create_desc& de) : desc(de) {
    // storage type in desc.precision, the rounding mode apart (as libfenix does)
    stochastic = de.precision == FX_PRECISION_BF16 || de.precision == FX_PRECISION_FP16 || de.precision == FX_PRECISION_U16 ||
                 de.precision == FX_PRECISION_UF16;
    if (de.precision == FX_PRECISION_BF16_RNE) desc.precision = FX_PRECISION_BF16;
    if (de.precision == FX_PRECISION_FP16_RNE) desc.precision = FX_PRECISION_FP16;
    if (de.precision == FX_PRECISION_U16_RNE) desc.precision = FX_PRECISION_U16;
    if (de.precision == FX_PRECISION_UF16_RNE) desc.precision = FX_PRECISION_UF16;
    d = Dims{de.width, de.height, de.depth};
    BX = (d.W + 7) / 8; BY = (d.H + 7) / 8; BZ = (d.D + 7) / 8;  // createSparseBrickLayout.ts:11-26
    NB = size_t(BX) * BY * BZ;
    size_t N = d.N();
    for (int s = 0; s < 2; ++s) for (int f = 0; f < 4; ++f) S[s][f].assign(N, T(0));
    for (int c = 0; c < 3; ++c) { u[c].assign(N, T(0)); us[c].assign(N, T(0)); w[c].assign(N, T(0)); }
    wmag.assign(N, T(0)); cmag.assign(N, T(0)); umag.assign(N, T(0));
    flagsInj.assign(NB, 0); flagsRender.assign(NB, 0);
    bact.assign(NB, sparse() ? 0 : 1);
    int L = de.solver == FX_SOLVER_VCYCLE ? std::max(2, de.vcycle.levels) : 3;  // PSP:32-35
    Dims ld = d;
    for (int l = 0; l < L; ++l) {
      Level<T> v;
      v.d = ld;
      v.BX = (ld.W + 7) / 8; v.BY = (ld.H + 7) / 8; v.BZ = (ld.D + 7) / 8;
      v.div.assign(ld.N(), T(0)); v.pA.assign(ld.N(), T(0)); v.pB.assign(ld.N(), T(0));
      v.bact.assign(size_t(v.BX) * v.BY * v.BZ, sparse() ? 0 : 1);
      v.shift = l;
      lv.push_back(std::move(v));
      ld = Dims{halve(ld.W), halve(ld.H), halve(ld.D)};
    }
    rt = de.runtime;
    q = clampQuality(de.quality);
    mode = de.advection_mode;
    src.assign(FX_FLOATS_PER_EMITTER, 0.f);
  }

  // ---------------------------------------------------------- params --
  // ORCH:236-266: the JS side computes in double and stores to a Float32Array.
  fx_sim_params make_params(double elapsed, double stepS) const {
    fx_sim_params p;
    double eff = std::min(stepS, 1.0 / 15.0);
    p.time = float(elapsed - startS);
    p.delta_time = float(eff);
    p.previous_time = float(lastS - startS);
    p.scalar_set_index = float(curSet);
    p.wind[0] = rt.wind[0]; p.wind[1] = rt.wind[1]; p.wind[2] = rt.wind[2]; p.wind[3] = rt.wind_strength;
    p.gravity[0] = rt.gravity[0]; p.gravity[1] = rt.gravity[1]; p.gravity[2] = rt.gravity[2]; p.gravity[3] = rt.gravity_strength;
    p.buoyancy = rt.buoyancy; p.vorticity_strength = rt.vorticity_strength;
    p.dx = float(double(rt.world_size) / double(std::max(d.W, d.D)));
    p.world_size = rt.world_size;
    p.vorticity_mask[0] = rt.vorticity_constant_mask; p.vorticity_mask[1] = rt.vorticity_velocity_mask;
    p.vorticity_mask[2] = rt.vorticity_heat_mask; p.vorticity_mask[3] = rt.vorticity_density_mask;
    return p;
  }

  // ------------------------------------------------------ injection --
  // Bricks that sparse-mode injection visits: bricks overlapping the voxel
  // bounding box (+1 voxel) of every gate-active source's influence sphere.
  std::vector<uint8_t> injection_bricks(const fx_sim_params& p) const {
    std::vector<uint8_t> m(NB, 0);
    uint32_t ns = std::max<uint32_t>(nsrc, 1);
    for (uint32_t i = 0; i < ns; ++i) {
      SrcView s{&src[size_t(i) * FX_FLOATS_PER_EMITTER]};
      float age = p.time - s.start();
      if (!(age >= 0.f && s.duration() >= age)) continue;
      float r = s.radius();
      if (s.kind() == 3.f) {
        float vw = p.dx / std::max(p.world_size, 0.0001f);
        r = r + std::max(r * 0.24f, vw * 3.f);
      }
      if (!(r > 0.f)) continue;
      const float c[3] = {s.px(), s.py(), s.pz()};
      const uint32_t n[3] = {d.W, d.H, d.D};
      int lo[3], hi[3];
      bool empty = false;
      for (int a = 0; a < 3; ++a) {
        double l = std::floor((double(c[a]) - r) * n[a] - 0.5) - 1.0;
        double h = std::ceil((double(c[a]) + r) * n[a] - 0.5) + 1.0;
        l = std::max(l, 0.0); h = std::min(h, double(n[a]) - 1.0);
        if (l > h) { empty = true; break; }
        lo[a] = int(l) >> 3; hi[a] = int(h) >> 3;
      }
      if (empty) continue;
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x) m[x + BX * (y + size_t(BY) * z)] = 1;
    }
    return m;
  }
  bool inject_here(const std::vector<uint8_t>& inj, uint32_t x, uint32_t y, uint32_t z) const {
    return !sparse() || inj[brickOf(x, y, z)];
  }

  V3<T> posOf(uint32_t x, uint32_t y, uint32_t z) const {
    return V3<T>{(T(x) + T(0.5)) / T(d.W), (T(y) + T(0.5)) / T(d.H), (T(z) + T(0.5)) / T(d.D)};
  }

  // SH/inject-scalar.wgsl.ts:32-70
  void inject_scalar(const Params<T>& P, const std::vector<uint8_t>& inj) {
    Emit<T> E{P};
    uint32_t ns = std::max<uint32_t>(nsrc, 1);
    auto& D_ = S[0][0]; auto& Tm = S[0][1]; auto& F = S[0][2];
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < d.D; ++z)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          if (!inject_here(inj, x, y, uint32_t(z))) continue;
          size_t i = d.idx(x, y, uint32_t(z));
          V3<T> pos = posOf(x, y, uint32_t(z));
          T den = D_[i], tem = Tm[i], fu = F[i];
          for (uint32_t k = 0; k < ns; ++k) {
            SrcView s{&src[size_t(k) * FX_FLOATS_PER_EMITTER]};
            bool burst = s.kind() == 3.f;
            if (!(s.kind() == 0.f || burst)) continue;
            T gate = E.activeGate(s);
            if (gate < T(0.0001)) continue;
            V3<T> off = pos - V3<T>{T(s.px()), T(s.py()), T(s.pz())};
            T dist = len3(off);
            T radius = s.radius();
            if (dist > (burst ? E.sourceOuterLimit(radius) : radius)) continue;
            T weight = burst ? E.eruptive(pos, dist, s, T(0.5)) : Emit<T>::sphereFalloff(dist, radius, T(0.5));
            T noise = Emit<T>::emitterNoise(pos, T(s.nscale()), T(s.seed()));
            T modulation = mixT(T(1), noise, clampT(T(s.nmix()), T(0), T(1)));
            T release = burst ? T(1) / Emit<T>::sourceDuration(s) : T(1);
            T contribution = weight * modulation * release * P.dt;
            den = clamp01(den + T(s.yd()) * contribution);
            tem = clamp01(tem + T(s.yt()) * contribution);
            fu = clamp01(fu + T(s.yf()) * contribution);
          }
          D_[i] = den; Tm[i] = tem; F[i] = fu;
        }
    round_set(0, true, 0);
  }

  // SH/inject-velocity.wgsl.ts:46-96
  void inject_velocity(const Params<T>& P, const std::vector<uint8_t>& inj) {
    Emit<T> E{P};
    uint32_t ns = std::max<uint32_t>(nsrc, 1);
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < d.D; ++z)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          if (!inject_here(inj, x, y, uint32_t(z))) continue;
          size_t i = d.idx(x, y, uint32_t(z));
          V3<T> pos = posOf(x, y, uint32_t(z));
          V3<T> vel{u[0][i], u[1][i], u[2][i]};
          for (uint32_t k = 0; k < ns; ++k) {
            SrcView s{&src[size_t(k) * FX_FLOATS_PER_EMITTER]};
            bool burst = s.kind() == 3.f;
            if (!(s.kind() == 1.f || burst)) continue;
            T gate = E.activeGate(s);
            if (gate < T(0.0001)) continue;
            V3<T> off = pos - V3<T>{T(s.px()), T(s.py()), T(s.pz())};
            T dist = len3(off);
            T radius = s.radius();
            if (dist > (burst ? E.sourceOuterLimit(radius) : radius)) continue;
            T weight = burst ? E.eruptive(pos, dist, s, T(s.falloff())) : Emit<T>::sphereFalloff(dist, radius, T(s.falloff()));
            T speed = s.speed();
            T mode_ = s.mode();
            V3<T> dir{0, 0, 0};
            V3<T> up{T(0), T(1), T(0)};
            if (burst) {
              V3<T> radial = dist > T(0.0001) ? off / std::max(dist, T(0.0001)) : up;
              V3<T> turb = Emit<T>::turbulentDir(pos + V3<T>{P.time * T(0.19), -P.time * T(0.13), P.time * T(0.07)}, T(s.seed()) + T(83));
              V3<T> bv = radial * speed + V3<T>{T(0), T(s.mode()), T(0)} + turb * T(s.vw());
              V3<T> impulse = bv * weight * P.dt / Emit<T>::sourceDuration(s);
              vel = vel + impulse;
              continue;
            } else if (mode_ == T(0)) {
              dir = dist > T(0.0001) ? off / std::max(dist, T(0.0001)) : up;
            } else if (mode_ == T(1)) {
              V3<T> dd{T(s.dx_()), T(s.dy_()), T(s.dz_())};
              dir = dot3(dd, dd) > T(0.0001) ? norm3(dd) : up;
            } else {
              dir = Emit<T>::turbulentDir(pos, T(s.seed()));
            }
            V3<T> force = dir * speed * weight * P.dt;
            T tight = clampT(T(s.vw()), T(0), T(1)) * weight;
            vel = mix3(vel + force, dir * speed, tight);
          }
          T lim = T(80) / std::max(P.dx, T(0.001));
          u[0][i] = clampT(vel.x, -lim, lim); u[1][i] = clampT(vel.y, -lim, lim); u[2][i] = clampT(vel.z, -lim, lim);
        }
  }

  // SH/inject-igniter.wgsl.ts:32-66
  void inject_igniter(const Params<T>& P, const std::vector<uint8_t>& inj) {
    Emit<T> E{P};
    uint32_t ns = std::max<uint32_t>(nsrc, 1);
    auto& R = S[0][3]; auto& Tm = S[0][1]; auto& F = S[0][2];
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < d.D; ++z)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          if (!inject_here(inj, x, y, uint32_t(z))) continue;
          size_t i = d.idx(x, y, uint32_t(z));
          V3<T> pos = posOf(x, y, uint32_t(z));
          T rea = R[i], tem = Tm[i];
          for (uint32_t k = 0; k < ns; ++k) {
            SrcView s{&src[size_t(k) * FX_FLOATS_PER_EMITTER]};
            bool burst = s.kind() == 3.f;
            if (!(s.kind() == 2.f || burst)) continue;
            T gate = E.activeGate(s);
            if (gate < T(0.0001)) continue;
            T dist = len3(pos - V3<T>{T(s.px()), T(s.py()), T(s.pz())});
            T radius = s.radius();
            if (dist > (burst ? E.sourceOuterLimit(radius) : radius)) continue;
            T weight = burst ? E.eruptive(pos, dist, s, T(0.6)) : Emit<T>::sphereFalloff(dist, radius, T(0.6));
            T intensity = s.intensity();
            T avail = burst ? T(1) : smooth(T(0.02), T(0.22), F[i]);
            T release = burst ? T(1) / Emit<T>::sourceDuration(s) : T(1.6);
            T seedv = intensity * weight * avail * P.dt * release;
            rea = clamp01(rea + seedv);
            tem = clamp01(tem + intensity * weight * P.dt * T(4));
          }
          R[i] = rea; Tm[i] = tem;
        }
    round_set(0, false, 1);
  }

  // ------------------------------------------------------ activation --
  // SH/compute-active-bricks.wgsl.ts:24-79 (REFERENCE) or the THRESHOLD rule.
  void activate(int set, const std::vector<T>* vel, std::vector<uint32_t>& flags) {
    std::fill(flags.begin(), flags.end(), 0u);  // ClearPass before ActiveBricksPass (ORCH:271-272)
    const auto& Dn = S[set][0]; const auto& Tm = S[set][1]; const auto& R = S[set][3];
    std::vector<int8_t> dil(NB, -1);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < int64_t(NB); ++b) {
      uint32_t bx = uint32_t(b % BX), by = uint32_t((b / BX) % BY), bz = uint32_t(b / (size_t(BX) * BY));
      uint32_t base[3] = {bx * 8, by * 8, bz * 8};
      if (desc.activation_rule == FX_ACTIVATION_REFERENCE) {
        bool active = false;
        T peak = T(0);
        for (uint32_t z = 0; z < 8; z += 2)
          for (uint32_t y = 0; y < 8; y += 2)
            for (uint32_t x = 0; x < 8; x += 2) {
              uint32_t cx = base[0] + x, cy = base[1] + y, cz = base[2] + z;
              if (cx < d.W && cy < d.H && cz < d.D) {
                size_t i = d.idx(cx, cy, cz);
                T e = Dn[i] + Tm[i] * T(0.7) + R[i] * T(0.9);
                peak = std::max(peak, e);
                if (e > T(0.012)) active = true;
              }
            }
        uint32_t cx = std::min(base[0] + 4, d.W - 1), cy = std::min(base[1] + 4, d.H - 1), cz = std::min(base[2] + 4, d.D - 1);
        size_t ci = d.idx(cx, cy, cz);
        T ce = Dn[ci] + Tm[ci] * T(0.7) + R[ci] * T(0.9);
        peak = std::max(peak, ce);
        if (ce > T(0.012)) active = true;
        if (active) dil[b] = (peak > T(0.05) || ce > T(0.035)) ? 1 : 0;
      } else {
        // Density-active bricks dilate by one brick; velocity-only bricks stay
        // active without dilating (velocity exists only inside the active set,
        // so dilating on it would grow the set by one ring every step).
        bool dens = false, fast = false;
        T dth = T(desc.activation_density), sth = T(desc.activation_speed);
        for (uint32_t z = 0; z < 8; ++z)
          for (uint32_t y = 0; y < 8; ++y)
            for (uint32_t x = 0; x < 8; ++x) {
              uint32_t cx = base[0] + x, cy = base[1] + y, cz = base[2] + z;
              if (cx >= d.W || cy >= d.H || cz >= d.D) continue;
              size_t i = d.idx(cx, cy, cz);
              V3<T> v{vel[0][i], vel[1][i], vel[2][i]};
              if (Dn[i] > dth) dens = true;
              if (len3(v) > sth) fast = true;
            }
        if (dens) dil[b] = 1;
        else if (fast) dil[b] = 0;
      }
    }
    for (size_t b = 0; b < NB; ++b) {
      if (dil[b] < 0) continue;
      int r = dil[b];
      int bx = int(b % BX), by = int((b / BX) % BY), bz = int(b / (size_t(BX) * BY));
      for (int z = -r; z <= r; ++z)
        for (int y = -r; y <= r; ++y)
          for (int x = -r; x <= r; ++x) {
            int nx = std::min(std::max(bx + x, 0), int(BX) - 1);
            int ny = std::min(std::max(by + y, 0), int(BY) - 1);
            int nz = std::min(std::max(bz + z, 0), int(BZ) - 1);
            flags[nx + BX * (ny + size_t(BY) * nz)] = 1u;
          }
    }
  }
  bool flagOn(const std::vector<uint32_t>& f, uint32_t x, uint32_t y, uint32_t z) const {
    // isActiveCoord (SC/active-brick-mask.wgsl.ts:13-18); brick coord clamped to counts-1.
    uint32_t bx = std::min(x / 8, BX - 1), by = std::min(y / 8, BY - 1), bz = std::min(z / 8, BZ - 1);
    return f[bx + BX * (by + size_t(BY) * bz)] != 0u;
  }

  // ------------------------------------------------------- advection --
  V3<T> readU(const std::vector<T>* v, size_t i) const { return V3<T>{v[0][i], v[1][i], v[2][i]}; }
  V3<T> sampleU(const std::vector<T>* v, V3<T> p) const {
    Tri<T> t = tri(d, p);
    return V3<T>{tri_eval(t, v[0].data()), tri_eval(t, v[1].data()), tri_eval(t, v[2].data())};
  }

  // SH/advect-velocity-semilagrangian.wgsl.ts:23-41
  void advect_velocity(const Params<T>& P) {
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < d.D; ++z)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          if (!on(x, y, uint32_t(z))) continue;
          size_t i = d.idx(x, y, uint32_t(z));
          if (d.boundary(x, y, uint32_t(z))) { us[0][i] = us[1][i] = us[2][i] = T(0); continue; }
          V3<T> coord{T(x) + T(0.5), T(y) + T(0.5), T(z) + T(0.5)};
          V3<T> half{T(0.5), T(0.5), T(0.5)};
          V3<T> back = coord - sampleU(u, coord - half) * P.dt;
          V3<T> adv = sampleU(u, back - half) * (T(1) - P.dt * T(0.025));
          us[0][i] = adv.x; us[1][i] = adv.y; us[2][i] = adv.z;
        }
  }

  // thermal noise shared by SH/advect-scalar-maccormack.wgsl.ts:52-57 and semilagrangian :69-74
  static T thermalNoise(V3<T> pos, T time) {
    V3<T> p = pos + V3<T>{time * T(1.7), -time * T(3.1), time * T(1.1)};
    T a = T(0.5) + T(0.5) * std::sin(dot3(p, V3<T>{T(0.097), T(0.061), T(-0.083)}));
    T b = T(0.5) + T(0.5) * std::sin(dot3(p, V3<T>{T(-0.043), T(0.137), T(0.071)}) + std::sin(p.y * T(0.079)) * T(1.6));
    return clampT(a * T(0.58) + b * T(0.42), T(0), T(1));
  }

  // decay terms, SH/advect-scalar-maccormack.wgsl.ts:75-100
  void decay_write(size_t i, uint32_t y, T den, T tem, T fu, T rea, T tn, const Params<T>& P) {
    T ny = (T(y) + T(0.5)) / T(d.H);
    T hot = smooth(T(0.22), T(0.78), tem + fu * T(0.2) + rea * T(0.26));
    T coolingPocket = smooth(T(0.44), T(0.9), T(1) - tn) * smooth(T(0.08), T(0.74), den) * (T(1) - hot * T(0.72));
    T cooling = P.dt * (T(0.016) + ny * T(0.052)) * mixT(T(1.28), T(0.58), hot) * mixT(T(0.86), T(1.9), coolingPocket);
    T smokeLoss = P.dt * (T(0.008) + ny * T(0.026)) * mixT(T(0.75), T(1.38), tn);
    T coldBreakup = P.dt * den * (T(1) - smooth(T(0.2), T(0.86), hot)) * smooth(T(0.18), T(0.72), den) *
                    smooth(T(0.34), T(0.92), tn) * T(0.34);
    T clearing = P.dt * den * hot * mixT(T(0.1), T(0.72), tn);
    S[1][0][i] = clamp01(std::max(T(0), den * (T(1) - smokeLoss) - clearing - coldBreakup));
    S[1][1][i] = clamp01(std::max(T(0), tem - cooling));
    S[1][2][i] = clamp01(std::max(T(0), fu - P.dt * mixT(T(0.012), T(0.034), tn)));
    S[1][3][i] = clamp01(std::max(T(0), rea - P.dt * mixT(T(0.46), T(0.82), tn)));
  }

  // SH/advect-scalar-maccormack.wgsl.ts:33-101 and SH/advect-scalar-semilagrangian.wgsl.ts:26-106:
  // set 0 -> set 1 advected by velocityScratch (Q10); inactive bricks (flagsInj) -> 0 (Q7).
  void advect_scalars(const Params<T>& P) {
    const std::vector<T>* vel = us;
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < d.D; ++z)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          if (!on(x, y, uint32_t(z))) continue;
          size_t i = d.idx(x, y, uint32_t(z));
          if (!flagOn(flagsInj, x, y, uint32_t(z))) {
            for (int f = 0; f < 4; ++f) S[1][f][i] = T(0);
            continue;
          }
          V3<T> coord{T(x) + T(0.5), T(y) + T(0.5), T(z) + T(0.5)};
          V3<T> half{T(0.5), T(0.5), T(0.5)};
          if (mode == FX_ADVECT_SEMI_LAGRANGIAN) {
            V3<T> back = coord - sampleU(vel, coord - half) * P.dt;
            V3<T> sp = back - half;
            Tri<T> t = tri(d, sp);
            T vals[4];
            for (int f = 0; f < 4; ++f) vals[f] = tri_eval(t, S[0][f].data());
            decay_write(i, y, vals[0], vals[1], vals[2], vals[3], thermalNoise(sp, P.time), P);
          } else {
            // advect<Name>(coord): identical positions for the four scalars.
            V3<T> velocity = sampleU(vel, coord - half);
            V3<T> back = coord - velocity * P.dt;
            V3<T> sp = back - half;
            Tri<T> tp = tri(d, sp);
            V3<T> fwd = sampleU(vel, sp);
            V3<T> rec = back + fwd * P.dt;
            Tri<T> tr = tri(d, rec - half);
            Tri<T> to = tri(d, coord - half);
            T vals[4];
            for (int f = 0; f < 4; ++f) {
              const T* F = S[0][f].data();
              T predicted = tri_eval(tp, F);
              T recovered = tri_eval(tr, F);
              T original = tri_eval(to, F);
              T corrected = predicted + T(0.5) * (original - recovered);
              vals[f] = clampT(corrected, std::min(original, predicted), std::max(original, predicted));
            }
            decay_write(i, y, vals[0], vals[1], vals[2], vals[3], thermalNoise(coord, P.time), P);
          }
        }
    round_set(1, true, 2);
  }

  // ------------------------------------------------------ combustion --
  // SH/combustion-update.wgsl.ts:27-93, read-old: set 1 -> set 0.
  static T thermalCluster(V3<T> c, T time) {
    V3<T> p = c + V3<T>{time * T(2.2), time * T(5.2), -time * T(1.7)};
    T body = T(0.5) + T(0.5) * std::sin(dot3(p, V3<T>{T(0.071), T(0.093), T(-0.064)}) + std::sin(p.y * T(0.057)) * T(1.8));
    T pocket = T(0.5) + T(0.5) * std::sin(dot3(p, V3<T>{T(0.143), T(-0.081), T(0.119)}) + std::sin(p.x * T(0.11) + p.z * T(0.07)) * T(1.2));
    T rising = T(0.5) + T(0.5) * std::sin(p.x * T(0.18) + p.y * T(0.13) - p.z * T(0.16) + time * T(1.4));
    return clampT(body * T(0.46) + pocket * T(0.34) + rising * T(0.2), T(0), T(1));
  }
  void combustion(const Params<T>& P) {
    const auto& Dn = S[1][0]; const auto& Tm = S[1][1]; const auto& F = S[1][2]; const auto& R = S[1][3];
    auto hotgas = [&](uint32_t x, uint32_t y, uint32_t z) {
      size_t k = d.cidx(x, y, z);
      return clampT(Tm[k] * T(1.55) + F[k] * T(0.26) + R[k] * T(0.44), T(0), T(1));
    };
    const T ilo = T(desc.ignition_lo), ihi = T(desc.ignition_hi), rate = T(desc.burn_rate);
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < d.D; ++zz)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!on(x, y, z)) continue;
          size_t i = d.idx(x, y, z);
          if (!flagOn(flagsRender, x, y, z)) {
            for (int f = 0; f < 4; ++f) S[0][f][i] = S[1][f][i];
            continue;
          }
          T dt = P.dt;
          T ny = (T(y) + T(0.5)) / T(d.H);
          T cluster = thermalCluster(V3<T>{T(x), T(y), T(z)}, P.time);
          T nh = (hotgas(inc(x, d.W - 1), y, z) + hotgas(dec(x), y, z) + hotgas(x, inc(y, d.H - 1), z) +
                  hotgas(x, dec(y), z) + hotgas(x, y, inc(z, d.D - 1)) + hotgas(x, y, dec(z))) * T(0.16666667);
          T localBefore = hotgas(x, y, z);
          T den = Dn[i], tem = Tm[i], fu = F[i], rea = R[i];
          T radiative = std::max(nh - localBefore * T(0.55), T(0)) * smooth(T(0.04), T(0.76), den) *
                        (T(0.45) + (T(1) - cluster) * T(0.55)) * dt * T(2.6);
          T hotPocket = smooth(T(0.58), T(0.92), cluster + rea * T(0.22));
          T coolingPocket = smooth(T(0.42), T(0.82), T(1) - cluster) * smooth(T(0.08), T(0.68), den) *
                            (T(1) - smooth(T(0.54), T(0.94), rea));
          T ignition = smooth(ilo, ihi, tem);
          T burn = std::min(fu, fu * ignition * dt * mixT(T(4.2), T(8.6), hotPocket) * rate);
          T hotSoot = smooth(T(0.24), T(0.88), tem + rea * T(0.2)) * smooth(T(0.36), T(0.9), T(1) - cluster + den * T(0.18));
          T smokeYield = burn * mixT(T(0.16), T(0.58), coolingPocket + (T(1) - ignition) * T(0.5) + hotSoot * T(0.45));
          T hotCooling = mixT(T(1.3), T(0.62), hotPocket);
          T heatLoss = dt * (T(0.018) + ny * T(0.052)) * hotCooling * mixT(T(1), T(2.25), coolingPocket);
          T densityLoss = dt * den * (T(0.02) + ny * T(0.07));
          T clearing = den * smooth(T(0.18), T(0.82), tem + rea * T(0.22) + fu * T(0.1)) * dt * mixT(T(0.55), T(2.45), hotPocket);
          T coolingSmoke = den * tem * (coolingPocket + hotSoot * T(0.32)) * dt * T(0.16);
          T reactionDecay = rea * (T(1) - dt * mixT(T(0.68), T(0.22), hotPocket));
          S[0][2][i] = clamp01(fu - burn);
          S[0][1][i] = clamp01(tem + burn * mixT(T(2.1), T(3.25), hotPocket) + radiative - heatLoss);
          S[0][0][i] = clamp01(std::max(T(0), den + smokeYield + coolingSmoke - densityLoss - clearing * T(0.72)));
          S[0][3][i] = clamp01(std::max(reactionDecay, burn * mixT(T(4), T(7.5), hotPocket) + radiative * T(1.8)));
        }
    round_set(0, true, 3);
  }

  // -------------------------------------------------------- buoyancy --
  // SH/apply-buoyancy.wgsl.ts:27-110 (set 0, velocityScratch in place, flagsRender).
  void buoyancy(const Params<T>& P) {
    const auto& Dn = S[0][0]; const auto& Tm = S[0][1]; const auto& R = S[0][3];
    auto hot = [&](uint32_t x, uint32_t y, uint32_t z) {
      size_t k = d.cidx(x, y, z);
      return clampT(Tm[k] * T(1) + R[k] * T(0.85), T(0), T(1));
    };
    auto rd = [&](uint32_t x, uint32_t y, uint32_t z) { return Dn[d.cidx(x, y, z)]; };
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < d.D; ++zz)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!on(x, y, z)) continue;
          if (!flagOn(flagsRender, x, y, z)) continue;
          size_t i = d.idx(x, y, z);
          T dt = P.dt;
          T den = Dn[i];
          T response = P.buoyancy * (T(1) + smooth(T(6), T(24), std::fabs(P.buoyancy)) * T(0.38));
          T hotGas = hot(x, y, z);
          T lift = hotGas * response;
          V3<T> g{P.gravity[0], P.gravity[1], P.gravity[2]};
          V3<T> gravityForce = g * den * P.gravity[3];
          T entrain = den * smooth(T(0.04), T(0.3), hotGas);
          T ny = (T(y) + T(0.5)) / T(d.H);
          T shear = std::sin(T(y) * T(0.073) + P.time * T(0.9)) * T(0.35) +
                    std::sin(T(x) * T(0.041) - T(z) * T(0.052) + P.time * T(0.43)) * T(0.22);
          T windAmount = P.wind[3] * (T(0.18) + den * T(0.55) + ny * T(0.45));
          V3<T> wdir{P.wind[0], P.wind[1], P.wind[2]};
          V3<T> wind = wdir * windAmount + V3<T>{-P.wind[2], T(0), P.wind[0]} * windAmount * shear;
          V3<T> tg = V3<T>{hot(inc(x, d.W - 1), y, z) - hot(dec(x), y, z), hot(x, inc(y, d.H - 1), z) - hot(x, dec(y), z),
                           hot(x, y, inc(z, d.D - 1)) - hot(x, y, dec(z))} * T(0.5);
          V3<T> dg = V3<T>{rd(inc(x, d.W - 1), y, z) - rd(dec(x), y, z), rd(x, inc(y, d.H - 1), z) - rd(x, dec(y), z),
                           rd(x, y, inc(z, d.D - 1)) - rd(x, y, dec(z))} * T(0.5);
          T dgl = std::max(len3(dg), T(0.0001));
          T coldSmoke = den * (T(1) - smooth(T(0.12), T(0.72), hotGas));
          T surfaceBreakup = smooth(T(0.008), T(0.06), dgl) * coldSmoke;
          V3<T> shred = norm3(V3<T>{std::sin(T(y) * T(0.117) + P.time * T(1.9)),
                                    std::sin(T(x) * T(0.071) - T(z) * T(0.083) + P.time * T(1.35)) * T(0.35),
                                    std::cos(T(x) * T(0.061) + T(y) * T(0.049) - P.time * T(1.6))});
          T collapse = den * (T(1) - smooth(T(0.16), T(0.64), hotGas));
          T smokeDrag = den * smooth(T(0.03), T(0.45), hotGas);
          V3<T> v = V3<T>{us[0][i], us[1][i], us[2][i]} * (T(1) - dt * T(0.08));
          v = v + wind * dt;
          v = v + gravityForce * dt;
          v = v + tg * collapse * dt * (T(3.2) + std::fabs(P.buoyancy) * T(0.08));
          v = v + (-dg / dgl) * surfaceBreakup * dt * T(2.8);
          v = v + shred * coldSmoke * dt * (T(0.72) + surfaceBreakup * T(2.4));
          v = v + wdir * coldSmoke * P.wind[3] * dt * T(0.9);
          v = v + V3<T>{T(0), smokeDrag * response * T(0.55), T(0)} * dt;
          T lim = T(80) / std::max(P.dx, T(0.001));
          v.y = clampT(v.y + (lift + entrain * response * T(0.24)) * dt, -lim, lim);
          if (x == 0 || x == d.W - 1 || z == 0 || z == d.D - 1) { v.x = T(0); v.z = T(0); }
          if (y == 0) v.y = std::max(v.y, T(0));
          us[0][i] = v.x; us[1][i] = v.y; us[2][i] = v.z;
        }
  }

  // ------------------------------------------------ vorticity / conf --
  // SH/compute-vorticity.wgsl.ts:23-50
  void vorticity() {
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < d.D; ++zz)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!on(x, y, z)) continue;
          size_t i = d.idx(x, y, z);
          if (d.boundary(x, y, z)) { w[0][i] = w[1][i] = w[2][i] = T(0); wmag[i] = T(0); continue; }
          V3<T> x0 = readU(us, d.cidx(dec(x), y, z)), x1 = readU(us, d.cidx(inc(x, d.W - 1), y, z));
          V3<T> y0 = readU(us, d.cidx(x, dec(y), z)), y1 = readU(us, d.cidx(x, inc(y, d.H - 1), z));
          V3<T> z0 = readU(us, d.cidx(x, y, dec(z))), z1 = readU(us, d.cidx(x, y, inc(z, d.D - 1)));
          V3<T> curl = V3<T>{(y1.z - y0.z) - (z1.y - z0.y), (z1.x - z0.x) - (x1.z - x0.z), (x1.y - x0.y) - (y1.x - y0.x)} * T(0.5);
          w[0][i] = curl.x; w[1][i] = curl.y; w[2][i] = curl.z;
          wmag[i] = len3(curl);
        }
  }
  // SH/apply-vorticity-confinement.wgsl.ts:30-66 (set 0)
  void confinement(const Params<T>& P) {
    const auto& Dn = S[0][0]; const auto& Tm = S[0][1]; const auto& R = S[0][3];
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < d.D; ++zz)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!on(x, y, z)) continue;
          size_t i = d.idx(x, y, z);
          if (d.boundary(x, y, z)) { cmag[i] = T(0); continue; }
          auto m = [&](uint32_t a, uint32_t b, uint32_t c) { return wmag[d.cidx(a, b, c)]; };
          V3<T> grad = V3<T>{m(inc(x, d.W - 1), y, z) - m(dec(x), y, z), m(x, inc(y, d.H - 1), z) - m(x, dec(y), z),
                             m(x, y, inc(z, d.D - 1)) - m(x, y, dec(z))} * T(0.5);
          V3<T> dir = grad / std::max(len3(grad), T(0.0001));
          T hotGas = clampT(Tm[i] + R[i] * T(0.85), T(0), T(1));
          T mask = clampT(P.vmask[0] + smooth(T(0.02), T(0.38), wmag[i]) * P.vmask[1] +
                              smooth(T(0.04), T(0.7), hotGas) * P.vmask[2] + smooth(T(0.04), T(0.65), Dn[i]) * P.vmask[3],
                          T(0), T(4));
          V3<T> force = cross3(dir, V3<T>{w[0][i], w[1][i], w[2][i]}) * P.vort * mask * T(1.45);
          V3<T> v = V3<T>{us[0][i], us[1][i], us[2][i]} + force * P.dt;
          T lim = T(80) / std::max(P.dx, T(0.001));
          us[0][i] = clampT(v.x, -lim, lim); us[1][i] = clampT(v.y, -lim, lim); us[2][i] = clampT(v.z, -lim, lim);
          cmag[i] = len3(force);
        }
  }

  // ----------------------------------------------- divergence/pressure --
  // SH/compute-divergence.wgsl.ts:22-42 into the fine level's divergence.
  void divergence() {
    auto& div = lv[0].div;
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < d.D; ++zz)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!on(x, y, z)) continue;
          size_t i = d.idx(x, y, z);
          if (d.boundary(x, y, z)) { div[i] = T(0); continue; }
          T dx = us[0][d.cidx(inc(x, d.W - 1), y, z)] - us[0][d.cidx(dec(x), y, z)];
          T dy = us[1][d.cidx(x, inc(y, d.H - 1), z)] - us[1][d.cidx(x, dec(y), z)];
          T dz = us[2][d.cidx(x, y, inc(z, d.D - 1))] - us[2][d.cidx(x, y, dec(z))];
          div[i] = (dx + dy + dz) * T(0.5);
        }
  }

  std::vector<T>& pbuf(Level<T>& L, int id) { return id == 0 ? L.pA : L.pB; }

  // SH/pressure-jacobi.wgsl.ts:15-36 (omega = 1) or its weighted form (V-cycle).
  void jacobi(Level<T>& L, const std::vector<T>& rhs, const std::vector<T>& s, std::vector<T>& t, bool weighted, T omega) {
    const Dims& g = L.d;
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < g.D; ++zz)
      for (uint32_t y = 0; y < g.H; ++y)
        for (uint32_t x = 0; x < g.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!L.on(x, y, z)) continue;
          size_t i = g.idx(x, y, z);
          if (g.boundary(x, y, z)) { t[i] = T(0); continue; }
          T sum = s[g.cidx(inc(x, g.W - 1), y, z)] + s[g.cidx(dec(x), y, z)] + s[g.cidx(x, inc(y, g.H - 1), z)] +
                  s[g.cidx(x, dec(y), z)] + s[g.cidx(x, y, inc(z, g.D - 1))] + s[g.cidx(x, y, dec(z))];
          T pj = (sum - rhs[i]) / T(6);
          t[i] = weighted ? mixT(s[i], pj, omega) : pj;
        }
  }
  // SH/pressure-residual.wgsl.ts:16-38
  double residual(Level<T>& L, const std::vector<T>& rhs, const std::vector<T>& s, std::vector<T>& t) {
    const Dims& g = L.d;
    double acc = 0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (int64_t zz = 0; zz < g.D; ++zz)
      for (uint32_t y = 0; y < g.H; ++y)
        for (uint32_t x = 0; x < g.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!L.on(x, y, z)) continue;
          size_t i = g.idx(x, y, z);
          if (g.boundary(x, y, z)) { t[i] = T(0); continue; }
          T sum = s[g.cidx(inc(x, g.W - 1), y, z)] + s[g.cidx(dec(x), y, z)] + s[g.cidx(x, inc(y, g.H - 1), z)] +
                  s[g.cidx(x, dec(y), z)] + s[g.cidx(x, y, inc(z, g.D - 1))] + s[g.cidx(x, y, dec(z))];
          T lap = sum - s[i] * T(6);
          t[i] = rhs[i] - lap;
          acc += double(t[i]) * double(t[i]);
        }
    return std::sqrt(acc);
  }
  // SH/pressure-restriction.wgsl.ts:27-50 (8-child mean) or the (1,3,3,1)/8 stencil scaled by 4.
  void restrict_(Level<T>& F, Level<T>& C, const std::vector<T>& fine, std::vector<T>& coarse, bool fullWeight) {
    const Dims& f = F.d; const Dims& c = C.d;
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < c.D; ++zz)
      for (uint32_t y = 0; y < c.H; ++y)
        for (uint32_t x = 0; x < c.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!C.on(x, y, z)) continue;
          size_t i = c.idx(x, y, z);
          if (c.boundary(x, y, z)) { coarse[i] = T(0); continue; }
          if (!fullWeight) {
            T sum = T(0);
            for (uint32_t k = 0; k < 2; ++k)
              for (uint32_t j = 0; j < 2; ++j)
                for (uint32_t a = 0; a < 2; ++a) sum += fine[f.cidx(2 * x + a, 2 * y + j, 2 * z + k)];
            coarse[i] = sum * T(0.125);
          } else {
            static const T wt[4] = {T(0.125), T(0.375), T(0.375), T(0.125)};
            T sum = T(0);
            for (int k = 0; k < 4; ++k) {
              uint32_t fz = uint32_t(std::min<int64_t>(std::max<int64_t>(int64_t(2 * z) - 1 + k, 0), f.D - 1));
              for (int j = 0; j < 4; ++j) {
                uint32_t fy = uint32_t(std::min<int64_t>(std::max<int64_t>(int64_t(2 * y) - 1 + j, 0), f.H - 1));
                for (int a = 0; a < 4; ++a) {
                  uint32_t fx = uint32_t(std::min<int64_t>(std::max<int64_t>(int64_t(2 * x) - 1 + a, 0), f.W - 1));
                  sum += (wt[k] * wt[j] * wt[a]) * fine[f.idx(fx, fy, fz)];
                }
              }
            }
            coarse[i] = sum * T(4);
          }
        }
  }
  // SH/pressure-prolongation.wgsl.ts:29-44: fine.pA += trilinear(coarse, (id+0.5)*0.5-0.5).
  void prolong(Level<T>& C, Level<T>& F, const std::vector<T>& coarse, std::vector<T>& fine) {
    const Dims& c = C.d; const Dims& f = F.d;
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < f.D; ++zz)
      for (uint32_t y = 0; y < f.H; ++y)
        for (uint32_t x = 0; x < f.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!F.on(x, y, z)) continue;
          size_t i = f.idx(x, y, z);
          if (f.boundary(x, y, z)) { fine[i] = T(0); continue; }
          V3<T> cp = V3<T>{(T(x) + T(0.5)) * T(0.5) - T(0.5), (T(y) + T(0.5)) * T(0.5) - T(0.5), (T(z) + T(0.5)) * T(0.5) - T(0.5)};
          Tri<T> t = tri(c, cp);
          fine[i] += tri_eval(t, coarse.data());
        }
  }

  int smooth_ref(Level<T>& L, int iters, int start) {  // PSP:129-143
    int active = start;
    for (int k = 0; k < iters; ++k) {
      jacobi(L, L.div, pbuf(L, active), pbuf(L, 1 - active), false, T(1));
      active = 1 - active;
    }
    return active;
  }
  void clear_pressure() {  // PSP:62-66 — pressureA of every level
    for (auto& L : lv) std::fill(L.pA.begin(), L.pA.end(), T(0));
  }
  int solve_reference() {  // PSP:68-84, aliasing kept (Q1-Q4)
    Level<T>& F = lv[0]; Level<T>& M = lv[1]; Level<T>& C = lv[2];
    int finePre = smooth_ref(F, q.fine_pre_iterations, 0);
    residual(F, F.div, pbuf(F, finePre), pbuf(F, 1 - finePre));
    restrict_(F, M, F.div, M.div, false);
    int midPre = smooth_ref(M, q.mid_pre_iterations, 0);
    residual(M, M.div, pbuf(M, midPre), pbuf(M, 1 - midPre));
    restrict_(M, C, M.div, C.div, false);
    int co = smooth_ref(C, q.coarse_iterations, 0);
    prolong(C, M, pbuf(C, co), M.pA);
    int mid = smooth_ref(M, q.mid_post_iterations, midPre);
    prolong(M, F, pbuf(M, mid), F.pA);
    return smooth_ref(F, q.fine_post_iterations, finePre);
  }
  // Weighted-Jacobi V-cycle. Result in fine pA (returns 0).
  int solve_vcycle() {
    int L = int(lv.size());
    T om = T(desc.vcycle.omega);
    resid.clear();
    auto smoothW = [&](Level<T>& V, int n) {
      for (int k = 0; k < n; ++k) { jacobi(V, V.div, V.pA, V.pB, true, om); std::swap(V.pA, V.pB); }
    };
    for (int c = 0; c < desc.vcycle.cycles; ++c) {
      for (int l = 0; l < L - 1; ++l) {
        if (l > 0) std::fill(lv[l].pA.begin(), lv[l].pA.end(), T(0));
        smoothW(lv[l], desc.vcycle.pre_sweeps);
        double n = residual(lv[l], lv[l].div, lv[l].pA, lv[l].pB);
        if (l == 0) resid.push_back(n);
        restrict_(lv[l], lv[l + 1], lv[l].pB, lv[l + 1].div, true);
      }
      std::fill(lv[L - 1].pA.begin(), lv[L - 1].pA.end(), T(0));
      smoothW(lv[L - 1], desc.vcycle.coarse_sweeps);
      for (int l = L - 2; l >= 0; --l) {
        prolong(lv[l + 1], lv[l], lv[l + 1].pA, lv[l].pA);
        smoothW(lv[l], desc.vcycle.post_sweeps);
      }
    }
    return 0;
  }
  int pressure_solve() {
    clear_pressure();
    if (desc.solver == FX_SOLVER_VCYCLE) return solve_vcycle();
    return solve_reference();
  }

  // SH/project-velocity.wgsl.ts:21-43: velocityScratch - grad p -> velocityCurrent.
  void projection() {
    Level<T>& F = lv[0];
    const auto& p = pbuf(F, lastPressure);
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < d.D; ++zz)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!on(x, y, z)) continue;
          size_t i = d.idx(x, y, z);
          if (d.boundary(x, y, z)) { u[0][i] = u[1][i] = u[2][i] = T(0); continue; }
          T gx = (p[d.cidx(inc(x, d.W - 1), y, z)] - p[d.cidx(dec(x), y, z)]) * T(0.5);
          T gy = (p[d.cidx(x, inc(y, d.H - 1), z)] - p[d.cidx(x, dec(y), z)]) * T(0.5);
          T gz = (p[d.cidx(x, y, inc(z, d.D - 1))] - p[d.cidx(x, y, dec(z))]) * T(0.5);
          u[0][i] = us[0][i] - gx; u[1][i] = us[1][i] - gy; u[2][i] = us[2][i] - gz;
        }
  }
  // SH/compute-debug-fields.wgsl.ts:18-25
  void debug_fields() {
#pragma omp parallel for schedule(static)
    for (int64_t zz = 0; zz < d.D; ++zz)
      for (uint32_t y = 0; y < d.H; ++y)
        for (uint32_t x = 0; x < d.W; ++x) {
          uint32_t z = uint32_t(zz);
          if (!on(x, y, z)) continue;
          size_t i = d.idx(x, y, z);
          umag[i] = len3(readU(u, i));
        }
  }

  // ----------------------------------------------------------- sparse --
  // Coarsen the fine brick set through the hierarchy and zero every field
  // outside the active bricks.
  void apply_sparse_set(const std::vector<uint32_t>& flags) {
    for (size_t b = 0; b < NB; ++b) bact[b] = flags[b] ? 1 : 0;
    lv[0].bact = bact;
    for (size_t l = 1; l < lv.size(); ++l) {
      Level<T>& F = lv[l - 1]; Level<T>& C = lv[l];
      std::fill(C.bact.begin(), C.bact.end(), 0);
      for (uint32_t z = 0; z < C.BZ; ++z)
        for (uint32_t y = 0; y < C.BY; ++y)
          for (uint32_t x = 0; x < C.BX; ++x) {
            uint8_t a = 0;
            for (uint32_t k = 0; k < 2; ++k)
              for (uint32_t j = 0; j < 2; ++j)
                for (uint32_t i = 0; i < 2; ++i) {
                  uint32_t fx = 2 * x + i, fy = 2 * y + j, fz = 2 * z + k;
                  if (fx < F.BX && fy < F.BY && fz < F.BZ) a |= F.bact[fx + F.BX * (fy + size_t(F.BY) * fz)];
                }
            C.bact[x + C.BX * (y + size_t(C.BY) * z)] = a;
          }
    }
    for (auto& L : lv) { L.fine = &bact; L.FBX = BX; L.FBY = BY; L.FBZ = BZ; }
    auto zero_out = [&](std::vector<T>& f, const Dims& g, const std::vector<uint8_t>& m, uint32_t bx, uint32_t by) {
      for (uint32_t z = 0; z < g.D; ++z)
        for (uint32_t y = 0; y < g.H; ++y)
          for (uint32_t x = 0; x < g.W; ++x)
            if (!m[(x >> 3) + bx * ((y >> 3) + size_t(by) * (z >> 3))]) f[g.idx(x, y, z)] = T(0);
    };
    for (int s = 0; s < 2; ++s) for (int f = 0; f < 4; ++f) zero_out(S[s][f], d, bact, BX, BY);
    for (int c = 0; c < 3; ++c) { zero_out(u[c], d, bact, BX, BY); zero_out(us[c], d, bact, BX, BY); zero_out(w[c], d, bact, BX, BY); }
    zero_out(wmag, d, bact, BX, BY); zero_out(cmag, d, bact, BX, BY); zero_out(umag, d, bact, BX, BY);
    for (auto& L : lv)
      for (uint32_t z = 0; z < L.d.D; ++z)
        for (uint32_t y = 0; y < L.d.H; ++y)
          for (uint32_t x = 0; x < L.d.W; ++x)
            if (!L.on(x, y, z)) { size_t i = L.d.idx(x, y, z); L.div[i] = L.pA[i] = L.pB[i] = T(0); }
  }

  bool debug_each_frame() const {
    if (desc.debug_fields == 0) return false;
    if (desc.debug_fields == 1) return true;
    return d.N() < 1800000ull;  // ORCH:193
  }

  // ------------------------------------------------------------- step --
  int step(double elapsed, double stepS) override {  // ORCH:230-305
    if (!initialized) { initialized = true; startS = elapsed; lastS = elapsed; }
    fx_sim_params fp = make_params(elapsed, stepS);
    Params<T> P = Params<T>::from(fp);
    std::vector<uint8_t> inj = injection_bricks(fp);
    inject_scalar(P, inj);
    inject_velocity(P, inj);
    inject_igniter(P, inj);
    if (sparse()) {
      activate(0, u, flagsInj);
      flagsRender = flagsInj;
      apply_sparse_set(flagsInj);
    } else {
      activate(0, u, flagsInj);
    }
    advect_velocity(P);
    advect_scalars(P);
    if (!sparse()) activate(1, us, flagsRender);
    combustion(P);
    buoyancy(P);
    if (stepIndex % uint64_t(q.vorticity_interval) == 0) { vorticity(); confinement(P); }
    if (stepIndex % uint64_t(q.pressure_interval) == 0) {
      clear_pressure();
      divergence();
      lastPressure = pressure_solve();
    }
    projection();
    if (debug_each_frame() || debugRequested) {  // ORCH:298-301
      debug_fields();
      debugRequested = false;
    }
    curSet = curSet == 0 ? 1 : 0;
    lastS = elapsed;
    stepIndex += 1;
    return 0;
  }

  int run_pass(int pass, const fx_sim_params* fp) override {
    Params<T> P = Params<T>::from(*fp);
    std::vector<uint8_t> inj = injection_bricks(*fp);
    switch (pass) {
      case FX_PASS_INJECT_SCALAR: inject_scalar(P, inj); break;
      case FX_PASS_INJECT_VELOCITY: inject_velocity(P, inj); break;
      case FX_PASS_INJECT_IGNITER: inject_igniter(P, inj); break;
      case FX_PASS_ACTIVATE_RENDER:
        activate(0, u, flagsInj);
        if (sparse()) { flagsRender = flagsInj; apply_sparse_set(flagsInj); }
        break;
      case FX_PASS_ACTIVATE_ADVECTED: activate(1, us, flagsRender); break;
      case FX_PASS_ADVECT_VELOCITY: advect_velocity(P); break;
      case FX_PASS_ADVECT_SCALARS: advect_scalars(P); break;
      case FX_PASS_COMBUSTION: combustion(P); break;
      case FX_PASS_BUOYANCY: buoyancy(P); break;
      case FX_PASS_VORTICITY: vorticity(); break;
      case FX_PASS_CONFINEMENT: confinement(P); break;
      case FX_PASS_DIVERGENCE: divergence(); break;
      case FX_PASS_PRESSURE: lastPressure = pressure_solve(); break;
      case FX_PASS_PROJECTION: projection(); break;
      case FX_PASS_DEBUG_FIELDS: debug_fields(); break;
      default: return FX_ERR_INVALID_ARG;
    }
    return 0;
  }

  int reset() override {  // ORCH:206-226, 381-383
    curSet = 0; stepIndex = 0; lastPressure = 0; initialized = false;
    startS = lastS;
    for (int s = 0; s < 2; ++s) for (int f = 0; f < 4; ++f) std::fill(S[s][f].begin(), S[s][f].end(), T(0));
    for (int c = 0; c < 3; ++c) { std::fill(u[c].begin(), u[c].end(), T(0)); std::fill(us[c].begin(), us[c].end(), T(0)); std::fill(w[c].begin(), w[c].end(), T(0)); }
    std::fill(wmag.begin(), wmag.end(), T(0)); std::fill(cmag.begin(), cmag.end(), T(0)); std::fill(umag.begin(), umag.end(), T(0));
    std::fill(flagsInj.begin(), flagsInj.end(), 0u); std::fill(flagsRender.begin(), flagsRender.end(), 0u);
    clear_pressure();
    if (sparse()) {
      std::fill(bact.begin(), bact.end(), 0);
      for (auto& L : lv) { std::fill(L.div.begin(), L.div.end(), T(0)); std::fill(L.pB.begin(), L.pB.end(), T(0)); std::fill(L.bact.begin(), L.bact.end(), 0); }
    }
    return 0;
  }

  int update_sources(const float* packed, uint32_t n) override {
    nsrc = n;
    src.assign(size_t(std::max<uint32_t>(n, 1)) * FX_FLOATS_PER_EMITTER, 0.f);
    if (n) std::memcpy(src.data(), packed, size_t(n) * FX_FLOATS_PER_EMITTER * sizeof(float));
    return 0;
  }
  int set_runtime(const fx_runtime_params* p, uint32_t m) override {  // ORCH:353-365
    if (m & FX_RP_WIND) { rt.wind[0] = p->wind[0]; rt.wind[1] = p->wind[1]; rt.wind[2] = p->wind[2]; }
    if (m & FX_RP_WIND_STRENGTH) rt.wind_strength = p->wind_strength;
    if (m & FX_RP_GRAVITY) { rt.gravity[0] = p->gravity[0]; rt.gravity[1] = p->gravity[1]; rt.gravity[2] = p->gravity[2]; }
    if (m & FX_RP_GRAVITY_STRENGTH) rt.gravity_strength = p->gravity_strength;
    if (m & FX_RP_BUOYANCY) rt.buoyancy = p->buoyancy;
    if (m & FX_RP_VORTICITY_STRENGTH) rt.vorticity_strength = p->vorticity_strength;
    if (m & FX_RP_VORTICITY_CONSTANT_MASK) rt.vorticity_constant_mask = p->vorticity_constant_mask;
    if (m & FX_RP_VORTICITY_VELOCITY_MASK) rt.vorticity_velocity_mask = p->vorticity_velocity_mask;
    if (m & FX_RP_VORTICITY_HEAT_MASK) rt.vorticity_heat_mask = p->vorticity_heat_mask;
    if (m & FX_RP_VORTICITY_DENSITY_MASK) rt.vorticity_density_mask = p->vorticity_density_mask;
    if (m & FX_RP_WORLD_SIZE) rt.world_size = p->world_size;
    return 0;
  }
  int set_quality(const fx_quality* qq) override { q = clampQuality(*qq); return 0; }
  int set_mode(int m) override { mode = m; return 0; }

  std::vector<T>* field(int id) {
    switch (id) {
      case FX_FIELD_DENSITY: case FX_FIELD_TEMPERATURE: case FX_FIELD_FUEL: case FX_FIELD_REACTION: return &S[0][id];
      case FX_FIELD_VELOCITY_X: case FX_FIELD_VELOCITY_Y: case FX_FIELD_VELOCITY_Z: return &u[id - FX_FIELD_VELOCITY_X];
      case FX_FIELD_SCRATCH_X: case FX_FIELD_SCRATCH_Y: case FX_FIELD_SCRATCH_Z: return &us[id - FX_FIELD_SCRATCH_X];
      case FX_FIELD_VORTICITY_X: case FX_FIELD_VORTICITY_Y: case FX_FIELD_VORTICITY_Z: return &w[id - FX_FIELD_VORTICITY_X];
      case FX_FIELD_VORTICITY_MAG: return &wmag;
      case FX_FIELD_CONFINEMENT_MAG: return &cmag;
      case FX_FIELD_VELOCITY_MAG: return &umag;
      case FX_FIELD_DIVERGENCE: return &lv[0].div;
      case FX_FIELD_PRESSURE: return &pbuf(lv[0], lastPressure);
      case FX_FIELD_ADV_DENSITY: case FX_FIELD_ADV_TEMPERATURE: case FX_FIELD_ADV_FUEL: case FX_FIELD_ADV_REACTION:
        return &S[1][id - FX_FIELD_ADV_DENSITY];
      default: return nullptr;
    }
  }
  int read_field(int id, double* out) override {
    auto* f = field(id);
    if (!f) return FX_ERR_INVALID_ARG;
    for (size_t i = 0; i < f->size(); ++i) out[i] = double((*f)[i]);
    return 0;
  }
  int write_field(int id, const double* in) override {
    auto* f = field(id);
    if (!f) return FX_ERR_INVALID_ARG;
    for (size_t i = 0; i < f->size(); ++i) (*f)[i] = T(in[i]);
    if ((id >= FX_FIELD_DENSITY && id <= FX_FIELD_FUEL) || (id >= FX_FIELD_ADV_DENSITY && id <= FX_FIELD_ADV_FUEL))
      round_store(*f);
    return 0;
  }
  int read_flags(int which, uint32_t* out) override {
    const auto& f = which == FX_FLAGS_INJECTED ? flagsInj : flagsRender;
    std::memcpy(out, f.data(), f.size() * 4);
    return 0;
  }
  int read_level(int level, int which, double* out) override {
    if (level < 0 || level >= int(lv.size())) return FX_ERR_INVALID_ARG;
    const auto& L = lv[level];
    const auto& f = which == FX_LEVEL_DIVERGENCE ? L.div : which == FX_LEVEL_PRESSURE_A ? L.pA : L.pB;
    for (size_t i = 0; i < f.size(); ++i) out[i] = double(f[i]);
    return 0;
  }
  int level_dims(int level, uint32_t* dims) override {
    if (level < 0 || level >= int(lv.size())) return FX_ERR_INVALID_ARG;
    dims[0] = lv[level].d.W; dims[1] = lv[level].d.H; dims[2] = lv[level].d.D;
    return 0;
  }
  int residuals(double* out, int cap) override {
    int n = std::min(cap, int(resid.size()));
    for (int i = 0; i < n; ++i) out[i] = resid[i];
    return int(resid.size());
  }
  std::vector<T>& lbuf(Level<T>& L, int which) { return which == FX_LEVEL_DIVERGENCE ? L.div : which == FX_LEVEL_PRESSURE_A ? L.pA : L.pB; }
  int write_level(int level, int which, const double* in) override {
    if (level < 0 || level >= int(lv.size()) || which < 0 || which > 2) return FX_ERR_INVALID_ARG;
    auto& f = lbuf(lv[level], which);
    for (size_t i = 0; i < f.size(); ++i) f[i] = T(in[i]);
    return 0;
  }
  int level_op(int op, int level, int src, int dst) override {
    const bool needs_coarse = op != OP_JACOBI;
    if (level < 0 || level + (needs_coarse ? 1 : 0) >= int(lv.size()) || src < 0 || src > 2 || dst < 0 || dst > 2) return FX_ERR_INVALID_ARG;
    switch (op) {
      case OP_PROLONG: prolong(lv[level + 1], lv[level], lbuf(lv[level + 1], src), lbuf(lv[level], dst)); return 0;
      case OP_RESTRICT_MEAN: restrict_(lv[level], lv[level + 1], lbuf(lv[level], src), lv[level + 1].div, false); return 0;
      case OP_RESTRICT_FULL: restrict_(lv[level], lv[level + 1], lbuf(lv[level], src), lv[level + 1].div, true); return 0;
      case OP_JACOBI: {
        std::vector<T> out = lbuf(lv[level], dst);
        jacobi(lv[level], lv[level].div, lbuf(lv[level], src), out, false, T(1));
        lbuf(lv[level], dst) = out;
        return 0;
      }
      default: return FX_ERR_INVALID_ARG;
    }
  }
  int request_debug() override { debugRequested = true; return 0; }
  int active_count() override {
    int n = 0;
    for (size_t b = 0; b < NB; ++b) n += bact[b] ? 1 : 0;
    return n;
  }

  // -------------------------------------------------------- raymarch --
  // RP:361-724 over the render set and render flags.
  int raymarch(const fx_camera* cam, const fx_render_params* rp, uint32_t Wpx, uint32_t Hpx, float* out) override;
};

// ---------------------------------------------------------- raymarch --
template <class T> int Sim<T>::raymarch(const fx_camera* cam, const fx_render_params* rp, uint32_t Wpx, uint32_t Hpx, float* out) {
  // Host packing, RP:83-88, 196-262; volume bounds volumeWorldBounds.ts:14-31.
  double maxH = double(std::max(d.W, d.D));
  float hx = float(2.05 * 13.5 * (d.W / maxH)), hy = float(2.05 * 13.5 * (d.H / maxH)), hz = float(2.05 * 13.5 * (d.D / maxH));
  float cy = float(double(hy) - 0.25);
  V3<T> camPos{cam->position[0], cam->position[1], cam->position[2]};
  V3<T> right{cam->right[0], cam->right[1], cam->right[2]};
  V3<T> upv{cam->up[0], cam->up[1], cam->up[2]};
  V3<T> fwd{cam->forward[0], cam->forward[1], cam->forward[2]};
  V3<T> center{T(0), T(cy), T(0)};
  V3<T> half{T(hx), T(hy), T(hz)};
  T aspect = cam->aspect, tanh_ = cam->tan_half_fov_y;
  T mode_ = T(float(rp->display_mode));
  uint32_t stepCount = uint32_t(float(rp->step_count));
  T lightCountF = T(float(rp->light_count));
  T sF = rp->scattering_forward, sB = rp->scattering_back;
  struct Lt { V3<T> dir; T kind; V3<T> pos; T intensity; V3<T> color; };
  Lt lights[3];
  for (int i = 0; i < 3; ++i) {
    if (i < rp->light_count) {
      const fx_light& l = rp->lights[i];
      lights[i] = Lt{V3<T>{l.direction[0], l.direction[1], l.direction[2]}, T(l.type == 1 ? 1 : 0),
                     V3<T>{l.position[0], l.position[1], l.position[2]}, T(l.intensity), V3<T>{l.color[0], l.color[1], l.color[2]}};
    } else {
      lights[i] = Lt{V3<T>{0, 0, 0}, T(0), V3<T>{0, 0, 0}, T(0), V3<T>{0, 0, 0}};
    }
  }
  const V3<T> dims{T(float(d.W)) - T(1), T(float(d.H)) - T(1), T(float(d.D)) - T(1)};  // sampleBounds, RP:361-363
  const auto& Dn = S[0][0]; const auto& Tm = S[0][1]; const auto& Fu = S[0][2]; const auto& Re = S[0][3];
  const auto& flags = flagsRender;
  auto fl = [&](uint32_t x, uint32_t y, uint32_t z) { return size_t(x) + size_t(d.W) * (size_t(y) + size_t(d.H) * z); };
  auto isActive = [&](V3<T> p) {  // RP:378-381
    uint32_t x = uint32_t(std::floor(clampT(p.x, T(0), dims.x))), y = uint32_t(std::floor(clampT(p.y, T(0), dims.y))),
             z = uint32_t(std::floor(clampT(p.z, T(0), dims.z)));
    return flagOn(flags, x, y, z);
  };
  auto distNext = [&](V3<T> vp, V3<T> rd) {  // RP:383-396
    T far = T(1e9), ext = T(8);
    V3<T> bc{std::floor(vp.x / ext), std::floor(vp.y / ext), std::floor(vp.z / ext)};
    V3<T> bnd{rd.x > T(0) ? (bc.x + T(1)) * ext : bc.x * ext, rd.y > T(0) ? (bc.y + T(1)) * ext : bc.y * ext,
              rd.z > T(0) ? (bc.z + T(1)) * ext : bc.z * ext};
    T tx = std::fabs(rd.x) > T(0.00001) ? (bnd.x - vp.x) / rd.x : far;
    T ty = std::fabs(rd.y) > T(0.00001) ? (bnd.y - vp.y) / rd.y : far;
    T tz = std::fabs(rd.z) > T(0.00001) ? (bnd.z - vp.z) / rd.z : far;
    return std::min(tx, std::min(ty, tz)) * len3(rd) + T(0.5);
  };
  auto sampleAll = [&](V3<T> p, T* o) {  // RP:398-441
    V3<T> c{clampT(p.x, T(0), dims.x), clampT(p.y, T(0), dims.y), clampT(p.z, T(0), dims.z)};
    uint32_t bx = uint32_t(std::floor(c.x)), by = uint32_t(std::floor(c.y)), bz = uint32_t(std::floor(c.z));
    uint32_t ux = std::min(bx + 1, uint32_t(dims.x)), uy = std::min(by + 1, uint32_t(dims.y)), uz = std::min(bz + 1, uint32_t(dims.z));
    T fx = fractT(c.x), fy = fractT(c.y), fz = fractT(c.z);
    size_t i000 = fl(bx, by, bz), i100 = fl(ux, by, bz), i010 = fl(bx, uy, bz), i110 = fl(ux, uy, bz);
    size_t i001 = fl(bx, by, uz), i101 = fl(ux, by, uz), i011 = fl(bx, uy, uz), i111 = fl(ux, uy, uz);
    const std::vector<T>* fs[4] = {&Dn, &Tm, &Fu, &Re};
    for (int k = 0; k < 4; ++k) {
      const auto& F = *fs[k];
      T x00 = mixT(F[i000], F[i100], fx), x10 = mixT(F[i010], F[i110], fx);
      T x01 = mixT(F[i001], F[i101], fx), x11 = mixT(F[i011], F[i111], fx);
      o[k] = mixT(mixT(x00, x10, fy), mixT(x01, x11, fy), fz);
    }
  };
  auto nearestD = [&](V3<T> p) {  // RP:443-446 (floor then clamp on u32)
    auto cl = [](T v, T mx) { T f = std::max(std::floor(v), T(0)); return std::min(uint32_t(f), uint32_t(mx)); };
    return Dn[fl(cl(p.x, dims.x), cl(p.y, dims.y), cl(p.z, dims.z))];
  };
  auto densGrad = [&](V3<T> p) {  // RP:448-463 (round = nearest even in WGSL)
    uint32_t mx = uint32_t(dims.x), my = uint32_t(dims.y), mz = uint32_t(dims.z);
    // WGSL f32 -> u32 saturates at 0; round() is half-to-even.
    auto rc = [](T v, uint32_t m) { T r = std::max(std::nearbyint(v), T(0)); return std::min(uint32_t(r), m); };
    uint32_t x = rc(p.x, mx), y = rc(p.y, my), z = rc(p.z, mz);
    size_t xp = fl(std::min(x + 1, mx), y, z), xn = fl(x == 0 ? 0 : x - 1, y, z);
    size_t yp = fl(x, std::min(y + 1, my), z), yn = fl(x, y == 0 ? 0 : y - 1, z);
    size_t zp = fl(x, y, std::min(z + 1, mz)), zn = fl(x, y, z == 0 ? 0 : z - 1);
    return V3<T>{Dn[xp] - Dn[xn], Dn[yp] - Dn[yn], Dn[zp] - Dn[zn]};
  };
  auto firePalette = [&](T t) { return fire_palette(t); };
  auto smokeAlbedo = [&](T den, T tem, T hot) { return smoke_albedo(den, tem, hot); };
  auto phaseHG = [&](T c, T g) { return phase_hg(c, g); };
  auto lightDir = [&](int i, V3<T> wp) { return light_dir(lights[i].kind, lights[i].dir, lights[i].pos, wp); };
  auto lightAtt = [&](int i, V3<T> wp) { return light_att(lights[i].kind, lights[i].pos, wp); };

#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t py = 0; py < int64_t(Hpx); ++py)
    for (uint32_t px = 0; px < Wpx; ++px) {
      float* o = out + (size_t(py) * Wpx + px) * 4;
      o[0] = o[1] = o[2] = o[3] = 0.f;
      // Fragment at the pixel centre; uv.y grows upwards (row 0 = top).
      T uvx = (T(px) + T(0.5)) / T(Wpx), uvy = T(1) - (T(py) + T(0.5)) / T(Hpx);
      T ndcx = uvx * T(2) - T(1), ndcy = uvy * T(2) - T(1);
      V3<T> rayDir = norm3(fwd + right * ndcx * aspect * tanh_ + upv * ndcy * tanh_);
      V3<T> bmin = center - half, bmax = center + half;
      V3<T> inv{T(1) / rayDir.x, T(1) / rayDir.y, T(1) / rayDir.z};
      V3<T> t0 = mul(bmin - camPos, inv), t1 = mul(bmax - camPos, inv);
      V3<T> tmn{std::min(t0.x, t1.x), std::min(t0.y, t1.y), std::min(t0.z, t1.z)};
      V3<T> tmx{std::max(t0.x, t1.x), std::max(t0.y, t1.y), std::max(t0.z, t1.z)};
      T bx = std::max(tmn.x, std::max(tmn.y, tmn.z)), by = std::min(tmx.x, std::min(tmx.y, tmx.z));
      if (bx > by) continue;  // discard
      bx = std::max(bx, T(0));
      V3<T> invExt{T(1) / (half.x * T(2)), T(1) / (half.y * T(2)), T(1) / (half.z * T(2))};
      V3<T> vrd = mul(mul(rayDir, invExt), dims);
      T voxDelta = len3(vrd) * (by - bx) / std::max(T(stepCount), T(1));
      T worldDelta = (by - bx) / std::max(T(stepCount), T(1));
      uint32_t nl = uint32_t(clampT(lightCountF, T(0), T(3)));
      V3<T> primary = nl > 0 ? lightDir(0, camPos) : V3<T>{T(0), T(1), T(0)};
      V3<T> lightStep = mul(primary, dims) * T(0.036);
      V3<T> vp = mul(mul(camPos + rayDir * bx - bmin, invExt), dims);
      T travelled = T(0);
      T rayLen = len3(vrd) * (by - bx);
      V3<T> acc{0, 0, 0};
      T alpha = T(0);
      V3<T> nrd = norm3(vrd);
      for (uint32_t k = 0; k < stepCount && travelled <= rayLen; ++k) {
        if (alpha > T(0.985)) break;
        T vstep = voxDelta;
        V3<T> sp = vp;
        V3<T> uvw{sp.x / dims.x, sp.y / dims.y, sp.z / dims.z};
        if (isActive(sp)) {
          T s[4];
          sampleAll(sp, s);
          T den = std::max(s[0] - T(0.004), T(0));
          T tem = s[1], fu = s[2], rea = s[3];
          T hot = clampT(tem * T(1.45) + fu * T(0.22) + rea * T(0.38), T(0), T(1));
          if (den > T(0.001) || hot > T(0.018)) {
            if (den < T(0.012) && hot < T(0.04)) {
              vstep = voxDelta * T(1.8);
            } else {
              V3<T> dg = densGrad(sp);
              T gl = std::max(len3(dg), T(0.0001));
              V3<T> normal = gl > T(0.00011) ? (-dg) / gl : V3<T>{T(0), T(1), T(0)};
              T detail = clampT(gl * T(5.2) + hot * T(0.28) + rea * T(0.18), T(0), T(1));
              T dd = std::max(std::pow(den, T(0.96)) * mixT(T(0.82), T(1.42), detail), T(0));
              V3<T> wp = V3<T>{mixT(bmin.x, bmax.x, uvw.x), mixT(bmin.y, bmax.y, uvw.y), mixT(bmin.z, bmax.z, uvw.z)};
              T topFade = T(1) - smooth(T(0.9), T(0.995), uvw.y);
              V3<T> total{0, 0, 0};
              T primT = T(1);
              if (nl > 0) {  // primaryLightTransmittance, RP:534-549
                auto cl3 = [&](V3<T> p) { return V3<T>{clampT(p.x, T(0), dims.x), clampT(p.y, T(0), dims.y), clampT(p.z, T(0), dims.z)}; };
                V3<T> p1 = cl3(sp + lightStep * T(0.9)), p2 = cl3(sp + lightStep * T(2)), p3 = cl3(sp + lightStep * T(3.6));
                T d1 = std::max(nearestD(p1) - T(0.004), T(0)), d2 = std::max(nearestD(p2) - T(0.004), T(0)),
                  d3 = std::max(nearestD(p3) - T(0.004), T(0));
                T od = dd * T(0.34) + d1 * T(1.15) + d2 * T(0.72) + d3 * T(0.4) - hot * T(0.06);
                primT = std::exp(-std::max(od, T(0)) * T(2.15));
              }
              for (uint32_t li = 0; li < 3; ++li) {
                if (li >= nl) break;
                V3<T> dl = lightDir(int(li), wp);
                T power = light_power(lights[li].intensity, lightAtt(int(li), wp));
                V3<T> color = lights[li].color;
                T lvc = clampT(dot3(dl, -rayDir), T(-1), T(1));
                T fs = phaseHG(lvc, sF) * T(5.2) * power;
                T bs = phaseHG(lvc, sB) * T(1.55) * power;
                T nf = clampT(dot3(normal, dl) * T(0.5) + T(0.5), T(0), T(1));
                T surface = std::pow(nf, T(1.35)) * mixT(T(1), T(0.56), smooth(T(0.16), T(0.9), dd));
                T rim = std::pow(T(1) - std::fabs(dot3(normal, -rayDir)), T(2.4)) *
                        std::pow(clampT(dot3(dl, -rayDir) * T(0.5) + T(0.5), T(0), T(1)), T(1.2));
                T directional = (surface * T(0.42) + rim * T(0.22)) * power;
                T scatter = (fs + bs) * mixT(T(0.36), T(0.74), hot);
                T trans = li > 0 ? T(1) : primT;
                total = total + color * trans * (directional + scatter);
              }
              V3<T> tl{T(1) - std::exp(-total.x * T(0.72)), T(1) - std::exp(-total.y * T(0.72)), T(1) - std::exp(-total.z * T(0.72))};
              V3<T> ambient{T(0.045), T(0.048), T(0.052)};
              V3<T> smoke = smokeAlbedo(dd, tem, hot);
              T absorb = mixT(T(0.92), T(0.34), smooth(T(0.26), T(1.05), dd));
              // thermalScatter, RP:551-556
              T heat = clampT(tem * T(1.25) + fu * T(0.18) + rea * T(0.34), T(0), T(1));
              T flame = clampT(tem * T(0.9) + fu * T(0.12) + rea * T(0.18), T(0), T(1));
              T catch_ = smooth(T(0.012), T(0.38), dd);
              V3<T> heatScatter = firePalette(flame) * heat * heat * catch_ * T(0.18);
              V3<T> lit = mul(smoke, ambient) + mul(smoke, tl) * absorb + tl * smooth(T(0.018), T(0.22), dd) * T(0.18) + heatScatter;
              T flameHeat = clampT(tem * T(0.9) + fu * T(0.14) + rea * T(0.16), T(0), T(1));
              V3<T> emissive = firePalette(flameHeat) * hot * (T(0.48) + fu * T(0.16) + rea * T(0.24));
              T contribution = clampT(dd * T(1.35) + hot * T(1.25), T(0), T(1));
              T stepDistance = worldDelta * mixT(T(1.02), T(0.76), contribution);
              vstep = voxDelta * mixT(T(1.02), T(0.76), contribution);
              V3<T> comp = lit + emissive;
              T opacity = (T(1) - std::exp(-(dd * T(4.75) + hot * T(0.42)) * stepDistance)) * topFade;
              if (mode_ > T(0.5) && mode_ < T(1.5)) {
                comp = mix3(V3<T>{T(0.01), T(0.01), T(0.012)}, V3<T>{T(0.88), T(0.9), T(0.94)}, clampT(dd * T(1.4), T(0), T(1)));
                opacity = T(1) - std::exp(-dd * stepDistance * T(8.6));
              } else if (mode_ >= T(1.5)) {
                comp = mix3(V3<T>{T(0.04), T(0.012), T(0)}, V3<T>{T(1), T(0.72), T(0.22)}, clampT(fu * T(1.3), T(0), T(1)));
                opacity = T(1) - std::exp(-fu * stepDistance * T(7));
              }
              acc = acc + comp * ((T(1) - alpha) * opacity);
              alpha = alpha + (T(1) - alpha) * opacity;
            }
          } else {
            vstep = voxDelta * T(1.35);
          }
        } else {
          vstep = std::max(distNext(sp, vrd), voxDelta);
        }
        vp = vp + nrd * vstep;
        travelled += vstep;
      }
      if (alpha < T(0.01)) continue;  // discard
      o[0] = float(acc.x); o[1] = float(acc.y); o[2] = float(acc.z); o[3] = float(alpha);
    }
  return 0;
}

}  // namespace orc

// Evaluates one raymarch shading helper in FP64 (bits 64) or FP32 (bits 32):
//   fn 0 phase_hg(in[0] = cos, in[1] = g) -> out[0]
//   fn 1 fire_palette(in[0]) -> out[0..2]
//   fn 2 smoke_albedo(in[0] den, in[1] tem, in[2] hot) -> out[0..2]
//   fn 3 light_dir(in[0] kind, in[1..3] dir, in[4..6] pos, in[7..9] world position) -> out[0..2]
//   fn 4 light_att(in[0] kind, in[1..3] pos, in[4..6] world position) -> out[0]
//   fn 5 light_power(in[0] intensity, in[1] attenuation) -> out[0]
//   fn 6 hash31(in[0..2]) -> out[0]
//   fn 7 emitterNoise(in[0..2] position, in[3] scale, in[4] seed) -> out[0]
//   fn 8 / 9 stochastic rounding to bf16 / f16 of in[0] with hash in[1] -> out[0]
//   fn 10 the rounding hash of (in[0] voxel, in[1] step, in[2] site) -> out[0]
//   fn 11 / 12 stochastic (hash in[1]) / nearest-even rounding of in[0] to 16-bit fixed point -> out[0]
namespace orc {
template <class T> int probe_t(int fn, const double* in, double* out) {
  using orc::V3;
  auto v3 = [&](int k) { return V3<T>{T(in[k]), T(in[k + 1]), T(in[k + 2])}; };
  auto put3 = [&](V3<T> v) { out[0] = double(v.x); out[1] = double(v.y); out[2] = double(v.z); };
  switch (fn) {
    case 0: out[0] = double(orc::phase_hg(T(in[0]), T(in[1]))); return 0;
    case 1: put3(orc::fire_palette(T(in[0]))); return 0;
    case 2: put3(orc::smoke_albedo(T(in[0]), T(in[1]), T(in[2]))); return 0;
    case 3: put3(orc::light_dir(T(in[0]), v3(1), v3(4), v3(7))); return 0;
    case 4: out[0] = double(orc::light_att(T(in[0]), v3(1), v3(4))); return 0;
    case 5: out[0] = double(orc::light_power(T(in[0]), T(in[1]))); return 0;
    case 6: out[0] = double(orc::Emit<T>::hash31(v3(0))); return 0;
    case 7: out[0] = double(orc::Emit<T>::emitterNoise(v3(0), T(in[3]), T(in[4]))); return 0;
    case 8: out[0] = double(orc::sr_bf16(float(in[0]), uint32_t(in[1]))); return 0;
    case 9: out[0] = double(orc::sr_f16(float(in[0]), uint32_t(in[1]))); return 0;
    case 10: out[0] = double(orc::sr_hash(uint32_t(in[0]), uint32_t(in[1]), uint32_t(in[2]))); return 0;
    case 11: out[0] = double(orc::sr_u16(float(in[0]), uint32_t(in[1]))); return 0;
    case 12: out[0] = double(orc::round_u16(float(in[0]))); return 0;
    case 13: out[0] = double(orc::sr_uf16(float(in[0]), uint32_t(in[1]))); return 0;
    case 14: out[0] = double(orc::round_uf16(float(in[0]))); return 0;
    default: return FX_ERR_INVALID_ARG;
  }
}
}  // namespace orc

// ------------------------------------------------------------- C ABI --
extern "C" {
void* orc_create(int bits, const fx_create_desc* desc) {
  if (!desc || desc->width < 4 || desc->height < 4 || desc->depth < 4) return nullptr;
  if (bits == 64) return static_cast<orc::Base*>(new orc::Sim<double>(*desc));
  return static_cast<orc::Base*>(new orc::Sim<float>(*desc));
}
void orc_destroy(void* h) { delete static_cast<orc::Base*>(h); }
int orc_update_sources(void* h, const float* p, uint32_t n) { return static_cast<orc::Base*>(h)->update_sources(p, n); }
int orc_set_runtime(void* h, const fx_runtime_params* p, uint32_t m) { return static_cast<orc::Base*>(h)->set_runtime(p, m); }
int orc_set_quality(void* h, const fx_quality* q) { return static_cast<orc::Base*>(h)->set_quality(q); }
int orc_set_advection_mode(void* h, int m) { return static_cast<orc::Base*>(h)->set_mode(m); }
int orc_step(void* h, double e, double s) { return static_cast<orc::Base*>(h)->step(e, s); }
int orc_reset(void* h) { return static_cast<orc::Base*>(h)->reset(); }
int orc_run_pass(void* h, int pass, const fx_sim_params* p) { return static_cast<orc::Base*>(h)->run_pass(pass, p); }
int orc_read_field(void* h, int id, double* out) { return static_cast<orc::Base*>(h)->read_field(id, out); }
int orc_write_field(void* h, int id, const double* in) { return static_cast<orc::Base*>(h)->write_field(id, in); }
int orc_read_flags(void* h, int which, uint32_t* out) { return static_cast<orc::Base*>(h)->read_flags(which, out); }
int orc_read_level(void* h, int l, int which, double* out) { return static_cast<orc::Base*>(h)->read_level(l, which, out); }
int orc_level_dims(void* h, int l, uint32_t* dims) { return static_cast<orc::Base*>(h)->level_dims(l, dims); }
int orc_residuals(void* h, double* out, int cap) { return static_cast<orc::Base*>(h)->residuals(out, cap); }
int orc_active_count(void* h) { return static_cast<orc::Base*>(h)->active_count(); }
int orc_raymarch(void* h, const fx_camera* c, const fx_render_params* rp, uint32_t w, uint32_t hh, float* out) {
  return static_cast<orc::Base*>(h)->raymarch(c, rp, w, hh, out);
}
int orc_request_debug(void* h) { return static_cast<orc::Base*>(h)->request_debug(); }
int orc_write_level(void* h, int l, int which, const double* in) { return static_cast<orc::Base*>(h)->write_level(l, which, in); }
int orc_level_op(void* h, int op, int l, int src, int dst) { return static_cast<orc::Base*>(h)->level_op(op, l, src, dst); }
int orc_probe(int bits, int fn, const double* in, double* out) {
  return bits == 64 ? orc::probe_t<double>(fn, in, out) : orc::probe_t<float>(fn, in, out);
}
}
