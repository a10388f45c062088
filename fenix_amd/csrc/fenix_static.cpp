// fenix_static.cpp — entry points that allocate no device memory.
//
// fx_bloom_device allocates its three half-resolution images on the first call
// and again whenever the canvas size changes. fx_bloom_device_noalloc runs the
// same kernels on a fixed, statically sized device array instead, for callers
// that must not allocate: code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "fenix_internal.h"

namespace {

// Half-resolution images of the bloom: bright pass, horizontal blur, vertical
// blur, RGBA float16 (8 bytes per pixel) each. 1024 x 1024 half-resolution
// pixels cover canvases up to 2048 x 2048: 24 MiB per GPU, held until the
// process ends.
constexpr size_t kHalfPixels = size_t(1024) * 1024;
constexpr size_t kImages = 3;
constexpr size_t kScratchBytes = kHalfPixels * 8 * kImages;

// The array's name is a host-side stand-in for the device symbol: host code uses
// only the per-device address that hipGetSymbolAddress returns for it.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

namespace fx {

// Resolves the scratch array's address on `device`, which must be current.
// Called once per device from fx_create; later calls return the cached address.
int static_scratch_setup(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return FX_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) return FX_ERR_DEVICE;
    g_scratch[device] = p;
  }
  if (out) *out = g_scratch[device];
  return FX_OK;
}

}  // namespace fx

using namespace fx;

extern "C" int fx_bloom_device_noalloc(fx_sim* s, const fx_bloom_params* p, uint32_t w, uint32_t h, const void* src,
                                       void* dst) {
  if (!s || !p || !src || !dst) return sim_fail(s, FX_ERR_INVALID_ARG, "null argument");
  if (fx_scene_count(s) > 1)  // scenes of a batch render on several streams
    return sim_fail(s, FX_ERR_UNSUPPORTED, "bloom runs on a simulation of one scene");
  if (w == 0 || h == 0) return sim_fail(s, FX_ERR_INVALID_ARG, "image size out of range");
  const uint32_t hw = w / 2 > 0 ? w / 2 : 1, hh = h / 2 > 0 ? h / 2 : 1;  // createBloomPass.ts:84-89
  if (size_t(hw) * hh > kHalfPixels) {
    char msg[160];
    std::snprintf(msg, sizeof msg,
                  "bloom without allocation: a %u x %u canvas needs %u x %u half-resolution pixels, above the "
                  "%zu of the static scratch (canvases up to 2048 x 2048)",
                  w, h, hw, hh, kHalfPixels);
    return sim_fail(s, FX_ERR_INVALID_ARG, msg);
  }
  const int device = sim_device(s);
  if (hipSetDevice(device) != hipSuccess) return sim_fail(s, FX_ERR_DEVICE, "hipSetDevice failed");
  void* base = nullptr;
  const int rc = static_scratch_setup(device, &base);  // cached since fx_create
  if (rc != FX_OK) return sim_fail(s, rc, "static bloom scratch unavailable on this device");
  BloomArgs a{p->enabled ? 1 : 0, p->threshold, p->strength, p->radius, w, h, hw, hh};
  char* b = static_cast<char*>(base);
  const size_t img = size_t(hw) * hh * 8;
  launch_bloom(sim_render_stream(s), a, src, b, b + img, b + 2 * img, dst);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sim_fail(s, FX_ERR_DEVICE, hipGetErrorString(e));
  return FX_OK;
}
