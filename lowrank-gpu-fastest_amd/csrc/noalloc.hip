// noalloc.hip -- the whole device path with its workspace in a static device
// array: for callers that cannot allocate device memory (code under stream
// capture, tight serving loops, harnesses that forbid hipMalloc). Host code
// plus one file-scope device array.
//
// The array `scratch` holds what mppca_denoise_device takes as its workspace
// (centred volume, means, Gram matrices, factors or projectors, weights, ranks,
// margins, status). Each GPU the library runs on holds its own copy, at its own
// address, from the first use until the process ends. In host code the name
// `scratch` is not a device address: its address on a device comes from
// hipGetSymbolAddress, once per device (with that device current), and only
// that pointer is passed to kernels.
#include <atomic>
#include <mutex>

#include "common.h"

namespace mppca {
namespace {

constexpr size_t kScratchBytes = (size_t)MPPCA_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::atomic<void*> g_scratch_ptr[kMaxScratchDevices];
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up once per device).
int scratch_on_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_err((int)e, "noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return set_err(MPPCA_EUNSUPPORTED, "noalloc: device %d beyond %d", dev, kMaxScratchDevices);
  void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
    if (!p) {
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return set_err(e != hipSuccess ? (int)e : MPPCA_EINVAL,
                       "noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return MPPCA_OK;
}

}  // namespace
}  // namespace mppca

using namespace mppca;

extern "C" {

int64_t mppca_noalloc_scratch_bytes(void) { return (int64_t)kScratchBytes; }

int mppca_denoise_device_noalloc(const float* vol, float* out, int X, int Y, int Z, int T, int px,
                                 int py, int pz, int stride, int form, int32_t* rank,
                                 float* margin, int32_t* info, void* stream) {
  Geom g;
  int rc = make_geom(&g, X, Y, Z, T, px, py, pz, stride);
  if (rc) return rc;
  if (!valid_form(form)) return set_err(MPPCA_EINVAL, "unknown form %d", form);
  if (!vol || !out) return set_err(MPPCA_EINVAL, "denoise (noalloc): null pointer");
  if (vol == out) return set_err(MPPCA_EINVAL, "denoise (noalloc): vol and out must be distinct");
  const size_t need = carve_workspace(nullptr, g, form).bytes;
  if (need > kScratchBytes)
    return set_err(MPPCA_ENOSPACE,
                   "denoise (noalloc): this geometry needs %zu bytes of workspace, above the "
                   "static scratch of %zu bytes; use mppca_denoise_device",
                   need, kScratchBytes);
  void* base = nullptr;
  if ((rc = scratch_on_current_device(&base))) return rc;
  const Workspace w = carve_workspace((char*)base, g, form);
  return run_device_path(vol, out, g, form, w, rank, margin, info, (hipStream_t)stream);
}

}  // extern "C"
