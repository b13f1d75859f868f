// Allocation-free variant of pii_form_factor_device.
//
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops, harnesses that forbid it): the same form factor,
// written into the caller's device buffer, with every intermediate buffer of
// the call -- descriptor, DoFs, world vertices, element and moment records,
// element-split partial sums, representative-node values and the expanded
// result -- carved from one fixed device array per GPU instead of the
// context's grow-only workspaces, and the descriptor / DoF upload staged
// through a fixed host array instead of pinned memory.  No hipMalloc,
// hipFree or hipHostMalloc is called.
//
// The array holds kScratchBytes per GPU until the process ends.  Calls on one
// GPU share it and keep the last call's contents; they are serialised by a
// per-device lock (calls from several threads on one GPU wait for each
// other).  Nothing in it needs to start at zero: every kernel writes what the
// next one reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pii_capi_internal.hpp"

namespace pii {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per GPU
constexpr size_t kStageBytes = size_t(1) << 20;      // 1 MiB host staging per thread

static __device__ __align__(256) char scratch[kScratchBytes];

// The device address of each GPU's copy of `scratch` (the name is a host-side
// stand-in in host code), looked up once per device with that device current.
struct DeviceScratch {
  std::once_flag once;
  char* ptr = nullptr;
  hipError_t err = hipSuccess;
  std::mutex mu;                 // serialises the calls that use this copy
};
DeviceScratch g_scratch[16];

DeviceScratch& device_scratch(int dev) {
  if (dev < 0 || dev >= 16) throw Error(PII_ERR_INVALID, "device id out of range");
  DeviceScratch& d = g_scratch[dev];
  std::call_once(d.once, [&] {
    void* p = nullptr;
    d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (d.err == hipSuccess) d.ptr = static_cast<char*>(p);
    else (void)hipGetLastError();
  });
  if (d.err != hipSuccess || !d.ptr)
    throw Error(PII_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                 hipGetErrorString(d.err));
  return d;
}

}  // namespace

}  // namespace pii

using namespace pii;

extern "C" {

int pii_form_factor_device_noalloc(pii_context* ctx, const pii_geometry* g, int dof_kind,
                                   const double* dofs, int ndofs, const pii_kgrid* kg,
                                   int precision, double* d_out) {
  return guarded([&] {
    if (!(ctx && g && kg && dofs && d_out))
      throw Error(PII_ERR_INVALID, "pii_form_factor_device_noalloc: NULL argument");
    std::lock_guard<std::mutex> lk(ctx->c.mu);
    PII_HIP(hipSetDevice(ctx->c.device));
    require_device_pointer(d_out, ctx->c.device, "pii_form_factor_device_noalloc");
    DeviceScratch& ds = device_scratch(ctx->c.device);
    std::lock_guard<std::mutex> lk2(ds.mu);
    alignas(256) static thread_local char stage[kStageBytes];
    Arena dev, host;
    dev.base = ds.ptr;
    dev.cap = kScratchBytes;
    dev.what = "device scratch";
    host.base = stage;
    host.cap = kStageBytes;
    host.what = "host staging";
    // a context sharing the caller context's stream and settings, whose
    // workspace takes its space from the two regions
    Context c;
    c.device = ctx->c.device;
    c.stream = ctx->c.stream;
    c.series = ctx->c.series;
    c.w.bind(&dev, &host);
    order_after_caller(ctx);
    try {
      form_factor_to(c, g, dof_kind, dofs, ndofs, kg, precision, d_out, hipMemcpyDeviceToDevice);
    } catch (...) {
      // kernels already queued must finish before another call reuses the array
      (void)hipStreamSynchronize(c.stream);
      throw;
    }
  });
}

}  // extern "C"
