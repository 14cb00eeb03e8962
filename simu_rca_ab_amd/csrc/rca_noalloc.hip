// Batched forward without any allocation by the caller or the library: the
// batch work list (header, per-half arguments, work items) lives in a
// file-scope device array instead of a workspace the caller allocates.
//
// Each GPU the library runs on holds its own copy of `scratch` until the
// process ends; its address on a device is looked up once per device with
// hipGetSymbolAddress (the array's name is only a host-side stand-in) and kept.
// A call copies the host-encoded work list into that device's array on the
// caller's stream and runs rca_batch_launch on it.  Calls that overlap on one
// GPU share the array and race; each call rewrites the whole work list it
// uses (header included), so nothing is carried over from a previous call.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/rca_das.h"

namespace rca_internal {
void set_error(const char* msg);
}

namespace {

constexpr size_t kScratchBytes = (size_t)4 << 20;   // 4 MiB: ~250k work items
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of this device's scratch array (looked up on first use).
int scratch_ptr(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    rca_internal::set_error(hipGetErrorString(e));
    return RCA_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    rca_internal::set_error("rca_batch_launch_noalloc: device index out of range");
    return RCA_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      rca_internal::set_error(e != hipSuccess ? hipGetErrorString(e) : "hipGetSymbolAddress returned NULL");
      return RCA_ERR_HIP;
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return RCA_OK;
}

}  // namespace

extern "C" {

int rca_batch_launch_noalloc(const void* host_ws, size_t ws_bytes, int max_halves, int max_items,
                             int interp, int n_sample_max, hipStream_t stream) {
  if (!host_ws || max_halves < 1 || max_items < 1) {
    rca_internal::set_error("rca_batch_launch_noalloc: bad arguments");
    return RCA_ERR_INVALID;
  }
  const size_t need = rca_batch_workspace_bytes(max_halves, max_items);
  if (ws_bytes < need) {
    rca_internal::set_error("rca_batch_launch_noalloc: host workspace smaller than the batch layout");
    return RCA_ERR_INVALID;
  }
  if (need > kScratchBytes) {
    rca_internal::set_error("rca_batch_launch_noalloc: the batch layout exceeds the 4 MiB scratch "
                            "array (fewer max_halves / max_items)");
    return RCA_ERR_UNSUPPORTED;
  }
  void* p = nullptr;
  int rc = scratch_ptr(&p);
  if (rc) return rc;
  const hipError_t e = hipMemcpyAsync(p, host_ws, need, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) {
    rca_internal::set_error(hipGetErrorString(e));
    return RCA_ERR_HIP;
  }
  return rca_batch_launch(p, max_halves, max_items, interp, n_sample_max, stream);
}

}  // extern "C"
