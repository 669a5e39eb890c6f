// sv_noalloc.hip -- the fixed device scratch behind sv_render_camera_noalloc
// and its per-device address.
//
// The array lives in this translation unit's code object; the runtime
// places one copy on every device that loads the library, zero-filled, and
// it stays there until the process ends. Host code never passes `scratch`
// itself to a launch or a copy (in host code the name is a host-side stand-in,
// not a device address): noalloc_scratch() resolves the device address once
// per device with hipGetSymbolAddress and every use goes through it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sv_internal.hpp"

namespace sv {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {nullptr};
bool g_done[kMaxDevices] = {false};
hipError_t g_err[kMaxDevices] = {hipSuccess};
}  // namespace

hipError_t noalloc_scratch(int device, void **out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_done[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipErrorInvalidDevice;  // the caller makes `device` current
    void *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_err[device] = e;
    g_done[device] = true;
  }
  *out = g_ptr[device];
  return g_err[device];
}

// Colour planes (c * P + i) -> interleaved pixels (3 i + c).
__global__ __launch_bounds__(256) void k_interleave3(const float *__restrict__ planes, size_t P,
                                                     float *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  out[3 * i + 0] = planes[i];
  out[3 * i + 1] = planes[P + i];
  out[3 * i + 2] = planes[2 * P + i];
}

void launch_interleave3(hipStream_t st, const float *planes, size_t P, float *out) {
  if (P == 0) return;
  hipLaunchKernelGGL(k_interleave3, dim3((uint32_t)((P + 255) / 256)), dim3(256), 0, st, planes, P, out);
}

}  // namespace sv
