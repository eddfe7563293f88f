// cs_implicit_gemm2_noalloc: implicit_gemm2 with its workspace in a file-scope
// device array instead of a stream-ordered allocation, for callers that cannot
// allocate.  Nothing here calls hipMalloc / hipFree or an allocating library
// routine.
//
// The array's name is a host-side stand-in, not a device address: each GPU
// holds its own copy at its own address, looked up once per device with
// hipGetSymbolAddress (with that device current) and used from then on.
// The kernels write every workspace word they read before reading it (row
// masks, pair counts and lists, split partials, transposed weights), so the
// array needs no clearing between calls; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/cumm_amd.h"

namespace cs_detail {
int implicit_gemm2_in(const cs_conv_params* params, void* mem, int64_t cap, int device);
int fail_runtime(const char* msg);  // sets cs_last_error, returns CS_ERR_RUNTIME
}

static __device__ __align__(256) unsigned char scratch[CS_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_mu;
std::unordered_map<int, void*> g_ptr;  // device ordinal -> address of that device's `scratch`

// Address of the current device's array; hipGetSymbolAddress runs once per
// device, on the device that is current for the call.
hipError_t scratch_on(int* device, void** ptr) {
  hipError_t e = hipGetDevice(device);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_ptr.find(*device);
  if (it != g_ptr.end()) {
    *ptr = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[*device] = p;
  *ptr = p;
  return hipSuccess;
}

}  // namespace

extern "C" int cs_implicit_gemm2_noalloc(const cs_conv_params* params) {
  int device = 0;
  void* ptr = nullptr;
  const hipError_t e = scratch_on(&device, &ptr);
  if (e != hipSuccess) {
    const std::string msg = std::string("cs_noalloc.hip: looking up the allocation-free scratch failed: HIP error ") +
                            std::to_string((int)e) + ": " + hipGetErrorString(e);
    return cs_detail::fail_runtime(msg.c_str());
  }
  return cs_detail::implicit_gemm2_in(params, ptr, CS_NOALLOC_SCRATCH_BYTES, device);
}
