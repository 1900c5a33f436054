// noalloc.hip -- y = A x for callers that cannot allocate device memory.
//
// femhip_operator_apply stages x and y in two device buffers it allocates on
// every call.  femhip_operator_apply_noalloc runs the same CSR kernel with the
// staging in `scratch`, a file-scope device array of fixed size that each GPU
// holds from the library's load until the process ends, so the call makes no
// hipMalloc or hipFree.  The host never names `scratch` in a kernel argument
// or copy: at context creation, once per device and with that device current,
// hipGetSymbolAddress gives the array's address on that device, and only that
// address is used afterwards.
#include <hip/hip_runtime.h>

#include <mutex>

#include "femhip_internal.h"

namespace femhip {
constexpr int32_t kNoallocMaxRows = 1 << 21;   // rows of the largest operator (2,097,152)
constexpr int kNoallocMaxDevices = 64;
}  // namespace femhip

// x staged in [0, kNoallocMaxRows), y in [kNoallocMaxRows, 2 kNoallocMaxRows): 32 MiB
static __device__ __align__(256) double scratch[2 * (size_t)femhip::kNoallocMaxRows];

namespace femhip {
namespace {

std::mutex g_setup_mu;
double *g_scratch[kNoallocMaxDevices] = {};   // address of `scratch` on each device
std::mutex g_call_mu[kNoallocMaxDevices];     // one call at a time per device within the process

}  // namespace

void noalloc_setup(int device) {
  FH_CHECK(device >= 0 && device < kNoallocMaxDevices, FEMHIP_ERR_INVALID, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_setup_mu);
  if (g_scratch[device]) return;
  int cur = -1;
  FH_HIP(hipGetDevice(&cur));
  FH_CHECK(cur == device, FEMHIP_ERR_STATE, "scratch set-up needs its device current");
  void *p = nullptr;
  FH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  FH_CHECK(p != nullptr, FEMHIP_ERR_HIP, "no device address for the scratch array");
  g_scratch[device] = static_cast<double *>(p);
}

}  // namespace femhip

using namespace femhip;

extern "C" int femhip_operator_apply_noalloc(femhip_operator_t *op, const double *x, double *y) {
  try {
    FH_CHECK(op && x && y, FEMHIP_ERR_INVALID, "null argument");
    FH_CHECK(op->computed, FEMHIP_ERR_STATE, "operator not computed");
    FH_CHECK(op->n <= kNoallocMaxRows, FEMHIP_ERR_INVALID,
             "femhip_operator_apply_noalloc: " + std::to_string(op->n) + " rows exceed its maximum of " +
                 std::to_string(kNoallocMaxRows) + " (use femhip_operator_apply)");
    const int dev = op->ctx->device;
    FH_CHECK(dev >= 0 && dev < kNoallocMaxDevices, FEMHIP_ERR_INVALID, "device index beyond the scratch table");
    FH_HIP(hipSetDevice(dev));
    double *base = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_setup_mu);
      base = g_scratch[dev];
    }
    FH_CHECK(base != nullptr, FEMHIP_ERR_STATE, "scratch not set up on this device");
    double *dx = base, *dy = base + kNoallocMaxRows;
    const size_t bytes = sizeof(double) * (size_t)op->n;
    hipStream_t st = op->ctx->stream;
    std::lock_guard<std::mutex> lk(g_call_mu[dev]);
    // every row of y is written by the kernel and x is copied in whole: the
    // previous call's contents need no clearing
    FH_HIP(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, st));
    spmv_full(*op, dx, dy);
    FH_HIP(hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, st));
    FH_HIP(hipStreamSynchronize(st));
    return FEMHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  }
}
