// Fixed device scratch of sba_solve_noalloc: the one entry point that allocates no device memory.
//
// The array lives in the code object, so each GPU the library runs on holds one copy of kScratchBytes from the
// moment the module is loaded until the process ends. Host code never passes the array's name to a kernel or a
// copy: its host-side symbol is not a device address. scratch_setup() looks up this device's address once (with
// the device current, from sba_create) and the runtime carves every stream of the solved window out of that
// address. All calls on one GPU share the array and it keeps the previous call's contents; calls that overlap
// on one GPU (two handles, two threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sba_device.h"

namespace sba_dev {
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
}  // namespace sba_dev

static __device__ __align__(256) unsigned char scratch[sba_dev::kScratchBytes];

namespace sba_dev {

// This device's address of the scratch array (the device must be current), looked up once per device.
// Returns the HIP error of the lookup, hipSuccess otherwise.
int scratch_setup(int device, void** ptr, size_t* bytes) {
  static std::mutex mu;
  static void* per_device[kMaxDevices] = {};
  *ptr = nullptr;
  *bytes = 0;
  if (device < 0 || device >= kMaxDevices) return static_cast<int>(hipErrorInvalidDevice);
  std::lock_guard<std::mutex> lock(mu);
  if (!per_device[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return static_cast<int>(e);
    per_device[device] = p;
  }
  *ptr = per_device[device];
  *bytes = kScratchBytes;
  return static_cast<int>(hipSuccess);
}

}  // namespace sba_dev
