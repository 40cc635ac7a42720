// k_scratch.hip - the fixed device scratch of the allocation-free internal
// force (fgpu_int_force_dev_noalloc).
//
// The array lives in this translation unit's code object: every device the
// library runs on holds its own copy, at its own address, until the process
// ends.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in for the symbol); scratch_setup looks up each device's
// address once, with that device current, and everything else uses that
// address.
#include <mutex>
#include <vector>

#include "fgpu_internal.h"

namespace fgpu {
namespace {

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // per device id: the array's device address

}  // namespace

void *scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0) throw Error("scratch: invalid device");
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw Error(std::string("scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

}  // namespace fgpu
