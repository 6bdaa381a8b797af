// Workspace of the allocation-free TT-SVD entry point (ttx_tt_svd_noalloc): one file-scope device array,
// TTX_NOALLOC_SCRATCH_BYTES long, of which every GPU that loads the library holds its own copy until
// the process ends.  Host code never passes the array's name to a kernel (on the host it is only a
// stand-in for the symbol): its device address is resolved once per device, with that device current,
// by noalloc_init (called from ttx_create), and only that address is used afterwards.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/ttx.h"

namespace ttx {

static __device__ __align__(256) unsigned char scratch[TTX_NOALLOC_SCRATCH_BYTES];

static constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// resolve (once per device) the device address of scratch on device dev, which must be current
hipError_t noalloc_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch[dev] = ptr;
  return hipSuccess;
}

// the address noalloc_init resolved for dev (nullptr before it ran)
void* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch[dev];
}

}  // namespace ttx
