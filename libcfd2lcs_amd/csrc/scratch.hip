// scratch.hip -- the fixed device workspace of the allocation-free entry
// point cfd2lcs_adjoint_forward_noalloc_c (api.cpp).
//
// `scratch` is a file-scope device array of kScratchBytes that every GPU the
// library runs on holds from module load until the process ends.  Host code
// never passes the array's name to a kernel or a copy: the name is a host
// stand-in, not a device address.  scratch_setup() looks the device address
// up once per device, with that device current (the library's init call does
// this), and scratch_ptr() returns the address for the current device.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "scratch.h"

namespace lcs {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

bool scratch_setup() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return false;
  if (g_scratch_ptr[dev]) return true;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return false;
  g_scratch_ptr[dev] = p;
  return true;
}

void* scratch_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  return g_scratch_ptr[dev];
}

}  // namespace lcs
