// Fixed device scratch for the entry point that allocates nothing
// (pmcd_dot_device_noalloc): callers under HIP graph capture or in loops that
// must not call hipMalloc pass their inputs and outputs; the partial sums and
// the result of the two-pass reduction live in this file-scope array instead
// of a per-call allocation.
//
// Each device holds its own copy of the array for the life of the process.
// Its address on a device comes from hipGetSymbolAddress, once per device
// with that device current, and is cached here; host code only ever passes
// that address to kernels and copies, never the array's name. Every call on
// one device shares the array, so calls that overlap on a device (other
// streams or threads) race; nothing in it needs to start at zero, since each
// launch writes every partial it reads.
#include <mutex>
#include <vector>

#include "device.hpp"
#include "kernels.hpp"

namespace pmcd {

namespace {

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // per device, null until first use

}  // namespace

double* noalloc_scratch() {
  int dev = 0;
  PMCD_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if(!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PMCD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PMCD_CHECK(p != nullptr, "no device address for the scratch array on device ", dev);
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace pmcd
