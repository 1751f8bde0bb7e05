// The exact left-to-right sum of device data without any device allocation
// (cs_ordered_sum_noalloc).  The reduction is the one of the PCG dot products
// and of cs_selftest_ordered_sum (cs_ordered_sum.hpp); its scratch -- one
// summary per tile of kSumTile addends and the finishing ticket -- lives in
// a file-scope device array instead of memory the library allocates.
//
// Each GPU holds its own copy of `scratch` at its own address.  Host code
// never passes the array's name to a kernel (in host code the name is a
// host-side stand-in, not a device address): the device address is looked up
// once per device with hipGetSymbolAddress, with that device current, and
// that pointer is what the kernels receive.  Every call on a GPU shares that
// GPU's array, which keeps the last call's contents; the ticket, the only
// field that must start at zero, is cleared by a kernel at the start of each
// call (a kernel, not a memset, so that the call can be captured in a graph).
// Calls that overlap on one GPU race on the array.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/clothsim.h"
#include "cs_device.hpp"
#include "cs_ordered_sum.hpp"

namespace cs {
namespace {

struct NoAllocScratch {
  double tsum[CS_NOALLOC_MAX_TILES];
  long long tot[CS_NOALLOC_MAX_TILES];
  long long pmin[CS_NOALLOC_MAX_TILES];
  long long pmax[CS_NOALLOC_MAX_TILES];
  int code[CS_NOALLOC_MAX_TILES];
  unsigned int ticket;
};

}  // namespace

// this GPU's scratch: a NoAllocScratch laid out in a byte array
static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];

namespace {

static_assert(CS_NOALLOC_MAX_TILES * (long long)kSumTile <= 2147483647LL,
              "the largest sum must fit an int count");

std::mutex g_mu;
std::vector<unsigned char*> g_ptr;  // per device; null until looked up

// the device address of this GPU's `scratch`, looked up on the first call on
// the device (with it current)
int scratch_on_device(unsigned char** out) {
  int dev = 0;
  CS_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_ptr.size() <= dev) g_ptr.resize(dev + 1, nullptr);
  if (!g_ptr[dev]) {
    unsigned char* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return kHipError;
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return kOk;
}

__global__ void k_clear_ticket(unsigned int* ticket) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *ticket = 0u;
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" int cs_ordered_sum_noalloc(int n, const double* t, double* out, void* stream) {
  if (n < 0 || (n > 0 && !t) || !out) {
    set_error("cs_ordered_sum_noalloc: n >= 0, t (n > 0) and out are required");
    return kBadArg;
  }
  if (sum_tiles(n) > CS_NOALLOC_MAX_TILES) {
    set_error("cs_ordered_sum_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
              std::to_string((long long)CS_NOALLOC_MAX_TILES * kSumTile) + " addends");
    return kCapacity;
  }
  unsigned char* base = nullptr;
  if (int st = scratch_on_device(&base)) return st;
  // device addresses of the fields inside this GPU's array (offsets from its
  // base; nothing is dereferenced on the host)
  SumScratch w{};
  w.tsum = reinterpret_cast<double*>(base + offsetof(NoAllocScratch, tsum));
  w.code = reinterpret_cast<int*>(base + offsetof(NoAllocScratch, code));
  w.tot = reinterpret_cast<long long*>(base + offsetof(NoAllocScratch, tot));
  w.pmin = reinterpret_cast<long long*>(base + offsetof(NoAllocScratch, pmin));
  w.pmax = reinterpret_cast<long long*>(base + offsetof(NoAllocScratch, pmax));
  w.ticket = reinterpret_cast<unsigned int*>(base + offsetof(NoAllocScratch, ticket));
  w.stats = nullptr;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_clear_ticket, dim3(1), dim3(64), 0, s, w.ticket);
  launch_ordered_sum(SumStore{SumJob{t, n, w, nullptr}, out}, n, 1, s);
  CS_HIP(hipGetLastError());
  return kOk;
}
