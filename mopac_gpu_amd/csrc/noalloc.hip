// Allocation-free variant of mhip_helect for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, hosts
// that forbid hipMalloc).  Inputs and output are device pointers the caller
// owns; the per-row partial sums the energy kernels need live in a
// file-scope device array of fixed size, one copy per GPU.
//
// The array's host-side name is only a stand-in for the symbol: its device
// address is looked up once per device with hipGetSymbolAddress (with that
// device current) and only that address is ever handed to kernels.
#include <mutex>

#include "common.h"
#include "k_fock.h"

namespace mhip {
namespace {

// Largest orbital count the variant accepts: one partial sum per row.
constexpr int kNoallocMaxRows = 65536;
// 512 KiB per GPU, held until the process ends.
static __device__ __align__(256) double scratch[kNoallocMaxRows];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call made on that device.
double* scratch_on_current_device() {
  int dev = -1;
  MHIP_CHECK(hipGetDevice(&dev));
  MHIP_REQUIRE(dev >= 0 && dev < kMaxDevices, "mhip_helect_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    MHIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MHIP_REQUIRE(ptr != nullptr, "mhip_helect_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace mhip

using namespace mhip;

extern "C" {

int mhip_helect_noalloc_max_rows(void) { return kNoallocMaxRows; }

int mhip_helect_noalloc(int norbs, const double* p, const double* h, const double* f, double* e,
                        void* stream) {
  try {
    MHIP_REQUIRE(p && h && f && e, "mhip_helect_noalloc: null pointer argument");
    MHIP_REQUIRE(norbs > 0, "mhip_helect_noalloc: norbs must be positive");
    MHIP_REQUIRE(norbs <= kNoallocMaxRows,
                 "mhip_helect_noalloc: norbs " + std::to_string(norbs) + " exceeds the maximum of " +
                     std::to_string(kNoallocMaxRows) + " rows of its fixed scratch");
    double* partial = scratch_on_current_device();
    // k_helect_rows writes every partial[0 .. norbs-1] before k_sum reads
    // them, so nothing left over from an earlier call is read.
    helect(static_cast<hipStream_t>(stream), norbs, p, h, f, 1.0, partial, e);
    return 0;
  } catch (const HipError& ex) {
    set_error(ex.msg);
    return 1;
  }
}

}  // extern "C"
