// The device scratch of asc_sparse_hip_csr_spmv_noalloc (the allocation-free
// compat SpMV): a file-scope device array, one copy per GPU the library runs
// on, held until the process ends.  The only device code of the C ABI, in a
// file of its own so that the rest of the ABI is plain host C++.
#include <hip/hip_runtime.h>

#include <mutex>

#include "abi_internal.h"
#include "asc_sparse_hip.h"

using namespace asc_hip_abi;

// In host code the array's name is not a device address: each device's copy
// is found once, with that device current (at context creation), by
// hipGetSymbolAddress, and only that pointer is used.
static __device__ __align__(256) unsigned char scratch[ASC_HIP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

namespace asc_hip_abi {
asc_hip_status noalloc_scratch(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices)
    return fail(ASC_HIP_UNSUPPORTED, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress of the allocation-free scratch");
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return ok();
}
}  // namespace asc_hip_abi
