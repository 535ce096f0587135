// Fixed device scratch of the allocation-free entry point (vh_matvec_noalloc).
//
// `scratch` is a file-scope device array: each GPU the library runs on holds
// its own copy, at its own address, from module load until the process ends.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in symbol, not a device address): scratch_init looks the device
// address up once per device with hipGetSymbolAddress, with that device
// current, and every user takes the address from scratch_ptr.
#include "vh_geom.h"
#include "../../include/voxhenry.h"
#include <mutex>

namespace vh {

namespace {
constexpr size_t kScratchBytes = VH_MATVEC_NOALLOC_BYTES;
constexpr int kMaxDevices = 64;
__device__ __align__(256) char scratch[kScratchBytes];
std::mutex scratch_mu;
char* scratch_dev[kMaxDevices] = {};
}  // namespace

void scratch_init(int device) {
  VH_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_dev[device]) return;
  void* p = nullptr;
  VH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  VH_REQUIRE(p != nullptr, "scratch array has no device address");
  scratch_dev[device] = (char*)p;
}

char* scratch_ptr(int device) {
  VH_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(scratch_mu);
  char* p = scratch_dev[device];
  if (!p) throw Error(ERR_STATE, "scratch of device " + std::to_string(device) + " not looked up (no context created on it)");
  return p;
}

size_t scratch_bytes() { return kScratchBytes; }

}  // namespace vh
