// Device scratch for the entry point that allocates nothing
// (lensed_loglike_batch_noalloc): one statically sized array per GPU, set up
// once per device and shared by every such call on that device.
//
// In host code the array's name is only a host-side stand-in: its device
// address comes from hipGetSymbolAddress, called once per device with that
// device current (lensed_create does it), and only that address is ever
// handed to kernels, copies or fills.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>

#include "scratch.h"

namespace lsd {
void set_error(const std::string& msg);
}

// per-candidate planes, partial sums, the render list, tables and the
// launch descriptor of one launch, carved at 256-byte boundaries
static __device__ __align__(256) unsigned char scratch[LENSED_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

namespace lsd {

int scratch_setup(int device, void** ptr) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("scratch_setup: device ordinal out of range");
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) {
            set_error("scratch_setup: the device must be current");
            return -1;
        }
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return -1;
        }
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    return 0;
}

}  // namespace lsd
