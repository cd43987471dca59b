// Device scratch of the entry point that allocates nothing.
#pragma once

#include <cstddef>

// bytes of the static device array per GPU (held until the process ends)
#define LENSED_SCRATCH_BYTES (128ull << 20)

namespace lsd {

// The device address of this GPU's copy of the scratch array, found once per
// device (hipGetSymbolAddress, with `device` current) and kept; 0 on success.
int scratch_setup(int device, void** ptr);

}  // namespace lsd
