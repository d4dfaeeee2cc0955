// Device-side helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace life {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Remap the hardware block index so that each XCD gets one
// contiguous range of blocks: neighbouring lattice rows (and the partial
// cache lines a push kernel writes at +-1 element) then meet in one L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nblocks) {
    const unsigned kX = 8;
    const unsigned q = nblocks / kX, r = nblocks % kX;
    const unsigned x = b % kX, k = b / kX;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

}  // namespace life
