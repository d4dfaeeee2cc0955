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

// Stream stores of one site per lane with LDS realignment.
//
// A push kernel stores population v of site (i, j) at ((i+cx) mod Nx,
// (j+cy) mod Ny).  For the six populations with cy = +-1 the stores of a
// wavefront are shifted by one element against the cache lines; measured on
// MI355X that costs ~20% of the achievable HBM rate.  Here a workgroup of
// kBlock consecutive sites first drops those values into LDS at their
// destination offset within the block, then every lane stores the value
// that lands on its own site offset, so all bulk stores are aligned.  Values
// whose destination leaves the block (first/last lane) or wraps around the
// row are stored directly.  Populations with cy = 0 are stored directly
// (they only change row, a multiple of Ny elements).
constexpr int kPushBlock = 256;

// Pointer to T in the global address space (for pointers read from device
// tables, which the compiler cannot otherwise place).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T, int BS = kPushBlock>
struct PushStage {
    T slot[6][BS + 2];
};

// P: T* or a pointer to T in the global address space.
template <typename T, bool NT, typename P>
__device__ __forceinline__ void put(P p, T v) {
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// fo: the nine post-collision values of site id = base + t (row i, column j).
template <typename T, bool NT, int BS = kPushBlock, typename DP>
__device__ __forceinline__ void push_store_realigned(DP __restrict__ dst, long long S, int Nx, int Ny, long long N,
                                                     long long base, int t, long long id, int i, int j,
                                                     const T fo[9], bool valid, PushStage<T, BS>& st) {
    const int im = i == 0 ? Nx - 1 : i - 1, ip = i == Nx - 1 ? 0 : i + 1;
    const long long r0 = (long long)i * Ny, rp = (long long)ip * Ny, rm = (long long)im * Ny;
    if (valid) {
        put<T, NT>(&dst[0 * S + r0 + j], fo[0]);
        put<T, NT>(&dst[1 * S + rp + j], fo[1]);
        put<T, NT>(&dst[2 * S + rm + j], fo[2]);
    }
    // cy = +1: v = 3 (cx 0), 5 (+1), 8 (-1); cy = -1: v = 4 (0), 6 (-1), 7 (+1)
    constexpr int kv[6] = {3, 5, 8, 4, 6, 7};
    constexpr int kcx[6] = {0, 1, -1, 0, -1, 1};
    constexpr int kcy[6] = {1, 1, 1, -1, -1, -1};
    const long long rowv[6] = {r0, rp, rm, r0, rm, rp};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int cy = kcy[q];
        const int jd = j + cy;
        const bool wrap = jd < 0 || jd >= Ny;
        const int to = t + cy;  // destination offset within the block
        if (valid) {
            if (!wrap && to >= 0 && to < BS) st.slot[q][to + 1] = fo[kv[q]];
            else dst[kv[q] * S + rowv[q] + (wrap ? (jd < 0 ? Ny - 1 : 0) : jd)] = fo[kv[q]];
        }
    }
    __syncthreads();
    // lane t now stores the value destined for site base + t, if it came
    // from inside the block without wrapping around the row
    if (valid) {
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int cy = kcy[q];
            const int from = t - cy;
            const bool src_in = from >= 0 && from < BS && base + from < N;
            // the source site j - cy must not have wrapped to reach j
            const bool src_nowrap = cy > 0 ? j != 0 : j != Ny - 1;
            if (src_in && src_nowrap) {
                const int cx = kcx[q];
                const int id_row = cx == 0 ? i : (cx > 0 ? ip : im);
                put<T, NT>(&dst[kv[q] * S + (long long)id_row * Ny + j], st.slot[q][t + 1]);
            }
        }
    }
}

}  // namespace life
