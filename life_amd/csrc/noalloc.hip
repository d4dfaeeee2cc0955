// Allocation-free form of life_ibm_epsilon_matrix for callers that may not
// allocate device memory (code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc).  The n x n matrix is assembled by the
// same kernel as life_ibm_epsilon_matrix, into a fixed device array of this
// file instead of the grid's growable buffer, then copied to the caller's
// host array.
//
// `scratch` holds kEpsScratchMax^2 doubles (8 MiB) on every device the
// library runs on, from module load to process exit.  Its device address is
// looked up once per device (hipGetSymbolAddress, with that device current)
// at the entry point's first call there; host code uses only that address,
// never the array's name.  Every call on a device shares its array: calls
// that overlap on one device race (the second assembly may overwrite the
// first before its copy-out), so callers serialise them.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "launch.h"
#include "runtime.h"

namespace life {

constexpr int kEpsScratchMax = 1024;  // largest system (markers) the form accepts

static __device__ __align__(256) double scratch[(size_t)kEpsScratchMax * kEpsScratchMax];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current).
double* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        LIFE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[device] = static_cast<double*>(p);
    }
    return g_scratch_dev[device];
}

}  // namespace
}  // namespace life

using namespace life;

extern "C" int life_ibm_epsilon_matrix_noalloc(life_grid* g, int first, int n, double* A) {
    try {
        if (!g || !A) throw std::runtime_error("null grid or output array");
        if (first < 0 || n < 0 || first + n > g->md.n) throw std::runtime_error("epsilon system out of range");
        if (n > kEpsScratchMax)
            throw std::runtime_error("epsilon system of " + std::to_string(n) +
                                     " markers exceeds the allocation-free limit of " +
                                     std::to_string(kEpsScratchMax) + " (use life_ibm_epsilon_matrix)");
        if (n == 0) return 0;
        LIFE_HIP(hipSetDevice(g->device));
        double* buf = scratch_address(g->device);
        launch_epsilon_matrix(g->md, g->g, first, n, buf, g->stream);
        LIFE_HIP(hipGetLastError());
        LIFE_HIP(hipMemcpyAsync(A, buf, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, g->stream));
        LIFE_HIP(hipStreamSynchronize(g->stream));
        return 0;
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return 2;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}
