// Host runtime: the device lattice (life_grid) and shared helpers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/life_hip.h"
#include "launch.h"
#include "life_dev.h"

namespace life {

void set_error(const std::string& msg);

struct HipError {
    hipError_t code;
    std::string where;
};

#define LIFE_HIP(expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) throw ::life::HipError{e_, std::string(#expr) + " @ " __FILE__}; \
    } while (0)

// Derived case scalars with the reference's rounding (Grid.cpp:1244-1260,
// params.h:89-90).
struct CaseScalars {
    double omega, tau, nu, Dx, Dt, Dm, Drho, cs, cs2, cs4;
};
CaseScalars derive_scalars(const life_params& p);

// GridClass::getRampCoefficient, Grid.cpp:548-556.
double ramp_coefficient(const life_params& p, double Dt, int t);

template <typename T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    LIFE_HIP(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T*>(p);
}
void dfree(void* p);
// The lattice-sized buffers of a grid (populations, rho0/u0, force_ibm):
// a freed block is kept, by device and size, for the next grid of that
// size (a serving engine creates and retires grids all the time, mostly of
// a few sizes), up to a cap; dfree_cached takes only blocks that no queued
// work uses any more.  A failed allocation first returns the kept blocks.
void* dalloc_cached_bytes(size_t bytes);
void dfree_cached(void* p);
template <typename T>
T* dalloc_cached(size_t n) {
    return static_cast<T*>(dalloc_cached_bytes((n ? n : 1) * sizeof(T)));
}

}  // namespace life

struct life_grid {
    int device = 0;
    hipStream_t stream = nullptr;
    life_params p{};
    int prec = 0;
    life::CaseScalars cs{};
    life::GridScalars g{};
    void* f[2] = {nullptr, nullptr};
    int cur = 0;  // buffer holding the latest populations
    double* rho0 = nullptr;
    double* u0 = nullptr;
    bool explicit_state = true;
    double* fibm = nullptr;
    double* fsite[2] = {nullptr, nullptr};  // per-site force_xy planes (WOMERSLEY with gravity)
    bool dense_clear_next = false;
    // boundaries
    std::vector<life::BCSite> bc_host;
    life::BCSite* bc = nullptr;
    int nbc = 0;
    bool convective = false;
    std::vector<double> u_in_host, rho_in_host;
    double* u_in = nullptr;
    double* rho_in = nullptr;
    double* delU = nullptr;
    double* ucol = nullptr;
    std::vector<int> type_host;
    // markers
    life::MarkerDev md{};
    int nbodies = 0;
    int* body_first = nullptr;
    int* body_n = nullptr;
    int* body_flex = nullptr;
    std::vector<int> body_first_h, body_n_h, body_flex_h;
    life::SpreadScratch sc{};
    int* flags = nullptr;  // [0] support overflow, [1] spread overflow
    double* epsA = nullptr;  // epsilon matrix scratch
    size_t epsA_cap = 0;
    // banded epsilon systems (flexible bodies)
    int eps_nsys = 0, eps_KB = 0;
    int* eps_first = nullptr;
    int* eps_n = nullptr;
    long long* eps_off = nullptr;
    int* eps_outside = nullptr;
    int* eps_outside_h = nullptr;  // page-locked copy of eps_outside
    double* eps_band = nullptr;
    double* eps_y = nullptr;
    std::vector<int> eps_first_h, eps_n_h;
    // timing
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_k0 = nullptr, ev_k1 = nullptr;
    // early bulk stream-collide of the next step on a second stream (one
    // CU-masked stream per device, shared by all grids; not owned)
    hipStream_t stream2 = nullptr;
    // With the coupling split (grid_split_coupling) `stream` is the
    // device's coupling stream and this is the grid's own stream from the
    // pool (the whole device), on which the plain-order lattice step runs;
    // given back at destroy
    hipStream_t stream_pool = nullptr;
    hipStream_t tail_stream = nullptr;  // where the pending step tail runs
    hipEvent_t ev_ready = nullptr, ev_bulk = nullptr, ev_tail = nullptr;
    bool bulk_pending = false;  // a bulk kernel of the next step is queued on stream2
    bool tail_pending = false;  // the last step's tail ran on tail_stream (stream not yet joined)
    int t_step = 0;             // time step of the last lbm_step
    // blow-up scan (GridClass::writeInfo's NaN test): every check_every
    // steps (0: never) the populations are scanned on the device; the
    // result reaches blow_h asynchronously and is read at the next scan or
    // at the end of the call
    int check_every = 100;
    unsigned long long* blow_d = nullptr;
    unsigned long long* blow_h = nullptr;  // page-locked
    hipEvent_t ev_blow = nullptr;
    int blow_t = -1;  // step of the scan in flight, -1 none

    void* f_cur() { return f[cur]; }
};

namespace life {
// Enqueue one lbmKernel on the grid stream; if ev_begin/ev_end are given,
// they are recorded around the fused stream-collide launch.
void lbm_step(life_grid* g, int t, hipEvent_t ev_begin, hipEvent_t ev_end);
// The same step in two parts.  lbm_bulk_early launches the bulk
// stream-collide of the next step on the second stream as soon as the
// populations of the current step are final (it reads only f and the
// uniform force); the coupling of the current step can run meanwhile.
// lbm_step then finishes that step (outlet speed, support sites,
// boundaries) after the bulk kernel.  Not for the first step after
// creation or upload (explicit rho/u).
void lbm_bulk_early(life_grid* g, hipEvent_t ev_begin, hipEvent_t ev_end);
// After lbm_bulk_early, lbm_step queues the rest of the step (outlet speed,
// support sites, boundaries) on the second stream behind the bulk kernel,
// once the first stream's work so far (the spread) is done, so that the
// next bulk kernel follows it there with no cross-stream wait; lbm_join
// then makes the grid's stream wait for that work before it reads the
// lattice (interpolation, downloads, the end of a call).
void lbm_join(life_grid* g);
// A second stream on `device` whose kernels leave reserve_cus CUs (at most
// one eighth of them, spread over the XCDs) to the other streams' work.
// One per (device, reserve) for the whole process, never destroyed: a
// CU-masked stream holds a hardware queue of its own.
hipStream_t masked_stream(int device, int reserve_cus);
// The grid's bulk stream-collide keeps cus CUs free for the coupling kernels
// (unless the environment sets LIFE_COUPLING_CUS); before any step.
void grid_reserve_coupling_cus(life_grid* g, int cus);
// The grid's stream becomes a stream on the device's first `xcds` XCDs, for
// the coupling chain, and its second stream (the early bulk
// stream-collide and the step tail) one on the other XCDs; a plain-order
// lattice step runs on the grid's own stream, on the whole device.  Falls
// back to grid_reserve_coupling_cus(g, fallback_cus) where the CUs do not
// split into eight equal XCDs.  Environment: LIFE_COUPLING_XCDS overrides
// xcds (0: the reserve instead); LIFE_COUPLING_CUS leaves the streams as
// they are.  Before any step.
void grid_split_coupling(life_grid* g, int xcds, int fallback_cus);
// Count of live grids and servers; the last release frees every pooled HIP
// resource (idle streams, kept device/pinned blocks, CU-masked streams).
void handle_acquire();
void handle_release();
// force_xy of time step t (uniform): the initial value (Grid.cpp:1042-1043),
// or under WOMERSLEY the oscillating pressure gradient of step t
// (Grid.cpp:52-61).
void body_force(const life_grid& g, int t, double F[2]);
// The grid's kernel scalars for time step t: force_xy of t and of t - 1.
GridScalars step_scalars(const life_grid& g, int t);
// Queue a blow-up scan of the current populations (step t) on stream s.
void blowup_scan(life_grid* g, int t, hipStream_t s);
// Read the scan in flight: with wait, after it completes; without, only if
// it already has.  Throws "Simulation blew up (t = .) at i = ., j = ." when
// it found a NaN velocity.
void blowup_check(life_grid* g, bool wait);
// Scan after step t when the cadence says so (reading the previous scan
// first if it is done).
void blowup_cadence(life_grid* g, int t, hipStream_t s);
void ibm_interpolate(life_grid* g);
void ibm_spread(life_grid* g);
// ibm_interpolate then ibm_spread, in one launch for a small case.
void ibm_interp_spread(life_grid* g);
void ibm_find_support(life_grid* g, int flexible_only);
void ibm_epsilon_matrix(life_grid* g, int first, int n, double* A_host);
// Declare the banded epsilon systems [first[k], first[k]+n[k]) once.
void ibm_epsilon_band_setup(life_grid* g, const std::vector<int>& first, const std::vector<int>& n, int KB);
// Assemble and solve them on the device (epsilon lands in md.eps) and, with
// copy_flags, queue the copy of their out-of-band flags to eps_outside_h
// (else the caller brings them over); does not wait.
void ibm_epsilon_band_launch(life_grid* g, bool copy_flags = true);
// After the grid stream has been synchronised past ibm_epsilon_band_launch:
// the indices of systems that were not banded (not solved; their epsilon
// still needs a solve).
std::vector<int> ibm_epsilon_band_dense(const life_grid* g);
void check_flags(life_grid* g);
}  // namespace life
