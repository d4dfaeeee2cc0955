// Internal: the host state of one LIFE case (life_case), shared by the
// single-case driver (coupling.cpp) and the serving engine (serve.cpp).
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "fem.h"
#include "fem_dev.h"
#include "runtime.h"

namespace life {

// Page-locked host memory for arrays that the single-case driver copies to
// and from the device every step: the copies are then DMA transfers that
// do not block the host, so the host must not write such an array while a
// copy out of it may still be queued.
//
// hipHostMalloc and hipHostFree cost far more than the copies they serve,
// and a serving engine creates and retires cases all the time, mostly of a
// few marker counts; freed blocks are therefore kept, by size, for the next
// array of that size (pinned_take / pinned_give, runtime.cpp), up to a cap,
// and stay allocated until the process ends.
void* pinned_take(size_t bytes);
void pinned_give(void* p, size_t bytes);

template <typename T>
struct PinnedAllocator {
    using value_type = T;
    PinnedAllocator() = default;
    template <typename U>
    PinnedAllocator(const PinnedAllocator<U>&) {}
    T* allocate(size_t n) { return static_cast<T*>(pinned_take((n ? n : 1) * sizeof(T))); }
    void deallocate(T* p, size_t n) { pinned_give(p, (n ? n : 1) * sizeof(T)); }
    template <typename U>
    bool operator==(const PinnedAllocator<U>&) const { return true; }
    template <typename U>
    bool operator!=(const PinnedAllocator<U>&) const { return false; }
};
using pinned_vector = std::vector<double, PinnedAllocator<double>>;

struct Body {
    int ID;
    int flex;  // 0 flexible, 1 rigid (eFlexibleType, defs.h:48)
    int kind;  // 0 circle, 1 filament
    int first = 0, n = 0;
    std::unique_ptr<FemBody> fem;
    // construction data kept until the markers are final
    std::vector<double> mpos;
    double fpos[2] = {0, 0}, fgeom[2] = {0, 0}, fangle = 0, frho = 0, fE = 0;
    std::string fnel, fbc;
};

// Optional phase timing (environment LIFE_TIMING=1): synchronises the
// stream at each phase boundary, so it is for diagnosis only.
struct PhaseTimes {
    bool on = std::getenv("LIFE_TIMING") != nullptr;
    double t[6] = {0, 0, 0, 0, 0, 0};  // lbm, recompute, interp+d2h, fem, spread, steps
    long subits = 0;
    std::chrono::steady_clock::time_point last;
    void start(life_grid* g) {
        if (!on) return;
        LIFE_HIP(hipStreamSynchronize(g->stream));
        last = std::chrono::steady_clock::now();
    }
    void lap(life_grid* g, int k) {
        if (!on) return;
        LIFE_HIP(hipStreamSynchronize(g->stream));
        const auto now = std::chrono::steady_clock::now();
        t[k] += std::chrono::duration<double, std::milli>(now - last).count();
        last = now;
    }
    void report() const {
        if (!on || t[5] == 0) return;
        std::fprintf(stderr, "life timing over %.0f steps (%ld subiterations), ms/step: lbm %.3f recompute %.3f "
                     "interp %.3f fem %.3f spread %.3f\n", t[5], subits, t[0] / t[5], t[1] / t[5], t[2] / t[5],
                     t[3] / t[5], t[4] / t[5]);
    }
};

}  // namespace life

struct life_case {
    life_params p{};
    life_grid* g = nullptr;
    life::CaseScalars cs{};
    int t = 0;
    std::vector<life::Body> bodies;
    std::vector<life::Body*> flex_bodies;  // flexible bodies in body order
    int nm = 0;
    life::pinned_vector pos, vel, force, eps, ds;
    // recorded after the last upload of pos/vel in a step; waited on before
    // the host next writes them
    hipEvent_t ev_upload = nullptr;
    hipEvent_t ev_forces = nullptr;  // marker forces/epsilon downloaded
    // banded epsilon systems were launched; their out-of-band flags are
    // read at the next synchronisation (the force download)
    bool eps_check_pending = false;
    std::vector<int> mbody;
    bool hasIBM = false, hasFlex = false;
    int subIt = 0, simDOFs = 0;
    double subRes = 0.0, subNum = 0.0, subDen = 0.0, relax = 1.0;
    std::vector<double> A;
    std::vector<double> luw;
    std::vector<int> luiw;
    std::vector<double> rhs, sol;
    bool forces_on_host = false;
    std::unique_ptr<life::PhaseTimes> timing{new life::PhaseTimes()};
    std::vector<hipEvent_t> kev;  // per-step stream-collide events of life_case_step_timed
    // FEM of the flexible bodies on the device (FP32-shifted populations;
    // empty otherwise).  While it runs the steps, the host's FemBody state
    // and marker arrays lag behind the device (host_stale) until a reader
    // calls case_sync_host.
    life::FemDevCase fdev;
    bool host_stale = false;
    // Device FEM: the spread is queued right behind the FEM of a
    // subiteration predicted to be the step's last (as many as the previous
    // step took), before the host reads the residual (LIFE_SPEC_SPREAD = 0
    // turns this off).
    bool spec_spread = true;
    int prev_subits = 0;

    life::MarkerView view(const life::Body& b) {
        return life::MarkerView{pos.data(), vel.data(), force.data(), eps.data(), b.first, b.n};
    }
};

namespace life {

constexpr int kMaxSubIt = 20;  // Objects.cpp:40 (MAXIT)

// Phases of ObjectsClass::objectKernel (Objects.cpp:26-60) that the
// serving engine interleaves across cases:
// host part of recomputeObjectVals (predictor / Aitken relaxation, new
// marker positions and velocities in c->pos, c->vel)
void case_recompute_host(life_case* c);
// femKernel with the forces and epsilon in c->force, c->eps; updates
// subRes/subNum/subDen and increments subIt
void case_fem(life_case* c);
// Aitken factor of subiteration subIt >= 1 (Objects.cpp:179-190), in c->relax
void case_aitken(life_case* c);
// host FEM state and marker arrays brought level with the device FEM
void case_sync_host(life_case* c);
// device FEM state level with the host's (after the host rewrote it)
void case_push_fem(life_case* c);
// another subiteration is due (Objects.cpp:59)
inline bool case_subiterate(const life_case* c) { return c->subIt < kMaxSubIt && c->subRes > c->p.subTol; }
// computeEpsilon for markers [first, first+n) by the dense path
void case_solve_eps_dense(life_case* c, int first, int n);
// After the marker arrays were restored from a restart file
// (ObjectsClass::readRestart, Objects.cpp:1186-1198): upload them, find the
// supports and spacing of every marker and recompute epsilon as at t > 0.
void case_restart_objects(life_case* c);

}  // namespace life
