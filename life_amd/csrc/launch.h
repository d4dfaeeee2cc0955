// Kernel launchers shared between the HIP translation units and the host
// runtime.  All launches are asynchronous on the given stream.
#pragma once

#include <hip/hip_runtime_api.h>

#include "life_dev.h"

namespace life {

// FP64 plain method (lbm_f64.hip)
void launch_push64(const double* src, double* dst, const GridScalars& g, const double* rho0, const double* u0,
                   const double* fibm, bool explicit_state, hipStream_t s);
void launch_support_push64(const double* src, double* dst, const GridScalars& g, const double* fibm,
                           const int* sites, const int* count, int max_sites, hipStream_t s);

// FP32-shifted populations (lbm_f32.hip)
void launch_push32(const float* src, float* dst, const GridScalars& g, const double* rho0, const double* u0,
                   const double* fibm, bool explicit_state, hipStream_t s);
void launch_support_push32(const float* src, float* dst, const GridScalars& g, const double* fibm,
                           const int* sites, const int* count, int max_sites, hipStream_t s);

// Boundary conditions, outlet, initialisation and materialisation
// (lbm_bc.hip), for T = double (plain) or float (shifted).
struct BCArgs {
    const BCSite* sites;
    int nbc;
    const double* u_in;   // Ny*2
    const double* rho_in; // Ny
    const double* delU;   // Ny*2
    double ramp;
};

template <typename T>
void launch_bc(const T* src, T* dst, const GridScalars& g, const BCArgs& a, const double* fibm, const double* rho0,
               const double* u0, bool explicit_state, hipStream_t s);
template <typename T>
void launch_convective_speed(const T* src, const GridScalars& g, const double* fibm, const double* u0,
                             bool explicit_state, double* ucol, double* delU, bool ordered, hipStream_t s);
template <typename T>
void launch_init_equilibrium(T* f, const GridScalars& g, const double* rho0, const double* u0, hipStream_t s);
template <typename T>
void launch_materialise(const T* f, const GridScalars& g, const double* fibm, const double* rho0, const double* u0,
                        bool explicit_state, double* f_aos, double* rho, double* u, double* fibm_aos,
                        hipStream_t s);
template <typename T>
void launch_upload(T* f, const GridScalars& g, const double* f_aos, hipStream_t s);
void launch_init_moments(double* rho0, double* u0, const GridScalars& g, int mode, double ux, double uy,
                         const double* u_in, const BCSite* bc, int nbc, hipStream_t s);
void launch_fibm_from_aos(double* fibm, const GridScalars& g, const double* fibm_aos, hipStream_t s);
// NaN velocity scan (GridClass::writeInfo): *flag = smallest lattice id
// whose |u| is NaN, or ~0 when there is none.
template <typename T>
void launch_blowup_scan(const T* f, const GridScalars& g, unsigned long long* flag, hipStream_t s);
void launch_fill_f64(double* p, double v, long long n, hipStream_t s);
void launch_fill_i32(int* p, int v, long long n, hipStream_t s);

// IBM (ibm.hip)
struct MarkerDev {
    int n;
    double* pos;       // n*2 (physical)
    double* vel;       // n*2 (physical)
    double* force;     // n*2 (lattice)
    double* eps;       // n
    double* ds;        // n
    int* body;         // n
    const int* body_first;
    const int* body_n;
    const int* body_flex;
    int* supp_cnt;     // n
    int* supp_site;    // n*9, lattice id or -1
    int* supp_ij;      // n*9*2
    double* supp_delta;// n*9
    int* overflow;     // 1: set when a marker finds more than 9 supports
};

struct SpreadScratch {
    int* site_count;   // N, zero between spreads
    int* site_uid;     // N
    int* uniq;         // cap, lattice ids of support sites of the last spread
    int* uniq_n;       // 1
    int* pair_slot;    // n*9
    int* list;         // cap*kMaxPerSite
    int* overflow;     // 1
    int cap;
};

void launch_find_support(const MarkerDev& m, const GridScalars& g, int flexible_only, hipStream_t s);
// launch_find_support + launch_compute_ds in one launch (one workgroup per body).
void launch_support_ds(const MarkerDev& m, const GridScalars& g, int nbodies, int flexible_only, hipStream_t s);
void launch_compute_ds(const MarkerDev& m, const GridScalars& g, int nbodies, int flexible_only, hipStream_t s);
void launch_epsilon_matrix(const MarkerDev& m, const GridScalars& g, int first, int n, double* A, hipStream_t s);
// Banded epsilon systems (one per entry of sys_first/sys_n), assembled and
// solved on the device into m.eps; outside[k] is written: 1 for a system whose
// matrix is not confined to the band (left for the dense host path).  cr:
// systems of up to kCrMaxRows rows by block cyclic reduction (FP32-shifted
// cases), else the pivoted elimination.
void launch_eps_band(const MarkerDev& m, const GridScalars& g, const int* sys_first, const int* sys_n,
                     const int* host_n, int nsys, int KB, double* band, const long long* band_off, int* outside,
                     double* ywork, bool cr, hipStream_t s);
template <typename T>
void launch_interpolate(const T* f, const MarkerDev& m, const GridScalars& g, hipStream_t s);
void launch_spread(const MarkerDev& m, const GridScalars& g, double* fibm, SpreadScratch& sc, bool dense_clear,
                   hipStream_t s);
// Interpolation (populations f, FP64 or FP32-shifted) and then the spread in
// one launch, for a case of at most 8192 marker-support pairs; false (and
// nothing launched) otherwise.
bool launch_interp_spread_small(const void* f, bool f64, const MarkerDev& m, const GridScalars& g, double* fibm,
                                SpreadScratch& sc, bool dense_clear, hipStream_t s);

// ---------------------------------------------------------------- serving
// Batched forms of the kernels above for the serving engine.  Every
// admitted case has one CaseDev entry in a device table that the host
// rewrites before each step; a launch covers the cases named by `list`
// (indices into the table; grid.y = list entry) or, for the stream-collide,
// a block table of (case, block) pairs.  Each case's values are computed
// exactly as the single-case kernels compute them.

// Banded epsilon systems of one case (device arrays of its life_grid).
struct EpsDev {
    int nsys;
    const int* first;
    const int* n;
    const long long* off;
    int* outside;
    double* band;
    double* y;
};

struct CaseDev {
    GridScalars g;
    const void* src;   // populations at the start of the step
    void* dst;         // populations after it (the current ones during the IBM phase)
    const double* rho0;
    const double* u0;
    double* fibm;
    int explicit_state;
    int convective;
    BCArgs bc;
    double* ucol;
    MarkerDev m;
    SpreadScratch sc;
    int nbodies;
    int moff;          // first marker of the case in the packed staging arrays
    EpsDev eps;
};

// Sites per block of the batched stream-collide: FP32 uses the block size
// measured best for the single-case kernel, FP64 the LDS kernel's 256.
constexpr int kServeBlock32 = 512;
constexpr int kServeBlock64 = 256;

struct CaseBlock {
    int c;    // table index
    int blk;  // block of that case (kServeBlock32 or kServeBlock64 sites)
};

struct CaseSys {
    int c;  // table index
    int k;  // system of that case
};

// stream-collide over a block table (all cases of one collision operator)
void launch_push32_batch(bool cm, const CaseDev* cases, const CaseBlock* blocks, int nblocks, hipStream_t s);
void launch_push64_batch(bool cm, const CaseDev* cases, const CaseBlock* blocks, int nblocks, hipStream_t s);
void launch_support_push32_batch(bool cm, const CaseDev* cases, const int* list, int n, int max_sites, hipStream_t s);
void launch_support_push64_batch(bool cm, const CaseDev* cases, const int* list, int n, int max_sites, hipStream_t s);
template <typename T>
void launch_bc_batch(bool cm, const CaseDev* cases, const int* list, int n, int max_nbc, hipStream_t s);
template <typename T>
void launch_convective_batch(const CaseDev* cases, const int* list, int n, int max_ny, bool ordered, hipStream_t s);
// Wave priority (s_setprio level 0..3) of the single-case coupling kernels
// on the current device.
hipError_t set_coupling_priority(int level);
// packed marker staging: in = [x, y, vx, vy] per marker, out = [Fx, Fy, eps]
void launch_markers_in_batch(const CaseDev* cases, const int* list, int n, int max_markers, const double* stage,
                             hipStream_t s);
void launch_markers_out_batch(const CaseDev* cases, const int* list, int n, int max_markers, double* stage,
                              hipStream_t s);
void launch_find_support_batch(const CaseDev* cases, const int* list, int n, int max_markers, int flexible_only,
                               hipStream_t s);
void launch_compute_ds_batch(const CaseDev* cases, const int* list, int n, int max_bodies, int flexible_only,
                             hipStream_t s);
// outside_out[q] = 1 for a listed system that is not banded (no solve done;
// outside_out may be page-locked host memory, readable after the stream is
// synchronised);
// cr as for launch_eps_band
void launch_eps_band_batch(const CaseDev* cases, const CaseSys* sys, int nsys, int max_n, int* outside_out,
                           bool cr, hipStream_t s);
// Supports and spacing of the listed systems' markers plus their banded
// epsilon systems, one workgroup per system (the work of
// launch_find_support_batch(flexible), launch_compute_ds_batch and
// launch_eps_band_batch in one launch).  Returns false, launching nothing,
// if a system has more markers than the fused kernel takes.
bool launch_recompute_fused_batch(const CaseDev* cases, const CaseSys* sys, int nsys, int max_n, int* outside_out,
                                  bool cr, hipStream_t s);
template <typename T>
void launch_interpolate_batch(const CaseDev* cases, const int* list, int n, int max_markers, hipStream_t s);
void launch_spread_batch(const CaseDev* cases, const int* list, int n, int max_markers, hipStream_t s);

}  // namespace life
