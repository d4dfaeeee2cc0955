// Shared definitions for the gfx950 kernels and the host runtime.
//
// Device lattice layout (HBM):
//   populations  f[v*S + id], v = 0..8, id = i*Ny + j (j fastest, as
//                Grid.cpp:70), S = N rounded up to 256 elements so every
//                plane starts on a 1 KiB (FP32) / 2 KiB (FP64) boundary.
//                Two such buffers (A/B) alternate every step.
//   force_ibm    fibm[id] (x) and fibm[S + id] (y), FP64, dense but only
//                non-zero at the support sites of the last spread.
//   rho0/u0      FP64 site moments valid only right after initialisation or
//                an upload (the reference keeps rho/u arrays; after the first
//                step they are recomputed from f, see the step kernels).
#pragma once

#include <stdint.h>

namespace life {

constexpr int kNV = 9;
constexpr int kSupp = 9;        // suppSize, defs.h:39
constexpr int kMaxPerSite = 64; // markers that may share one support site

// D2Q9 tables (Grid.cpp:1247-1250).
constexpr int kCX[kNV] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
constexpr int kCY[kNV] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
constexpr int kOpp[kNV] = {0, 2, 1, 4, 3, 6, 5, 8, 7};

enum Prec : int { kFP64 = 0, kFP32 = 1 };

// One boundary site, with its normal resolved on the host
// (GridClass::getNormalVector, Grid.cpp:498-545).
struct BCSite {
    int32_t id;    // i*Ny + j
    int16_t i16;   // unused padding kept for alignment
    int8_t type;   // eLatType
    int8_t nx, ny; // normal vector
    int8_t nd;     // normal direction (0 = x, 1 = y)
    int8_t pad[2];
    int32_t i, j;
};

// Scalars the kernels need, set once per grid.
struct GridScalars {
    int Nx, Ny;
    long long N, S;
    int cm;                 // central moments (else BGK)
    double omega;
    double cs2, cs4;        // c_s*c_s and c_s*c_s*c_s*c_s with the reference's rounding
    double Fx, Fy;          // force_xy (uniform)
    double Fx0, Fy0;        // force_xy + 0.0 (force_xy + force_ibm at non-support sites)
    double Fxm, Fym;        // force_xy of the previous step: the moments of the incoming populations
                            // (equal to Fx, Fy unless the forcing varies in time, WOMERSLEY)
    double Dx, velScale;    // Dx and Dt/Dx (IBMNode.cpp:54)
    // force_xy per site (WOMERSLEY with gravity, Grid.cpp:52-61): null when
    // the force is uniform.  Planes x at [id], y at [S + id].  fprev: the
    // force of the previous step (the moments of the incoming populations);
    // fcur: this step's, (rho_n Drho g + dp cos) Dx^2 Dt^2 / Dm from the
    // incoming density rho_n, written by the bulk stream-collide and read by
    // the kernels after it in the step.
    const double* fprev;
    double* fcur;
    double fDrho, fgx, fgy;  // Drho and gravity (physical)
    double fdx, fdy;         // dpdx, dpdy times this step's cosine
    double fsc, fDm;         // (Dx Dt)^2 and Dm
};

}  // namespace life
