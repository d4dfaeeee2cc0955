// FP64 D2Q9 site operators with the reference's operation order.
//
// Every function here reproduces, operation for operation, the double
// expression trees of the reference so that the plain method can be
// bit-compatible with it.  Translation units including this header are
// compiled with -ffp-contract=off (no fused multiply-add) and without any
// fast-math flag.  Integer lattice velocities are kept as int and promoted
// exactly where the reference promotes them; the compiler only folds
// value-preserving identities (x*1.0, x*-1.0).
#pragma once

#include <hip/hip_runtime.h>

#include "life_dev.h"

namespace life {

#define LIFE_DI __device__ __forceinline__

// Weights w_v (Grid.cpp:1248), rounded once from the exact fractions.
LIFE_DI double w64(int v) {
    return v == 0 ? 4.0 / 9.0 : (v < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
}

// GridClass::equilibrium, Grid.cpp:249-264.
template <bool CM>
LIFE_DI double feq64(double rho, double ux, double uy, int v) {
    const int cx = kCX[v], cy = kCY[v];
    if (CM) {
        double ax = 9.0 * (cx * cx) * (ux * ux) + 6.0 * cx * ux - 3.0 * (ux * ux) + 2.0;
        double ay = 9.0 * (cy * cy) * (uy * uy) + 6.0 * cy * uy - 3.0 * (uy * uy) + 2.0;
        return 0.25 * rho * w64(v) * ax * ay;
    } else {
        double cu = cx * ux + cy * uy;
        double q = ux * ux * (cx * cx - 1.0 / 3.0) + uy * uy * (cy * cy - 1.0 / 3.0);
        return rho * w64(v) * (1.0 + 3.0 * cu + 4.5 * q + 9.0 * cx * cy * ux * uy);
    }
}

// GridClass::latticeForce (Guo), Grid.cpp:267-279.
LIFE_DI double guo64(double ux, double uy, double Fx, double Fy, int v) {
    const int cx = kCX[v], cy = kCY[v];
    return 3.0 * w64(v) * (Fx * (cx - ux + cx * 3.0 * (cx * ux + cy * uy)) +
                           (Fy * (cy - uy + cy * 3.0 * (cx * ux + cy * uy))));
}

// GridClass::macroscopic (Grid.cpp:282-299) and
// IBMNodeClass::updateMacroscopic (IBMNode.cpp:114-122): sums in v order,
// then (momentum + F/2) / rho.  Fx/Fy is force_xy, or force_xy + force_ibm.
LIFE_DI void moments64(const double fs[kNV], double Fx, double Fy, double& rho, double& ux, double& uy) {
    double r = 0.0, mx = 0.0, my = 0.0;
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        r += fs[v];
        mx += kCX[v] * fs[v];
        my += kCY[v] * fs[v];
    }
    rho = r;
    ux = (mx + 0.5 * Fx) / r;
    uy = (my + 0.5 * Fy) / r;
}

// force_xy of one site from its density (Grid.cpp:59-60), same operation
// order: (rho Drho g + dp cos) (Dx Dt)^2 / Dm.
LIFE_DI double site_force_x(const GridScalars& g, double rho) { return (rho * g.fDrho * g.fgx + g.fdx) * g.fsc / g.fDm; }
LIFE_DI double site_force_y(const GridScalars& g, double rho) { return (rho * g.fDrho * g.fgy + g.fdy) * g.fsc / g.fDm; }

LIFE_DI double density64(const double fs[kNV]) {
    double r = 0.0;
#pragma unroll
    for (int v = 0; v < kNV; v++) r += fs[v];
    return r;
}

// BGK collision + Guo forcing (Grid.cpp:236-244).
LIFE_DI void collide_bgk64(double omega, const double fs[kNV], double rho, double ux, double uy, double Fx,
                           double Fy, double fo[kNV]) {
    const double c2 = 1.0 - 0.5 * omega;
#pragma unroll
    for (int v = 0; v < kNV; v++)
        fo[v] = fs[v] + omega * (feq64<false>(rho, ux, uy, v) - fs[v]) + c2 * guo64(ux, uy, Fx, Fy, v);
}

// Central-moments collision (Grid.cpp:106-223, De Rosis 2019).  The nine
// back-transformation rows keep the reference's term order and grouping.
LIFE_DI void collide_cm64(double omega, double cs2, double cs4, const double fs[kNV], double rho, double ux,
                          double uy, double Fx, double Fy, double fo[kNV]) {
    double k4p = 0.0, k5p = 0.0;
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        const double dx = kCX[v] - ux, dy = kCY[v] - uy;
        k4p += fs[v] * (dx * dx - dy * dy);
        k5p += fs[v] * dx * dy;
    }
    const double k0 = rho, k1 = 0.5 * Fx, k2 = 0.5 * Fy, k3 = 2.0 * rho * cs2;
    const double k4 = (1.0 - omega) * k4p, k5 = (1.0 - omega) * k5p;
    const double k6 = 0.5 * Fy * cs2, k7 = 0.5 * Fx * cs2, k8 = rho * cs4;
    const double X = ux, Y = uy, XX = ux * ux, YY = uy * uy;
    const double XXYY = XX * YY, XYY = X * YY, XXY = XX * Y, XY = X * Y;

    fo[0] = (XX * YY - XX - YY + 1.0) * k0 + (2.0 * X * YY - 2.0 * X) * k1 + (2.0 * Y * XX - 2.0 * Y) * k2 +
            (0.5 * XX + 0.5 * YY - 1.0) * k3 + (0.5 * YY - 0.5 * XX) * k4 + 4.0 * X * Y * k5 + 2.0 * Y * k6 +
            2.0 * X * k7 + k8;
    fo[1] = (0.5 * XX - 0.5 * XXYY - 0.5 * XYY + 0.5 * X) * k0 + (X - X * YY - 0.5 * YY + 0.5) * k1 +
            (-Y * XX - Y * X) * k2 + (-0.25 * XX - 0.25 * X - 0.25 * YY + 0.25) * k3 +
            (0.25 * XX + 0.25 * X - 0.25 * YY + 0.25) * k4 + (-Y - 2.0 * X * Y) * k5 + (-Y) * k6 +
            (-X - 0.5) * k7 - 0.5 * k8;
    fo[2] = (-0.5 * XXYY + 0.5 * XX + 0.5 * XYY - 0.5 * X) * k0 + (X - X * YY + 0.5 * YY - 0.5) * k1 +
            (-Y * XX + Y * X) * k2 + (-0.25 * XX + 0.25 * X - 0.25 * YY + 0.25) * k3 +
            (0.25 * XX - 0.25 * X - 0.25 * YY + 0.25) * k4 + (Y - 2.0 * X * Y) * k5 + (-Y) * k6 +
            (0.5 - X) * k7 - 0.5 * k8;
    fo[3] = (0.5 * YY - 0.5 * XXY - 0.5 * XXYY + 0.5 * Y) * k0 + (-X * YY - X * Y) * k1 +
            (Y - XX * Y - 0.5 * XX + 0.5) * k2 + (-0.25 * XX - 0.25 * YY - 0.25 * Y + 0.25) * k3 +
            (0.25 * XX - 0.25 * YY - 0.25 * Y - 0.25) * k4 + (-X - 2.0 * X * Y) * k5 + (-Y - 0.5) * k6 +
            (-X) * k7 - 0.5 * k8;
    fo[4] = (-0.5 * XXYY + 0.5 * XXY + 0.5 * YY - 0.5 * Y) * k0 + (-X * YY + X * Y) * k1 +
            (Y - XX * Y + 0.5 * XX - 0.5) * k2 + (-0.25 * XX - 0.25 * YY + 0.25 * Y + 0.25) * k3 +
            (0.25 * XX - 0.25 * YY + 0.25 * Y - 0.25) * k4 + (X - 2.0 * X * Y) * k5 + (0.5 - Y) * k6 +
            (-X) * k7 - 0.5 * k8;
    fo[5] = (0.25 * XXYY + 0.25 * XXY + 0.25 * XYY + 0.25 * XY) * k0 +
            (0.25 * Y + 0.5 * XY + 0.5 * XYY + 0.25 * YY) * k1 + (0.25 * X + 0.5 * XY + 0.5 * XXY + 0.25 * XX) * k2 +
            (0.125 * XX + 0.125 * X + 0.125 * YY + 0.125 * Y) * k3 +
            (-0.125 * XX - 0.125 * X + 0.125 * YY + 0.125 * Y) * k4 + (0.5 * X + 0.5 * Y + X * Y + 0.25) * k5 +
            (0.5 * Y + 0.25) * k6 + (0.5 * X + 0.25) * k7 + 0.25 * k8;
    fo[6] = (0.25 * XXYY - 0.25 * XXY - 0.25 * XYY + 0.25 * XY) * k0 +
            (0.25 * Y - 0.5 * XY + 0.5 * XYY - 0.25 * YY) * k1 + (0.25 * X - 0.5 * XY + 0.5 * XXY - 0.25 * XX) * k2 +
            (0.125 * XX - 0.125 * X + 0.125 * YY - 0.125 * Y) * k3 +
            (-0.125 * XX + 0.125 * X + 0.125 * YY - 0.125 * Y) * k4 + (X * Y - 0.5 * Y - 0.5 * X + 0.25) * k5 +
            (0.5 * Y - 0.25) * k6 + (0.5 * X - 0.25) * k7 + 0.25 * k8;
    fo[7] = (0.25 * XXYY - 0.25 * XXY + 0.25 * XYY - 0.25 * XY) * k0 +
            (0.5 * XYY - 0.5 * XY - 0.25 * Y + 0.25 * YY) * k1 + (0.5 * XY - 0.25 * X + 0.5 * XXY - 0.25 * XX) * k2 +
            (0.125 * XX + 0.125 * X + 0.125 * YY - 0.125 * Y) * k3 +
            (-0.125 * XX - 0.125 * X + 0.125 * YY - 0.125 * Y) * k4 + (0.5 * Y - 0.5 * X + X * Y - 0.25) * k5 +
            (0.5 * Y - 0.25) * k6 + (0.5 * X + 0.25) * k7 + 0.25 * k8;
    fo[8] = (0.25 * XXYY + 0.25 * XXY - 0.25 * XYY - 0.25 * XY) * k0 +
            (0.5 * XY - 0.25 * Y + 0.5 * XYY - 0.25 * YY) * k1 + (0.5 * XXY - 0.5 * XY - 0.25 * X + 0.25 * XX) * k2 +
            (0.125 * XX - 0.125 * X + 0.125 * YY + 0.125 * Y) * k3 +
            (-0.125 * XX + 0.125 * X + 0.125 * YY + 0.125 * Y) * k4 + (0.5 * X - 0.5 * Y + X * Y - 0.25) * k5 +
            (0.5 * Y + 0.25) * k6 + (0.5 * X - 0.25) * k7 + 0.25 * k8;
}

template <bool CM>
LIFE_DI void collide64(const GridScalars& g, const double fs[kNV], double rho, double ux, double uy, double Fx,
                       double Fy, double fo[kNV]) {
    if (CM) collide_cm64(g.omega, g.cs2, g.cs4, fs, rho, ux, uy, Fx, Fy, fo);
    else collide_bgk64(g.omega, fs, rho, ux, uy, Fx, Fy, fo);
}

// 3-point regularised delta (Utils::diracDelta, Utils.h:220-232).
LIFE_DI double dirac64(double dist) {
    double a = fabs(dist);
    if (a > 1.5) return 0.0;
    if (a > 0.5) return (5.0 - 3.0 * a - sqrt(-3.0 * ((1.0 - a) * (1.0 - a)) + 1.0)) / 6.0;
    return (1.0 + sqrt(1.0 - 3.0 * (a * a))) / 3.0;
}

// Population storage: FP64 as is, FP32 as the shift f - w (w rounded to
// float), read back into FP64 as stored + w.
template <typename T>
struct Pops;

template <>
struct Pops<double> {
    static LIFE_DI double get(const double* f, long long S, int v, long long id) { return f[v * S + id]; }
    static LIFE_DI void put(double* f, long long S, int v, long long id, double x) { f[v * S + id] = x; }
};

template <>
struct Pops<float> {
    static LIFE_DI double get(const float* f, long long S, int v, long long id) {
        return (double)f[v * S + id] + w64(v);
    }
    static LIFE_DI void put(float* f, long long S, int v, long long id, double x) {
        f[v * S + id] = (float)(x - w64(v));
    }
};

template <typename T>
LIFE_DI void load_site(const T* f, long long S, long long id, double fs[kNV]) {
#pragma unroll
    for (int v = 0; v < kNV; v++) fs[v] = Pops<T>::get(f, S, v, id);
}

}  // namespace life
