// FP32-shifted D2Q9 site arithmetic shared by the single-case and batched
// stream-collide kernels.  Stored value g_v = f_v - w_v (rho0 = 1) in FP32.
// Include only from translation units compiled with -ffp-contract=fast (the
// fused multiply-adds are part of the method's rounding).
#pragma once

#include <hip/hip_runtime.h>

#include "life_dev.h"

namespace life {

#define LIFE_DF __device__ __forceinline__

LIFE_DF float w32(int v) { return v == 0 ? 4.0f / 9.0f : (v < 5 ? 1.0f / 9.0f : 1.0f / 36.0f); }

// x * w_v without the weights' FP32 rounding bias.  w_v = k_v / 36 with
// k_v = 16, 4, 1, and float(1/36) is 7.5e-9 (relative) above 1/36, the
// same for every v (the k_v are powers of two).  Multiplying by it makes
// every equilibrium 7.5e-9 too large, so each collision adds that fraction
// of the site's mass and momentum: a body force proportional to u that
// accumulates over thousands of steps.  Here 1/36 = kW36hi + kW36lo to
// about 2^-50, k_v * x is exact, and one fused multiply-add rounds once.
constexpr float kW36hi = 1.0f / 36.0f;
constexpr float kW36lo = (float)(1.0 / 36.0 - (double)(1.0f / 36.0f));
LIFE_DF float wmul(int v, float x) {
    const float t = (v == 0 ? 16.0f : (v < 5 ? 4.0f : 1.0f)) * x;
    return __builtin_fmaf(t, kW36hi, t * kW36lo);
}

// BGK + Guo on deviations.
LIFE_DF void collide_bgk32(float omega, const float g[kNV], float drho, float ux, float uy, float Fx,
                           float Fy, float go[kNV]) {
    const float usq = 1.5f * (ux * ux + uy * uy);
    const float cF = 1.0f - 0.5f * omega;
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        const float cx = (float)kCX[v], cy = (float)kCY[v];
        const float cu = cx * ux + cy * uy;
        // rho * P as P + drho * P: 1 + drho rounded to float would carry
        // the same error over every site of a region of uniform density
        const float P = 3.0f * cu + 4.5f * cu * cu - usq;
        const float eq = wmul(v, drho + P + drho * P);
        const float fv = wmul(v, 3.0f * ((cx - ux) * Fx + (cy - uy) * Fy + 3.0f * cu * (cx * Fx + cy * Fy)));
        go[v] = g[v] + omega * (eq - g[v]) + cF * fv;
    }
}

// Central moments on deviations (De Rosis 2019 basis, Grid.cpp:106-223).
LIFE_DF void collide_cm32(float omega, float cs2, float cs4, const float g[kNV], float drho, float ux, float uy,
                          float Fx, float Fy, float go[kNV]) {
    float k4p = ux * ux - uy * uy, k5p = ux * uy;  // sum_v w_v (...) in closed form
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        const float dx = (float)kCX[v] - ux, dy = (float)kCY[v] - uy;
        k4p += g[v] * (dx * dx - dy * dy);
        k5p += g[v] * dx * dy;
    }
    const float k1 = 0.5f * Fx, k2 = 0.5f * Fy;
    const float k4 = (1.0f - omega) * k4p, k5 = (1.0f - omega) * k5p;
    const float k6 = 0.5f * Fy * cs2, k7 = 0.5f * Fx * cs2;
    const float X = ux, Y = uy, XX = ux * ux, YY = uy * uy, XY = ux * uy;
    const float XYY = X * YY, XXY = XX * Y;
    // M^-1 columns for k1, k2, k4, k5, k6, k7
    const float m1[kNV] = {2.0f * XYY - 2.0f * X,
                           X - XYY - 0.5f * YY + 0.5f,
                           X - XYY + 0.5f * YY - 0.5f,
                           -XYY - XY,
                           -XYY + XY,
                           0.25f * Y + 0.5f * XY + 0.5f * XYY + 0.25f * YY,
                           0.25f * Y - 0.5f * XY + 0.5f * XYY - 0.25f * YY,
                           0.5f * XYY - 0.5f * XY - 0.25f * Y + 0.25f * YY,
                           0.5f * XY - 0.25f * Y + 0.5f * XYY - 0.25f * YY};
    const float m2[kNV] = {2.0f * XXY - 2.0f * Y,
                           -XXY - XY,
                           -XXY + XY,
                           Y - XXY - 0.5f * XX + 0.5f,
                           Y - XXY + 0.5f * XX - 0.5f,
                           0.25f * X + 0.5f * XY + 0.5f * XXY + 0.25f * XX,
                           0.25f * X - 0.5f * XY + 0.5f * XXY - 0.25f * XX,
                           0.5f * XY - 0.25f * X + 0.5f * XXY - 0.25f * XX,
                           0.5f * XXY - 0.5f * XY - 0.25f * X + 0.25f * XX};
    const float m4[kNV] = {0.5f * YY - 0.5f * XX,
                           0.25f * XX + 0.25f * X - 0.25f * YY + 0.25f,
                           0.25f * XX - 0.25f * X - 0.25f * YY + 0.25f,
                           0.25f * XX - 0.25f * YY - 0.25f * Y - 0.25f,
                           0.25f * XX - 0.25f * YY + 0.25f * Y - 0.25f,
                           -0.125f * XX - 0.125f * X + 0.125f * YY + 0.125f * Y,
                           -0.125f * XX + 0.125f * X + 0.125f * YY - 0.125f * Y,
                           -0.125f * XX - 0.125f * X + 0.125f * YY - 0.125f * Y,
                           -0.125f * XX + 0.125f * X + 0.125f * YY + 0.125f * Y};
    const float m5[kNV] = {4.0f * XY,
                           -Y - 2.0f * XY,
                           Y - 2.0f * XY,
                           -X - 2.0f * XY,
                           X - 2.0f * XY,
                           0.5f * X + 0.5f * Y + XY + 0.25f,
                           XY - 0.5f * Y - 0.5f * X + 0.25f,
                           0.5f * Y - 0.5f * X + XY - 0.25f,
                           0.5f * X - 0.5f * Y + XY - 0.25f};
    const float m6[kNV] = {2.0f * Y, -Y, -Y, -Y - 0.5f, 0.5f - Y,
                           0.5f * Y + 0.25f, 0.5f * Y - 0.25f, 0.5f * Y - 0.25f, 0.5f * Y + 0.25f};
    const float m7[kNV] = {2.0f * X, -X - 0.5f, 0.5f - X, -X, -X,
                           0.5f * X + 0.25f, 0.5f * X - 0.25f, 0.5f * X + 0.25f, 0.5f * X - 0.25f};
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        const float cx = (float)kCX[v], cy = (float)kCY[v];
        const float bx = 9.0f * (cx * cx) * XX + 6.0f * cx * X - 3.0f * XX;
        const float by = 9.0f * (cy * cy) * YY + 6.0f * cy * Y - 3.0f * YY;
        const float dev = wmul(v, 0.5f * (bx + by) + 0.25f * bx * by);  // feqCM(1,u) - w
        // drho * feqCM(1,u) = drho * w + drho * dev
        go[v] = dev + wmul(v, drho) + drho * dev + m1[v] * k1 + m2[v] * k2 + m4[v] * k4 + m5[v] * k5 + m6[v] * k6 +
                m7[v] * k7;
    }
}

template <bool CM>
LIFE_DF void collide32(const GridScalars& g, const float gs[kNV], float drho, float ux, float uy, float Fx, float Fy,
                       float go[kNV]) {
    if (CM) collide_cm32((float)g.omega, (float)g.cs2, (float)g.cs4, gs, drho, ux, uy, Fx, Fy, go);
    else collide_bgk32((float)g.omega, gs, drho, ux, uy, Fx, Fy, go);
}

LIFE_DF void moments32(const float gs[kNV], float Fx, float Fy, float& drho, float& ux, float& uy) {
    const float dr = ((gs[0] + gs[1]) + (gs[2] + gs[3])) + ((gs[4] + gs[5]) + (gs[6] + gs[7])) + gs[8];
    const float jx = (gs[1] - gs[2]) + (gs[5] - gs[6]) + (gs[7] - gs[8]);
    const float jy = (gs[3] - gs[4]) + (gs[5] - gs[6]) + (gs[8] - gs[7]);
    // 1 / (1 + dr) with the rounding error e of the sum 1 + dr put back
    // (|dr| < 1, so e is exact): a bare 1 + dr carries the same error over
    // every site of a region of uniform density, a velocity bias of up to
    // 6e-8 shared by the whole region
    const float r = 1.0f + dr;
    const float e = dr - (r - 1.0f);
    const float inv0 = 1.0f / r;
    const float inv = inv0 - e * inv0 * inv0;
    drho = dr;
    ux = (jx + 0.5f * Fx) * inv;
    uy = (jy + 0.5f * Fy) * inv;
}

// Push of one site (any site): the nine post-collision values go to the
// neighbours with periodic wrap (Grid.cpp:226-244).
LIFE_DF void push_site32(float* __restrict__ dst, const GridScalars& g, long long id, const float go[kNV]) {
    const long long S = g.S;
    const int Ny = g.Ny, Nx = g.Nx;
    const int i = (int)(id / Ny), j = (int)(id - (long long)i * Ny);
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        int ii = i + kCX[v], jj = j + kCY[v];
        ii = ii < 0 ? Nx - 1 : (ii >= Nx ? 0 : ii);
        jj = jj < 0 ? Ny - 1 : (jj >= Ny ? 0 : jj);
        dst[v * S + (long long)ii * Ny + jj] = go[v];
    }
}

// First step after initialisation or upload: rho/u/force_ibm as stored.
template <bool CM>
LIFE_DF void push32_explicit_site(const float* __restrict__ src, float* __restrict__ dst, const GridScalars& g,
                                  const double* __restrict__ rho0, const double* __restrict__ u0,
                                  const double* __restrict__ fibm, long long id) {
    const long long S = g.S;
    float gs[kNV], go[kNV];
#pragma unroll
    for (int v = 0; v < kNV; v++) gs[v] = src[v * S + id];
    const float drho = (float)(rho0[id] - 1.0);
    const float ux = (float)u0[2 * id], uy = (float)u0[2 * id + 1];
    const float Fx = (float)(g.Fx + fibm[id]), Fy = (float)(g.Fy + fibm[S + id]);
    collide32<CM>(g, gs, drho, ux, uy, Fx, Fy, go);
    push_site32(dst, g, id, go);
}

// Support site of the last spread: moments and forcing with force_xy +
// force_ibm (IBMNode.cpp:97-136).
template <bool CM>
LIFE_DF void support_push32_site(const float* __restrict__ src, float* __restrict__ dst, const GridScalars& g,
                                 const double* __restrict__ fibm, long long id) {
    const long long S = g.S;
    float gs[kNV], go[kNV];
#pragma unroll
    for (int v = 0; v < kNV; v++) gs[v] = src[v * S + id];
    // moments as stored at the end of the previous step (its force_xy),
    // forcing of this step
    const float Fxm = (float)(g.Fxm + fibm[id]), Fym = (float)(g.Fym + fibm[S + id]);
    const float Fx = (float)(g.Fx + fibm[id]), Fy = (float)(g.Fy + fibm[S + id]);
    float drho, ux, uy;
    moments32(gs, Fxm, Fym, drho, ux, uy);
    collide32<CM>(g, gs, drho, ux, uy, Fx, Fy, go);
    push_site32(dst, g, id, go);
}

}  // namespace life
