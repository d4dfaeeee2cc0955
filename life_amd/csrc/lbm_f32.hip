// FP32-shifted populations: fused push stream-collide kernels.
//
// Stored value g_v = f_v - w_v (rho0 = 1) in FP32, so that the O(1) part of
// each population never enters the rounding.  Moments and collision run in
// FP32 on the deviations:
//   drho = sum g_v, rho = 1 + drho, j = sum c_v g_v, u = (j + F/2) / rho
//   BGK:  feq_v - w_v = w_v (drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u.u))
//   CM:   f*_v - w_v = (feqCM_v(1,u) - w_v) + drho feqCM_v(1,u)
//                      + sum_m Minv_vm(u) k_m over the non-equilibrium
//                      central moments m = 1,2,4,5,6,7 (Grid.cpp:125-223)
// Same push/wrap pattern and same support-site treatment as the FP64 path
// (lbm_f64.hip).  72 B of HBM traffic per site update.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include "dev_util.h"
#include "launch.h"

namespace life {

#define LIFE_DF __device__ __forceinline__

using v4f = float __attribute__((ext_vector_type(4)));

LIFE_DF float w32(int v) { return v == 0 ? 4.0f / 9.0f : (v < 5 ? 1.0f / 9.0f : 1.0f / 36.0f); }

// BGK + Guo on deviations.
LIFE_DF void collide_bgk32(float omega, const float g[kNV], float drho, float rho, float ux, float uy, float Fx,
                           float Fy, float go[kNV]) {
    const float usq = 1.5f * (ux * ux + uy * uy);
    const float cF = 1.0f - 0.5f * omega;
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        const float cx = (float)kCX[v], cy = (float)kCY[v];
        const float cu = cx * ux + cy * uy;
        const float eq = w32(v) * (drho + rho * (3.0f * cu + 4.5f * cu * cu - usq));
        const float fv = 3.0f * w32(v) * ((cx - ux) * Fx + (cy - uy) * Fy + 3.0f * cu * (cx * Fx + cy * Fy));
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
        const float dev = w32(v) * (0.5f * (bx + by) + 0.25f * bx * by);  // feqCM(1,u) - w
        const float feq1 = w32(v) + dev;
        go[v] = dev + drho * feq1 + m1[v] * k1 + m2[v] * k2 + m4[v] * k4 + m5[v] * k5 + m6[v] * k6 + m7[v] * k7;
    }
}

template <bool CM>
LIFE_DF void collide32(const GridScalars& g, const float gs[kNV], float drho, float ux, float uy, float Fx, float Fy,
                       float go[kNV]) {
    if (CM) collide_cm32((float)g.omega, (float)g.cs2, (float)g.cs4, gs, drho, ux, uy, Fx, Fy, go);
    else collide_bgk32((float)g.omega, gs, drho, 1.0f + drho, ux, uy, Fx, Fy, go);
}

LIFE_DF void moments32(const float gs[kNV], float Fx, float Fy, float& drho, float& ux, float& uy) {
    const float dr = ((gs[0] + gs[1]) + (gs[2] + gs[3])) + ((gs[4] + gs[5]) + (gs[6] + gs[7])) + gs[8];
    const float jx = (gs[1] - gs[2]) + (gs[5] - gs[6]) + (gs[7] - gs[8]);
    const float jy = (gs[3] - gs[4]) + (gs[5] - gs[6]) + (gs[8] - gs[7]);
    const float inv = 1.0f / (1.0f + dr);
    drho = dr;
    ux = (jx + 0.5f * Fx) * inv;
    uy = (jy + 0.5f * Fy) * inv;
}

// Store with or without the non-temporal (streaming) hint.
template <bool NT>
LIFE_DF void st32(float* p, float v) {
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// One site per lane.  NT selects the streaming hint: bit 0 for the loads,
// bit 1 for the stores that stay in their row and column (c = 0, +-x),
// bit 2 for the stores shifted by one site along the row.  XCD = remap
// blocks to XCD-contiguous ranges.
template <bool CM, bool EXPLICIT, int NT = 3, bool XCD = true>
__global__ __launch_bounds__(256) void k_push32(const float* __restrict__ src, float* __restrict__ dst, GridScalars g,
                                                const double* __restrict__ rho0, const double* __restrict__ u0,
                                                const double* __restrict__ fibm, unsigned nblocks) {
    const unsigned blk = XCD ? xcd_block(blockIdx.x, nblocks) : blockIdx.x;
    const long long id = (long long)blk * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
    const long long S = g.S;
    const int Ny = g.Ny, Nx = g.Nx;
    const int i = (int)(id / Ny), j = (int)(id - (long long)i * Ny);
    float gs[kNV], go[kNV];
#pragma unroll
    for (int v = 0; v < kNV; v++) gs[v] = (NT & 1) ? __builtin_nontemporal_load(&src[v * S + id]) : src[v * S + id];
    float drho, ux, uy, Fx = (float)g.Fx, Fy = (float)g.Fy;
    if (EXPLICIT) {
        drho = (float)(rho0[id] - 1.0);
        ux = (float)u0[2 * id];
        uy = (float)u0[2 * id + 1];
        Fx = (float)(g.Fx + fibm[id]);
        Fy = (float)(g.Fy + fibm[S + id]);
    } else {
        moments32(gs, Fx, Fy, drho, ux, uy);
    }
    collide32<CM>(g, gs, drho, ux, uy, Fx, Fy, go);
    const int im = i == 0 ? Nx - 1 : i - 1, ip = i == Nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? Ny - 1 : j - 1, jp = j == Ny - 1 ? 0 : j + 1;
    const long long r0 = (long long)i * Ny, rp = (long long)ip * Ny, rm = (long long)im * Ny;
    constexpr bool A = (NT & 2) != 0, B = (NT & 4) != 0;
    st32<A>(&dst[0 * S + r0 + j], go[0]);
    st32<A>(&dst[1 * S + rp + j], go[1]);
    st32<A>(&dst[2 * S + rm + j], go[2]);
    st32<B>(&dst[3 * S + r0 + jp], go[3]);
    st32<B>(&dst[4 * S + r0 + jm], go[4]);
    st32<B>(&dst[5 * S + rp + jp], go[5]);
    st32<B>(&dst[6 * S + rm + jm], go[6]);
    st32<B>(&dst[7 * S + rp + jm], go[7]);
    st32<B>(&dst[8 * S + rm + jp], go[8]);
}

// Bulk stream-collide with LDS-realigned stores (see push_store_realigned).
template <bool CM, bool XCD = true>
__global__ __launch_bounds__(kPushBlock) void k_push32_lds(const float* __restrict__ src, float* __restrict__ dst,
                                                           GridScalars g, unsigned nblocks) {
    __shared__ PushStage<float> st;
    const long long base = (long long)(XCD ? xcd_block(blockIdx.x, nblocks) : blockIdx.x) * kPushBlock;
    const int t = threadIdx.x;
    const long long id = base + t;
    const bool valid = id < g.N;
    const long long idc = valid ? id : 0;
    const long long S = g.S;
    const int Ny = g.Ny;
    const int i = (int)(idc / Ny), j = (int)(idc - (long long)i * Ny);
    float gs[kNV], go[kNV];
#pragma unroll
    for (int v = 0; v < kNV; v++) gs[v] = valid ? __builtin_nontemporal_load(&src[v * S + idc]) : 0.0f;
    float drho, ux, uy;
    const float Fx = (float)g.Fx, Fy = (float)g.Fy;
    moments32(gs, Fx, Fy, drho, ux, uy);
    collide32<CM>(g, gs, drho, ux, uy, Fx, Fy, go);
    push_store_realigned<float, true>(dst, S, g.Nx, Ny, g.N, base, t, idc, i, j, go, valid, st);
}

// Vector form of k_push32 for Ny % 4 == 0: each lane owns four consecutive
// sites of one row and moves each population plane with one aligned 16-byte
// load and one aligned 16-byte store.  Populations with c_y = +-1 land one
// site up/down the row; the element that crosses into the neighbouring
// lane's float4 is exchanged with a wave shuffle.  A lane at the start of a
// wave or row ("head") and one at the end ("tail") cannot complete that
// exchange and write the affected elements one by one instead, including
// the periodic wrap of the row (Grid.cpp:229, 240).
template <bool CM>
__global__ __launch_bounds__(256) void k_push32x4(const float* __restrict__ src, float* __restrict__ dst,
                                                  GridScalars g, unsigned nblocks) {
    const long long q = (long long)xcd_block(blockIdx.x, nblocks) * blockDim.x + threadIdx.x;
    const long long id0 = 4 * q;
    const bool valid = id0 < g.N;
    const long long S = g.S;
    const int Ny = g.Ny, Nx = g.Nx;
    const int lane = threadIdx.x & 63;
    const long long idc = valid ? id0 : 0;
    const int i = (int)(idc / Ny), j0 = (int)(idc - (long long)i * Ny);
    v4f in[kNV];
#pragma unroll
    for (int v = 0; v < kNV; v++)
        in[v] = valid ? __builtin_nontemporal_load(reinterpret_cast<const v4f*>(src + v * S + idc))
                      : v4f{0.f, 0.f, 0.f, 0.f};
    float o[kNV][4];
    const float Fx = (float)g.Fx, Fy = (float)g.Fy;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        float gs[kNV], go[kNV];
#pragma unroll
        for (int v = 0; v < kNV; v++) gs[v] = s == 0 ? in[v].x : (s == 1 ? in[v].y : (s == 2 ? in[v].z : in[v].w));
        float drho, ux, uy;
        moments32(gs, Fx, Fy, drho, ux, uy);
        collide32<CM>(g, gs, drho, ux, uy, Fx, Fy, go);
#pragma unroll
        for (int v = 0; v < kNV; v++) o[v][s] = go[v];
    }
    const bool head = lane == 0 || j0 == 0;
    const bool tail = lane == 63 || j0 + 4 == Ny;
    const int im = i == 0 ? Nx - 1 : i - 1, ip = i == Nx - 1 ? 0 : i + 1;
    const long long r0 = (long long)i * Ny, rp = (long long)ip * Ny, rm = (long long)im * Ny;
    // c_y = 0: whole float4 to the same j
    if (valid) {
        *reinterpret_cast<v4f*>(dst + 0 * S + r0 + j0) = v4f{o[0][0], o[0][1], o[0][2], o[0][3]};
        *reinterpret_cast<v4f*>(dst + 1 * S + rp + j0) = v4f{o[1][0], o[1][1], o[1][2], o[1][3]};
        *reinterpret_cast<v4f*>(dst + 2 * S + rm + j0) = v4f{o[2][0], o[2][1], o[2][2], o[2][3]};
    }
    // c_y = +1 (v = 3, 5, 8): site j -> j+1
    const int vup[3] = {3, 5, 8};
    const long long rup[3] = {r0, rp, rm};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int v = vup[k];
        const float prev = __shfl_up(o[v][3], 1);
        float* base = dst + v * S + rup[k];
        if (valid) {
            if (head) {
                base[j0 + 1] = o[v][0];
                base[j0 + 2] = o[v][1];
                base[j0 + 3] = o[v][2];
            } else {
                *reinterpret_cast<v4f*>(base + j0) = v4f{prev, o[v][0], o[v][1], o[v][2]};
            }
            if (tail) base[j0 + 4 == Ny ? 0 : j0 + 4] = o[v][3];
        }
    }
    // c_y = -1 (v = 4, 6, 7): site j -> j-1
    const int vdn[3] = {4, 6, 7};
    const long long rdn[3] = {r0, rm, rp};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int v = vdn[k];
        const float next = __shfl_down(o[v][0], 1);
        float* base = dst + v * S + rdn[k];
        if (valid) {
            if (tail) {
                base[j0] = o[v][1];
                base[j0 + 1] = o[v][2];
                base[j0 + 2] = o[v][3];
            } else {
                *reinterpret_cast<v4f*>(base + j0) = v4f{o[v][1], o[v][2], o[v][3], next};
            }
            if (head) base[j0 == 0 ? Ny - 1 : j0 - 1] = o[v][0];
        }
    }
}

template <bool CM>
__global__ __launch_bounds__(256) void k_support_push32(const float* __restrict__ src, float* __restrict__ dst,
                                                        GridScalars g, const double* __restrict__ fibm,
                                                        const int* __restrict__ sites,
                                                        const int* __restrict__ count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *count) return;
    const long long S = g.S, id = sites[k];
    const int Ny = g.Ny, Nx = g.Nx;
    const int i = (int)(id / Ny), j = (int)(id - (long long)i * Ny);
    float gs[kNV], go[kNV];
#pragma unroll
    for (int v = 0; v < kNV; v++) gs[v] = src[v * S + id];
    const float Fx = (float)(g.Fx + fibm[id]), Fy = (float)(g.Fy + fibm[S + id]);
    float drho, ux, uy;
    moments32(gs, Fx, Fy, drho, ux, uy);
    collide32<CM>(g, gs, drho, ux, uy, Fx, Fy, go);
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        int ii = i + kCX[v], jj = j + kCY[v];
        ii = ii < 0 ? Nx - 1 : (ii >= Nx ? 0 : ii);
        jj = jj < 0 ? Ny - 1 : (jj >= Ny ? 0 : jj);
        dst[v * S + (long long)ii * Ny + jj] = go[v];
    }
}

// Launch configuration of the bulk stream-collide kernel, read once from
// the environment for tuning sweeps: LIFE_PUSH32 = scalar | x4 (kernel),
// LIFE_PUSH32_NT = 0..7 (streaming-hint mask of k_push32), LIFE_PUSH32_XCD
// = 0 | 1 (block remap), LIFE_PUSH32_BS = 64 | 128 | 256 (block size).
struct PushConfig {
    bool x4 = false;
    bool lds = true;
    int nt = 3;
    bool xcd = true;
    int bs = 256;
    PushConfig() {
        if (const char* e = std::getenv("LIFE_PUSH32")) {
            x4 = std::strcmp(e, "x4") == 0;
            lds = std::strcmp(e, "lds") == 0;
        }
        if (const char* e = std::getenv("LIFE_PUSH32_NT")) nt = std::atoi(e) & 7;
        if (const char* e = std::getenv("LIFE_PUSH32_XCD")) xcd = std::atoi(e) != 0;
        if (const char* e = std::getenv("LIFE_PUSH32_BS")) bs = std::atoi(e);
        if (bs != 64 && bs != 128 && bs != 256) bs = 256;
    }
};

static const PushConfig& push_config() {
    static const PushConfig c;
    return c;
}

template <bool CM, int NT>
static void launch_push32_nt(const float* src, float* dst, const GridScalars& g, const PushConfig& c, hipStream_t s) {
    const unsigned nb = (unsigned)((g.N + c.bs - 1) / c.bs);
    if (c.xcd) k_push32<CM, false, NT, true><<<nb, c.bs, 0, s>>>(src, dst, g, nullptr, nullptr, nullptr, nb);
    else k_push32<CM, false, NT, false><<<nb, c.bs, 0, s>>>(src, dst, g, nullptr, nullptr, nullptr, nb);
}

template <bool CM>
static void launch_push32_bulk(const float* src, float* dst, const GridScalars& g, hipStream_t s) {
    const PushConfig& c = push_config();
    if (c.lds) {
        const unsigned nb = (unsigned)((g.N + kPushBlock - 1) / kPushBlock);
        if (c.xcd) k_push32_lds<CM, true><<<nb, kPushBlock, 0, s>>>(src, dst, g, nb);
        else k_push32_lds<CM, false><<<nb, kPushBlock, 0, s>>>(src, dst, g, nb);
        return;
    }
    if (c.x4 && g.Ny % 4 == 0) {
        const long long nq = g.N / 4;
        const unsigned nb = (unsigned)((nq + 255) / 256);
        k_push32x4<CM><<<nb, 256, 0, s>>>(src, dst, g, nb);
        return;
    }
    switch (c.nt) {
        case 0: launch_push32_nt<CM, 0>(src, dst, g, c, s); break;
        case 1: launch_push32_nt<CM, 1>(src, dst, g, c, s); break;
        case 2: launch_push32_nt<CM, 2>(src, dst, g, c, s); break;
        case 3: launch_push32_nt<CM, 3>(src, dst, g, c, s); break;
        case 4: launch_push32_nt<CM, 4>(src, dst, g, c, s); break;
        case 5: launch_push32_nt<CM, 5>(src, dst, g, c, s); break;
        case 6: launch_push32_nt<CM, 6>(src, dst, g, c, s); break;
        default: launch_push32_nt<CM, 7>(src, dst, g, c, s); break;
    }
}

void launch_push32(const float* src, float* dst, const GridScalars& g, const double* rho0, const double* u0,
                   const double* fibm, bool explicit_state, hipStream_t s) {
    if (!explicit_state) {
        if (g.cm) launch_push32_bulk<true>(src, dst, g, s);
        else launch_push32_bulk<false>(src, dst, g, s);
        return;
    }
    const unsigned nb = (unsigned)((g.N + 255) / 256);
    if (g.cm) k_push32<true, true><<<nb, 256, 0, s>>>(src, dst, g, rho0, u0, fibm, nb);
    else k_push32<false, true><<<nb, 256, 0, s>>>(src, dst, g, rho0, u0, fibm, nb);
}

void launch_support_push32(const float* src, float* dst, const GridScalars& g, const double* fibm, const int* sites,
                           const int* count, int max_sites, hipStream_t s) {
    if (max_sites <= 0) return;
    const unsigned nb = (unsigned)((max_sites + 255) / 256);
    if (g.cm) k_support_push32<true><<<nb, 256, 0, s>>>(src, dst, g, fibm, sites, count);
    else k_support_push32<false><<<nb, 256, 0, s>>>(src, dst, g, fibm, sites, count);
}

}  // namespace life
