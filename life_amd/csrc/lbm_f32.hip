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

#include "d2q9_f32.h"
#include "dev_util.h"
#include "launch.h"

namespace life {

using v4f = float __attribute__((ext_vector_type(4)));

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
        moments32(gs, (float)g.Fxm, (float)g.Fym, drho, ux, uy);
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

// Bulk stream-collide body of one block of kPushBlock sites from `base`.
// SP/DP: float pointers, generic or in the global address space.  BS:
// sites per block; NTS: streaming hint on the stores.
template <bool CM, typename SP, typename DP, int BS = kPushBlock, bool NTS = true>
__device__ __forceinline__ void push32_lds_block(SP __restrict__ src, DP __restrict__ dst, const GridScalars& g,
                                                 long long base, int t, PushStage<float, BS>& st) {
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
    moments32(gs, (float)g.Fxm, (float)g.Fym, drho, ux, uy);
    collide32<CM>(g, gs, drho, ux, uy, Fx, Fy, go);
    push_store_realigned<float, NTS, BS>(dst, S, g.Nx, Ny, g.N, base, t, idc, i, j, go, valid, st);
}

// Bulk stream-collide with LDS-realigned stores (see push_store_realigned).
template <bool CM, bool XCD = true, int BS = kPushBlock, bool NTS = true>
__global__ __launch_bounds__(BS) void k_push32_lds(const float* __restrict__ src, float* __restrict__ dst,
                                                   GridScalars g, unsigned nblocks) {
    __shared__ PushStage<float, BS> st;
    const long long base = (long long)(XCD ? xcd_block(blockIdx.x, nblocks) : blockIdx.x) * BS;
    push32_lds_block<CM, const float*, float*, BS, NTS>(src, dst, g, base, threadIdx.x, st);
}

// Serving: one launch over the blocks of every admitted case of one
// collision operator.  A case's first step after admission takes its
// rho/u/force_ibm as stored (the explicit path of k_push32).
template <bool CM>
__global__ __launch_bounds__(kServeBlock32) void k_push32_batch(const CaseDev* __restrict__ cases,
                                                                const CaseBlock* __restrict__ blocks) {
    __shared__ PushStage<float, kServeBlock32> st;
    // neighbouring blocks of a case on one XCD (the L2 that merges their
    // shared edge lines)
    const CaseBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const CaseDev& c = cases[b.c];
    const long long base = (long long)b.blk * kServeBlock32;
    if (c.explicit_state) {
        const long long id = base + threadIdx.x;
        if (id < c.g.N)
            push32_explicit_site<CM>(static_cast<const float*>(c.src), static_cast<float*>(c.dst), c.g, c.rho0, c.u0,
                                     c.fibm, id);
    } else {
        // table pointers are global memory: typed so, the bulk loads and
        // stores are global (not flat) instructions
        push32_lds_block<CM, gptr<const float>, gptr<float>, kServeBlock32>((gptr<const float>)c.src,
                                                                           (gptr<float>)c.dst, c.g, base,
                                                                           threadIdx.x, st);
    }
}

template <bool CM>
__global__ __launch_bounds__(256) void k_support_push32_batch(const CaseDev* __restrict__ cases,
                                                              const int* __restrict__ list) {
    const CaseDev& c = cases[list[blockIdx.y]];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (c.explicit_state || c.m.n <= 0 || k >= *c.sc.uniq_n) return;
    support_push32_site<CM>(static_cast<const float*>(c.src), static_cast<float*>(c.dst), c.g, c.fibm, c.sc.uniq[k]);
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
        moments32(gs, (float)g.Fxm, (float)g.Fym, drho, ux, uy);
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

// Four sites per lane with LDS realignment, for Ny % 4 == 0 (a lane's four
// sites then lie in one row, and every float4 of a plane is 16-byte
// aligned).  Each population plane moves with one 16-byte load per lane;
// the c_y = 0 populations go out with one 16-byte store (their
// destination is the same four sites one row over).  The six c_y = +-1
// populations land one site up or down the row: a workgroup drops them
// into LDS at their destination offset within its block of 4*BS sites and
// every lane then stores the float4 that lands on its own four sites.
// Values that leave the block or wrap around the row (Grid.cpp:229, 240)
// are stored directly by their source lane; a lane whose float4 holds a
// destination with no source in the block (the block's first site for
// c_y = +1, its last for c_y = -1) or a wrapped row end (j = 0 for
// c_y = +1, j = Ny - 1 for c_y = -1) stores its other three elements one
// by one.  The arithmetic is k_push32's, site by site.
template <int BS>
struct PushStage4 {
    // 4 floats of padding at each end keep the lane reads 16-byte aligned
    float slot[6][4 * BS + 8];
};

template <bool CM, int BS>
__global__ __launch_bounds__(BS) void k_push32x4_lds(const float* __restrict__ src, float* __restrict__ dst,
                                                     GridScalars g, unsigned nblocks) {
    __shared__ __attribute__((aligned(16))) PushStage4<BS> st;
    const int t = threadIdx.x;
    const long long base = (long long)xcd_block(blockIdx.x, nblocks) * (4 * BS);
    const long long id0 = base + 4 * t;
    const bool valid = id0 < g.N;  // N is a multiple of 4 (Ny is)
    const long long S = g.S;
    const int Ny = g.Ny, Nx = g.Nx;
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
        for (int v = 0; v < kNV; v++) gs[v] = in[v][s];
        float drho, ux, uy;
        moments32(gs, (float)g.Fxm, (float)g.Fym, drho, ux, uy);
        collide32<CM>(g, gs, drho, ux, uy, Fx, Fy, go);
#pragma unroll
        for (int v = 0; v < kNV; v++) o[v][s] = go[v];
    }
    const int im = i == 0 ? Nx - 1 : i - 1, ip = i == Nx - 1 ? 0 : i + 1;
    const long long r0 = (long long)i * Ny, rp = (long long)ip * Ny, rm = (long long)im * Ny;
    if (valid) {
        __builtin_nontemporal_store(v4f{o[0][0], o[0][1], o[0][2], o[0][3]},
                                    reinterpret_cast<v4f*>(dst + 0 * S + r0 + j0));
        __builtin_nontemporal_store(v4f{o[1][0], o[1][1], o[1][2], o[1][3]},
                                    reinterpret_cast<v4f*>(dst + 1 * S + rp + j0));
        __builtin_nontemporal_store(v4f{o[2][0], o[2][1], o[2][2], o[2][3]},
                                    reinterpret_cast<v4f*>(dst + 2 * S + rm + j0));
    }
    // cy = +1: v = 3 (cx 0), 5 (+1), 8 (-1); cy = -1: v = 4 (0), 6 (-1), 7 (+1)
    constexpr int kv[6] = {3, 5, 8, 4, 6, 7};
    constexpr int kcy[6] = {1, 1, 1, -1, -1, -1};
    const long long rowv[6] = {r0, rp, rm, r0, rm, rp};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int cy = kcy[q];
        if (!valid) continue;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int jd = j0 + s + cy;
            const int to = 4 * t + s + cy;  // destination offset within the block
            if (jd < 0 || jd >= Ny) dst[kv[q] * S + rowv[q] + (jd < 0 ? Ny - 1 : 0)] = o[kv[q]][s];
            else if (to < 0 || to >= 4 * BS) dst[kv[q] * S + rowv[q] + jd] = o[kv[q]][s];
            else st.slot[q][to + 4] = o[kv[q]][s];
        }
    }
    __syncthreads();
    if (!valid) return;
    // lane t stores the values destined for sites 4t .. 4t+3 of the block;
    // the one (at most) without a source in the block is left out
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int cy = kcy[q];
        const v4f w = *reinterpret_cast<const v4f*>(&st.slot[q][4 * t + 4]);
        float* row = dst + kv[q] * S + rowv[q];
        // element without a source: first site of the block or a row start
        // (cy = +1); last site of the block (or of the lattice) or a row end
        // (cy = -1)
        const bool gap = cy > 0 ? (t == 0 || j0 == 0) : (t == BS - 1 || id0 + 4 >= g.N || j0 + 4 == Ny);
        if (!gap) {
            __builtin_nontemporal_store(w, reinterpret_cast<v4f*>(row + j0));
        } else if (cy > 0) {
            row[j0 + 1] = w[1];
            row[j0 + 2] = w[2];
            row[j0 + 3] = w[3];
        } else {
            row[j0] = w[0];
            row[j0 + 1] = w[1];
            row[j0 + 2] = w[2];
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
    support_push32_site<CM>(src, dst, g, fibm, sites[k]);
}

// Launch configuration of the bulk stream-collide kernel, read once from
// the environment for tuning sweeps: LIFE_PUSH32 = scalar | x4 | lds |
// x4lds (kernel; lds is the default: on the 8192 x 4096 canopy it measured
// 0.438 ms alone against 0.452-0.459 ms for x4lds at every block size, and
// the small x4lds blocks delay the coupling kernels beside it; x4lds needs
// Ny % 4 == 0 and falls back to lds otherwise), LIFE_PUSH32_X4BS = 64 |
// 128 | 256 | 512 (lanes per block of
// k_push32x4_lds), LIFE_PUSH32_NT = 0..7 (streaming-hint mask of k_push32),
// LIFE_PUSH32_XCD = 0 | 1 (block remap), LIFE_PUSH32_BS = 64 | 128 | 256
// (block size of k_push32), LIFE_PUSH32_LDSBS = 256 | 512 | 1024 (block size
// of k_push32_lds; 512 measured best on MI355X), LIFE_PUSH32_LDSNT = 0 | 1
// (streaming stores in k_push32_lds).
struct PushConfig {
    bool x4 = false;
    bool x4lds = false;
    int x4lds_bs = 256;
    bool lds = true;
    int nt = 3;
    bool xcd = true;
    int bs = 256;
    int lds_bs = 512;
    bool lds_nt = true;
    PushConfig() {
        if (const char* e = std::getenv("LIFE_PUSH32")) {
            x4 = std::strcmp(e, "x4") == 0;
            x4lds = std::strcmp(e, "x4lds") == 0;
            lds = x4lds || std::strcmp(e, "lds") == 0;  // x4lds falls back to lds
        }
        if (const char* e = std::getenv("LIFE_PUSH32_X4BS")) {
            const int v = std::atoi(e);
            x4lds_bs = v == 64 || v == 128 || v == 512 ? v : 256;
        }
        if (const char* e = std::getenv("LIFE_PUSH32_NT")) nt = std::atoi(e) & 7;
        if (const char* e = std::getenv("LIFE_PUSH32_XCD")) xcd = std::atoi(e) != 0;
        if (const char* e = std::getenv("LIFE_PUSH32_BS")) bs = std::atoi(e);
        if (bs != 64 && bs != 128 && bs != 256) bs = 256;
        if (const char* e = std::getenv("LIFE_PUSH32_LDSBS")) {
            const int v = std::atoi(e);
            lds_bs = v == 256 || v == 1024 ? v : 512;
        }
        if (const char* e = std::getenv("LIFE_PUSH32_LDSNT")) lds_nt = std::atoi(e) != 0;
    }
};

// Read at every launch, so that a test can switch kernels between steps.
static PushConfig push_config() { return PushConfig(); }

template <bool CM, int NT>
static void launch_push32_nt(const float* src, float* dst, const GridScalars& g, const PushConfig& c, hipStream_t s) {
    const unsigned nb = (unsigned)((g.N + c.bs - 1) / c.bs);
    if (c.xcd) k_push32<CM, false, NT, true><<<nb, c.bs, 0, s>>>(src, dst, g, nullptr, nullptr, nullptr, nb);
    else k_push32<CM, false, NT, false><<<nb, c.bs, 0, s>>>(src, dst, g, nullptr, nullptr, nullptr, nb);
}

template <bool CM>
static void launch_push32_bulk(const float* src, float* dst, const GridScalars& g, hipStream_t s) {
    const PushConfig c = push_config();
    if (c.x4lds && g.Ny % 4 == 0) {
        const long long sites = 4LL * c.x4lds_bs;
        const unsigned nb = (unsigned)((g.N + sites - 1) / sites);
        switch (c.x4lds_bs) {
            case 64: k_push32x4_lds<CM, 64><<<nb, 64, 0, s>>>(src, dst, g, nb); break;
            case 128: k_push32x4_lds<CM, 128><<<nb, 128, 0, s>>>(src, dst, g, nb); break;
            case 512: k_push32x4_lds<CM, 512><<<nb, 512, 0, s>>>(src, dst, g, nb); break;
            default: k_push32x4_lds<CM, 256><<<nb, 256, 0, s>>>(src, dst, g, nb); break;
        }
        return;
    }
    if (c.lds) {
        const unsigned nb = (unsigned)((g.N + c.lds_bs - 1) / c.lds_bs);
        if (c.lds_bs == 512) {
            if (c.lds_nt) k_push32_lds<CM, true, 512, true><<<nb, 512, 0, s>>>(src, dst, g, nb);
            else k_push32_lds<CM, true, 512, false><<<nb, 512, 0, s>>>(src, dst, g, nb);
        } else if (c.lds_bs == 1024) {
            k_push32_lds<CM, true, 1024, true><<<nb, 1024, 0, s>>>(src, dst, g, nb);
        } else if (!c.lds_nt) {
            k_push32_lds<CM, true, kPushBlock, false><<<nb, kPushBlock, 0, s>>>(src, dst, g, nb);
        } else if (c.xcd) {
            k_push32_lds<CM, true><<<nb, kPushBlock, 0, s>>>(src, dst, g, nb);
        } else {
            k_push32_lds<CM, false><<<nb, kPushBlock, 0, s>>>(src, dst, g, nb);
        }
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

void launch_push32_batch(bool cm, const CaseDev* cases, const CaseBlock* blocks, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return;
    if (cm) k_push32_batch<true><<<(unsigned)nblocks, kServeBlock32, 0, s>>>(cases, blocks);
    else k_push32_batch<false><<<(unsigned)nblocks, kServeBlock32, 0, s>>>(cases, blocks);
}

void launch_support_push32_batch(bool cm, const CaseDev* cases, const int* list, int n, int max_sites, hipStream_t s) {
    if (n <= 0 || max_sites <= 0) return;
    const dim3 grid((unsigned)((max_sites + 255) / 256), (unsigned)n);
    if (cm) k_support_push32_batch<true><<<grid, 256, 0, s>>>(cases, list);
    else k_support_push32_batch<false><<<grid, 256, 0, s>>>(cases, list);
}

}  // namespace life
