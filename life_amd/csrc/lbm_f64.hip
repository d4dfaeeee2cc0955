// Plain method (FP64 populations): fused push stream-collide kernels.
//
// Reference step (Grid.cpp:36-100): every site collides with the rho/u/F
// stored at the end of the previous step and pushes its nine post-collision
// populations to its neighbours with periodic wrap (Grid.cpp:226-244).  The
// stored rho/u are the moments of the stored populations, with force_xy
// (Grid.cpp:282-299) or, at sites that received an IBM force,
// force_xy + force_ibm (IBMNode.cpp:97-136).  This kernel therefore reads
// only f (72 B/site) and recomputes the moments; the few support sites of
// the last spread are redone by k_support_push with their force_ibm.
//
// Compiled with -ffp-contract=off: bit-compatible with the reference.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include "d2q9_f64.h"
#include "dev_util.h"
#include "launch.h"

namespace life {

template <bool CM, bool EXPLICIT>
__global__ __launch_bounds__(256) void k_push64(const double* __restrict__ src, double* __restrict__ dst,
                                                GridScalars g, const double* __restrict__ rho0,
                                                const double* __restrict__ u0, const double* __restrict__ fibm,
                                                unsigned nblocks) {
    const long long id = (long long)xcd_block(blockIdx.x, nblocks) * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
    const long long S = g.S;
    const int Ny = g.Ny, Nx = g.Nx;
    const int i = (int)(id / Ny), j = (int)(id - (long long)i * Ny);
    double fs[kNV], fo[kNV];
    load_site(src, S, id, fs);
    double rho, ux, uy, Fx = g.Fx0, Fy = g.Fy0;
    if (EXPLICIT) {
        // rho/u/force_ibm exactly as stored by initialisation or a restart
        rho = rho0[id];
        ux = u0[2 * id];
        uy = u0[2 * id + 1];
        double sx = g.Fx, sy = g.Fy;
        if (g.fcur) {
            sx = site_force_x(g, rho);
            sy = site_force_y(g, rho);
            g.fcur[id] = sx;
            g.fcur[S + id] = sy;
        }
        Fx = sx + fibm[id];
        Fy = sy + fibm[S + id];
    } else if (g.fcur) {
        moments64(fs, g.fprev[id], g.fprev[S + id], rho, ux, uy);
        const double sx = site_force_x(g, rho), sy = site_force_y(g, rho);
        g.fcur[id] = sx;
        g.fcur[S + id] = sy;
        Fx = sx + 0.0;
        Fy = sy + 0.0;
    } else {
        moments64(fs, g.Fxm, g.Fym, rho, ux, uy);
    }
    collide64<CM>(g, fs, rho, ux, uy, Fx, Fy, fo);
    const int im = i == 0 ? Nx - 1 : i - 1, ip = i == Nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? Ny - 1 : j - 1, jp = j == Ny - 1 ? 0 : j + 1;
    const long long r0 = (long long)i * Ny, rp = (long long)ip * Ny, rm = (long long)im * Ny;
    dst[0 * S + r0 + j] = fo[0];
    dst[1 * S + rp + j] = fo[1];
    dst[2 * S + rm + j] = fo[2];
    dst[3 * S + r0 + jp] = fo[3];
    dst[4 * S + r0 + jm] = fo[4];
    dst[5 * S + rp + jp] = fo[5];
    dst[6 * S + rm + jm] = fo[6];
    dst[7 * S + rp + jm] = fo[7];
    dst[8 * S + rm + jp] = fo[8];
}

// Bulk step with LDS-realigned stores (push_store_realigned): the same
// values reach the same addresses as k_push64, with aligned wavefront stores.
// SP/DP: double pointers, generic or in the global address space.  SITEF:
// force_xy per site (g.fprev / g.fcur), else the uniform g.Fxm / g.Fx0.
template <bool CM, bool SITEF, int BS, bool NTS, typename SP, typename DP>
__device__ __forceinline__ void push64_lds_block(SP __restrict__ src, DP __restrict__ dst, const GridScalars& g,
                                                 long long base, int t, PushStage<double, BS>& st) {
    const long long id = base + t;
    const bool valid = id < g.N;
    const long long idc = valid ? id : 0;
    const long long S = g.S;
    const int Ny = g.Ny;
    const int i = (int)(idc / Ny), j = (int)(idc - (long long)i * Ny);
    double fs[kNV], fo[kNV];
#pragma unroll
    for (int v = 0; v < kNV; v++) fs[v] = valid ? __builtin_nontemporal_load(&src[v * S + idc]) : 1.0;
    double rho, ux, uy;
    if (SITEF) {
        moments64(fs, g.fprev[idc], g.fprev[S + idc], rho, ux, uy);
        const double sx = site_force_x(g, rho), sy = site_force_y(g, rho);
        if (valid) {
            g.fcur[idc] = sx;
            g.fcur[S + idc] = sy;
        }
        collide64<CM>(g, fs, rho, ux, uy, sx + 0.0, sy + 0.0, fo);
    } else {
        moments64(fs, g.Fxm, g.Fym, rho, ux, uy);
        collide64<CM>(g, fs, rho, ux, uy, g.Fx0, g.Fy0, fo);
    }
    push_store_realigned<double, NTS, BS>(dst, S, g.Nx, Ny, g.N, base, t, idc, i, j, fo, valid, st);
}

template <bool CM, bool SITEF, int BS = kPushBlock, bool NTS = true>
__global__ __launch_bounds__(BS) void k_push64_lds(const double* __restrict__ src, double* __restrict__ dst,
                                                   GridScalars g, unsigned nblocks) {
    __shared__ PushStage<double, BS> st;
    const long long base = (long long)xcd_block(blockIdx.x, nblocks) * BS;
    push64_lds_block<CM, SITEF, BS, NTS>(src, dst, g, base, threadIdx.x, st);
}

// Push of one site with periodic wrap (Grid.cpp:226-244).
__device__ __forceinline__ void push_site64(double* __restrict__ dst, const GridScalars& g, long long id,
                                            const double fo[kNV]) {
    const long long S = g.S;
    const int Ny = g.Ny, Nx = g.Nx;
    const int i = (int)(id / Ny), j = (int)(id - (long long)i * Ny);
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        int ii = i + kCX[v], jj = j + kCY[v];
        ii = ii < 0 ? Nx - 1 : (ii >= Nx ? 0 : ii);
        jj = jj < 0 ? Ny - 1 : (jj >= Ny ? 0 : jj);
        dst[v * S + (long long)ii * Ny + jj] = fo[v];
    }
}

// Serving: one launch over the blocks of every admitted case of one
// collision operator; a case's first step takes rho/u/force_ibm as stored.
template <bool CM>
__global__ __launch_bounds__(kServeBlock64) void k_push64_batch(const CaseDev* __restrict__ cases,
                                                             const CaseBlock* __restrict__ blocks) {
    __shared__ PushStage<double, kServeBlock64> st;
    // neighbouring blocks of a case on one XCD (the L2 that merges their
    // shared edge lines)
    const CaseBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const CaseDev& c = cases[b.c];
    const long long base = (long long)b.blk * kServeBlock64;
    const double* src = static_cast<const double*>(c.src);
    double* dst = static_cast<double*>(c.dst);
    if (c.explicit_state) {
        const long long id = base + threadIdx.x;
        if (id >= c.g.N) return;
        double fs[kNV], fo[kNV];
        load_site(src, c.g.S, id, fs);
        const double Fx = c.g.Fx + c.fibm[id], Fy = c.g.Fy + c.fibm[c.g.S + id];
        collide64<CM>(c.g, fs, c.rho0[id], c.u0[2 * id], c.u0[2 * id + 1], Fx, Fy, fo);
        push_site64(dst, c.g, id, fo);
    } else {
        push64_lds_block<CM, false, kServeBlock64, true>((gptr<const double>)c.src, (gptr<double>)c.dst, c.g, base,
                                                         threadIdx.x, st);
    }
}

// Redo the push of the support sites of the last spread, where the stored
// u includes force_ibm and the Guo/central-moment forcing uses
// force_xy + force_ibm.  sites[0..*count) are lattice ids.
template <bool CM>
__device__ __forceinline__ void support_push64_site(const double* __restrict__ src, double* __restrict__ dst,
                                                    const GridScalars& g, const double* __restrict__ fibm,
                                                    long long id) {
    const long long S = g.S;
    double fs[kNV], fo[kNV];
    load_site(src, S, id, fs);
    // moments as stored at the end of the previous step (its force_xy),
    // forcing of this step
    const double Fxm = (g.fprev ? g.fprev[id] : g.Fxm) + fibm[id];
    const double Fym = (g.fprev ? g.fprev[S + id] : g.Fym) + fibm[S + id];
    double rho, ux, uy;
    moments64(fs, Fxm, Fym, rho, ux, uy);
    const double Fx = (g.fcur ? site_force_x(g, rho) : g.Fx) + fibm[id];
    const double Fy = (g.fcur ? site_force_y(g, rho) : g.Fy) + fibm[S + id];
    collide64<CM>(g, fs, rho, ux, uy, Fx, Fy, fo);
    push_site64(dst, g, id, fo);
}

template <bool CM>
__global__ __launch_bounds__(256) void k_support_push64(const double* __restrict__ src, double* __restrict__ dst,
                                                        GridScalars g, const double* __restrict__ fibm,
                                                        const int* __restrict__ sites,
                                                        const int* __restrict__ count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < *count) support_push64_site<CM>(src, dst, g, fibm, sites[k]);
}

template <bool CM>
__global__ __launch_bounds__(256) void k_support_push64_batch(const CaseDev* __restrict__ cases,
                                                              const int* __restrict__ list) {
    const CaseDev& c = cases[list[blockIdx.y]];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (c.explicit_state || c.m.n <= 0 || k >= *c.sc.uniq_n) return;
    support_push64_site<CM>(static_cast<const double*>(c.src), static_cast<double*>(c.dst), c.g, c.fibm,
                            c.sc.uniq[k]);
}

// Bulk kernel configuration, read at every launch (tuning sweeps):
// LIFE_PUSH64 = scalar selects k_push64; LIFE_PUSH64_BS = 256 | 512 | 1024
// lanes per block of k_push64_lds (512 by default: on the 8192 x 4096
// lattice a step took 0.967-1.023 ms against 1.044-1.048 ms at 256 and
// 0.970-0.982 at 1024); LIFE_PUSH64_NT = 0 | 1 streaming stores (off:
// 13-16 % slower).
struct Push64Config {
    bool lds = true;
    int bs = 512;
    bool nt = true;
    Push64Config() {
        if (const char* e = std::getenv("LIFE_PUSH64")) lds = std::strcmp(e, "scalar") != 0;
        if (const char* e = std::getenv("LIFE_PUSH64_BS")) {
            const int v = std::atoi(e);
            bs = v == 256 || v == 1024 ? v : 512;
        }
        if (const char* e = std::getenv("LIFE_PUSH64_NT")) nt = std::atoi(e) != 0;
    }
};

template <bool CM, bool SITEF>
static void launch_push64_lds(const double* src, double* dst, const GridScalars& g, const Push64Config& c,
                              hipStream_t s) {
    const unsigned nb = (unsigned)((g.N + c.bs - 1) / c.bs);
    if (c.bs == 512) {
        if (c.nt) k_push64_lds<CM, SITEF, 512, true><<<nb, 512, 0, s>>>(src, dst, g, nb);
        else k_push64_lds<CM, SITEF, 512, false><<<nb, 512, 0, s>>>(src, dst, g, nb);
    } else if (c.bs == 1024) {
        if (c.nt) k_push64_lds<CM, SITEF, 1024, true><<<nb, 1024, 0, s>>>(src, dst, g, nb);
        else k_push64_lds<CM, SITEF, 1024, false><<<nb, 1024, 0, s>>>(src, dst, g, nb);
    } else {
        if (c.nt) k_push64_lds<CM, SITEF, kPushBlock, true><<<nb, kPushBlock, 0, s>>>(src, dst, g, nb);
        else k_push64_lds<CM, SITEF, kPushBlock, false><<<nb, kPushBlock, 0, s>>>(src, dst, g, nb);
    }
}

void launch_push64(const double* src, double* dst, const GridScalars& g, const double* rho0, const double* u0,
                   const double* fibm, bool explicit_state, hipStream_t s) {
    const Push64Config c;
    if (!explicit_state && c.lds) {
        if (g.fcur) {
            if (g.cm) launch_push64_lds<true, true>(src, dst, g, c, s);
            else launch_push64_lds<false, true>(src, dst, g, c, s);
        } else {
            if (g.cm) launch_push64_lds<true, false>(src, dst, g, c, s);
            else launch_push64_lds<false, false>(src, dst, g, c, s);
        }
        return;
    }
    const unsigned nb = (unsigned)((g.N + 255) / 256);
    if (g.cm) {
        if (explicit_state) k_push64<true, true><<<nb, 256, 0, s>>>(src, dst, g, rho0, u0, fibm, nb);
        else k_push64<true, false><<<nb, 256, 0, s>>>(src, dst, g, rho0, u0, fibm, nb);
    } else {
        if (explicit_state) k_push64<false, true><<<nb, 256, 0, s>>>(src, dst, g, rho0, u0, fibm, nb);
        else k_push64<false, false><<<nb, 256, 0, s>>>(src, dst, g, rho0, u0, fibm, nb);
    }
}

void launch_support_push64(const double* src, double* dst, const GridScalars& g, const double* fibm,
                           const int* sites, const int* count, int max_sites, hipStream_t s) {
    if (max_sites <= 0) return;
    const unsigned nb = (unsigned)((max_sites + 255) / 256);
    if (g.cm) k_support_push64<true><<<nb, 256, 0, s>>>(src, dst, g, fibm, sites, count);
    else k_support_push64<false><<<nb, 256, 0, s>>>(src, dst, g, fibm, sites, count);
}

void launch_push64_batch(bool cm, const CaseDev* cases, const CaseBlock* blocks, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return;
    if (cm) k_push64_batch<true><<<(unsigned)nblocks, kServeBlock64, 0, s>>>(cases, blocks);
    else k_push64_batch<false><<<(unsigned)nblocks, kServeBlock64, 0, s>>>(cases, blocks);
}

void launch_support_push64_batch(bool cm, const CaseDev* cases, const int* list, int n, int max_sites, hipStream_t s) {
    if (n <= 0 || max_sites <= 0) return;
    const dim3 grid((unsigned)((max_sites + 255) / 256), (unsigned)n);
    if (cm) k_support_push64_batch<true><<<grid, 256, 0, s>>>(cases, list);
    else k_support_push64_batch<false><<<grid, 256, 0, s>>>(cases, list);
}

}  // namespace life
