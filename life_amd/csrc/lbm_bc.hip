// Boundary conditions, convective outlet, initialisation and state
// materialisation.  These touch O(Nx+Ny) sites per step, so they always
// compute in FP64 with the reference's operation order; for the FP32-shifted
// storage they read stored + w and write back value - w.
//
// Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "d2q9_f64.h"
#include "launch.h"

namespace life {

// GridClass::regularisedBC (Grid.cpp:387-465, Latt et al. 2008) on one
// site's nine post-stream populations fs, with rho_n/u_n as set by applyBCs.
// rhoNb1/rhoNb2 are the current densities at site + n and site + 2n, used
// by the corner extrapolation (Utils.h:150-162).
template <bool CM>
__device__ void regularised64(const GridScalars& g, int type, int nx, int ny, int nd, double fs[kNV], double& rn,
                              double& uxn, double& uyn, double rhoNb1, double rhoNb2) {
    const bool velocity_like = (type == 1 || type == 2 || type == 3);  // wall, velocity, free-slip
    const bool corner = (nx != 0 && ny != 0);
    if (corner) {
        if (velocity_like) rn = 2.0 * rhoNb1 - rhoNb2;
    } else {
        const int nn = nd == 0 ? nx : ny;
        double fplus = 0.0, fzero = 0.0;
#pragma unroll
        for (int v = 0; v < kNV; v++) {
            const int cn = nd == 0 ? kCX[v] : kCY[v];
            if (cn == -nn) fplus += fs[v];
            else if (cn == 0) fzero += fs[v];
        }
        if (velocity_like) rn = (2.0 * fplus + fzero) / (1.0 - nn * (nd == 0 ? uxn : uyn));
        else if (type == 4) {
            const double un = nn * (1.0 - (2.0 * fplus + fzero) / rn);
            if (nd == 0) uxn = un;
            else uyn = un;
        }
    }
    const double rho = rn, ux = uxn, uy = uyn;
    double Sxx = 0.0, Syy = 0.0, Sxy = 0.0;
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        const double feq = feq64<CM>(rho, ux, uy, v);
        bool unknown;
        if (corner) unknown = (kCX[v] == nx || kCY[v] == ny);
        else unknown = ((nd == 0 ? kCX[v] : kCY[v]) == (nd == 0 ? nx : ny));
        if (unknown) {
            if (corner && nx * kCX[v] + ny * kCY[v] == 0) fs[v] = feq;
            else fs[v] = feq + (fs[kOpp[v]] - feq64<CM>(rho, ux, uy, kOpp[v]));
        }
        const double fneq = fs[v] - feq;
        Sxx += kCX[v] * kCX[v] * fneq;
        Syy += kCY[v] * kCY[v] * fneq;
        Sxy += kCX[v] * kCY[v] * fneq;
    }
#pragma unroll
    for (int v = 0; v < kNV; v++)
        fs[v] = feq64<CM>(rho, ux, uy, v) +
                (w64(v) / (2.0 * g.cs4)) * (((kCX[v] * kCX[v] - g.cs2) * Sxx) + ((kCY[v] * kCY[v] - g.cs2) * Syy) +
                                            (2.0 * kCX[v] * kCY[v] * Sxy));
}

// One thread per boundary site: GridClass::applyBCs + regularisedBC /
// convectiveBC (Grid.cpp:302-474) on the populations just streamed into
// dst.  rho_n/u_n start as the site's end-of-previous-step moments (the
// reference swaps its rho/u arrays at the start of the step, Grid.cpp:43-45).
template <typename T, bool CM>
__device__ inline void bc_site(const T* __restrict__ src, T* __restrict__ dst, const GridScalars& g, const BCArgs& a,
                               const double* __restrict__ fibm, const double* __restrict__ rho0,
                               const double* __restrict__ u0, int explicit_state, int k) {
    if (k >= a.nbc) return;
    const BCSite b = a.sites[k];
    const long long S = g.S, id = b.id;
    const int Ny = g.Ny, j = b.j;
    if (b.type == 5) {  // convective: f[2,6,8] = f_n + 3 w (delU . c)
        const double dux = a.delU[2 * j], duy = a.delU[2 * j + 1];
        const int vs[3] = {2, 6, 8};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int v = vs[q];
            Pops<T>::put(dst, S, v, id, Pops<T>::get(src, S, v, id) + 3.0 * w64(v) * (dux * kCX[v] + duy * kCY[v]));
        }
        return;
    }
    if (b.type == 0) return;
    // the populations regularisedBC rebuilds from the others (the ones that
    // entered from outside the domain) are never read, so only the known
    // ones are loaded: six of nine on an edge
    const bool corner = b.nx != 0 && b.ny != 0;
    // Every load the site needs is issued before the arithmetic (one round
    // trip to memory, not up to three): its known populations, the two
    // inward neighbours' populations (the zero gradient of free-slip and
    // pressure sites, the density extrapolation of corners) and, at a
    // pressure corner, its previous populations.
    const bool nbr = b.type == 3 || b.type == 4 || corner;
    const bool prev = b.type == 4 && corner && !explicit_state;
    const long long id1 = (long long)(b.i + b.nx) * Ny + (j + b.ny);
    const long long id2 = (long long)(b.i + 2 * b.nx) * Ny + (j + 2 * b.ny);
    double fs[kNV], f1[kNV], f2[kNV], fp[kNV];
#pragma unroll
    for (int v = 0; v < kNV; v++) {
        const bool unknown = corner ? (kCX[v] == b.nx || kCY[v] == b.ny)
                                    : ((b.nd == 0 ? kCX[v] : kCY[v]) == (b.nd == 0 ? b.nx : b.ny));
        fs[v] = unknown ? 0.0 : Pops<T>::get(dst, S, v, id);
    }
    if (nbr) {
        load_site(dst, S, id1, f1);
        load_site(dst, S, id2, f2);
    }
    if (prev) load_site(src, S, id, fp);
    // rho_n/u_n of the previous step outlive applyBCs + regularisedBC only
    // at a pressure corner: everywhere else all three are overwritten below
    // (wall/velocity set u, free-slip and pressure take u from the normal
    // condition and the zero gradient, rho comes from the post-stream
    // populations, the corner extrapolation or rho_in).  The previous
    // populations, nine scattered loads per site, are read only there.
    double rn = 0.0, uxn = 0.0, uyn = 0.0;
    if (b.type == 4 && corner) {
        if (explicit_state) {
            rn = rho0[id];
            uxn = u0[2 * id];
            uyn = u0[2 * id + 1];
        } else {
            moments64(fp, (g.fprev ? g.fprev[id] : g.Fxm) + fibm[id], (g.fprev ? g.fprev[S + id] : g.Fym) + fibm[S + id],
                      rn, uxn, uyn);
        }
    }
    double rho1 = 0.0, rho2 = 0.0;
    switch (b.type) {
        case 1:  // wall
            uxn = 0.0;
            uyn = 0.0;
            break;
        case 2:  // velocity
            uxn = a.u_in[2 * j] * a.ramp;
            uyn = a.u_in[2 * j + 1] * a.ramp;
            break;
        case 3:    // free slip
        case 4: {  // pressure
            if (b.type == 3) {
                if (b.nd == 0) uxn = 0.0;
                else uyn = 0.0;
            } else {
                rn = a.rho_in[j];
            }
            // Utils::zeroGradient order 2 on the current velocity (Utils.h:200-210)
            double r1, r2, u1x, u1y, u2x, u2y;
            moments64(f1, g.fcur ? g.fcur[id1] : g.Fx, g.fcur ? g.fcur[S + id1] : g.Fy, r1, u1x, u1y);
            moments64(f2, g.fcur ? g.fcur[id2] : g.Fx, g.fcur ? g.fcur[S + id2] : g.Fy, r2, u2x, u2y);
            if (b.nd != 0) uxn = (4.0 / 3.0) * u1x - (1.0 / 3.0) * u2x;
            if (b.nd != 1) uyn = (4.0 / 3.0) * u1y - (1.0 / 3.0) * u2y;
            break;
        }
        default:
            break;
    }
    if (corner) {
        rho1 = density64(f1);
        rho2 = density64(f2);
    }
    regularised64<CM>(g, b.type, b.nx, b.ny, b.nd, fs, rn, uxn, uyn, rho1, rho2);
#pragma unroll
    for (int v = 0; v < kNV; v++) Pops<T>::put(dst, S, v, id, fs[v]);
}

template <typename T, bool CM>
__global__ __launch_bounds__(256) void k_bc(const T* __restrict__ src, T* __restrict__ dst, GridScalars g,
                                            BCArgs a, const double* __restrict__ fibm,
                                            const double* __restrict__ rho0, const double* __restrict__ u0,
                                            int explicit_state) {
    bc_site<T, CM>(src, dst, g, a, fibm, rho0, u0, explicit_state, blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T, bool CM>
__global__ __launch_bounds__(256) void k_bc_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list) {
    const CaseDev& c = cases[list[blockIdx.y]];
    bc_site<T, CM>(static_cast<const T*>(c.src), static_cast<T*>(c.dst), c.g, c.bc, c.fibm, c.rho0, c.u0,
                   c.explicit_state, blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
void launch_bc(const T* src, T* dst, const GridScalars& g, const BCArgs& a, const double* fibm, const double* rho0,
               const double* u0, bool explicit_state, hipStream_t s) {
    if (a.nbc <= 0) return;
    const unsigned nb = (unsigned)((a.nbc + 255) / 256);
    if (g.cm) k_bc<T, true><<<nb, 256, 0, s>>>(src, dst, g, a, fibm, rho0, u0, explicit_state ? 1 : 0);
    else k_bc<T, false><<<nb, 256, 0, s>>>(src, dst, g, a, fibm, rho0, u0, explicit_state ? 1 : 0);
}

// GridClass::convectiveSpeed (Grid.cpp:477-495) on the end-of-previous-step
// velocity of the last three columns.  One workgroup.  ordered = true keeps
// the reference's serial left-to-right mean (plain method); otherwise an
// FP64 tree reduction is used.
constexpr int kConvThreads = 1024;

// Outlet velocity of column c = 0..2 from the right wall, row j: stored in
// ucol (one thread per (c, j), many workgroups).
template <typename T>
__device__ inline void convective_column(const T* __restrict__ src, const GridScalars& g,
                                         const double* __restrict__ fibm, const double* __restrict__ u0,
                                         int explicit_state, double* __restrict__ ucol, int t) {
    const int Ny = g.Ny, Nx = g.Nx;
    if (t >= 3 * Ny) return;
    const int c = t / Ny, j = t - c * Ny;
    const long long S = g.S;
    const long long id = (long long)(Nx - 1 - c) * Ny + j;
    double ux, uy;
    if (explicit_state) {
        ux = u0[2 * id];
        uy = u0[2 * id + 1];
    } else {
        double fs[kNV], r;
        load_site(src, S, id, fs);
        moments64(fs, (g.fprev ? g.fprev[id] : g.Fxm) + fibm[id], (g.fprev ? g.fprev[S + id] : g.Fym) + fibm[S + id],
                  r, ux, uy);
    }
    ucol[(c * 2 + 0) * (long long)Ny + j] = ux;
    ucol[(c * 2 + 1) * (long long)Ny + j] = uy;
}

// uOut and delU from ucol, one workgroup of kConvThreads.  ordered: the
// reference's serial left-to-right mean; otherwise each thread sums its
// rows j = tid, tid + kConvThreads, ... in increasing j and a fixed tree
// combines the threads.
__device__ inline void convective_finish(const GridScalars& g, const double* __restrict__ ucol,
                                         double* __restrict__ delU, int ordered, double* red, double& uOutS) {
    const int Ny = g.Ny;
    double part = 0.0;
    for (int j = threadIdx.x; j < Ny; j += blockDim.x) part += ucol[j];
    red[threadIdx.x] = part;
    __syncthreads();
    if (ordered) {
        if (threadIdx.x == 0) {
            double uOut = 0.0;
            for (int j = 0; j < Ny; j++) uOut += ucol[j];
            uOutS = uOut / (double)Ny;
        }
    } else {
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) uOutS = red[0] / (double)Ny;
    }
    __syncthreads();
    const double uOut = uOutS;
    for (int j = threadIdx.x; j < Ny; j += blockDim.x) {
        for (int d = 0; d < 2; d++) {
            const double u1 = ucol[(0 * 2 + d) * (long long)Ny + j];
            const double u2 = ucol[(1 * 2 + d) * (long long)Ny + j];
            const double u3 = ucol[(2 * 2 + d) * (long long)Ny + j];
            delU[2 * j + d] = (-uOut / 2.0) * (3.0 * u1 - 4.0 * u2 + u3);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_convective_cols(const T* __restrict__ src, GridScalars g,
                                                         const double* __restrict__ fibm,
                                                         const double* __restrict__ u0, int explicit_state,
                                                         double* __restrict__ ucol) {
    convective_column(src, g, fibm, u0, explicit_state, ucol, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(kConvThreads) void k_convective_finish(GridScalars g, const double* __restrict__ ucol,
                                                                    double* __restrict__ delU, int ordered) {
    __shared__ double red[kConvThreads];
    __shared__ double uOutS;
    convective_finish(g, ucol, delU, ordered, red, uOutS);
}

// Serving: grid.y = listed case (cases with a convective outlet).
template <typename T>
__global__ __launch_bounds__(256) void k_convective_cols_batch(const CaseDev* __restrict__ cases,
                                                               const int* __restrict__ list) {
    const CaseDev& c = cases[list[blockIdx.y]];
    convective_column(static_cast<const T*>(c.src), c.g, c.fibm, c.u0, c.explicit_state, c.ucol,
                      blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(kConvThreads) void k_convective_finish_batch(const CaseDev* __restrict__ cases,
                                                                          const int* __restrict__ list, int ordered) {
    __shared__ double red[kConvThreads];
    __shared__ double uOutS;
    const CaseDev& c = cases[list[blockIdx.x]];
    convective_finish(c.g, c.ucol, const_cast<double*>(c.bc.delU), ordered, red, uOutS);
}

template <typename T>
void launch_convective_speed(const T* src, const GridScalars& g, const double* fibm, const double* u0,
                             bool explicit_state, double* ucol, double* delU, bool ordered, hipStream_t s) {
    k_convective_cols<T><<<(unsigned)((3 * g.Ny + 255) / 256), 256, 0, s>>>(src, g, fibm, u0, explicit_state ? 1 : 0,
                                                                           ucol);
    k_convective_finish<<<1, kConvThreads, 0, s>>>(g, ucol, delU, ordered ? 1 : 0);
}

template <typename T>
void launch_bc_batch(bool cm, const CaseDev* cases, const int* list, int n, int max_nbc, hipStream_t s) {
    if (n <= 0 || max_nbc <= 0) return;
    const dim3 grid((unsigned)((max_nbc + 255) / 256), (unsigned)n);
    if (cm) k_bc_batch<T, true><<<grid, 256, 0, s>>>(cases, list);
    else k_bc_batch<T, false><<<grid, 256, 0, s>>>(cases, list);
}

template <typename T>
void launch_convective_batch(const CaseDev* cases, const int* list, int n, int max_ny, bool ordered, hipStream_t s) {
    if (n <= 0 || max_ny <= 0) return;
    k_convective_cols_batch<T><<<dim3((unsigned)((3 * max_ny + 255) / 256), (unsigned)n), 256, 0, s>>>(cases, list);
    k_convective_finish_batch<<<(unsigned)n, kConvThreads, 0, s>>>(cases, list, ordered ? 1 : 0);
}

// f = f_eq(rho, u) at every site (Grid.cpp:1048-1061).
template <typename T, bool CM>
__global__ __launch_bounds__(256) void k_init_eq(T* __restrict__ f, GridScalars g, const double* __restrict__ rho0,
                                                 const double* __restrict__ u0) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
    const double r = rho0[id], ux = u0[2 * id], uy = u0[2 * id + 1];
#pragma unroll
    for (int v = 0; v < kNV; v++) Pops<T>::put(f, g.S, v, id, feq64<CM>(r, ux, uy, v));
}

template <typename T>
void launch_init_equilibrium(T* f, const GridScalars& g, const double* rho0, const double* u0, hipStream_t s) {
    const unsigned nb = (unsigned)((g.N + 255) / 256);
    if (g.cm) k_init_eq<T, true><<<nb, 256, 0, s>>>(f, g, rho0, u0);
    else k_init_eq<T, false><<<nb, 256, 0, s>>>(f, g, rho0, u0);
}

// Reference-layout state: f AoS, rho, u (macroscopic with force_xy +
// force_ibm, as left by IBMNodeClass::updateMacroscopic), force_ibm AoS.
template <typename T>
__global__ __launch_bounds__(256) void k_materialise(const T* __restrict__ f, GridScalars g,
                                                     const double* __restrict__ fibm, const double* __restrict__ rho0,
                                                     const double* __restrict__ u0, int explicit_state,
                                                     double* __restrict__ f_aos, double* __restrict__ rho,
                                                     double* __restrict__ u, double* __restrict__ fibm_aos) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
    double fs[kNV];
    load_site(f, g.S, id, fs);
    if (f_aos)
#pragma unroll
        for (int v = 0; v < kNV; v++) f_aos[id * kNV + v] = fs[v];
    if (rho || u) {
        double r, ux, uy;
        if (explicit_state) {
            r = rho0[id];
            ux = u0[2 * id];
            uy = u0[2 * id + 1];
        } else {
            moments64(fs, (g.fcur ? g.fcur[id] : g.Fx) + fibm[id], (g.fcur ? g.fcur[g.S + id] : g.Fy) + fibm[g.S + id],
                      r, ux, uy);
        }
        if (rho) rho[id] = r;
        if (u) {
            u[2 * id] = ux;
            u[2 * id + 1] = uy;
        }
    }
    if (fibm_aos) {
        fibm_aos[2 * id] = fibm[id];
        fibm_aos[2 * id + 1] = fibm[g.S + id];
    }
}

template <typename T>
void launch_materialise(const T* f, const GridScalars& g, const double* fibm, const double* rho0, const double* u0,
                        bool explicit_state, double* f_aos, double* rho, double* u, double* fibm_aos,
                        hipStream_t s) {
    const unsigned nb = (unsigned)((g.N + 255) / 256);
    k_materialise<T><<<nb, 256, 0, s>>>(f, g, fibm, rho0, u0, explicit_state ? 1 : 0, f_aos, rho, u, fibm_aos);
}

template <typename T>
__global__ __launch_bounds__(256) void k_upload(T* __restrict__ f, GridScalars g, const double* __restrict__ f_aos) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
#pragma unroll
    for (int v = 0; v < kNV; v++) Pops<T>::put(f, g.S, v, id, f_aos[id * kNV + v]);
}

template <typename T>
void launch_upload(T* f, const GridScalars& g, const double* f_aos, hipStream_t s) {
    const unsigned nb = (unsigned)((g.N + 255) / 256);
    k_upload<T><<<nb, 256, 0, s>>>(f, g, f_aos);
}

__global__ void k_fibm_from_aos(double* __restrict__ fibm, GridScalars g, const double* __restrict__ a) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
    fibm[id] = a[2 * id];
    fibm[g.S + id] = a[2 * id + 1];
}

void launch_fibm_from_aos(double* fibm, const GridScalars& g, const double* fibm_aos, hipStream_t s) {
    const unsigned nb = (unsigned)((g.N + 255) / 256);
    k_fibm_from_aos<<<nb, 256, 0, s>>>(fibm, g, fibm_aos);
}

__global__ void k_fill_f64(double* p, double v, long long n) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = v;
}
__global__ void k_fill_i32(int* p, int v, long long n) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = v;
}

void launch_fill_f64(double* p, double v, long long n, hipStream_t s) {
    if (n <= 0) return;
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    k_fill_f64<<<(unsigned)nb, 256, 0, s>>>(p, v, n);
}
void launch_fill_i32(int* p, int v, long long n, hipStream_t s) {
    if (n <= 0) return;
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    k_fill_i32<<<(unsigned)nb, 256, 0, s>>>(p, v, n);
}


// Initial moments (Grid.cpp:1004-1033): rho = 1, u uniform (mode 0) or the
// inlet profile u_in[j] (mode 1); wall sites are set to zero afterwards.
__global__ void k_init_moments(double* __restrict__ rho0, double* __restrict__ u0, GridScalars g, int mode, double ux,
                               double uy, const double* __restrict__ u_in) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
    rho0[id] = 1.0;
    if (mode == 1) {
        const int j = (int)(id % g.Ny);
        u0[2 * id] = u_in[2 * j];
        u0[2 * id + 1] = u_in[2 * j + 1];
    } else {
        u0[2 * id] = ux;
        u0[2 * id + 1] = uy;
    }
}

__global__ void k_zero_wall_u(double* __restrict__ u0, const BCSite* __restrict__ bc, int nbc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbc) return;
    if (bc[k].type == 1) {
        u0[2 * (long long)bc[k].id] = 0.0;
        u0[2 * (long long)bc[k].id + 1] = 0.0;
    }
}

void launch_init_moments(double* rho0, double* u0, const GridScalars& g, int mode, double ux, double uy,
                         const double* u_in, const BCSite* bc, int nbc, hipStream_t s) {
    const unsigned nb = (unsigned)((g.N + 255) / 256);
    k_init_moments<<<nb, 256, 0, s>>>(rho0, u0, g, mode, ux, uy, u_in);
    if (nbc > 0) k_zero_wall_u<<<(nbc + 255) / 256, 256, 0, s>>>(u0, bc, nbc);
}

// The blow-up test of GridClass::writeInfo (Grid.cpp:562-581): the
// velocity magnitude of every site from its populations, and the smallest
// lattice id (i-major, the reference's scan order) whose magnitude is NaN
// into *flag (left untouched when there is none).  u = j / rho without the
// half-force terms: a finite force changes no NaN.
template <typename T>
__global__ __launch_bounds__(256) void k_blowup_scan(const T* __restrict__ f, long long S, long long N,
                                                     unsigned long long* __restrict__ flag) {
    for (long long id = (long long)blockIdx.x * 256 + threadIdx.x; id < N; id += (long long)gridDim.x * 256) {
        double fs[kNV];
        load_site(f, S, id, fs);
        double r = 0.0, jx = 0.0, jy = 0.0;
#pragma unroll
        for (int v = 0; v < kNV; v++) {
            r += fs[v];
            jx += kCX[v] * fs[v];
            jy += kCY[v] * fs[v];
        }
        const double ux = jx / r, uy = jy / r;
        if (__builtin_isnan(sqrt(ux * ux + uy * uy))) atomicMin(flag, (unsigned long long)id);
    }
}

__global__ void k_flag_reset(unsigned long long* flag) { *flag = ~0ull; }

template <typename T>
void launch_blowup_scan(const T* f, const GridScalars& g, unsigned long long* flag, hipStream_t s) {
    k_flag_reset<<<1, 1, 0, s>>>(flag);
    long long nb = (g.N + 255) / 256;
    if (nb > 8192) nb = 8192;
    k_blowup_scan<T><<<(unsigned)nb, 256, 0, s>>>(f, g.S, g.N, flag);
}

#define LIFE_INST(T)                                                                                                 \
    template void launch_blowup_scan<T>(const T*, const GridScalars&, unsigned long long*, hipStream_t);             \
    template void launch_bc<T>(const T*, T*, const GridScalars&, const BCArgs&, const double*, const double*,        \
                               const double*, bool, hipStream_t);                                                    \
    template void launch_convective_speed<T>(const T*, const GridScalars&, const double*, const double*, bool,       \
                                             double*, double*, bool, hipStream_t);                                   \
    template void launch_init_equilibrium<T>(T*, const GridScalars&, const double*, const double*, hipStream_t);     \
    template void launch_materialise<T>(const T*, const GridScalars&, const double*, const double*, const double*,   \
                                        bool, double*, double*, double*, double*, hipStream_t);                      \
    template void launch_upload<T>(T*, const GridScalars&, const double*, hipStream_t);                             \
    template void launch_bc_batch<T>(bool, const CaseDev*, const int*, int, int, hipStream_t);                       \
    template void launch_convective_batch<T>(const CaseDev*, const int*, int, int, bool, hipStream_t);
LIFE_INST(double)
LIFE_INST(float)

}  // namespace life
