// Device FEM: FemBody (fem.cpp) for every flexible body at once, one
// workgroup of kFemThreads lanes per body.  Every floating-point
// expression keeps the host FEM's operands and order (compiled with
// -ffp-contract=off); the work is split over lanes only where the host's
// results do not depend on the split:
//  - per-node, per-element, per-DOF and per-marker loops whose iterations
//    write disjoint values;
//  - the global matrices and force vector, where an entry gets at most two
//    element terms onto +0.0 (even elements, then odd ones: a + b = b + a);
//  - the load vector: a node's three DOFs take the terms of the element on
//    its left (its second node), then of the one on its right, each in
//    load-map order, exactly the host's accumulation order;
//  - the banded elimination: one wavefront, the row updates of a step in
//    parallel (each entry is updated once per step), the pivot search and
//    the back substitution in the host's order;
//  - dot products: one lane, in index order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <stdexcept>

#include "fem.h"
#include "fem_dev.h"
#include "runtime.h"

namespace life {

namespace {

#define FDI __device__ __forceinline__

constexpr int kNZd[6][2] = {{0, 1}, {0, 1}, {2, -1}, {3, 4}, {3, 4}, {5, -1}};

// Utils::shiftAngle (Utils.h:235-247), as wrap_angle in fem.cpp.
FDI double wrap_angle_d(double a) {
    const double x = a + M_PI, y = 2.0 * M_PI;
    a = (x > -y && x < y) ? x : fmod(x, y);
    if (a < 0.0) a += 2.0 * M_PI;
    return a - M_PI;
}

FDI void make_T(double ca, double sa, double T[6][6]) {
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) T[i][j] = 0.0;
    T[0][0] = T[1][1] = T[3][3] = T[4][4] = ca;
    T[0][1] = T[3][4] = sa;
    T[1][0] = T[4][3] = -sa;
    T[2][2] = T[5][5] = 1.0;
}

// Transpose(T) * A * T over the structurally non-zero terms (rotate6, fem.cpp)
FDI void rotate6_d(const double T[6][6], const double A[6][6], double out[6][6]) {
    double tmp[6][6];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 2; q++)
                if (kNZd[i][q] >= 0) s += T[kNZd[i][q]][i] * A[kNZd[i][q]][j];
            tmp[i][j] = s;
        }
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 2; q++)
                if (kNZd[j][q] >= 0) s += tmp[i][kNZd[j][q]] * T[kNZd[j][q]][j];
            out[i][j] = s;
        }
}

FDI void mul_t6_d(const double T[6][6], const double* v, double* out) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (kNZd[i][q] >= 0) s += T[kNZd[i][q]][i] * v[kNZd[i][q]];
        out[i] = s;
    }
}

FDI void mul6_d(const double T[6][6], const double* v, double* out) {
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (kNZd[i][q] >= 0) s += T[i][kNZd[i][q]] * v[kNZd[i][q]];
        out[i] = s;
    }
}

// Lanes of one wavefront exchange values through LDS: order the writes of
// some lanes before the later reads of others.
FDI void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

FDI double* vecp(const FemDevBody& B, int k) { return B.vec + (size_t)k * B.ndof; }
FDI double& band(double* A, int r, int c) { return A[(size_t)r * kFemBandW + (c - r + kFemKL)]; }
FDI const double* elp(const FemDevBody& B, int e) { return B.el + (size_t)e * kFemElStride; }

// FemBody::update_fem (FEMBody.cpp:198-223)
FDI void update_fem_d(const FemDevBody& B) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const double* U = vecp(B, kFU);
    for (int n = tid; n < B.nnode; n += nt) {
        double* nd = B.node + 6 * n;
        nd[2] = nd[0] + U[3 * n + 0];
        nd[3] = nd[1] + U[3 * n + 1];
        nd[5] = nd[4] + U[3 * n + 2];
    }
    __syncthreads();
    for (int e = tid; e < B.nel; e += nt) {
        const double* n0 = B.node + 6 * e;
        const double* n1 = B.node + 6 * (e + 1);
        const double ev[2] = {n1[2] - n0[2], n1[3] - n0[3]};
        double* el = B.el + (size_t)e * kFemElStride;
        const double angle = atan2(ev[1], ev[0]);
        double d = 0.0;
        d += ev[0] * ev[0];
        d += ev[1] * ev[1];
        el[1] = sqrt(d);
        el[2] = angle;
        el[7] = cos(angle);
        el[8] = sin(angle);
    }
    __syncthreads();
}

// FemBody::finish_newmark (FEMBody.cpp:129-144)
FDI void finish_newmark_d(const FemDevBody& B) {
    const double Dt = B.s.Dt, al = B.s.alpha, de = B.s.delta;
    const double a6 = 1.0 / (al * (Dt * Dt));
    const double a7 = -1.0 / (al * Dt);
    const double a8 = -(1.0 / (2.0 * al) - 1.0);
    const double a9 = Dt * (1.0 - de);
    const double a10 = de * Dt;
    double *U = vecp(B, kFU), *Un = vecp(B, kFUn), *Udn = vecp(B, kFUdn), *Uddn = vecp(B, kFUddn);
    double *Ud = vecp(B, kFUd), *Udd = vecp(B, kFUdd);
    for (int k = threadIdx.x; k < B.ndof; k += blockDim.x) {
        const double udd = a6 * (U[k] - Un[k]) + a7 * Udn[k] + a8 * Uddn[k];
        Udd[k] = udd;
        Ud[k] = Udn[k] + a9 * Uddn[k] + a10 * udd;
    }
    __syncthreads();
}

// FemBody::shape (FEMElement.cpp:141-160)
FDI void shape_d(double L, const double v[6], double zeta, double out[2]) {
    const double z = (zeta + 1.0) / 2.0;
    const double N0 = 1.0 - (zeta + 1.0) / 2.0;
    const double N1 = 1.0 - 3.0 * (z * z) + 2.0 * (z * z * z);
    const double N2 = (z - 2.0 * (z * z) + (z * z * z)) * L;
    const double N3 = z;
    const double N4 = 3.0 * (z * z) - 2.0 * (z * z * z);
    const double N5 = (-(z * z) + (z * z * z)) * L;
    out[0] = v[0] * N0 + v[3] * N3;
    out[1] = v[1] * N1 + v[2] * N2 + v[4] * N4 + v[5] * N5;
}

// FemBody::update_ibm (FEMBody.cpp:147-195): markers follow the beam.
// elu: LDS, 12 per element.
FDI void update_ibm_d(const FemDevBody& B, double* elu) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const double *U = vecp(B, kFU), *Ud = vecp(B, kFUd);
    for (int q = tid; q < B.nel; q += nt) {
        const double* e = elp(B, q);
        double T[6][6];
        make_T(e[7], e[8], T);
        const int d0 = 3 * q;
        double dU[6], dUd[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            dU[k] = U[d0 + k];
            dUd[k] = Ud[d0 + k];
        }
        const double* nn0 = B.node + 6 * q;
#pragma unroll
        for (int n = 0; n < 2; n++) {
            const double* nn = B.node + 6 * (q + n);
            dU[n * 3 + 0] += nn[0] - nn0[2];
            dU[n * 3 + 1] += nn[1] - nn0[3];
            dU[n * 3 + 2] += nn[4] - e[2];
            dU[n * 3 + 2] = wrap_angle_d(dU[n * 3 + 2]);
        }
        mul6_d(T, dU, elu + 12 * q);
        mul6_d(T, dUd, elu + 12 * q + 6);
    }
    __syncthreads();
    for (int i = tid; i < B.nm; i += nt) {
        const int q = B.pm_el[i];
        const double* e = elp(B, q);
        const double* lU = elu + 12 * q;
        double s[2], sd[2];
        shape_d(e[1], lU, B.pm_zeta[i], s);
        shape_d(e[1], lU + 6, B.pm_zeta[i], sd);
        const double* p0 = B.node + 6 * q + 2;
        const double T00 = e[7], T01 = e[8], T10 = -e[8], T11 = e[7];
        // gs_r = 0.0 + T[0][r] s0 + T[1][r] s1
        const double gs0 = 0.0 + T00 * s[0] + T10 * s[1];
        const double gs1 = 0.0 + T01 * s[0] + T11 * s[1];
        const double gd0 = 0.0 + T00 * sd[0] + T10 * sd[1];
        const double gd1 = 0.0 + T01 * sd[0] + T11 * sd[1];
        B.mpos[2 * i] = p0[0] + gs0;
        B.mpos[2 * i + 1] = p0[1] + gs1;
        B.mvel[2 * i] = gd0;
        B.mvel[2 * i + 1] = gd1;
    }
    __syncthreads();
}

// One load-map entry of element q (FemBody::load_vector, FEMElement.cpp:27-69):
// its global load terms RG.
FDI void load_entry_d(const FemDevBody& B, int q, int entry, double RG[6]) {
    const double* e = elp(B, q);
    const double L = e[1], A = e[3], rho = e[6];
    double T[6][6];
    make_T(e[7], e[8], T);
    const double wt0 = rho * A * B.s.gx, wt1 = rho * A * B.s.gy;
    const double fs = B.s.Dm / (B.s.Dt * B.s.Dt);
    const int k = B.map_node[entry];
    const double a = B.map_z[2 * entry], b = B.map_z[2 * entry + 1];
    const double sc = -B.meps[k] * 1.0 * fs;
    const double vv[2] = {sc * B.mforce[2 * k] + wt0, sc * B.mforce[2 * k + 1] + wt1};
    double Fl[2];
#pragma unroll
    for (int r = 0; r < 2; r++) Fl[r] = 0.0 + T[r][0] * vv[0] + T[r][1] * vv[1];
    const double a2 = a * a, b2 = b * b, a3 = a * a * a, b3 = b * b * b;
    double R[6];
    R[0] = Fl[0] * 0.5 * L * (0.5 * b - 0.5 * a + 0.25 * a2 - 0.25 * b2);
    R[1] = Fl[1] * 0.5 * L * (0.5 * b - 0.5 * a - a2 * a2 / 16.0 + b2 * b2 / 16.0 + 3.0 * a2 / 8.0 - 3.0 * b2 / 8.0);
    R[2] = Fl[1] * 0.5 * L *
           (L * (-a2 * a2 + b2 * b2) / 32.0 - L * (-a3 + b3) / 24.0 - L * (-a2 + b2) / 16.0 + L * (b - a) / 8.0);
    R[3] = Fl[0] * 0.5 * L * (-0.25 * a2 + 0.25 * b2 + 0.5 * b - 0.5 * a);
    R[4] = Fl[1] * 0.5 * L * (0.5 * b - 0.5 * a + a2 * a2 / 16.0 - b2 * b2 / 16.0 - 3.0 * a2 / 8.0 + 3.0 * b2 / 8.0);
    R[5] = Fl[1] * 0.5 * L *
           (L * (-a2 * a2 + b2 * b2) / 32.0 + L * (-a3 + b3) / 24.0 - L * (-a2 + b2) / 16.0 - L * (b - a) / 8.0);
    mul_t6_d(T, R, RG);
}

// R (FEMBodyClass::constructRVector): node n's DOFs take the terms of
// element n-1 (its second node) then of element n (its first node), each
// in load-map order, onto 0.0.
FDI void load_vector_d(const FemDevBody& B) {
    double* R = vecp(B, kFR);
    for (int n = threadIdx.x; n < B.nnode; n += blockDim.x) {
        double acc[3] = {0.0, 0.0, 0.0};
        if (n >= 1) {
            const int q = n - 1;
            for (int en = B.map_start[q]; en < B.map_start[q + 1]; en++) {
                double RG[6];
                load_entry_d(B, q, en, RG);
                acc[0] += RG[3];
                acc[1] += RG[4];
                acc[2] += RG[5];
            }
        }
        if (n < B.nel) {
            for (int en = B.map_start[n]; en < B.map_start[n + 1]; en++) {
                double RG[6];
                load_entry_d(B, n, en, RG);
                acc[0] += RG[0];
                acc[1] += RG[1];
                acc[2] += RG[2];
            }
        }
        R[3 * n] = acc[0];
        R[3 * n + 1] = acc[1];
        R[3 * n + 2] = acc[2];
    }
    __syncthreads();
}

// force_vector + mass_stiff of element q (FEMElement.cpp:72-138): the
// element's terms added onto the global F, M and K.
FDI void element_terms_d(const FemDevBody& B, int q) {
    double* e = B.el + (size_t)q * kFemElStride;
    const double L0 = e[0], L = e[1], angle = e[2], A = e[3], I = e[4], E = e[5];
    double T[6][6];
    make_T(e[7], e[8], T);
    const double* nd0 = B.node + 6 * q;
    const double* nd1 = B.node + 6 * (q + 1);
    const double u = (L * L - L0 * L0) / (L + L0);
    const double th1 = wrap_angle_d(nd0[5] - angle);
    const double th2 = wrap_angle_d(nd1[5] - angle);
    const double F0 = (E * A / L0) * u;
    const double M1 = (2 * E * I / L0) * (2.0 * th1 + th2);
    const double M2 = (2 * E * I / L0) * (th1 + 2.0 * th2);
    double Fe[6];
    Fe[0] = -F0;
    Fe[1] = (1.0 / L0) * (M1 + M2);
    Fe[2] = M1;
    Fe[3] = F0;
    Fe[4] = -(1.0 / L0) * (M1 + M2);
    Fe[5] = M2;
    double FG[6];
    mul_t6_d(T, Fe, FG);
    const int d0 = 3 * q;
    double* Fg = vecp(B, kFF);
#pragma unroll
    for (int k = 0; k < 6; k++) Fg[d0 + k] += FG[k];
    double Al[6][6], G[6][6], N[6][6];
    const double* Ml = B.Mloc + 36 * (size_t)q;
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) Al[i][j] = Ml[6 * i + j];
    rotate6_d(T, Al, G);
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) band(B.M, d0 + i, d0 + j) += G[i][j];
    const double* Kl = B.KLloc + 36 * (size_t)q;
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) Al[i][j] = Kl[6 * i + j];
    rotate6_d(T, Al, G);
    const double Fa = -Fe[0], V0 = Fe[4];
    double KNL[6][6];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) KNL[i][j] = 0.0;
    KNL[0][1] = -V0 / L0;
    KNL[0][4] = V0 / L0;
    KNL[1][0] = -V0 / L0;
    KNL[1][1] = Fa / L0;
    KNL[1][3] = V0 / L0;
    KNL[1][4] = -Fa / L0;
    KNL[3][1] = V0 / L0;
    KNL[3][4] = -V0 / L0;
    KNL[4][0] = V0 / L0;
    KNL[4][1] = -Fa / L0;
    KNL[4][3] = -V0 / L0;
    KNL[4][4] = Fa / L0;
    rotate6_d(T, KNL, N);
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) band(B.K, d0 + i, d0 + j) += G[i][j] + N[i][j];
}

// FemBody::build_global (FEMBody.cpp:90-109)
FDI void build_global_d(const FemDevBody& B) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nb = B.ndof * kFemBandW;
    for (int k = tid; k < nb; k += nt) {
        B.M[k] = 0.0;
        B.K[k] = 0.0;
    }
    double* F = vecp(B, kFF);
    for (int k = tid; k < B.ndof; k += nt) F[k] = 0.0;
    __syncthreads();
    for (int ph = 0; ph < 2; ph++) {
        for (int q = ph + 2 * tid; q < B.nel; q += 2 * nt) element_terms_d(B, q);
        __syncthreads();
    }
}

// lu_solve_band (lu.cpp) on wavefront 0: K x = F over the trailing block
// (rows/columns >= nbc), x into delU.  LDS: W (m x kFemBandW), y (m),
// last (m ints).
FDI void lu_solve_d(const FemDevBody& B, double* W, double* y, int* last) {
    const int L = threadIdx.x;
    const int n = B.ndof, bc = B.nbc, m = n - bc, kl = kFemKL, ku = kFemKL, kext = kl + ku;
    const double* Kg = B.K;
    const double* b = vecp(B, kFF);
    double* x = vecp(B, kFDelU);
    if (L < 64) {
        for (int r = L; r < m; r += 64) {
            const int lo = r - kl < 0 ? 0 : r - kl;
            const int hi = r + ku > m - 1 ? m - 1 : r + ku;
            const int top = r + kext > m - 1 ? m - 1 : r + kext;
            for (int c = lo; c <= hi; c++)
                W[(size_t)r * kFemBandW + (c - r + kl)] = Kg[(size_t)(r + bc) * kFemBandW + (c - r + kl)];
            for (int c = hi + 1; c <= top; c++) W[(size_t)r * kFemBandW + (c - r + kl)] = 0.0;
            last[r] = hi;
            y[r] = b[r + bc];
        }
        wave_fence();
        for (int k = 0; k < m; k++) {
            const int ilast = k + kl > m - 1 ? m - 1 : k + kl;
            int p = k;
            double best = fabs(W[(size_t)k * kFemBandW + kl]);
            for (int i = k + 1; i <= ilast; i++) {
                const double a = fabs(W[(size_t)i * kFemBandW + (k - i + kl)]);
                if (a > best) {
                    best = a;
                    p = i;
                }
            }
            if (best == 0.0) continue;
            if (p != k) {
                const int jl = last[k] > last[p] ? last[k] : last[p];
                for (int j = k + L; j <= jl; j += 64) {
                    double& a = W[(size_t)k * kFemBandW + (j - k + kl)];
                    double& c = W[(size_t)p * kFemBandW + (j - p + kl)];
                    const double t = a;
                    a = c;
                    c = t;
                }
                wave_fence();
                if (L == 0) {
                    const int tl = last[k];
                    last[k] = last[p];
                    last[p] = tl;
                    const double t = y[k];
                    y[k] = y[p];
                    y[p] = t;
                }
                wave_fence();
            }
            const double piv = W[(size_t)k * kFemBandW + kl];
            const int jlast = last[k];
            const int nr = ilast - k;
            if (L >= 1 && L <= nr) {
                double& a = W[(size_t)(k + L) * kFemBandW + (k - (k + L) + kl)];
                a = a / piv;
            }
            wave_fence();
            const int ncol = jlast - k;
            for (int idx = L; idx < nr * ncol; idx += 64) {
                const int r = 1 + idx / ncol, c = k + 1 + idx % ncol;
                const int i = k + r;
                const double l = W[(size_t)i * kFemBandW + (k - i + kl)];
                double& a = W[(size_t)i * kFemBandW + (c - i + kl)];
                a = a - l * W[(size_t)k * kFemBandW + (c - k + kl)];
            }
            if (L >= 1 && L <= nr) {
                const int i = k + L;
                const double l = W[(size_t)i * kFemBandW + (k - i + kl)];
                y[i] = y[i] - l * y[k];
                if (last[i] < jlast) last[i] = jlast;
            }
            wave_fence();
        }
        if (L == 0) {
            for (int i = m - 1; i >= 0; i--) {
                const double* ri = W + (size_t)i * kFemBandW + kl - i;
                double s = y[i];
                const int jlast = last[i];
                for (int j = i + 1; j <= jlast; j++) s = s - ri[j] * x[j + bc];
                x[i + bc] = s / ri[i];
            }
            for (int i = 0; i < bc; i++) x[i] = 0.0;
        }
    }
    __syncthreads();
}

// FemBody::newton_dynamic (FEMBody.cpp:71-87, 112-126) and the residual
// test of dynamic(): returns resNR (every lane).
FDI double newton_d(const FemDevBody& B, double* lds) {
    const int tid = threadIdx.x, nt = blockDim.x;
    build_global_d(B);
    const double Dt = B.s.Dt, al = B.s.alpha;
    const double a0 = 1.0 / (al * (Dt * Dt));
    const double a2 = 1.0 / (al * Dt);
    const double a3 = 1.0 / (2.0 * al) - 1.0;
    const int nd = B.ndof;
    double *U = vecp(B, kFU), *Un = vecp(B, kFUn), *Ud = vecp(B, kFUd), *Udd = vecp(B, kFUdd);
    double *vec = vecp(B, kFVec), *F = vecp(B, kFF), *R = vecp(B, kFR), *delU = vecp(B, kFDelU);
    for (int k = tid; k < nd; k += nt) vec[k] = a0 * (Un[k] - U[k]) + a2 * Ud[k] + a3 * Udd[k];
    __syncthreads();
    const int kb = kFemKL;
    for (int i = tid; i < nd; i += nt) {
        double s = 0.0;
        const int lo = i - kb < 0 ? 0 : i - kb, hi = i + kb > nd - 1 ? nd - 1 : i + kb;
        for (int j = lo; j <= hi; j++) s += band(B.M, i, j) * vec[j];
        F[i] = R[i] - F[i] + s;
        for (int j = lo; j <= hi; j++) band(B.K, i, j) = band(B.K, i, j) + a0 * band(B.M, i, j);
    }
    __syncthreads();
    const int m = nd - B.nbc;
    lu_solve_d(B, lds, lds + (size_t)m * kFemBandW, reinterpret_cast<int*>(lds + (size_t)m * (kFemBandW + 1)));
    for (int k = tid; k < nd; k += nt) U[k] = U[k] + delU[k];
    __syncthreads();
    update_fem_d(B);
    __shared__ double res_sh;
    if (tid == 0) {
        double r = 0.0;
        for (int k = 0; k < nd; k++) r += delU[k] * delU[k];
        res_sh = sqrt(r) / (B.s.ref_L * sqrt((double)nd));
    }
    __syncthreads();
    return res_sh;
}

FDI void copy_vec(const FemDevBody& B, int to, int from) {
    const double* s = vecp(B, from);
    double* d = vecp(B, to);
    for (int k = threadIdx.x; k < B.ndof; k += blockDim.x) d[k] = s[k];
}

// list: the bodies of this launch (null: 0 .. nb-1); param: per listed
// body, its step t (predict) or relaxation factor (relax); null: value.
FDI int body_index(const int* list) { return list ? list[blockIdx.x] : (int)blockIdx.x; }

__global__ __launch_bounds__(kFemThreads) void k_fem_predict(const FemDevBody* __restrict__ bodies,
                                                            const int* __restrict__ list, int t_all,
                                                            const double* __restrict__ t_d) {
    extern __shared__ double lds[];
    const FemDevBody B = bodies[body_index(list)];
    const int t = t_d ? (int)t_d[blockIdx.x] : t_all;
    // reset_values (FEMBody.cpp:341-349): the history moves back one step
    copy_vec(B, kFUnm2, kFUnm1);
    __syncthreads();
    copy_vec(B, kFUnm1, kFUn);
    __syncthreads();
    copy_vec(B, kFUn, kFU);
    copy_vec(B, kFUdn, kFUd);
    copy_vec(B, kFUddn, kFUdd);
    __syncthreads();
    // predictor (FEMBody.cpp:259-289)
    double *U = vecp(B, kFU), *Un = vecp(B, kFUn), *Unm1 = vecp(B, kFUnm1), *Unm2 = vecp(B, kFUnm2);
    for (int k = threadIdx.x; k < B.ndof; k += blockDim.x) {
        if (t > 2) U[k] = 2.5 * Un[k] - 2.0 * Unm1[k] + 0.5 * Unm2[k];
        else if (t == 2) U[k] = 2.0 * Un[k] - Unm1[k];
        else if (t == 1) U[k] = Un[k];
    }
    __syncthreads();
    update_fem_d(B);
    finish_newmark_d(B);
    update_ibm_d(B, lds);
    copy_vec(B, kFUkm1, kFU);
}

// ObjectsClass::recomputeObjectVals relaxation (Objects.cpp:198-210)
__global__ __launch_bounds__(kFemThreads) void k_fem_relax(const FemDevBody* __restrict__ bodies,
                                                          const int* __restrict__ list, double omega,
                                                          const double* __restrict__ omega_d) {
    extern __shared__ double lds[];
    const FemDevBody B = bodies[body_index(list)];
    const double w = omega_d ? omega_d[blockIdx.x] : omega;
    double *U = vecp(B, kFU), *Ukm1 = vecp(B, kFUkm1);
    for (int k = threadIdx.x; k < B.ndof; k += blockDim.x) U[k] = Ukm1[k] + w * (U[k] - Ukm1[k]);
    __syncthreads();
    update_fem_d(B);
    finish_newmark_d(B);
    update_ibm_d(B, lds);
    copy_vec(B, kFUkm1, kFU);
}

// FemBody::dynamic (FEMBody.cpp:26-68) and subResidual (FEMBody.cpp:244-256)
// redo: the same solve again after epsilon changed (the dense fallback of
// computeEpsilon): Rkm1 already holds the previous subiteration's Rk.
__global__ __launch_bounds__(kFemThreads) void k_fem_dynamic(const FemDevBody* __restrict__ bodies,
                                                            const int* __restrict__ list, int redo) {
    extern __shared__ double lds[];
    const FemDevBody B = bodies[body_index(list)];
    copy_vec(B, kFU, kFUn);
    copy_vec(B, kFUd, kFUdn);
    copy_vec(B, kFUdd, kFUddn);
    __syncthreads();
    update_fem_d(B);
    load_vector_d(B);
    int it = 0;
    double res;
    do {
        res = newton_d(B, lds);
        it++;
    } while (res > 1e-10 && it < 20);
    finish_newmark_d(B);
    update_ibm_d(B, lds);
    // Rkm1 takes Rk, Rk = U - Ukm1
    if (!redo) copy_vec(B, kFRkm1, kFRk);
    __syncthreads();
    double *Rk = vecp(B, kFRk), *Rkm1 = vecp(B, kFRkm1), *U = vecp(B, kFU), *Ukm1 = vecp(B, kFUkm1);
    for (int k = threadIdx.x; k < B.ndof; k += blockDim.x) Rk[k] = U[k] - Ukm1[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, num = 0.0, den = 0.0;
        for (int k = 0; k < B.ndof; k++) r += Rk[k] * Rk[k];
        for (int k = 0; k < B.ndof; k++) num += Rkm1[k] * (Rk[k] - Rkm1[k]);
        for (int k = 0; k < B.ndof; k++) den += (Rk[k] - Rkm1[k]) * (Rk[k] - Rkm1[k]);
        B.out[0] = r;
        B.out[1] = num;
        B.out[2] = den;
        B.out[3] = (double)it;
        B.out[4] = res;
    }
}

// out[0..2] of every listed body, packed (3 per body)
__global__ void k_fem_outs(const FemDevBody* __restrict__ bodies, const int* __restrict__ list, int nb,
                           double* __restrict__ outs) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nb) return;
    const double* o = bodies[list ? list[q] : q].out;
    outs[3 * q] = o[0];
    outs[3 * q + 1] = o[1];
    outs[3 * q + 2] = o[2];
}

__global__ void k_fem_sums(const FemDevBody* __restrict__ bodies, int nb, double* __restrict__ sums) {
    double r = 0.0, num = 0.0, den = 0.0;
    for (int b = 0; b < nb; b++) {
        const double* o = bodies[b].out;
        r += o[0];
        num += o[1];
        den += o[2];
    }
    sums[0] = r;
    sums[1] = num;
    sums[2] = den;
}

}  // namespace

size_t fem_dev_lds_bytes(int ndof) {
    // LU: band + right-hand side + extents; update_ibm: 12 per element
    const size_t lu = sizeof(double) * (size_t)ndof * (kFemBandW + 2);
    const size_t ibm = sizeof(double) * 12 * (size_t)(ndof / 3);
    return lu > ibm ? lu : ibm;
}

void launch_fem_predict(const FemDevBody* b, const int* list, int nb, int t, const double* t_d, size_t lds,
                        hipStream_t st) {
    if (nb > 0) k_fem_predict<<<nb, kFemThreads, lds, st>>>(b, list, t, t_d);
}

void launch_fem_relax(const FemDevBody* b, const int* list, int nb, double omega, const double* omega_d, size_t lds,
                      hipStream_t st) {
    if (nb > 0) k_fem_relax<<<nb, kFemThreads, lds, st>>>(b, list, omega, omega_d);
}

void launch_fem_dynamic(const FemDevBody* b, const int* list, int nb, size_t lds, bool redo, hipStream_t st) {
    if (nb > 0) k_fem_dynamic<<<nb, kFemThreads, lds, st>>>(b, list, redo ? 1 : 0);
}

void launch_fem_outs(const FemDevBody* b, const int* list, int nb, double* outs, hipStream_t st) {
    if (nb > 0) k_fem_outs<<<(nb + 63) / 64, 64, 0, st>>>(b, list, nb, outs);
}

void launch_fem_sums(const FemDevBody* b, int nb, double* sums, hipStream_t st) {
    if (nb > 0) k_fem_sums<<<1, 1, 0, st>>>(b, nb, sums);
}

// ---------------------------------------------------------------- host side

namespace {
template <typename T>
T* upload_vec(FemDevCase& fc, const std::vector<T>& v, hipStream_t st) {
    T* p = dalloc<T>(v.size());
    fc.blocks.push_back(p);
    if (!v.empty()) LIFE_HIP(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
    return p;
}

// Host state of a body in the device layout: vectors, node and element
// state (the element's ca/sa follow from its angle).
void pack_state(const FemBody& fb, const FemBody::Layout& lay, std::vector<double>& vecs, std::vector<double>& el) {
    const int nd = fb.ndof();
    vecs.assign((size_t)kFNVec * nd, 0.0);
    const std::vector<double>* v[] = {&fb.U, &fb.Ud, &fb.Udd, &fb.Un, &fb.Unm1, &fb.Unm2,
                                      &fb.Udn, &fb.Uddn, &fb.Ukm1, &fb.Rk, &fb.Rkm1};
    for (int k = 0; k < kFemState; k++) std::memcpy(&vecs[(size_t)k * nd], v[k]->data(), sizeof(double) * nd);
    const int nel = (int)lay.el.size() / 7;
    el.assign((size_t)nel * kFemElStride, 0.0);
    for (int e = 0; e < nel; e++) {
        for (int k = 0; k < 7; k++) el[(size_t)e * kFemElStride + k] = lay.el[7 * e + k];
        const double angle = lay.el[7 * e + 2];
        el[(size_t)e * kFemElStride + 7] = std::cos(angle);
        el[(size_t)e * kFemElStride + 8] = std::sin(angle);
    }
}
}  // namespace

void fem_dev_create(FemDevCase& fc, const std::vector<const FemBody*>& bodies, const std::vector<int>& first,
                    const std::vector<int>& nm, double* pos, double* vel, const double* force, const double* eps,
                    const FemDevScalars& s, hipStream_t st) {
    fem_dev_destroy(fc);
    fc.nb = (int)bodies.size();
    if (fc.nb == 0) return;
    size_t lds = 0;
    for (int b = 0; b < fc.nb; b++) {
        const FemBody& fb = *bodies[b];
        FemBody::Layout lay;
        fb.layout(lay);
        std::vector<double> vecs, el;
        pack_state(fb, lay, vecs, el);
        FemDevBody d{};
        d.ndof = fb.ndof();
        d.nbc = fb.bc_dofs();
        d.nel = fb.num_elements();
        d.nnode = fb.num_nodes();
        d.nm = nm[b];
        d.s = s;
        d.node = upload_vec(fc, lay.node, st);
        d.el = upload_vec(fc, el, st);
        d.Mloc = upload_vec(fc, lay.Mloc, st);
        d.KLloc = upload_vec(fc, lay.KLloc, st);
        d.map_start = upload_vec(fc, lay.map_start, st);
        d.map_node = upload_vec(fc, lay.map_node, st);
        d.map_z = upload_vec(fc, lay.map_z, st);
        d.pm_el = upload_vec(fc, lay.pm_el, st);
        d.pm_zeta = upload_vec(fc, lay.pm_zeta, st);
        d.vec = upload_vec(fc, vecs, st);
        std::vector<double> zeros((size_t)d.ndof * kFemBandW, 0.0);
        d.M = upload_vec(fc, zeros, st);
        d.K = upload_vec(fc, zeros, st);
        d.out = upload_vec(fc, std::vector<double>(8, 0.0), st);
        d.mpos = pos + 2 * (size_t)first[b];
        d.mvel = vel + 2 * (size_t)first[b];
        d.mforce = force + 2 * (size_t)first[b];
        d.meps = eps + first[b];
        fc.host.push_back(d);
        const size_t need = fem_dev_lds_bytes(d.ndof);
        if (need > lds) lds = need;
    }
    if (lds > kFemMaxLds) throw std::runtime_error("flexible body too large for the device FEM");
    fc.lds = lds;
    fc.d_bodies = upload_vec(fc, fc.host, st);
    fc.sums = dalloc<double>(4);
    fc.blocks.push_back(fc.sums);
    LIFE_HIP(hipHostMalloc((void**)&fc.sums_h, 4 * sizeof(double), hipHostMallocDefault));
    LIFE_HIP(hipStreamSynchronize(st));
}

void fem_dev_destroy(FemDevCase& fc) {
    for (void* p : fc.blocks) dfree(p);
    fc.blocks.clear();
    fc.host.clear();
    fc.d_bodies = nullptr;
    fc.sums = nullptr;
    if (fc.sums_h) (void)hipHostFree(fc.sums_h);
    fc.sums_h = nullptr;
    fc.nb = 0;
}

void fem_dev_download(FemDevCase& fc, const std::vector<FemBody*>& bodies, hipStream_t st) {
    std::vector<std::vector<double>> vecs(fc.nb), node(fc.nb), el(fc.nb), out(fc.nb);
    for (int b = 0; b < fc.nb; b++) {
        const FemDevBody& d = fc.host[b];
        vecs[b].resize((size_t)kFemState * d.ndof);
        node[b].resize((size_t)6 * d.nnode);
        el[b].resize((size_t)kFemElStride * d.nel);
        out[b].resize(8);
        LIFE_HIP(hipMemcpyAsync(vecs[b].data(), d.vec, sizeof(double) * vecs[b].size(), hipMemcpyDeviceToHost, st));
        LIFE_HIP(hipMemcpyAsync(node[b].data(), d.node, sizeof(double) * node[b].size(), hipMemcpyDeviceToHost, st));
        LIFE_HIP(hipMemcpyAsync(el[b].data(), d.el, sizeof(double) * el[b].size(), hipMemcpyDeviceToHost, st));
        LIFE_HIP(hipMemcpyAsync(out[b].data(), d.out, sizeof(double) * 8, hipMemcpyDeviceToHost, st));
    }
    LIFE_HIP(hipStreamSynchronize(st));
    for (int b = 0; b < fc.nb; b++) {
        const int nel = fc.host[b].nel;
        std::vector<double> el7((size_t)7 * nel);
        for (int e = 0; e < nel; e++)
            for (int k = 0; k < 7; k++) el7[7 * e + k] = el[b][(size_t)e * kFemElStride + k];
        bodies[b]->import_state(vecs[b].data(), node[b].data(), el7.data(), out[b][0], out[b][1], out[b][2],
                                (int)out[b][3], out[b][4]);
    }
}

void fem_dev_upload(FemDevCase& fc, const std::vector<const FemBody*>& bodies, hipStream_t st) {
    for (int b = 0; b < fc.nb; b++) {
        const FemBody& fb = *bodies[b];
        FemBody::Layout lay;
        fb.layout(lay);
        std::vector<double> vecs, el;
        pack_state(fb, lay, vecs, el);
        const FemDevBody& d = fc.host[b];
        LIFE_HIP(hipMemcpyAsync(d.vec, vecs.data(), sizeof(double) * vecs.size(), hipMemcpyHostToDevice, st));
        LIFE_HIP(hipMemcpyAsync(d.node, lay.node.data(), sizeof(double) * lay.node.size(), hipMemcpyHostToDevice, st));
        LIFE_HIP(hipMemcpyAsync(d.el, el.data(), sizeof(double) * el.size(), hipMemcpyHostToDevice, st));
    }
    LIFE_HIP(hipStreamSynchronize(st));
}

}  // namespace life
