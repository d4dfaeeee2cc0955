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

#include <atomic>
#include <chrono>
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
// kNZd[i][q] for a row index known only at run time
FDI int kNZdRow(int i, int q) {
    if (i == 2 || i == 5) return q == 0 ? i : -1;
    return q == 0 ? (i < 3 ? 0 : 3) : (i < 3 ? 1 : 4);
}

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

// Loads through the global address space: they cannot alias LDS, so the
// compiler may keep many in flight across LDS stores.
typedef const __attribute__((address_space(1))) double g_double;
typedef const __attribute__((address_space(1))) int g_int;
FDI double gload(const double* p) { return *(g_double*)p; }
FDI int gload(const int* p) { return *(g_int*)p; }

FDI double* vecp(const FemDevBody& B, int k) { return B.vec + (size_t)k * B.ndof; }
FDI double& band(double* A, int r, int c) { return A[(size_t)r * kFemBandW + (c - r + kFemKL)]; }

// Position of entry (r, c) of an element's mass or linear stiffness matrix
// among its upper-triangle non-zeros (the layout of FemDevBody::mats), -1
// for a structural zero; both matrices are symmetric with the same pattern.
// Branch-free: 4 bits per position lo * 6 + hi of the upper triangle in
// three words, 15 for a zero.
constexpr int kMatNZPos[kFemMatNZ] = {0, 3, 7, 8, 10, 11, 14, 16, 17, 21, 28, 29, 35};
constexpr unsigned long long mat_nz_word(int w) {
    unsigned long long v = ~0ull;
    for (int z = 0; z < kFemMatNZ; z++)
        if (kMatNZPos[z] / 16 == w) {
            const int sh = 4 * (kMatNZPos[z] % 16);
            v = (v & ~(15ull << sh)) | ((unsigned long long)z << sh);
        }
    return v;
}
__host__ __device__ inline int mat_nz(int r, int c) {
    constexpr unsigned long long w0 = mat_nz_word(0), w1 = mat_nz_word(1), w2 = mat_nz_word(2);
    const int lo = r < c ? r : c, hi = r < c ? c : r, p = lo * 6 + hi;
    const unsigned long long w = p < 16 ? w0 : (p < 32 ? w1 : w2);
    const int z = (int)((w >> (4 * (p & 15))) & 15ull);
    return z == 15 ? -1 : z;
}
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

// LDS of the dynamic solve (doubles): K band (ndof x kFemBandW), M band
// (ndof x kMBand), the force vector F, the load vector R, the Newmark
// vector, the increment delU, the elimination's row extents, and the
// body's marker forces and epsilon (3 per marker).  update_ibm's element
// terms reuse the K band.
constexpr int kMBand = 2 * kFemKL + 1;
struct FemLds {
    double *K, *M, *F, *R, *vec, *x, *mk;
    int* last;
};
FDI FemLds fem_lds(const FemDevBody& B, double* lds) {
    FemLds L;
    const size_t n = (size_t)B.ndof;
    L.K = lds;
    L.M = L.K + n * kFemBandW;
    L.F = L.M + n * kMBand;
    L.R = L.F + n;
    L.vec = L.R + n;
    L.x = L.vec + n;
    L.last = reinterpret_cast<int*>(L.x + n);
    L.mk = L.x + 2 * n;
    return L;
}
FDI double& mband(double* M, int r, int c) { return M[(size_t)r * kMBand + (c - r + kFemKL)]; }

// One load-map entry of element q (FemBody::load_vector, FEMElement.cpp:27-69):
// its global load terms RG.  mk: the body's marker forces (x, y) and
// epsilon, 3 per marker; k: the entry's marker, a and b: its zeta range.
FDI void load_entry_d(const FemDevBody& B, const double* mk, int q, int k, double a, double b, double RG[6]) {
    const double* e = elp(B, q);
    const double L = e[1], A = e[3], rho = e[6];
    double T[6][6];
    make_T(e[7], e[8], T);
    const double wt0 = rho * A * B.s.gx, wt1 = rho * A * B.s.gy;
    const double fs = B.s.Dm / (B.s.Dt * B.s.Dt);
    const double sc = -mk[3 * k + 2] * 1.0 * fs;
    const double vv[2] = {sc * mk[3 * k] + wt0, sc * mk[3 * k + 1] + wt1};
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

// The load-map entries en0 .. en1-1 of element q, added in map order onto
// acc (the node's terms of the element: its first node's RG[0..2] or its
// second node's RG[3..5]).  The entries' map values are loaded eight at a
// time, ahead of the arithmetic.
FDI void add_entries_d(const FemDevBody& B, const double* mk, int q, int en0, int en1, int half, double acc[3]) {
    constexpr int U = 8;
    for (int base = en0; base < en1; base += U) {
        int kk[U];
        double za[U], zb[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int en = base + u;
            kk[u] = en < en1 ? gload(B.map_node + en) : 0;
            za[u] = en < en1 ? gload(B.map_z + 2 * en) : 0.0;
            zb[u] = en < en1 ? gload(B.map_z + 2 * en + 1) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (base + u < en1) {
                double RG[6];
                load_entry_d(B, mk, q, kk[u], za[u], zb[u], RG);
                acc[0] += RG[3 * half];
                acc[1] += RG[3 * half + 1];
                acc[2] += RG[3 * half + 2];
            }
        }
    }
}

// R (FEMBodyClass::constructRVector) into LDS: node n's DOFs take the terms
// of element n-1 (its second node) then of element n (its first node),
// each in load-map order, onto 0.0.
FDI void load_vector_d(const FemDevBody& B, const FemLds& S) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = tid; k < B.nm; k += nt) {
        S.mk[3 * k] = gload(B.mforce + 2 * k);
        S.mk[3 * k + 1] = gload(B.mforce + 2 * k + 1);
        S.mk[3 * k + 2] = gload(B.meps + k);
    }
    __syncthreads();
    for (int n = tid; n < B.nnode; n += nt) {
        double acc[3] = {0.0, 0.0, 0.0};
        const int s0 = n >= 1 ? gload(B.map_start + n - 1) : 0;
        const int s1 = gload(B.map_start + n);
        const int s2 = n < B.nel ? gload(B.map_start + n + 1) : 0;
        if (n >= 1) add_entries_d(B, S.mk, n - 1, s0, s1, 1, acc);
        if (n < B.nel) add_entries_d(B, S.mk, n, s1, s2, 0, acc);
        S.R[3 * n] = acc[0];
        S.R[3 * n + 1] = acc[1];
        S.R[3 * n + 2] = acc[2];
    }
    __syncthreads();
}

// Entry (r, c) of the element's rotation matrix make_T(ca, sa).
FDI double t_entry(double ca, double sa, int r, int c) {
    if (r / 3 != c / 3) return 0.0;
    const int a = r % 3, b = c % 3;
    if (a == 2 || b == 2) return a == b ? 1.0 : 0.0;
    return a == b ? ca : (a == 0 ? sa : -sa);
}

// Row i of Transpose(T) A T, with the terms of rotate6_d in its order:
// the rows of A that row i of Transpose(T) A reads are a0 and a1 (a1
// unused when row i reads one row), weighted by T[ra][i] and T[rb][i].
FDI void rotate_row_d(double ca, double sa, int i, const double a0[6], const double a1[6], double out[6]) {
    const int ra = kNZdRow(i, 0), rb = kNZdRow(i, 1);
    const double t0 = t_entry(ca, sa, ra, i), t1 = rb >= 0 ? t_entry(ca, sa, rb, i) : 0.0;
    double tmp[6];
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double s = 0.0;
        s += t0 * a0[j];
        if (rb >= 0) s += t1 * a1[j];
        tmp[j] = s;
    }
    double T[6][6];
    make_T(ca, sa, T);
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 2; q++)
            if (kNZd[j][q] >= 0) s += tmp[kNZd[j][q]] * T[kNZd[j][q]][j];
        out[j] = s;
    }
}

// Row r of the element's geometric stiffness K_NL (FEMElement.cpp:119-131).
FDI void knl_row(int r, double Fa, double V0, double L0, double out[6]) {
#pragma unroll
    for (int j = 0; j < 6; j++) out[j] = 0.0;
    switch (r) {
        case 0: out[1] = -V0 / L0; out[4] = V0 / L0; break;
        case 1: out[0] = -V0 / L0; out[1] = Fa / L0; out[3] = V0 / L0; out[4] = -Fa / L0; break;
        case 3: out[1] = V0 / L0; out[4] = -V0 / L0; break;
        case 4: out[0] = V0 / L0; out[1] = -Fa / L0; out[3] = -V0 / L0; out[4] = Fa / L0; break;
        default: break;
    }
}

// force_vector + mass_stiff of element q (FEMElement.cpp:72-138), row i of
// the element's terms (one lane per row): added onto the global F, M and K
// in LDS.  Each entry takes the same operations as the whole-element form.
// Every value the row reads from device memory is loaded first, so the
// loads share one round trip.
FDI void element_row_d(const FemDevBody& B, const FemLds& S, int q, int i) {
    const int ra = kNZdRow(i, 0), rb = kNZdRow(i, 1);
    const int rb1 = rb >= 0 ? rb : ra;
    const double* e = B.el + (size_t)q * kFemElStride;
    const double* Mc = B.mats + 2 * kFemMatNZ * (size_t)q;
    const double* Kc = Mc + kFemMatNZ;
    double ev[9], m0[6], m1[6], k0[6], k1[6];
#pragma unroll
    for (int k = 0; k < 9; k++) ev[k] = e[k];
    const double nth0 = B.node[6 * q + 5], nth1 = B.node[6 * (q + 1) + 5];
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const int za = mat_nz(ra, j), zb = mat_nz(rb1, j);
        m0[j] = za >= 0 ? Mc[za] : 0.0;
        m1[j] = zb >= 0 ? Mc[zb] : 0.0;
        k0[j] = za >= 0 ? Kc[za] : 0.0;
        k1[j] = zb >= 0 ? Kc[zb] : 0.0;
    }
    const double L0 = ev[0], L = ev[1], angle = ev[2], A = ev[3], I = ev[4], E = ev[5];
    const double ca = ev[7], sa = ev[8];
    const double u = (L * L - L0 * L0) / (L + L0);
    const double th1 = wrap_angle_d(nth0 - angle);
    const double th2 = wrap_angle_d(nth1 - angle);
    const double F0 = (E * A / L0) * u;
    const double M1 = (2 * E * I / L0) * (2.0 * th1 + th2);
    const double M2 = (2 * E * I / L0) * (th1 + 2.0 * th2);
    // Fe (the element's local forces): the two entries row i reads
    const double Fsum = (1.0 / L0) * (M1 + M2);
    auto fe = [&](int r) {
        switch (r) {
            case 0: return -F0;
            case 1: return Fsum;
            case 2: return M1;
            case 3: return F0;
            case 4: return -((1.0 / L0) * (M1 + M2));
            default: return M2;
        }
    };
    // row i of Transpose(T) Fe (mul_t6_d)
    double fg = 0.0;
    fg += t_entry(ca, sa, ra, i) * fe(ra);
    if (rb >= 0) fg += t_entry(ca, sa, rb, i) * fe(rb);
    const int d0 = 3 * q, row = d0 + i;
    double G[6], N[6], n0[6], n1[6];
    rotate_row_d(ca, sa, i, m0, m1, G);
    S.F[row] += fg;
#pragma unroll
    for (int j = 0; j < 6; j++) mband(S.M, row, d0 + j) += G[j];
    rotate_row_d(ca, sa, i, k0, k1, G);
    const double Fa = F0, V0 = -((1.0 / L0) * (M1 + M2));
    knl_row(ra, Fa, V0, L0, n0);
    knl_row(rb1, Fa, V0, L0, n1);
    rotate_row_d(ca, sa, i, n0, n1, N);
#pragma unroll
    for (int j = 0; j < 6; j++) band(S.K, row, d0 + j) += G[j] + N[j];
}

// FemBody::build_global (FEMBody.cpp:90-109) into LDS.  Elements q and q+1
// share a node, so the even elements go first, then the odd ones; an entry
// takes at most two terms, whose sum does not depend on their order.
FDI void build_global_d(const FemDevBody& B, const FemLds& S) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = tid; k < B.ndof * kFemBandW; k += nt) S.K[k] = 0.0;
    for (int k = tid; k < B.ndof * kMBand; k += nt) S.M[k] = 0.0;
    for (int k = tid; k < B.ndof; k += nt) S.F[k] = 0.0;
    __syncthreads();
    const int nph = (B.nel + 1) / 2;  // elements of one parity, at most
    for (int ph = 0; ph < 2; ph++) {
        for (int t = tid; t < 6 * nph; t += nt) {
            const int q = ph + 2 * (t / 6);
            if (q < B.nel) element_row_d(B, S, q, t % 6);
        }
        __syncthreads();
    }
}

// Inverse of a 3x3 block (row-major, in place) by its adjugate.
FDI void inv3_d(double* D) {
    const double a = D[0], b = D[1], c = D[2], d = D[3], e = D[4], f = D[5], g = D[6], h = D[7], i = D[8];
    const double c00 = e * i - f * h, c01 = f * g - d * i, c02 = d * h - e * g;
    const double det = a * c00 + b * c01 + c * c02;
    const double r = 1.0 / det;
    D[0] = c00 * r;
    D[1] = (c * h - b * i) * r;
    D[2] = (b * f - c * e) * r;
    D[3] = c01 * r;
    D[4] = (a * i - c * g) * r;
    D[5] = (c * d - a * f) * r;
    D[6] = c02 * r;
    D[7] = (b * g - a * h) * r;
    D[8] = (a * e - b * d) * r;
}

// Node blocks of the cyclic reduction: sub-diagonal A, diagonal D (its
// inverse once the node is eliminated), super-diagonal C, right-hand side f.
constexpr int kBlkA = 0, kBlkD = 9, kBlkC = 18, kBlkF = 27, kBlk = 30;

// K x = F (K + a0 M from newton_d, band in LDS) with x = 0 on the nbc
// clamped DOFs.  An element couples two neighbouring nodes, so K is block
// tridiagonal in 3 x 3 node blocks: A_n x_{n-1} + D_n x_n + C_n x_{n+1} =
// f_n.  The clamped DOFs become identity rows and columns, which leaves the
// other unknowns' equations as lu_solve_band's trailing block.  Block cyclic
// reduction: at stride s = 1, 2, 4, ... the nodes at odd multiples of s are
// eliminated from their neighbours' equations (the nodes at multiples of
// 2s), so the chain of dependent steps is log2(nnode) levels long instead
// of nnode; back substitution then runs the levels in reverse.  K is
// symmetric positive definite (the Newmark mass term dominates), so no
// pivoting is needed.  The values agree with the host's scalar elimination
// to rounding, not bit for bit.  Blocks: kBlk doubles per node in the M
// band's space (free once K + a0 M is formed); a level's new blocks are
// staged in the K band's space (free once the blocks are read out of it).
FDI void block_solve_d(const FemDevBody& B, const FemLds& S) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int N = B.nnode, nbc = B.nbc;
    // constrained DOFs: identity rows and columns, zero right-hand side
    for (int t = tid; t < nbc * (2 * kFemKL + 1); t += nt) {
        const int d = t / (2 * kFemKL + 1), o = t % (2 * kFemKL + 1) - kFemKL;  // o = column - row
        const int c = d + o;
        if (c >= 0 && c < B.ndof) {
            band(S.K, d, c) = (o == 0) ? 1.0 : 0.0;  // row d
            if (o != 0) band(S.K, c, d) = 0.0;        // column d
        }
    }
    for (int d = tid; d < nbc; d += nt) S.F[d] = 0.0;
    __syncthreads();
    double* blk = S.M;
    for (int t = tid; t < N * kBlk; t += nt) {
        const int n = t / kBlk, q = t % kBlk;
        double v;
        if (q < kBlkF) {
            const int cn = n + q / 9 - 1, i = (q % 9) / 3, j = q % 3;  // column node n-1, n, n+1
            v = (cn < 0 || cn >= N) ? 0.0 : band(S.K, 3 * n + i, 3 * cn + j);
        } else {
            v = S.F[3 * n + q - kBlkF];
        }
        blk[t] = v;
    }
    __syncthreads();
    double* stage = S.K;
    int s = 1;
    for (; s < N; s *= 2) {
        // nodes j = s, 3s, 5s, ... leave: D_j becomes its inverse
        for (int j = s + 2 * s * tid; j < N; j += 2 * s * nt) inv3_d(blk + kBlk * j + kBlkD);
        __syncthreads();
        // nodes i = 0, 2s, 4s, ... take them out, entry (r, c) per lane:
        //   al = A_i inv(D_{i-s}), ga = C_i inv(D_{i+s})
        //   D_i - al C_{i-s} - ga A_{i+s},  -al A_{i-s},  -ga C_{i+s},
        //   f_i - al f_{i-s} - ga f_{i+s}
        const int nr = (N - 1) / (2 * s) + 1;
        for (int t = tid; t < 9 * nr; t += nt) {
            const int k = t / 9, r = (t % 9) / 3, c = t % 3, i = 2 * s * k;
            const double* bi = blk + kBlk * i;
            double Dn = bi[kBlkD + 3 * r + c], An = 0.0, Cn = 0.0, fn = bi[kBlkF + r];
            if (i - s >= 0) {
                const double* bj = blk + kBlk * (i - s);
                double al[3];
#pragma unroll
                for (int q = 0; q < 3; q++)
                    al[q] = bi[kBlkA + 3 * r] * bj[kBlkD + q] + bi[kBlkA + 3 * r + 1] * bj[kBlkD + 3 + q] +
                            bi[kBlkA + 3 * r + 2] * bj[kBlkD + 6 + q];
                Dn = Dn - (al[0] * bj[kBlkC + c] + al[1] * bj[kBlkC + 3 + c] + al[2] * bj[kBlkC + 6 + c]);
                An = -(al[0] * bj[kBlkA + c] + al[1] * bj[kBlkA + 3 + c] + al[2] * bj[kBlkA + 6 + c]);
                fn = fn - (al[0] * bj[kBlkF] + al[1] * bj[kBlkF + 1] + al[2] * bj[kBlkF + 2]);
            }
            if (i + s < N) {
                const double* bj = blk + kBlk * (i + s);
                double ga[3];
#pragma unroll
                for (int q = 0; q < 3; q++)
                    ga[q] = bi[kBlkC + 3 * r] * bj[kBlkD + q] + bi[kBlkC + 3 * r + 1] * bj[kBlkD + 3 + q] +
                            bi[kBlkC + 3 * r + 2] * bj[kBlkD + 6 + q];
                Dn = Dn - (ga[0] * bj[kBlkA + c] + ga[1] * bj[kBlkA + 3 + c] + ga[2] * bj[kBlkA + 6 + c]);
                Cn = -(ga[0] * bj[kBlkC + c] + ga[1] * bj[kBlkC + 3 + c] + ga[2] * bj[kBlkC + 6 + c]);
                fn = fn - (ga[0] * bj[kBlkF] + ga[1] * bj[kBlkF + 1] + ga[2] * bj[kBlkF + 2]);
            }
            double* o = stage + kBlk * k;
            o[kBlkA + 3 * r + c] = An;
            o[kBlkD + 3 * r + c] = Dn;
            o[kBlkC + 3 * r + c] = Cn;
            if (c == 0) o[kBlkF + r] = fn;
        }
        __syncthreads();
        for (int t = tid; t < kBlk * nr; t += nt) blk[kBlk * 2 * s * (t / kBlk) + t % kBlk] = stage[t];
        __syncthreads();
    }
    // node 0 alone: x_0 = inv(D_0) f_0
    if (tid == 0) inv3_d(blk + kBlkD);
    __syncthreads();
    if (tid < 3) {
        const double* b0 = blk;
        S.x[tid] = b0[kBlkD + 3 * tid] * b0[kBlkF] + b0[kBlkD + 3 * tid + 1] * b0[kBlkF + 1] +
                   b0[kBlkD + 3 * tid + 2] * b0[kBlkF + 2];
    }
    __syncthreads();
    // back substitution, levels in reverse: x_j = inv(D_j) (f_j - A_j x_{j-s} - C_j x_{j+s})
    for (s /= 2; s >= 1; s /= 2) {
        const int ne = (N + s - 1) / (2 * s);
        for (int t = tid; t < 3 * ne; t += nt) {
            const int j = s + 2 * s * (t / 3), r = t % 3;
            const double* bj = blk + kBlk * j;
            const double* xl = S.x + 3 * (j - s);
            double v[3];
#pragma unroll
            for (int a = 0; a < 3; a++)
                v[a] = bj[kBlkF + a] - (bj[kBlkA + 3 * a] * xl[0] + bj[kBlkA + 3 * a + 1] * xl[1] +
                                        bj[kBlkA + 3 * a + 2] * xl[2]);
            if (j + s < N) {
                const double* xr = S.x + 3 * (j + s);
#pragma unroll
                for (int a = 0; a < 3; a++)
                    v[a] = v[a] - (bj[kBlkC + 3 * a] * xr[0] + bj[kBlkC + 3 * a + 1] * xr[1] +
                                   bj[kBlkC + 3 * a + 2] * xr[2]);
            }
            S.x[3 * j + r] = bj[kBlkD + 3 * r] * v[0] + bj[kBlkD + 3 * r + 1] * v[1] + bj[kBlkD + 3 * r + 2] * v[2];
        }
        __syncthreads();
    }
    for (int d = tid; d < nbc; d += nt) S.x[d] = 0.0;
    __syncthreads();
}

// FemBody::newton_dynamic (FEMBody.cpp:71-87, 112-126) and the residual
// test of dynamic(): returns resNR (every lane).
FDI double newton_d(const FemDevBody& B, const FemLds& S) {
    const int tid = threadIdx.x, nt = blockDim.x;
    build_global_d(B, S);
    const double Dt = B.s.Dt, al = B.s.alpha;
    const double a0 = 1.0 / (al * (Dt * Dt));
    const double a2 = 1.0 / (al * Dt);
    const double a3 = 1.0 / (2.0 * al) - 1.0;
    const int nd = B.ndof;
    double *U = vecp(B, kFU), *Un = vecp(B, kFUn), *Ud = vecp(B, kFUd), *Udd = vecp(B, kFUdd);
    for (int k = tid; k < nd; k += nt) S.vec[k] = a0 * (Un[k] - U[k]) + a2 * Ud[k] + a3 * Udd[k];
    __syncthreads();
    const int kb = kFemKL;
    for (int i = tid; i < nd; i += nt) {
        double s = 0.0;
        const int lo = i - kb < 0 ? 0 : i - kb, hi = i + kb > nd - 1 ? nd - 1 : i + kb;
        for (int j = lo; j <= hi; j++) s += mband(S.M, i, j) * S.vec[j];
        S.F[i] = S.R[i] - S.F[i] + s;
        for (int j = lo; j <= hi; j++) band(S.K, i, j) = band(S.K, i, j) + a0 * mband(S.M, i, j);
    }
    __syncthreads();
    block_solve_d(B, S);
    for (int k = tid; k < nd; k += nt) U[k] = U[k] + S.x[k];
    __syncthreads();
    update_fem_d(B);
    __shared__ double res_sh;
    if (tid == 0) {
        double r = 0.0;
        for (int k = 0; k < nd; k++) r += S.x[k] * S.x[k];
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

// LDS of a launch: the solve's scratch (fem_lds), then the body mirrored
// from device memory: its state vectors U .. Rkm1, nodes, elements and
// element matrices.  The kernels run on the mirror (one round trip to
// device memory in, one out) and write the markers and out[] directly.
__host__ __device__ inline size_t scratch_doubles(int ndof, int nm) { return (size_t)ndof * (kFemBandW + kMBand + 5) + 3 * (size_t)nm; }

FDI FemDevBody mirror_in(const FemDevBody& G, double* lds, bool mats) {
    FemDevBody B = G;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nv = kFemState * G.ndof, nn = 6 * G.nnode, ne = kFemElStride * G.nel,
              nk = mats ? 2 * kFemMatNZ * G.nel : 0;
    double* m = lds + scratch_doubles(G.ndof, G.nm);
    B.vec = m;
    B.node = m + nv;
    B.el = B.node + nn;
    B.mats = B.el + ne;
    // vec, node, el, mats lie in this order in one device block:
    // one contiguous copy, sixteen loads in flight per lane
    const int total = nv + nn + ne + nk;
    constexpr int U = 16;
    for (int base = 0; base < total; base += U * nt) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = gload(G.vec + min(base + u * nt + tid, total - 1));
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = base + u * nt + tid;
            if (k < total) m[k] = v[u];
        }
    }
    __syncthreads();
    return B;
}

FDI void mirror_out(const FemDevBody& G, const FemDevBody& B) {
    __syncthreads();
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nv = kFemState * G.ndof, nn = 6 * G.nnode, ne = kFemElStride * G.nel;
    for (int k = tid; k < nv; k += nt) G.vec[k] = B.vec[k];
    for (int k = tid; k < nn; k += nt) G.node[k] = B.node[k];
    for (int k = tid; k < ne; k += nt) G.el[k] = B.el[k];
}

// list: the bodies of this launch (null: 0 .. nb-1); param: per listed
// body, its step t (predict) or relaxation factor (relax); null: value.
FDI int body_index(const int* list) { return list ? list[blockIdx.x] : (int)blockIdx.x; }

__global__ __launch_bounds__(kFemThreads) void k_fem_predict(const FemDevBody* __restrict__ bodies,
                                                            const int* __restrict__ list, int t_all,
                                                            const double* __restrict__ t_d) {
    extern __shared__ double lds[];
    const FemDevBody G = bodies[body_index(list)];
    const FemDevBody B = mirror_in(G, lds, false);
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
    mirror_out(G, B);
}

// ObjectsClass::recomputeObjectVals relaxation (Objects.cpp:198-210)
__global__ __launch_bounds__(kFemThreads) void k_fem_relax(const FemDevBody* __restrict__ bodies,
                                                          const int* __restrict__ list, double omega,
                                                          const double* __restrict__ omega_d) {
    extern __shared__ double lds[];
    const FemDevBody G = bodies[body_index(list)];
    const FemDevBody B = mirror_in(G, lds, false);
    const double w = omega_d ? omega_d[blockIdx.x] : omega;
    double *U = vecp(B, kFU), *Ukm1 = vecp(B, kFUkm1);
    for (int k = threadIdx.x; k < B.ndof; k += blockDim.x) U[k] = Ukm1[k] + w * (U[k] - Ukm1[k]);
    __syncthreads();
    update_fem_d(B);
    finish_newmark_d(B);
    update_ibm_d(B, lds);
    copy_vec(B, kFUkm1, kFU);
    mirror_out(G, B);
}

// The workgroup that finishes last (an atomic count of finished bodies,
// with device-scope fences on both sides; nothing waits) adds the bodies'
// subiteration terms out[0..2] in body order, as case_fem does, and writes
// them and the epsilon systems' out-of-band flags into page-locked host
// memory, re-arms the count for the next launch, and last (after a
// system-scope fence) writes the launch's sequence number beside the sums.
FDI void sums_by_last(const FemDevBody* __restrict__ bodies, const FemSumsOut& so) {
    __shared__ int last;
    __shared__ double t[3 * kFemThreads];
    __syncthreads();  // every lane's writes of this body precede the count
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(so.done, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    const int nb = gridDim.x;
    for (int b = threadIdx.x; b < nb && b < kFemThreads; b += blockDim.x) {
        const double* o = bodies[b].out;
        t[3 * b] = o[0];
        t[3 * b + 1] = o[1];
        t[3 * b + 2] = o[2];
    }
    for (int k = threadIdx.x; k < so.nflags; k += blockDim.x) so.flags_h[k] = so.flags[k];
    // each lane's flag copies reach the host before lane 0 writes seq
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, num = 0.0, den = 0.0;
        for (int b = 0; b < nb; b++) {
            const double* o = b < kFemThreads ? t + 3 * b : bodies[b].out;
            r += o[0];
            num += o[1];
            den += o[2];
        }
        so.sums_h[0] = r;
        so.sums_h[1] = num;
        so.sums_h[2] = den;
        atomicExch(so.done, 0);
        __threadfence_system();
        *(volatile double*)&so.sums_h[3] = so.seq;
    }
}

// FemBody::dynamic (FEMBody.cpp:26-68) and subResidual (FEMBody.cpp:244-256)
// redo: the same solve again after epsilon changed (the dense fallback of
// computeEpsilon): Rkm1 already holds the previous subiteration's Rk.
__global__ __launch_bounds__(kFemThreads) void k_fem_dynamic(const FemDevBody* __restrict__ bodies,
                                                            const int* __restrict__ list, int redo,
                                                            FemSumsOut so) {
    extern __shared__ double lds[];
    const FemDevBody G = bodies[body_index(list)];
    const FemDevBody B = mirror_in(G, lds, true);
    copy_vec(B, kFU, kFUn);
    copy_vec(B, kFUd, kFUdn);
    copy_vec(B, kFUdd, kFUddn);
    __syncthreads();
    update_fem_d(B);
    const FemLds S = fem_lds(B, lds);
    load_vector_d(B, S);
    int it = 0;
    double res;
    do {
        res = newton_d(B, S);
        it++;
    } while (res > 1e-10 && it < 20);
    finish_newmark_d(B);
    update_ibm_d(B, lds);
    // Rkm1 takes Rk, Rk = U - Ukm1
    if (!redo) copy_vec(B, kFRkm1, kFRk);
    __syncthreads();
    double *Rk = vecp(B, kFRk), *Rkm1 = vecp(B, kFRkm1), *U = vecp(B, kFU), *Ukm1 = vecp(B, kFUkm1);
    // Rk and Rkm1 also into LDS (F and R are free now) for the sums below
    for (int k = threadIdx.x; k < B.ndof; k += blockDim.x) {
        const double rk = U[k] - Ukm1[k];
        Rk[k] = rk;
        S.F[k] = rk;
        S.R[k] = Rkm1[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // the three sums in DOF order each, side by side
        double r = 0.0, num = 0.0, den = 0.0;
        for (int k = 0; k < B.ndof; k++) {
            const double rk = S.F[k], rm = S.R[k];
            r += rk * rk;
            num += rm * (rk - rm);
            den += (rk - rm) * (rk - rm);
        }
        B.out[0] = r;
        B.out[1] = num;
        B.out[2] = den;
        B.out[3] = (double)it;
        B.out[4] = res;
    }
    mirror_out(G, B);
    if (so.done) sums_by_last(bodies, so);
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

}  // namespace

size_t fem_dev_lds_bytes(int ndof, int nm) {
    // the dynamic solve's scratch (fem_lds; update_ibm's 12 per element fit
    // in its K band), then the mirror of a chain of ndof / 3 nodes
    const size_t nnode = (size_t)ndof / 3, nel = nnode - 1;
    const size_t mirror = kFemState * (size_t)ndof + 6 * nnode + (kFemElStride + 2 * kFemMatNZ) * nel;
    return sizeof(double) * (scratch_doubles(ndof, nm) + mirror);
}

void launch_fem_predict(const FemDevBody* b, const int* list, int nb, int t, const double* t_d, size_t lds,
                        hipStream_t st) {
    if (nb == 0) return;
    k_fem_predict<<<nb, kFemThreads, lds, st>>>(b, list, t, t_d);
    LIFE_HIP(hipGetLastError());
}

void launch_fem_relax(const FemDevBody* b, const int* list, int nb, double omega, const double* omega_d, size_t lds,
                      hipStream_t st) {
    if (nb == 0) return;
    k_fem_relax<<<nb, kFemThreads, lds, st>>>(b, list, omega, omega_d);
    LIFE_HIP(hipGetLastError());
}

void launch_fem_dynamic(const FemDevBody* b, const int* list, int nb, size_t lds, bool redo, hipStream_t st,
                        const FemSumsOut& so) {
    if (nb == 0) return;
    if (so.done && list) throw std::logic_error("device FEM: sums of a listed launch");
    k_fem_dynamic<<<nb, kFemThreads, lds, st>>>(b, list, redo ? 1 : 0, so);
    LIFE_HIP(hipGetLastError());
}

void launch_fem_outs(const FemDevBody* b, const int* list, int nb, double* outs, hipStream_t st) {
    if (nb > 0) k_fem_outs<<<(nb + 63) / 64, 64, 0, st>>>(b, list, nb, outs);
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
    vecs.assign((size_t)kFemState * nd, 0.0);
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
        // the block solve takes a chain of 3 x 3 node blocks
        if (d.ndof != 3 * d.nnode || d.nel != d.nnode - 1 || d.nbc > d.ndof || d.nbc < 0)
            throw std::runtime_error("device FEM: a body must be a chain of nodes with 3 DOFs each");
        if (lay.node.size() != (size_t)6 * d.nnode || lay.Mloc.size() != (size_t)36 * d.nel ||
            lay.KLloc.size() != (size_t)36 * d.nel)
            throw std::runtime_error("device FEM: unexpected body layout");
        // one block: vectors, nodes, elements, element matrices (the kernels
        // copy it into LDS in one run); of M and KL only the upper-triangle
        // non-zeros, the host's values exactly
        std::vector<double> slab(vecs);
        slab.insert(slab.end(), lay.node.begin(), lay.node.end());
        slab.insert(slab.end(), el.begin(), el.end());
        for (int q = 0; q < d.nel; q++)
            for (const std::vector<double>* m : {&lay.Mloc, &lay.KLloc}) {
                double nz[kFemMatNZ];
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) {
                        const double v = (*m)[(size_t)36 * q + 6 * i + j];
                        const int z = mat_nz(i, j);
                        if (z < 0 ? v != 0.0 : (i <= j ? false : v != nz[z]))
                            throw std::runtime_error("device FEM: element matrix outside the beam pattern");
                        if (z >= 0 && i <= j) nz[z] = v;
                    }
                slab.insert(slab.end(), nz, nz + kFemMatNZ);
            }
        d.vec = upload_vec(fc, slab, st);
        d.node = d.vec + (size_t)kFemState * d.ndof;
        d.el = d.node + (size_t)6 * d.nnode;
        d.mats = d.el + (size_t)kFemElStride * d.nel;
        d.map_start = upload_vec(fc, lay.map_start, st);
        d.map_node = upload_vec(fc, lay.map_node, st);
        d.map_z = upload_vec(fc, lay.map_z, st);
        d.pm_el = upload_vec(fc, lay.pm_el, st);
        d.pm_zeta = upload_vec(fc, lay.pm_zeta, st);
        d.out = upload_vec(fc, std::vector<double>(8, 0.0), st);
        d.mpos = pos + 2 * (size_t)first[b];
        d.mvel = vel + 2 * (size_t)first[b];
        d.mforce = force + 2 * (size_t)first[b];
        d.meps = eps + first[b];
        fc.host.push_back(d);
        const size_t need = fem_dev_lds_bytes(d.ndof, d.nm);
        if (need > lds) lds = need;
    }
    if (lds > kFemMaxLds) throw std::runtime_error("flexible body too large for the device FEM");
    fc.lds = lds;
    // dynamic LDS beyond 64 KiB per workgroup
    for (const void* k : {(const void*)k_fem_predict, (const void*)k_fem_relax, (const void*)k_fem_dynamic})
        LIFE_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFemMaxLds));
    fc.d_bodies = upload_vec(fc, fc.host, st);
    fc.done = dalloc<int>(1);
    fc.blocks.push_back(fc.done);
    LIFE_HIP(hipMemsetAsync(fc.done, 0, sizeof(int), st));
    LIFE_HIP(hipHostMalloc((void**)&fc.sums_h, 4 * sizeof(double), hipHostMallocDefault));
    LIFE_HIP(hipStreamSynchronize(st));
}

void fem_wait_sums(const FemSumsOut& so, hipEvent_t ev, double spin_us) {
    const volatile double* w = so.sums_h + 3;
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        if (*w == so.seq) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return;
        }
        const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
        if (us > spin_us) break;
    }
    // not seen within the spin: the event (and any error of the launch)
    LIFE_HIP(hipEventSynchronize(ev));
    std::atomic_thread_fence(std::memory_order_acquire);
}

void fem_dev_destroy(FemDevCase& fc) {
    for (void* p : fc.blocks) dfree(p);
    fc.blocks.clear();
    fc.host.clear();
    fc.d_bodies = nullptr;
    fc.done = nullptr;
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
