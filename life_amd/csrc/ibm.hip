// Immersed-boundary kernels: support search, marker spacing, epsilon
// matrix, interpolation + direct forcing, and the ordered spread.
//
// All arithmetic is FP64 with the reference's operation order
// (IBMNode.cpp, Objects.cpp:235-321), for both population storages.
// Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <stdexcept>

#include "d2q9_f64.h"
#include "launch.h"

namespace life {

// The single-case coupling kernels run beside the next step's bulk
// stream-collide (launched early on another stream): their waves take the
// highest issue priority on a SIMD they share with stream-collide waves,
// since they are usually the critical path of the step.  The level (0..3)
// is set per device by set_coupling_priority (LIFE_COUPLING_PRIO, default 3).
__device__ int g_coupling_prio = 3;
__device__ __forceinline__ void coupling_priority() {
    const int p = g_coupling_prio;
    if (p >= 3) __builtin_amdgcn_s_setprio(3);
    else if (p == 2) __builtin_amdgcn_s_setprio(2);
    else if (p == 1) __builtin_amdgcn_s_setprio(1);
}

hipError_t set_coupling_priority(int level) {
    return hipMemcpyToSymbol(HIP_SYMBOL(g_coupling_prio), &level, sizeof(int));
}

// IBMNodeClass::findSupport (IBMNode.cpp:139-179): lattice sites within 1.5
// spacings of the marker in both directions, scanned i-major over the 5x5
// block around the nearest site, kept in scan order (at most 9).
__device__ inline void find_support_xy(const MarkerDev& m, const GridScalars& g, int k, double x, double y) {
    const double Dx = g.Dx;
    const int inear = (int)round(x / Dx), jnear = (int)round(y / Dx);
    int n = 0;
    for (int i = inear - 2; i <= inear + 2; i++) {
        const double dx = fabs(x / Dx - i);
        for (int j = jnear - 2; j <= jnear + 2; j++) {
            const double dy = fabs(y / Dx - j);
            if (dx < 1.5 && dy < 1.5 && i >= 0 && i <= g.Nx - 1 && j >= 0 && j <= g.Ny - 1) {
                if (n == kSupp) {
                    *m.overflow = 1;
                    continue;
                }
                m.supp_site[k * kSupp + n] = i * g.Ny + j;
                m.supp_ij[(k * kSupp + n) * 2] = i;
                m.supp_ij[(k * kSupp + n) * 2 + 1] = j;
                m.supp_delta[k * kSupp + n] = dirac64(dx) * dirac64(dy);
                n++;
            }
        }
    }
    for (int s = n; s < kSupp; s++) {
        m.supp_site[k * kSupp + s] = -1;
        m.supp_ij[(k * kSupp + s) * 2] = 0;
        m.supp_ij[(k * kSupp + s) * 2 + 1] = 0;
        m.supp_delta[k * kSupp + s] = 0.0;
    }
    m.supp_cnt[k] = n;
}

__device__ inline void find_support_one(const MarkerDev& m, const GridScalars& g, int flexible_only, int k) {
    if (flexible_only && m.body_flex[m.body[k]] != 0) return;
    find_support_xy(m, g, k, m.pos[2 * k], m.pos[2 * k + 1]);
}

__global__ void k_find_support(MarkerDev m, GridScalars g, int flexible_only) {
    coupling_priority();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m.n) find_support_one(m, g, flexible_only, k);
}

__global__ void k_find_support_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list,
                                     int flexible_only) {
    const CaseDev& c = cases[list[blockIdx.y]];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < c.m.n) find_support_one(c.m, c.g, flexible_only, k);
}

// IBMNodeClass::computeDs (IBMNode.cpp:182-204): distance (lattice units)
// to the nearest other marker of the same body, starting from 10.  One
// workgroup per body; the body's positions are staged in LDS when they fit.
constexpr int kDsLdsMarkers = 2048;

// The positions of body markers [first, first + cnt) staged in LDS (spos)
// when they fit, else read from device memory; all lanes of the workgroup.
__device__ inline const double* stage_body_positions(const MarkerDev& m, int first, int cnt, double* spos) {
    const bool lds = cnt <= kDsLdsMarkers;
    if (lds)
        for (int t = threadIdx.x; t < 2 * cnt; t += blockDim.x) spos[t] = m.pos[2 * first + t];
    __syncthreads();
    return lds ? spos : m.pos + 2 * first;
}

__device__ inline void ds_from_positions(const MarkerDev& m, const GridScalars& g, int first, int cnt,
                                         const double* P) {
    const double2* P2 = reinterpret_cast<const double2*>(P);
    // sqrt(.)/Dx is monotone, so the minimum of sqrt(d.d)/Dx over the other
    // markers is sqrt(min d.d)/Dx: one square root per marker, same value.
    // The minimum of a set does not depend on the order of the comparisons,
    // so eight running minima take eight markers per round, their positions
    // loaded (16 bytes each) before the comparisons.
    constexpr int U = 8;
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const double x = P[2 * k], y = P[2 * k + 1];
        double mn[U];
#pragma unroll
        for (int u = 0; u < U; u++) mn[u] = INFINITY;
        int q0 = 0;
        for (; q0 + U <= cnt; q0 += U) {
            double2 pq[U];
#pragma unroll
            for (int u = 0; u < U; u++) pq[u] = P2[q0 + u];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const double dx = x - pq[u].x, dy = y - pq[u].y;
                const double d2 = 0.0 + dx * dx + dy * dy;
                if (q0 + u != k && d2 < mn[u]) mn[u] = d2;
            }
        }
        for (int q = q0; q < cnt; q++) {
            const double dx = x - P[2 * q], dy = y - P[2 * q + 1];
            const double d2 = 0.0 + dx * dx + dy * dy;
            if (q != k && d2 < mn[0]) mn[0] = d2;
        }
        double d2min = mn[0];
#pragma unroll
        for (int u = 1; u < U; u++)
            if (mn[u] < d2min) d2min = mn[u];
        const double mag = sqrt(d2min) / g.Dx;
        m.ds[first + k] = mag < 10.0 ? mag : 10.0;
    }
}

__device__ inline void compute_ds_body(const MarkerDev& m, const GridScalars& g, int flexible_only, int b,
                                       double* spos) {
    if (flexible_only && m.body_flex[b] != 0) return;
    const int first = m.body_first[b], cnt = m.body_n[b];
    ds_from_positions(m, g, first, cnt, stage_body_positions(m, first, cnt, spos));
}

__global__ void k_compute_ds(MarkerDev m, GridScalars g, int flexible_only) {
    coupling_priority();
    extern __shared__ double spos[];
    compute_ds_body(m, g, flexible_only, blockIdx.x, spos);
}

// Supports and spacing of one body per workgroup in one launch: the two
// are independent (spacing reads only positions), so the coupling chain
// has one kernel fewer to wait for.
// The body's positions are read once (into LDS), and its kind once per
// workgroup, so beside a bandwidth-bound kernel the chain of dependent
// device-memory reads is two long (body table, positions).
__global__ void k_support_ds(MarkerDev m, GridScalars g, int flexible_only) {
    coupling_priority();
    extern __shared__ double spos[];
    const int b = blockIdx.x;
    const int first = m.body_first[b], cnt = m.body_n[b];
    if (flexible_only && m.body_flex[b] != 0) return;  // the whole workgroup
    const double* P = stage_body_positions(m, first, cnt, spos);
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) find_support_xy(m, g, first + k, P[2 * k], P[2 * k + 1]);
    ds_from_positions(m, g, first, cnt, P);
}

__global__ void k_compute_ds_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list, int flexible_only) {
    extern __shared__ double spos[];
    const CaseDev& c = cases[list[blockIdx.y]];
    if ((int)blockIdx.x < c.nbodies) compute_ds_body(c.m, c.g, flexible_only, blockIdx.x, spos);
}

// A_ij of ObjectsClass::computeEpsilon (Objects.cpp:271-300) for the system
// of markers [first, first+n): sum over the supports s of marker i of
// delta_i(s) * delta(x_j - s), times ds_j.
// The sum over the supports of marker mi of delta_s * d(xj - i_s) d(yj - j_s)
// in support order (an epsilon matrix entry before the spacing factor,
// Objects.cpp:271-300).  find_support keeps the sites with |x - i| < 1.5 and
// |y - j| < 1.5 inside the lattice, scanned i-major in ascending order: a
// block of at most three consecutive columns by three consecutive rows, the
// first support being the lowest of each.  So the kernel values of the x
// and y distances are evaluated once per column and per row (six
// evaluations instead of eighteen) and picked per support: the same values,
// products and sum as evaluating them support by support.
__device__ __forceinline__ double support_sum(const MarkerDev& m, int mi, double xj, double yj) {
    const int cnt = m.supp_cnt[mi];
    int sij[2 * kSupp];
    double sd[kSupp];
#pragma unroll
    for (int s = 0; s < 2 * kSupp; s++) sij[s] = m.supp_ij[mi * kSupp * 2 + s];
#pragma unroll
    for (int s = 0; s < kSupp; s++) sd[s] = m.supp_delta[mi * kSupp + s];
    const int i0 = sij[0], j0 = sij[1];
    const double x0 = dirac64(fabs(xj - i0)), x1 = dirac64(fabs(xj - (i0 + 1))), x2 = dirac64(fabs(xj - (i0 + 2)));
    const double y0 = dirac64(fabs(yj - j0)), y1 = dirac64(fabs(yj - (j0 + 1))), y2 = dirac64(fabs(yj - (j0 + 2)));
    double a = 0.0;
#pragma unroll
    for (int s = 0; s < kSupp; s++) {
        if (s < cnt) {
            const int kx = sij[2 * s] - i0, ky = sij[2 * s + 1] - j0;
            const double dx = kx == 0 ? x0 : (kx == 1 ? x1 : x2);
            const double dy = ky == 0 ? y0 : (ky == 1 ? y1 : y2);
            a += sd[s] * (dx * dy);
        }
    }
    return a;
}

__global__ void k_epsilon_matrix(MarkerDev m, GridScalars g, int first, int n, double* __restrict__ A) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * n) return;
    const int r = (int)(t / n), c = (int)(t - (long long)r * n);
    const int mi = first + r, mj = first + c;
    const double Dx = g.Dx;
    const double xj = m.pos[2 * mj] / Dx, yj = m.pos[2 * mj + 1] / Dx;
    A[t] = support_sum(m, mi, xj, yj) * m.ds[mj];
}

// ---------------------------------------------------------------- epsilon, banded
// Batched form of computeEpsilon for systems whose matrix is banded (every
// filament: A_ij = 0 unless markers i and j are within 3 lattice spacings).
// Band storage: row r holds columns r-KB .. r+2KB (width 3KB+1, the upper
// part leaves room for the fill-in of partial pivoting).  A system with a
// coupling pair outside the band is flagged and solved on the host with the
// dense path instead.

// One thread per band entry; grid.y = system.
__device__ inline void eps_band_entry(const MarkerDev& m, const GridScalars& g, int first, int n, int KB,
                                      double* __restrict__ band, long long t) {
    const int W = 3 * KB + 1;
    if (t >= (long long)n * W) return;
    const int r = (int)(t / W), q = (int)(t - (long long)r * W);
    const int c = r - KB + q;
    double a = 0.0;
    if (q <= 2 * KB && c >= 0 && c < n) {
        const int mi = first + r, mj = first + c;
        const double Dx = g.Dx;
        // every read issued before the arithmetic: the support slots past
        // the count hold zeros and are read but not used
        const double xj = m.pos[2 * mj] / Dx, yj = m.pos[2 * mj + 1] / Dx, dsj = m.ds[mj];
        a = support_sum(m, mi, xj, yj) * dsj;
    }
    band[t] = a;
}

// Pairs outside the band that could couple (support of i within 1.5 of
// x_i, delta of j within 1.5 of the support: |x_i - x_j| < 3 spacings).
// One thread per pair r < c - KB (the test is symmetric); grid.y = system.
__device__ inline void eps_far_pair(const MarkerDev& m, const GridScalars& g, int first, int n, int KB,
                                    int* __restrict__ outside, long long t) {
    if (t >= (long long)n * n) return;
    const int r = (int)(t / n), c = (int)(t - (long long)r * n);
    if (c <= r + KB) return;
    const double lim = 3.01 * g.Dx;
    const int mi = first + r, mj = first + c;
    if (fabs(m.pos[2 * mi] - m.pos[2 * mj]) < lim && fabs(m.pos[2 * mi + 1] - m.pos[2 * mj + 1]) < lim)
        *outside = 1;
}

// The out-of-band check of one system (the pairs of eps_far_pair) by one
// workgroup, which writes the system's flag (1: a pair outside the band
// couples, 0: none).  Positions are staged in LDS (P: 2n doubles, or null to
// read them from device memory).  Eight lanes share a row r and take every
// eighth column beyond the band, four columns at a time; rows are dealt out
// in steps of blockDim / 8, so every lane sees about the same number of
// pairs.
__device__ inline void eps_far_rows(const MarkerDev& m, const GridScalars& g, int first, int n, int KB,
                                    int* __restrict__ outside, double* __restrict__ P) {
    __shared__ int far_sh;
    const double* X = m.pos + 2 * first;
    if (threadIdx.x == 0) far_sh = 0;
    if (P)
        for (int k = threadIdx.x; k < 2 * n; k += blockDim.x) P[k] = X[k];
    __syncthreads();
    if (P) X = P;
    const double lim = 3.01 * g.Dx;
    constexpr int CL = 8, U = 4;
    const int rl = threadIdx.x / CL, cl = threadIdx.x % CL, nrl = blockDim.x / CL;
    int far = 0;
    for (int r = rl; r < n; r += nrl) {
        const double xr = X[2 * r], yr = X[2 * r + 1];
        for (int c0 = r + KB + 1 + cl; c0 < n; c0 += U * CL) {
            double xs[U], ys[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int c = min(c0 + u * CL, n - 1);
                xs[u] = X[2 * c];
                ys[u] = X[2 * c + 1];
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (c0 + u * CL < n && fabs(xr - xs[u]) < lim && fabs(yr - ys[u]) < lim) far = 1;
        }
    }
    if (far) atomicOr(&far_sh, 1);
    __syncthreads();
    if (threadIdx.x == 0) *outside = far_sh;
}

// Launches whose systems have up to this many markers stage the positions
// for the out-of-band check in LDS.
constexpr int kFarLdsMarkers = 1024;

// Band entries (the first be workgroups of grid.x) and the out-of-band
// check (the last workgroup, which also writes the flag) of each system in
// one launch: the two are independent and the solve needs both, so the
// coupling chain waits for one kernel.
__global__ void k_eps_band_prep(MarkerDev m, GridScalars g, const int* __restrict__ sys_first,
                                const int* __restrict__ sys_n, int KB, double* __restrict__ band,
                                const long long* __restrict__ band_off, int* __restrict__ outside, int be,
                                int stage_far) {
    extern __shared__ double P[];
    coupling_priority();
    const int sys = blockIdx.y;
    const int bx = (int)blockIdx.x;
    if (bx < be)
        eps_band_entry(m, g, sys_first[sys], sys_n[sys], KB, band + band_off[sys],
                       (long long)bx * blockDim.x + threadIdx.x);
    else
        eps_far_rows(m, g, sys_first[sys], sys_n[sys], KB, outside + sys, stage_far ? P : nullptr);
}

// Batched assembly: grid.y = entry of the system table.
__global__ void k_eps_band_entries_batch(const CaseDev* __restrict__ cases, const CaseSys* __restrict__ sys, int KB) {
    const CaseSys q = sys[blockIdx.y];
    const CaseDev& c = cases[q.c];
    eps_band_entry(c.m, c.g, c.eps.first[q.k], c.eps.n[q.k], KB, c.eps.band + c.eps.off[q.k],
                   (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_eps_far_check_batch(const CaseDev* __restrict__ cases, const CaseSys* __restrict__ sys, int KB) {
    const CaseSys q = sys[blockIdx.y];
    const CaseDev& c = cases[q.c];
    eps_far_pair(c.m, c.g, c.eps.first[q.k], c.eps.n[q.k], KB, c.eps.outside + q.k,
                 (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Canonical elimination (partial pivoting, first maximum) of one banded
// system per wavefront, right-hand side of ones, writing epsilon.
//
// With lower and upper bandwidth KB, every non-zero that step k of the
// dense elimination reads or writes lies in rows k..k+KB and columns
// k..k+2KB (row swaps widen the upper band by KB).  That window stays in
// registers: register slot r holds row k+r, and lane L < C holds the
// window column j with j = L (mod C), so columns never move between lanes.
// Step k:
//  - the lane holding column k finds the pivot (first row of maximum
//    magnitude; rows past n hold zeros) in its own registers, and one
//    v_readlane makes the row index uniform;
//  - rows k and k+p are swapped in every column;
//  - column k goes through LDS to lanes 1..KB, lane r forms the multiplier
//    l_r = a(k+r, k) / a(k, k) (one division for the whole wavefront), and
//    the multipliers come back through LDS;
//  - every other column gets a_ij - l_i a_kj once, from the same operands
//    as the dense algorithm (rows past n get l = 0: no change), and row k
//    of U is emitted.
// The slots then shift by one row (row k+R enters), and column k's lane
// takes column k+C, whose entries in rows k+1..k+KB are structural zeros.
// The skipped dense operations all involve an exact zero, so the values
// equal the dense elimination's.  A single wavefront is latency-bound
// here: every step is one dependent chain (search, division, update).

// Orders LDS writes of some lanes before later reads by other lanes of the
// same wavefront (LDS operations of a wavefront complete in order; the
// fence stops the compiler from reordering them across lanes).
__device__ __forceinline__ void lane_exchange_fence() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Lane `lane`'s value as a wavefront-uniform (SGPR) value.
__device__ inline double uniform64(double v, int lane) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Back substitution on one lane in increasing column order, two rows per
// iteration (no register shuffling between them); rows of U stored with
// half-bandwidth KBS, window half-bandwidth KB.  Terms past n are 0 * 0
// (structural zeros of U times the zero-initialised x window), which
// leave s unchanged, so rows need no bounds tests.
template <int KB, int KBS>
__device__ __forceinline__ void band_back_substitute(int first, int n, const double* __restrict__ B,
                                                     const double* __restrict__ y, double* __restrict__ eps) {
    constexpr int W = 3 * KBS + 1;
    double xa[2 * KB + 1];  // xa[q] = x[i+1+q] for the row pair (i, i-1)
#pragma unroll
    for (int q = 0; q <= 2 * KB; q++) xa[q] = 0.0;
    int i = n - 1;
    for (; i >= 1; i -= 2) {
        const double* U1 = B + i * W + KBS;
        const double* U0 = B + (i - 1) * W + KBS;
        double u1[2 * KB + 1], u0[2 * KB + 1];
#pragma unroll
        for (int q = 0; q <= 2 * KB; q++) {
            u1[q] = U1[q];
            u0[q] = U0[q];
        }
        double s = y[i];
#pragma unroll
        for (int q = 0; q < 2 * KB; q++) s = s - u1[1 + q] * xa[q];
        const double x1 = s / u1[0];
        eps[first + i] = x1;
        double t = y[i - 1];
        t = t - u0[1] * x1;
#pragma unroll
        for (int q = 1; q < 2 * KB; q++) t = t - u0[1 + q] * xa[q - 1];
        const double x0 = t / u0[0];
        eps[first + i - 1] = x0;
#pragma unroll
        for (int q = 2 * KB; q > 1; q--) xa[q] = xa[q - 2];
        xa[1] = x1;
        xa[0] = x0;
    }
    if (i == 0) {
        const double* U = B + KBS;
        double s = y[0];
#pragma unroll
        for (int q = 0; q < 2 * KB; q++) s = s - U[1 + q] * xa[q];
        eps[first] = s / U[0];
    }
}

// B: the band (rows of U replace it), y: the right-hand side; both either
// in LDS or in global memory (LDS_ selects the instantiation, so that every
// access is compiled for its own address space).  xb: 2R doubles of LDS.
// The band is stored with half-bandwidth KBS (W = 3KBS + 1 per row); the
// elimination window has half-bandwidth KB <= KBS, which must bound every
// non-zero of the matrix (|column - row| <= KB).
template <int KB, int KBS, bool LDS_, bool ONE_WAVE = false>
__device__ __forceinline__ void eps_band_solve_impl(int first, int n, double* __restrict__ B, double* __restrict__ y,
                                                    double* __restrict__ eps, double* __restrict__ xb) {
    static_assert(KB <= KBS, "window wider than the stored band");
    constexpr int R = KB + 1, C = 2 * KB + 1, W = 3 * KBS + 1;
    const int L = threadIdx.x;
    // lanes 0..C-1 hold the window columns; lane C holds the right-hand
    // side, which the row operations treat as one more column
    const bool rhs_lane = L == C;
    double w[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int q = L - r + KBS;
        const bool in = L < C && r < n && L < n && q >= 0 && q < W;
        const double a = B[in ? r * W + q : 0];
        w[r] = in ? a : (rhs_lane && r < n ? 1.0 : 0.0);
    }
    int pl = 0;   // lane holding column k
    int off = L;  // this lane holds column k + off
    for (int k = 0; k < n; k++) {
        // row k+R enters this lane's column after the step (band offset
        // q = column - row + KBS, always inside the stored band here); the
        // load is unconditional (clamped address) so that its latency
        // overlaps the step
        const int cn = L == pl ? C : off;
        const bool in = L < C && k + R < n && k + cn < n;
        const double a_in = B[in ? (k + R) * W + (cn - R + KBS) : 0];
        double best = fabs(w[0]);
        int pr = 0;
#pragma unroll
        for (int r = 1; r < R; r++) {
            const double a = fabs(w[r]);
            const bool gt = a > best;
            best = gt ? a : best;
            pr = gt ? r : pr;
        }
        const int p = __builtin_amdgcn_readlane(pr, pl);
        if (__builtin_expect(p != 0, 0)) {
            const double tw = w[0];
#pragma unroll
            for (int r = 1; r < R; r++)
                if (r == p) {
                    w[0] = w[r];
                    w[r] = tw;
                }
        }
        double l[R];
        if constexpr (KB <= 4) {
            // narrow window: column k and the multipliers as wavefront-
            // uniform values (v_readlane), no LDS round trip
            const double piv0 = uniform64(w[0], pl);
            const double piv = piv0 != 0.0 ? piv0 : 1.0;  // zero column: all multipliers 0
            double num = 0.0;
#pragma unroll
            for (int r = 1; r < R; r++) {
                const double a = uniform64(w[r], pl);
                num = L == r ? a : num;
            }
            const double lm = num / piv;
#pragma unroll
            for (int r = 1; r < R; r++) l[r] = uniform64(lm, r);
        } else {
            if (L == pl) {
#pragma unroll
                for (int r = 0; r < R; r++) xb[r] = w[r];
            }
            // other lanes read what lane pl wrote: a wavefront-scope fence
            // keeps the compiler from moving their reads above the write
            lane_exchange_fence();
            const double piv0 = xb[0];
            const double piv = piv0 != 0.0 ? piv0 : 1.0;  // zero column: all multipliers 0
            const double num = (L >= 1 && L < R) ? xb[L] : 0.0;
            if (L >= 1 && L < R) xb[R + L] = num / piv;
            lane_exchange_fence();
#pragma unroll
            for (int r = 1; r < R; r++) l[r] = xb[R + r];
        }
#pragma unroll
        for (int r = 1; r < R; r++) {
            const double upd = w[r] - l[r] * w[0];
            w[r] = L == pl ? 0.0 : upd;
        }
        // row k of U (in place of row k of the band) and the final y_k
        if (L < C && k + off < n) B[k * W + (off + KBS)] = w[0];
        if (rhs_lane) y[k] = w[0];
        // drop row k, take in row k+R; column k's lane takes column k+C
#pragma unroll
        for (int r = 0; r < R - 1; r++) w[r] = w[r + 1];
        w[R - 1] = in ? a_in : (rhs_lane && k + R < n ? 1.0 : 0.0);
        off = L == pl ? C - 1 : off - 1;
        pl = pl == C - 1 ? 0 : pl + 1;
        if constexpr (KB > 4) lane_exchange_fence();  // xb is rewritten next step
        else __builtin_amdgcn_wave_barrier();
    }
    // lane 0 reads the rows of U and y the other lanes wrote
    if (ONE_WAVE) lane_exchange_fence();
    else __syncthreads();
    if (L == 0) band_back_substitute<KB, KBS>(first, n, B, y, eps);
}

// Largest |column - row| over the non-zero entries of rows [r0, r1) of a
// band stored with half-bandwidth KBS (entries q = 0..2KBS of each row).
template <int KBS>
__device__ __forceinline__ int band_width_rows(const double* __restrict__ B, long long t0, long long t1, int stride) {
    constexpr int W = 3 * KBS + 1;
    int bw = 0;
    for (long long t = t0; t < t1; t += stride) {
        const int q = (int)(t % W);
        if (q <= 2 * KBS && B[t] != 0.0) bw = max(bw, q < KBS ? KBS - q : q - KBS);
    }
    return bw;
}

// Elimination with the narrowest window that holds the matrix's non-zeros
// (bandwidth bw): the same operations on the non-zero entries, fewer on
// structural zeros, so the same values.
template <int KBS, bool LDS_, bool ONE_WAVE = false>
__device__ __forceinline__ void eps_band_solve_bw(int bw, int first, int n, double* __restrict__ B,
                                                  double* __restrict__ y, double* __restrict__ eps,
                                                  double* __restrict__ xb) {
    if (bw <= 2) eps_band_solve_impl<2, KBS, LDS_, ONE_WAVE>(first, n, B, y, eps, xb);
    else if (bw == 3) eps_band_solve_impl<3, KBS, LDS_, ONE_WAVE>(first, n, B, y, eps, xb);
    else if (bw == 4) eps_band_solve_impl<4, KBS, LDS_, ONE_WAVE>(first, n, B, y, eps, xb);
    else eps_band_solve_impl<KBS, KBS, LDS_, ONE_WAVE>(first, n, B, y, eps, xb);
}

// One system per 64-lane workgroup: stage the band in LDS when it fits,
// measure its bandwidth on the way, and solve with the matching window.
// bw_sh: one int of LDS.
template <int KB>
__device__ __forceinline__ void eps_band_solve_body(int first, int n, double* __restrict__ G, double* __restrict__ ywork,
                                                    double* __restrict__ eps, int lds_rows, double* sm,
                                                    double* xb, int* bw_sh) {
    constexpr int W = 3 * KB + 1;
    const int nw = n * W;
    if (threadIdx.x == 0) *bw_sh = 0;
    __syncthreads();
    int bw = 0;
    if (n <= lds_rows) {
#pragma unroll 8
        for (int t = threadIdx.x; t < nw; t += 64) {
            const double v = G[t];
            sm[t] = v;
            const int q = t % W;
            if (q <= 2 * KB && v != 0.0) bw = max(bw, q < KB ? KB - q : q - KB);
        }
    } else {
        bw = band_width_rows<KB>(G, threadIdx.x, nw, 64);
    }
    atomicMax(bw_sh, bw);
    __syncthreads();
    bw = *bw_sh;
    if (n <= lds_rows) eps_band_solve_bw<KB, true>(bw, first, n, sm, sm + (long long)nw, eps, xb);
    else eps_band_solve_bw<KB, false>(bw, first, n, G, ywork + first, eps, xb);
}

template <int KB>
__global__ __launch_bounds__(64) void k_eps_band_solve(const int* __restrict__ sys_first, const int* __restrict__ sys_n,
                                                       const int* __restrict__ outside, double* __restrict__ band,
                                                       const long long* __restrict__ band_off,
                                                       double* __restrict__ ywork, double* __restrict__ eps,
                                                       int lds_rows, int skip_upto) {
    extern __shared__ double sm[];
    __shared__ double xb[2 * (KB + 1)];
    __shared__ int bw;
    const int sys = blockIdx.x;
    coupling_priority();
    // skip_upto: systems of up to that many rows are the reduction's
    if (outside[sys] || sys_n[sys] <= skip_upto) return;
    eps_band_solve_body<KB>(sys_first[sys], sys_n[sys], band + band_off[sys], ywork, eps, lds_rows, sm, xb, &bw);
}

template <int KB>
__global__ __launch_bounds__(64) void k_eps_band_solve_batch(const CaseDev* __restrict__ cases,
                                                             const CaseSys* __restrict__ sys, int lds_rows,
                                                             int skip_upto) {
    extern __shared__ double sm[];
    __shared__ double xb[2 * (KB + 1)];
    __shared__ int bw;
    const CaseSys q = sys[blockIdx.x];
    const CaseDev& c = cases[q.c];
    if (c.eps.outside[q.k] || c.eps.n[q.k] <= skip_upto) return;
    eps_band_solve_body<KB>(c.eps.first[q.k], c.eps.n[q.k], c.eps.band + c.eps.off[q.k], c.eps.y, c.m.eps, lds_rows,
                            sm, xb, &bw);
}

// ---------------------------------------------------------------- epsilon by cyclic reduction
// FP32-shifted cases: the banded epsilon system A eps = 1 of a body solved by
// block cyclic reduction instead of the pivoted elimination above.  With
// half-bandwidth bw <= Bk, A is block tridiagonal in Bk x Bk blocks (block
// row I: rows I*Bk .. I*Bk+Bk-1; padding rows beyond n are identity rows
// with a zero right-hand side).  At stride s = 1, 2, 4, ... the blocks at
// odd multiples of s are eliminated from their neighbours' equations, one
// lane per entry of each surviving block, and back substitution runs the
// levels in reverse: log2(n / Bk) dependent levels instead of n pivot
// steps.  A_ij = S_ij ds_j with S the Gram matrix of the markers' delta
// functions over the lattice (symmetric positive definite for distinct
// markers), so every Schur complement the reduction forms is S's Schur
// complement scaled by ds: no pivoting is needed.  The values agree with the
// pivoted elimination to rounding (the plain method keeps that one, whose
// operations are the reference's).  Systems of up to kCrMaxRows rows and
// half-bandwidth up to kCrMaxBw (blocks of 2 x 2 or 4 x 4); a wider band is
// solved by the pivoted elimination on wavefront 0 of the same kernel, and
// larger systems by the pivoted kernel.  The markers of a filament lie one
// lattice spacing apart, which bounds the half-bandwidth at 4; a lower limit
// (cr_max_bw) sends narrower bands to the fallback as well.
constexpr int kCrMaxRows = 384;
constexpr int kCrMaxBw = 4;

// Widest half-bandwidth the reduction takes: environment LIFE_EPS_CR_MAX_BW
// (0 .. 4, default 4; 0: every system by the pivoted elimination), read at
// each launch.
static int cr_max_bw() {
    const char* e = std::getenv("LIFE_EPS_CR_MAX_BW");
    const int v = e && *e ? std::atoi(e) : kCrMaxBw;
    return v < 0 ? 0 : (v > kCrMaxBw ? kCrMaxBw : v);
}

// LDS (doubles) of the reduction with Bk x Bk blocks over n rows: blocks,
// staged new blocks of a level, solution.
__host__ __device__ constexpr size_t cr_lds_doubles(int Bk, int n) {
    return (size_t)((n + Bk - 1) / Bk) * (3 * Bk * Bk + Bk) + (size_t)(((n + Bk - 1) / Bk + 1) / 2) * (3 * Bk * Bk + Bk) +
           (size_t)((n + Bk - 1) / Bk) * Bk;
}

// In-place inverse of a Bk x Bk block (row-major) by Gauss-Jordan
// elimination without pivoting, in registers.
template <int Bk>
__device__ __forceinline__ void cr_inverse(double* __restrict__ D) {
    double a[Bk][Bk];
#pragma unroll
    for (int i = 0; i < Bk; i++)
#pragma unroll
        for (int j = 0; j < Bk; j++) a[i][j] = D[Bk * i + j];
#pragma unroll
    for (int k = 0; k < Bk; k++) {
        const double p = 1.0 / a[k][k];
        a[k][k] = 1.0;
#pragma unroll
        for (int j = 0; j < Bk; j++) a[k][j] = a[k][j] * p;
#pragma unroll
        for (int i = 0; i < Bk; i++) {
            if (i == k) continue;
            const double f = a[i][k];
            a[i][k] = 0.0;
#pragma unroll
            for (int j = 0; j < Bk; j++) a[i][j] = a[i][j] - f * a[k][j];
        }
    }
#pragma unroll
    for (int i = 0; i < Bk; i++)
#pragma unroll
        for (int j = 0; j < Bk; j++) D[Bk * i + j] = a[i][j];
}

// The reduction of one system by the whole workgroup.  G: the band (KB = 8
// storage, LDS or global memory), read once into the blocks; blk: LDS for
// the blocks; work: LDS for the staged blocks and the solution (it may
// alias G when G is in LDS: G is read only before the first level).
template <int Bk>
__device__ __forceinline__ void eps_cr_solve(const double* __restrict__ G, int n, double* __restrict__ eps,
                                             double* __restrict__ blk, double* __restrict__ work) {
    constexpr int KB = 8, W = 3 * KB + 1;
    constexpr int BB = Bk * Bk, BLK = 3 * BB + Bk;  // A (sub), D (diagonal), C (super), f
    constexpr int oA = 0, oD = BB, oC = 2 * BB, oF = 3 * BB;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nb = (n + Bk - 1) / Bk;
    // blocks from the band: eight entries per lane at a time, their loads
    // (unconditional, clamped address) issued before any is used
    constexpr int U = 8;
    const int total = nb * BLK;
    for (int base = 0; base < total; base += U * nt) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = base + u * nt + tid;
            const int I = t / BLK, q = t - I * BLK;
            const int part = q / BB, i = (q % BB) / Bk, j = q % Bk;
            const int r = I * Bk + i, c = (I + part - 1) * Bk + j, o = c - r;
            const bool band_entry = q < oF && r < n && c >= 0 && c < n && o >= -KB && o <= KB;
            const double a = G[band_entry ? (size_t)r * W + (o + KB) : 0];
            double val;
            if (q >= oF) val = I * Bk + (q - oF) < n ? 1.0 : 0.0;
            else if (r >= n) val = (part == 1 && i == j) ? 1.0 : 0.0;
            else val = band_entry ? a : 0.0;
            v[u] = val;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = base + u * nt + tid;
            if (t < total) blk[t] = v[u];
        }
    }
    __syncthreads();
    double* stage = work;
    double* x = work + (size_t)((nb + 1) / 2) * BLK;
    int s = 1;
    for (; s < nb; s *= 2) {
        for (int j = s + 2 * s * tid; j < nb; j += 2 * s * nt) cr_inverse<Bk>(blk + (size_t)BLK * j + oD);
        __syncthreads();
        const int nr = (nb - 1) / (2 * s) + 1;
        for (int t = tid; t < BB * nr; t += nt) {
            const int k = t / BB, r = (t % BB) / Bk, c = t % Bk, i = 2 * s * k;
            const double* bi = blk + (size_t)BLK * i;
            double Dn = bi[oD + Bk * r + c], An = 0.0, Cn = 0.0, fn = bi[oF + r];
            if (i - s >= 0) {
                const double* bj = blk + (size_t)BLK * (i - s);
                double al[Bk];
#pragma unroll
                for (int q = 0; q < Bk; q++) {
                    double a = 0.0;
#pragma unroll
                    for (int m = 0; m < Bk; m++) a += bi[oA + Bk * r + m] * bj[oD + Bk * m + q];
                    al[q] = a;
                }
                double dc = 0.0, da = 0.0, df = 0.0;
#pragma unroll
                for (int q = 0; q < Bk; q++) {
                    dc += al[q] * bj[oC + Bk * q + c];
                    da += al[q] * bj[oA + Bk * q + c];
                    df += al[q] * bj[oF + q];
                }
                Dn = Dn - dc;
                An = -da;
                fn = fn - df;
            }
            if (i + s < nb) {
                const double* bj = blk + (size_t)BLK * (i + s);
                double ga[Bk];
#pragma unroll
                for (int q = 0; q < Bk; q++) {
                    double a = 0.0;
#pragma unroll
                    for (int m = 0; m < Bk; m++) a += bi[oC + Bk * r + m] * bj[oD + Bk * m + q];
                    ga[q] = a;
                }
                double da = 0.0, dc = 0.0, df = 0.0;
#pragma unroll
                for (int q = 0; q < Bk; q++) {
                    da += ga[q] * bj[oA + Bk * q + c];
                    dc += ga[q] * bj[oC + Bk * q + c];
                    df += ga[q] * bj[oF + q];
                }
                Dn = Dn - da;
                Cn = -dc;
                fn = fn - df;
            }
            double* o = stage + (size_t)BLK * k;
            o[oA + Bk * r + c] = An;
            o[oD + Bk * r + c] = Dn;
            o[oC + Bk * r + c] = Cn;
            if (c == 0) o[oF + r] = fn;
        }
        __syncthreads();
        for (int t = tid; t < BLK * nr; t += nt) blk[(size_t)BLK * 2 * s * (t / BLK) + t % BLK] = stage[t];
        __syncthreads();
    }
    // block 0 alone: x_0 = inv(D_0) f_0
    if (tid == 0) cr_inverse<Bk>(blk + oD);
    __syncthreads();
    if (tid < Bk) {
        double a = 0.0;
#pragma unroll
        for (int m = 0; m < Bk; m++) a += blk[oD + Bk * tid + m] * blk[oF + m];
        x[tid] = a;
    }
    __syncthreads();
    // back substitution: x_j = inv(D_j) (f_j - A_j x_{j-s} - C_j x_{j+s})
    for (s /= 2; s >= 1; s /= 2) {
        const int ne = (nb + s - 1) / (2 * s);
        for (int t = tid; t < Bk * ne; t += nt) {
            const int j = s + 2 * s * (t / Bk), r = t % Bk;
            const double* bj = blk + (size_t)BLK * j;
            const double* xl = x + (size_t)Bk * (j - s);
            const bool right = j + s < nb;
            const double* xr = x + (size_t)Bk * (right ? j + s : j - s);
            double v[Bk];
#pragma unroll
            for (int a = 0; a < Bk; a++) {
                double ax = 0.0, cx = 0.0;
#pragma unroll
                for (int m = 0; m < Bk; m++) {
                    ax += bj[oA + Bk * a + m] * xl[m];
                    cx += bj[oC + Bk * a + m] * xr[m];
                }
                v[a] = bj[oF + a] - ax - (right ? cx : 0.0);
            }
            double xv = 0.0;
#pragma unroll
            for (int m = 0; m < Bk; m++) xv += bj[oD + Bk * r + m] * v[m];
            x[(size_t)Bk * j + r] = xv;
        }
        __syncthreads();
    }
    for (int k = tid; k < n; k += nt) eps[k] = x[k];
}

// Block size for a measured half-bandwidth (at most kCrMaxBw).
__device__ __forceinline__ void eps_cr_dispatch(int bw, const double* G, int n, double* eps, double* blk,
                                                double* work) {
    if (bw <= 2) eps_cr_solve<2>(G, n, eps, blk, work);
    else eps_cr_solve<4>(G, n, eps, blk, work);
}

// Half-bandwidth of a system's band in global memory (largest |column - row|
// over its non-zero entries), reduced over the workgroup.
__device__ __forceinline__ int band_width_wg(const double* __restrict__ G, int n, int* bw_sh) {
    constexpr int KB = 8, W = 3 * KB + 1;
    if (threadIdx.x == 0) *bw_sh = 0;
    __syncthreads();
    // sixteen loads in flight per lane (clamped addresses)
    constexpr int U = 16;
    const int total = n * W, nt = blockDim.x;
    int bw = 0;
    for (int base = 0; base < total; base += U * nt) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = G[min(base + u * nt + (int)threadIdx.x, total - 1)];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int t = base + u * nt + (int)threadIdx.x;
            const int q = t % W;
            if (t < total && q <= 2 * KB && v[u] != 0.0) bw = max(bw, q < KB ? KB - q : q - KB);
        }
    }
    atomicMax(bw_sh, bw);
    __syncthreads();
    return *bw_sh;
}

// LDS of the reduction kernels for systems of up to max_n rows (blocks of
// the widest size, 4).
static size_t eps_cr_lds(int max_n) {
    const int n = max_n < kCrMaxRows ? max_n : kCrMaxRows;
    return sizeof(double) * cr_lds_doubles(kCrMaxBw, n);
}

// Dynamic LDS beyond 64 KiB per workgroup (gfx950 has 160 KiB) for a kernel,
// allowed once per process.
static void allow_lds(const void* kernel, size_t bytes) {
    if (bytes > 64 * 1024 &&
        hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
        throw std::runtime_error("hipFuncSetAttribute: dynamic LDS of the epsilon solve");
}

// One system per workgroup: the reduction, or for a band wider than
// kCrMaxBw the pivoted elimination on the band in device memory (ywork: its
// right-hand side).  That elimination's lanes are those of wavefront 0; the
// other wavefronts run along without storing anything and meet it at the
// workgroup barrier before the back substitution.
__device__ __forceinline__ void eps_cr_or_pivot(double* __restrict__ G, int first, int n, double* __restrict__ ywork,
                                                double* __restrict__ eps, double* sm, double* xb, int* bw_sh,
                                                int maxbw) {
    const int bw = band_width_wg(G, n, bw_sh);
    if (bw > maxbw) {
        eps_band_solve_bw<8, false>(bw, first, n, G, ywork + first, eps, xb);
        return;
    }
    const int nb = (n + kCrMaxBw - 1) / kCrMaxBw;
    eps_cr_dispatch(bw, G, n, eps + first, sm, sm + (size_t)nb * (3 * kCrMaxBw * kCrMaxBw + kCrMaxBw));
}

__global__ __launch_bounds__(256) void k_eps_band_cr(const int* __restrict__ sys_first, const int* __restrict__ sys_n,
                                                     const int* __restrict__ outside, double* __restrict__ band,
                                                     const long long* __restrict__ band_off,
                                                     double* __restrict__ ywork, double* __restrict__ eps, int maxbw) {
    extern __shared__ double sm[];
    __shared__ double xb[2 * (8 + 1)];
    __shared__ int bw_sh;
    const int sys = blockIdx.x, n = sys_n[sys];
    coupling_priority();
    if (outside[sys] || n > kCrMaxRows) return;
    eps_cr_or_pivot(band + band_off[sys], sys_first[sys], n, ywork, eps, sm, xb, &bw_sh, maxbw);
}

__global__ __launch_bounds__(256) void k_eps_band_cr_batch(const CaseDev* __restrict__ cases,
                                                           const CaseSys* __restrict__ sys, int maxbw) {
    extern __shared__ double sm[];
    __shared__ double xb[2 * (8 + 1)];
    __shared__ int bw_sh;
    const CaseSys q = sys[blockIdx.x];
    const CaseDev& c = cases[q.c];
    const int n = c.eps.n[q.k];
    if (c.eps.outside[q.k] || n > kCrMaxRows) return;
    eps_cr_or_pivot(c.eps.band + c.eps.off[q.k], c.eps.first[q.k], n, c.eps.y, c.m.eps, sm, xb, &bw_sh, maxbw);
}

// ---------------------------------------------------------------- fused recompute
// The device half of recomputeObjectVals (Objects.cpp:217-227 and the
// computeEpsilon of flexible bodies, Objects.cpp:235-321) for one flexible
// body per workgroup, without global round trips between the stages:
// supports of its markers (to global memory, as k_find_support), spacing
// from positions staged in LDS (as k_compute_ds), the band of its epsilon
// system built in LDS (as eps_band_entry), the out-of-band check (as
// eps_far_pair) and, if the system is banded, the solve by wavefront 0
// (as k_eps_band_solve).  Same operations in the same order as those
// kernels.  LDS: n * (3KB + 5) doubles; bodies of up to kFusedMaxMarkers
// markers.  Used by the serving engine.  A large workgroup needs that many
// of a CU's wave slots free at once, so beside a concurrent bulk
// stream-collide (the single-case step schedule, and the serving engine's
// early bulk kernel) it waits for CUs to drain; the serving engine launches
// it with kFusedThreads lanes, and the single-case step uses the separate
// small-block kernels.
constexpr int kFusedMaxMarkers = 256;

template <int KB>
__device__ __forceinline__ void recompute_body(const MarkerDev& m, const GridScalars& g, int first, int n,
                                               int* __restrict__ outside, double* __restrict__ sm,
                                               double* __restrict__ xb, int* __restrict__ flag,
                                               int* __restrict__ bw_sh, int cr_bw) {
    constexpr int W = 3 * KB + 1;
    double* P = sm;                // 2n positions
    double* dsl = sm + 2 * n;      // n spacings
    double* Bl = dsl + n;          // n * W band
    double* yl = Bl + n * W;       // n right-hand side
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) {
        *flag = 0;
        *bw_sh = 0;
    }
    for (int k = tid; k < n; k += nt) {
        // the position and the body kind are read side by side
        const double x = m.pos[2 * (first + k)], y = m.pos[2 * (first + k) + 1];
        if (m.body_flex[m.body[first + k]] == 0) find_support_xy(m, g, first + k, x, y);
        P[2 * k] = x;
        P[2 * k + 1] = y;
    }
    __syncthreads();
    // computeDs (IBMNode.cpp:182-204), as compute_ds_body
    for (int k = tid; k < n; k += nt) {
        const double x = P[2 * k], y = P[2 * k + 1];
        double d2min = INFINITY;
        for (int q = 0; q < n; q++) {
            if (q == k) continue;
            const double dx = x - P[2 * q], dy = y - P[2 * q + 1];
            const double d2 = 0.0 + dx * dx + dy * dy;
            if (d2 < d2min) d2min = d2;
        }
        const double mag = sqrt(d2min) / g.Dx;
        const double ds = mag < 10.0 ? mag : 10.0;
        m.ds[first + k] = ds;
        dsl[k] = ds;
    }
    __syncthreads();
    // band entries (as eps_band_entry, reading the staged positions and ds)
    const double Dx = g.Dx;
    int bw = 0;
    for (int t = tid; t < n * W; t += nt) {
        const int r = t / W, q = t - r * W;
        const int c = r - KB + q;
        double a = 0.0;
        if (q <= 2 * KB && c >= 0 && c < n) {
            const int mi = first + r;
            const double xj = P[2 * c] / Dx, yj = P[2 * c + 1] / Dx;
            a = support_sum(m, mi, xj, yj) * dsl[c];
        }
        Bl[t] = a;
        if (a != 0.0) bw = max(bw, q < KB ? KB - q : q - KB);
    }
    atomicMax(bw_sh, bw);
    // pairs outside the band that could couple (as eps_far_pair)
    const double lim = 3.01 * g.Dx;
    for (int t = tid; t < n * n; t += nt) {
        const int r = t / n, c = t - r * n;
        if (c <= r + KB) continue;
        if (fabs(P[2 * r] - P[2 * c]) < lim && fabs(P[2 * r + 1] - P[2 * c + 1]) < lim) *flag = 1;
    }
    __syncthreads();
    const int far = *flag;
    if (tid == 0) *outside = far;
    if (far) return;
    if (*bw_sh <= cr_bw) {
        // the reduction by the whole workgroup, its blocks after the band
        double* blk = yl + n;
        const int nb = (n + kCrMaxBw - 1) / kCrMaxBw;
        eps_cr_dispatch(*bw_sh, Bl, n, m.eps + first, blk,
                        blk + (size_t)nb * (3 * kCrMaxBw * kCrMaxBw + kCrMaxBw));
        return;
    }
    if (tid >= 64) return;
    eps_band_solve_bw<KB, true, true>(*bw_sh, first, n, Bl, yl, m.eps, xb);
}

// Lanes per body workgroup (at least one wavefront for the solve).  Beside
// the early bulk kernel, 256, 512 and 1024 measured within 3% of each other
// on MI355X and 128 8% slower (environment LIFE_SERVE_FUSED_THREADS = 64 |
// 128 | 256 | 512 for sweeps, 0 for the separate batched kernels).
constexpr int kFusedThreads = 512;

template <int KB>
__global__ __launch_bounds__(512) void k_recompute_body_batch(const CaseDev* __restrict__ cases,
                                                              const CaseSys* __restrict__ sys, int cr_bw) {
    extern __shared__ double sm[];
    __shared__ double xb[2 * (KB + 1)];
    __shared__ int flag, bw;
    const CaseSys q = sys[blockIdx.x];
    const CaseDev& c = cases[q.c];
    recompute_body<KB>(c.m, c.g, c.eps.first[q.k], c.eps.n[q.k], c.eps.outside + q.k, sm, xb, &flag, &bw, cr_bw);
}

// ObjectsClass::ibmKernelInterp body (IBMNode.cpp:26-58): rho and rho*u at
// the supports (GridClass::macroscopic values: force_xy only) interpolated
// with the support deltas, then F = 2 (Dt/Dx rho_I U - (rho u)_I).  One
// lane per (marker, support slot): each lane takes its support's moments
// and weighted terms, and the marker's first lane adds them up in support
// order, as the reference's loop does, so the nine gathers of nine
// populations are in flight at once.  kInterpMarkers markers per workgroup
// of kSupp lanes each.
constexpr int kInterpMarkers = 28;
constexpr int kInterpThreads = kInterpMarkers * kSupp;  // 252

template <typename T>
__device__ inline void interpolate_slots(const T* __restrict__ f, const MarkerDev& m, const GridScalars& g,
                                         int k0, double* __restrict__ sh, int t) {
    const int km = t / kSupp, s = t - km * kSupp;
    const int k = k0 + km;
    double* a = sh;
    double* bx = sh + kInterpThreads;
    double* by = sh + 2 * kInterpThreads;
    if (t < kInterpThreads && k < m.n && s < m.supp_cnt[k]) {
        const long long id = m.supp_site[k * kSupp + s];
        double fs[kNV], rho, ux, uy;
        load_site(f, g.S, id, fs);
        moments64(fs, g.fcur ? g.fcur[id] : g.Fx, g.fcur ? g.fcur[g.S + id] : g.Fy, rho, ux, uy);
        const double d = m.supp_delta[k * kSupp + s];
        a[t] = rho * d * 1.0 * 1.0;
        bx[t] = rho * ux * d * 1.0 * 1.0;
        by[t] = rho * uy * d * 1.0 * 1.0;
    }
    __syncthreads();
    if (t < kInterpThreads && s == 0 && k < m.n) {
        double iR = 0.0, mx = 0.0, my = 0.0;
        const int cnt = m.supp_cnt[k];
        for (int q = 0; q < cnt; q++) {
            iR += a[t + q];
            mx += bx[t + q];
            my += by[t + q];
        }
        const double vs = g.velScale;
        m.force[2 * k] = 2.0 * (vs * iR * m.vel[2 * k] - mx);
        m.force[2 * k + 1] = 2.0 * (vs * iR * m.vel[2 * k + 1] - my);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_interpolate(const T* __restrict__ f, MarkerDev m, GridScalars g) {
    __shared__ double sh[3 * kInterpThreads];
    coupling_priority();
    interpolate_slots(f, m, g, blockIdx.x * kInterpMarkers, sh, (int)threadIdx.x);
}

// Batched: the populations after this step's stream-collide.
template <typename T>
__global__ __launch_bounds__(256) void k_interpolate_batch(const CaseDev* __restrict__ cases,
                                                           const int* __restrict__ list) {
    __shared__ double sh[3 * kInterpThreads];
    const CaseDev& c = cases[list[blockIdx.y]];
    interpolate_slots(static_cast<const T*>(c.dst), c.m, c.g, blockIdx.x * kInterpMarkers, sh, (int)threadIdx.x);
}

// ---------------------------------------------------------------- spread
// The reference adds marker contributions to each site in marker order
// (ORDERED, IBMNode.cpp:67-83).  Here every (marker, support) pair takes a
// slot at its site, each touched site gets a compact index, the pairs are
// scattered into per-site lists, and one thread per site sorts its list by
// marker and sums it from zero: the same additions in the same order.

__device__ inline void spread_clear_one(double* __restrict__ fibm, long long S, const int* __restrict__ uniq,
                                        const int* __restrict__ uniq_n, int k) {
    if (k >= *uniq_n) return;
    const long long id = uniq[k];
    fibm[id] = 0.0;
    fibm[S + id] = 0.0;
}

__device__ inline void spread_count_one(const MarkerDev& m, const SpreadScratch& sc, int p) {
    if (p >= m.n * kSupp) return;
    const int k = p / kSupp, s = p - k * kSupp;
    if (s >= m.supp_cnt[k]) return;
    const int site = m.supp_site[p];
    const int slot = atomicAdd(&sc.site_count[site], 1);
    sc.pair_slot[p] = slot;
    if (slot == 0) {
        const int u = atomicAdd(sc.uniq_n, 1);
        sc.uniq[u] = site;
        sc.site_uid[site] = u;
    }
}

__device__ inline void spread_scatter_one(const MarkerDev& m, const SpreadScratch& sc, int p) {
    if (p >= m.n * kSupp) return;
    const int k = p / kSupp, s = p - k * kSupp;
    if (s >= m.supp_cnt[k]) return;
    const int site = m.supp_site[p];
    const int slot = sc.pair_slot[p];
    if (slot >= kMaxPerSite) {
        *sc.overflow = 1;
        return;
    }
    sc.list[(long long)sc.site_uid[site] * kMaxPerSite + slot] = p;
}

__device__ inline void spread_reduce_one(const MarkerDev& m, const SpreadScratch& sc, double* __restrict__ fibm,
                                         long long S, int u) {
    if (u >= *sc.uniq_n) return;
    const int site = sc.uniq[u];
    int n = sc.site_count[site];
    if (n > kMaxPerSite) n = kMaxPerSite;
    int pairs[kMaxPerSite];
    for (int q = 0; q < n; q++) {
        const int p = sc.list[(long long)u * kMaxPerSite + q];
        int r = q - 1;
        while (r >= 0 && pairs[r] > p) {
            pairs[r + 1] = pairs[r];
            r--;
        }
        pairs[r + 1] = p;
    }
    double fx = 0.0, fy = 0.0;
    for (int q = 0; q < n; q++) {
        const int p = pairs[q];
        const int k = p / kSupp;
        const double d = m.supp_delta[p];
        const double e = m.eps[k], ds = m.ds[k];
        fx += m.force[2 * k] * e * ds * 1.0 * d;
        fy += m.force[2 * k + 1] * e * ds * 1.0 * d;
    }
    fibm[site] = fx;
    fibm[S + site] = fy;
    sc.site_count[site] = 0;
}

__global__ void k_spread_clear(double* __restrict__ fibm, long long S, const int* __restrict__ uniq,
                               const int* __restrict__ uniq_n) {
    spread_clear_one(fibm, S, uniq, uniq_n, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_spread_reset(int* uniq_n) { *uniq_n = 0; }

__global__ void k_spread_count(MarkerDev m, SpreadScratch sc) {
    coupling_priority();
    spread_count_one(m, sc, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_spread_scatter(MarkerDev m, SpreadScratch sc) {
    coupling_priority();
    spread_scatter_one(m, sc, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_spread_reduce(MarkerDev m, SpreadScratch sc, double* __restrict__ fibm, long long S) {
    coupling_priority();
    spread_reduce_one(m, sc, fibm, S, blockIdx.x * blockDim.x + threadIdx.x);
}

// The whole spread of a case with few marker-support pairs in one
// workgroup: the clear, reset, count, scatter and reduce passes of the
// separate launches in turn, each the same per-pair or per-site code,
// separated by workgroup barriers (which make one pass's writes to the
// scratch in device memory visible to the next).  Same pairs, same order at
// each site, same sums; one launch instead of five, which is most of the
// spread's time for a case of a few hundred markers.
constexpr int kSpreadOneBlock = 1024;
constexpr int kSpreadOneMaxPairs = 8192;

__device__ inline void spread_passes_wg(const MarkerDev& m, const SpreadScratch& sc, double* __restrict__ fibm,
                                        long long S, int clear) {
    const int t = threadIdx.x, nt = blockDim.x;
    const int np = m.n * kSupp;
    if (clear)
        for (int k = t; k < sc.cap; k += nt) spread_clear_one(fibm, S, sc.uniq, sc.uniq_n, k);
    __syncthreads();
    if (t == 0) *sc.uniq_n = 0;
    __syncthreads();
    for (int p = t; p < np; p += nt) spread_count_one(m, sc, p);
    __syncthreads();
    for (int p = t; p < np; p += nt) spread_scatter_one(m, sc, p);
    __syncthreads();
    for (int u = t; u < sc.cap; u += nt) spread_reduce_one(m, sc, fibm, S, u);
}

__global__ __launch_bounds__(kSpreadOneBlock) void k_spread_one(MarkerDev m, SpreadScratch sc,
                                                                double* __restrict__ fibm, long long S, int clear) {
    coupling_priority();
    spread_passes_wg(m, sc, fibm, S, clear);
}

// Rigid bodies of a small case, where the spread follows the interpolation
// with nothing between them: both in one workgroup.  Four groups of 256
// lanes take kInterpMarkers markers each per round with the interpolation's
// own per-marker code (kInterpThreads of each group's lanes); a barrier;
// then the spread passes as in k_spread_one.
template <typename T>
__global__ __launch_bounds__(kSpreadOneBlock) void k_interp_spread_one(const T* __restrict__ f, MarkerDev m,
                                                                       GridScalars g, SpreadScratch sc,
                                                                       double* __restrict__ fibm, int clear) {
    __shared__ double sh[kSpreadOneBlock / 256][3 * kInterpThreads];
    coupling_priority();
    constexpr int kGroups = kSpreadOneBlock / 256;
    const int grp = (int)threadIdx.x / 256, tl = (int)threadIdx.x % 256;
    for (int k0 = 0; k0 < m.n; k0 += kGroups * kInterpMarkers)
        interpolate_slots(f, m, g, k0 + grp * kInterpMarkers, sh[grp], tl);
    __syncthreads();
    spread_passes_wg(m, sc, fibm, g.S, clear);
}

// Batched spread, grid.y = case (the reset: one thread per case).
__global__ void k_spread_clear_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list) {
    const CaseDev& c = cases[list[blockIdx.y]];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < c.sc.cap; k += gridDim.x * blockDim.x)
        spread_clear_one(c.fibm, c.g.S, c.sc.uniq, c.sc.uniq_n, k);
}

__global__ void k_spread_reset_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) *cases[list[q]].sc.uniq_n = 0;
}

__global__ void k_spread_count_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list) {
    const CaseDev& c = cases[list[blockIdx.y]];
    spread_count_one(c.m, c.sc, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_spread_scatter_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list) {
    const CaseDev& c = cases[list[blockIdx.y]];
    spread_scatter_one(c.m, c.sc, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_spread_reduce_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list) {
    const CaseDev& c = cases[list[blockIdx.y]];
    spread_reduce_one(c.m, c.sc, c.fibm, c.g.S, blockIdx.x * blockDim.x + threadIdx.x);
}

// Packed marker staging <-> per-case marker arrays.
__global__ void k_markers_in_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list,
                                   const double* __restrict__ stage) {
    const CaseDev& c = cases[list[blockIdx.y]];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= c.m.n) return;
    const double* q = stage + 4 * (long long)(c.moff + k);
    c.m.pos[2 * k] = q[0];
    c.m.pos[2 * k + 1] = q[1];
    c.m.vel[2 * k] = q[2];
    c.m.vel[2 * k + 1] = q[3];
}

__global__ void k_markers_out_batch(const CaseDev* __restrict__ cases, const int* __restrict__ list,
                                    double* __restrict__ stage) {
    const CaseDev& c = cases[list[blockIdx.y]];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= c.m.n) return;
    double* q = stage + 3 * (long long)(c.moff + k);
    q[0] = c.m.force[2 * k];
    q[1] = c.m.force[2 * k + 1];
    q[2] = c.m.eps[k];
}

__global__ void k_eps_outside_clear_batch(const CaseDev* __restrict__ cases, const CaseSys* __restrict__ sys,
                                          int nsys) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nsys) cases[sys[q].c].eps.outside[sys[q].k] = 0;
}

__global__ void k_eps_outside_gather_batch(const CaseDev* __restrict__ cases, const CaseSys* __restrict__ sys,
                                           int nsys, int* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nsys) out[q] = cases[sys[q].c].eps.outside[sys[q].k];
}

static int eps_lds_rows(int KB, int max_n) {
    const int W = 3 * KB + 1;
    const int lds_rows = (int)((60 * 1024) / (sizeof(double) * (W + 1)));
    return max_n < lds_rows ? max_n : lds_rows;
}

void launch_eps_band(const MarkerDev& m, const GridScalars& g, const int* sys_first, const int* sys_n,
                     const int* host_n, int nsys, int KB, double* band, const long long* band_off, int* outside,
                     double* ywork, bool cr, hipStream_t s) {
    if (nsys <= 0 || KB != 8) return;
    int max_n = 0;
    for (int k = 0; k < nsys; k++) max_n = max_n > host_n[k] ? max_n : host_n[k];
    const int W = 3 * KB + 1;
    const long long ne = (long long)max_n * W;
    const int be = (int)((ne + 255) / 256);
    const bool stage_far = max_n <= kFarLdsMarkers;
    const size_t far_lds = stage_far ? sizeof(double) * 2 * (size_t)max_n : 0;
    k_eps_band_prep<<<dim3((unsigned)(be + 1), nsys), 256, far_lds, s>>>(m, g, sys_first, sys_n, KB, band, band_off,
                                                                          outside, be, stage_far ? 1 : 0);
    if (cr) {
        const size_t lds = eps_cr_lds(max_n);
        static const bool once = (allow_lds((const void*)k_eps_band_cr, eps_cr_lds(kCrMaxRows)), true);
        (void)once;
        k_eps_band_cr<<<nsys, 256, lds, s>>>(sys_first, sys_n, outside, band, band_off, ywork, m.eps, cr_max_bw());
        if (max_n <= kCrMaxRows) return;
    }
    // band, then right-hand side, in LDS for systems of up to lds_rows rows
    const int rows = eps_lds_rows(KB, max_n);
    k_eps_band_solve<8><<<nsys, 64, sizeof(double) * (size_t)rows * (W + 1), s>>>(
        sys_first, sys_n, outside, band, band_off, ywork, m.eps, rows, cr ? kCrMaxRows : 0);
}

void launch_find_support(const MarkerDev& m, const GridScalars& g, int flexible_only, hipStream_t s) {
    if (m.n <= 0) return;
    k_find_support<<<(m.n + 127) / 128, 128, 0, s>>>(m, g, flexible_only);
}

void launch_support_ds(const MarkerDev& m, const GridScalars& g, int nbodies, int flexible_only, hipStream_t s) {
    if (m.n <= 0 || nbodies <= 0) return;
    k_support_ds<<<nbodies, 256, sizeof(double) * 2 * kDsLdsMarkers, s>>>(m, g, flexible_only);
}

void launch_compute_ds(const MarkerDev& m, const GridScalars& g, int nbodies, int flexible_only, hipStream_t s) {
    if (m.n <= 0 || nbodies <= 0) return;
    k_compute_ds<<<nbodies, 256, sizeof(double) * 2 * kDsLdsMarkers, s>>>(m, g, flexible_only);
}

void launch_epsilon_matrix(const MarkerDev& m, const GridScalars& g, int first, int n, double* A, hipStream_t s) {
    const long long t = (long long)n * n;
    if (t <= 0) return;
    k_epsilon_matrix<<<(unsigned)((t + 255) / 256), 256, 0, s>>>(m, g, first, n, A);
}

template <typename T>
void launch_interpolate(const T* f, const MarkerDev& m, const GridScalars& g, hipStream_t s) {
    if (m.n <= 0) return;
    k_interpolate<T><<<(m.n + kInterpMarkers - 1) / kInterpMarkers, kInterpThreads, 0, s>>>(f, m, g);
}
template void launch_interpolate<double>(const double*, const MarkerDev&, const GridScalars&, hipStream_t);
template void launch_interpolate<float>(const float*, const MarkerDev&, const GridScalars&, hipStream_t);

void launch_spread(const MarkerDev& m, const GridScalars& g, double* fibm, SpreadScratch& sc, bool dense_clear,
                   hipStream_t s) {
    const int np = m.n * kSupp;
    const int cap = sc.cap;
    if (np > 0 && np <= kSpreadOneMaxPairs && cap >= np) {
        if (dense_clear) launch_fill_f64(fibm, 0.0, g.S * 2, s);
        k_spread_one<<<1, kSpreadOneBlock, 0, s>>>(m, sc, fibm, g.S, dense_clear ? 0 : 1);
        return;
    }
    if (dense_clear) {
        launch_fill_f64(fibm, 0.0, g.S * 2, s);
    } else if (cap > 0) {
        k_spread_clear<<<(cap + 255) / 256, 256, 0, s>>>(fibm, g.S, sc.uniq, sc.uniq_n);
    }
    k_spread_reset<<<1, 1, 0, s>>>(sc.uniq_n);
    if (np <= 0) return;
    k_spread_count<<<(np + 255) / 256, 256, 0, s>>>(m, sc);
    k_spread_scatter<<<(np + 255) / 256, 256, 0, s>>>(m, sc);
    k_spread_reduce<<<(cap + 127) / 128, 128, 0, s>>>(m, sc, fibm, g.S);
}

bool launch_interp_spread_small(const void* f, bool f64, const MarkerDev& m, const GridScalars& g, double* fibm,
                                SpreadScratch& sc, bool dense_clear, hipStream_t s) {
    const int np = m.n * kSupp;
    if (np <= 0 || np > kSpreadOneMaxPairs || sc.cap < np) return false;
    if (dense_clear) launch_fill_f64(fibm, 0.0, g.S * 2, s);
    if (f64)
        k_interp_spread_one<double><<<1, kSpreadOneBlock, 0, s>>>(static_cast<const double*>(f), m, g, sc, fibm,
                                                                  dense_clear ? 0 : 1);
    else
        k_interp_spread_one<float><<<1, kSpreadOneBlock, 0, s>>>(static_cast<const float*>(f), m, g, sc, fibm,
                                                                 dense_clear ? 0 : 1);
    return true;
}

// ---------------------------------------------------------------- batched launchers
void launch_find_support_batch(const CaseDev* cases, const int* list, int n, int max_markers, int flexible_only,
                               hipStream_t s) {
    if (n <= 0 || max_markers <= 0) return;
    k_find_support_batch<<<dim3((unsigned)((max_markers + 127) / 128), (unsigned)n), 128, 0, s>>>(cases, list,
                                                                                                 flexible_only);
}

void launch_compute_ds_batch(const CaseDev* cases, const int* list, int n, int max_bodies, int flexible_only,
                             hipStream_t s) {
    if (n <= 0 || max_bodies <= 0) return;
    k_compute_ds_batch<<<dim3((unsigned)max_bodies, (unsigned)n), 256, sizeof(double) * 2 * kDsLdsMarkers, s>>>(
        cases, list, flexible_only);
}

void launch_eps_band_batch(const CaseDev* cases, const CaseSys* sys, int nsys, int max_n, int* outside_out,
                           bool cr, hipStream_t s) {
    constexpr int KB = 8, W = 3 * KB + 1;
    if (nsys <= 0 || max_n <= 0) return;
    const unsigned nb = (unsigned)((nsys + 255) / 256);
    k_eps_outside_clear_batch<<<nb, 256, 0, s>>>(cases, sys, nsys);
    const long long ne = (long long)max_n * W, np = (long long)max_n * max_n;
    k_eps_band_entries_batch<<<dim3((unsigned)((ne + 255) / 256), (unsigned)nsys), 256, 0, s>>>(cases, sys, KB);
    k_eps_far_check_batch<<<dim3((unsigned)((np + 255) / 256), (unsigned)nsys), 256, 0, s>>>(cases, sys, KB);
    if (cr) {
        static const bool once = (allow_lds((const void*)k_eps_band_cr_batch, eps_cr_lds(kCrMaxRows)), true);
        (void)once;
        k_eps_band_cr_batch<<<(unsigned)nsys, 256, eps_cr_lds(max_n), s>>>(cases, sys, cr_max_bw());
    }
    if (!cr || max_n > kCrMaxRows) {
        const int rows = eps_lds_rows(KB, max_n);
        k_eps_band_solve_batch<KB><<<(unsigned)nsys, 64, sizeof(double) * (size_t)rows * (W + 1), s>>>(
            cases, sys, rows, cr ? kCrMaxRows : 0);
    }
    k_eps_outside_gather_batch<<<nb, 256, 0, s>>>(cases, sys, nsys, outside_out);
}

template <typename T>
void launch_interpolate_batch(const CaseDev* cases, const int* list, int n, int max_markers, hipStream_t s) {
    if (n <= 0 || max_markers <= 0) return;
    k_interpolate_batch<T><<<dim3((unsigned)((max_markers + kInterpMarkers - 1) / kInterpMarkers), (unsigned)n),
                             kInterpThreads, 0, s>>>(cases, list);
}
template void launch_interpolate_batch<double>(const CaseDev*, const int*, int, int, hipStream_t);
template void launch_interpolate_batch<float>(const CaseDev*, const int*, int, int, hipStream_t);

void launch_spread_batch(const CaseDev* cases, const int* list, int n, int max_markers, hipStream_t s) {
    if (n <= 0) return;
    const int np = max_markers * kSupp;
    const unsigned nc = (unsigned)((np + 255) / 256);
    if (np > 0) k_spread_clear_batch<<<dim3(nc, (unsigned)n), 256, 0, s>>>(cases, list);
    k_spread_reset_batch<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(cases, list, n);
    if (np <= 0) return;
    k_spread_count_batch<<<dim3(nc, (unsigned)n), 256, 0, s>>>(cases, list);
    k_spread_scatter_batch<<<dim3(nc, (unsigned)n), 256, 0, s>>>(cases, list);
    k_spread_reduce_batch<<<dim3((unsigned)((np + 127) / 128), (unsigned)n), 128, 0, s>>>(cases, list);
}

void launch_markers_in_batch(const CaseDev* cases, const int* list, int n, int max_markers, const double* stage,
                             hipStream_t s) {
    if (n <= 0 || max_markers <= 0) return;
    k_markers_in_batch<<<dim3((unsigned)((max_markers + 127) / 128), (unsigned)n), 128, 0, s>>>(cases, list, stage);
}

void launch_markers_out_batch(const CaseDev* cases, const int* list, int n, int max_markers, double* stage,
                              hipStream_t s) {
    if (n <= 0 || max_markers <= 0) return;
    k_markers_out_batch<<<dim3((unsigned)((max_markers + 127) / 128), (unsigned)n), 128, 0, s>>>(cases, list, stage);
}

// Fused recompute (one workgroup per flexible body); false when a body has
// more markers than the LDS layout takes (the caller uses the separate
// kernels then).
static size_t fused_recompute_lds(int max_n, bool cr) {
    return sizeof(double) * ((size_t)max_n * (3 * 8 + 5) + (cr ? cr_lds_doubles(kCrMaxBw, max_n) : 0));
}

bool launch_recompute_fused_batch(const CaseDev* cases, const CaseSys* sys, int nsys, int max_n, int* outside_out,
                                  bool cr, hipStream_t s) {
    if (max_n > kFusedMaxMarkers) return false;
    if (nsys <= 0) return true;
    static const int nt = [] {
        const char* e = std::getenv("LIFE_SERVE_FUSED_THREADS");
        const int v = e ? std::atoi(e) : kFusedThreads;
        return v == 0 || v == 64 || v == 128 || v == 256 ? v : kFusedThreads;
    }();
    if (nt == 0) return false;  // the separate batched kernels
    static const bool once =
        (allow_lds((const void*)k_recompute_body_batch<8>, fused_recompute_lds(kFusedMaxMarkers, true)), true);
    (void)once;
    // cr_bw -1: the pivoted elimination for every system (the plain method)
    k_recompute_body_batch<8><<<(unsigned)nsys, (unsigned)nt, fused_recompute_lds(max_n, cr), s>>>(
        cases, sys, cr ? cr_max_bw() : -1);
    k_eps_outside_gather_batch<<<(unsigned)((nsys + 255) / 256), 256, 0, s>>>(cases, sys, nsys, outside_out);
    return true;
}

}  // namespace life
