// Immersed-boundary kernels: support search, marker spacing, epsilon
// matrix, interpolation + direct forcing, and the ordered spread.
//
// All arithmetic is FP64 with the reference's operation order
// (IBMNode.cpp, Objects.cpp:235-321), for both population storages.
// Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "d2q9_f64.h"
#include "launch.h"

namespace life {

// IBMNodeClass::findSupport (IBMNode.cpp:139-179): lattice sites within 1.5
// spacings of the marker in both directions, scanned i-major over the 5x5
// block around the nearest site, kept in scan order (at most 9).
__global__ void k_find_support(MarkerDev m, GridScalars g, int flexible_only) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n) return;
    if (flexible_only && m.body_flex[m.body[k]] != 0) return;
    const double Dx = g.Dx;
    const double x = m.pos[2 * k], y = m.pos[2 * k + 1];
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

// IBMNodeClass::computeDs (IBMNode.cpp:182-204): distance (lattice units)
// to the nearest other marker of the same body, starting from 10.  One
// workgroup per body; the body's positions are staged in LDS when they fit.
constexpr int kDsLdsMarkers = 2048;
__global__ void k_compute_ds(MarkerDev m, GridScalars g, int flexible_only) {
    extern __shared__ double spos[];
    const int b = blockIdx.x;
    if (flexible_only && m.body_flex[b] != 0) return;
    const int first = m.body_first[b], cnt = m.body_n[b];
    const bool lds = cnt <= kDsLdsMarkers;
    if (lds)
        for (int t = threadIdx.x; t < 2 * cnt; t += blockDim.x) spos[t] = m.pos[2 * first + t];
    __syncthreads();
    const double* P = lds ? spos : m.pos + 2 * first;
    // sqrt(.)/Dx is monotone, so the minimum of sqrt(d.d)/Dx over the other
    // markers is sqrt(min d.d)/Dx: one square root per marker, same value
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const double x = P[2 * k], y = P[2 * k + 1];
        double d2min = INFINITY;
        for (int q = 0; q < cnt; q++) {
            if (q == k) continue;
            const double dx = x - P[2 * q], dy = y - P[2 * q + 1];
            const double d2 = 0.0 + dx * dx + dy * dy;
            if (d2 < d2min) d2min = d2;
        }
        const double mag = sqrt(d2min) / g.Dx;
        m.ds[first + k] = mag < 10.0 ? mag : 10.0;
    }
}

// A_ij of ObjectsClass::computeEpsilon (Objects.cpp:271-300) for the system
// of markers [first, first+n): sum over the supports s of marker i of
// delta_i(s) * delta(x_j - s), times ds_j.
__global__ void k_epsilon_matrix(MarkerDev m, GridScalars g, int first, int n, double* __restrict__ A) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * n) return;
    const int r = (int)(t / n), c = (int)(t - (long long)r * n);
    const int mi = first + r, mj = first + c;
    const double Dx = g.Dx;
    const double xj = m.pos[2 * mj] / Dx, yj = m.pos[2 * mj + 1] / Dx;
    double a = 0.0;
    const int cnt = m.supp_cnt[mi];
    for (int s = 0; s < cnt; s++) {
        const double dx = fabs(xj - m.supp_ij[(mi * kSupp + s) * 2]);
        const double dy = fabs(yj - m.supp_ij[(mi * kSupp + s) * 2 + 1]);
        a += m.supp_delta[mi * kSupp + s] * (dirac64(dx) * dirac64(dy));
    }
    A[t] = a * m.ds[mj];
}

// ---------------------------------------------------------------- epsilon, banded
// Batched form of computeEpsilon for systems whose matrix is banded (every
// filament: A_ij = 0 unless markers i and j are within 3 lattice spacings).
// Band storage: row r holds columns r-KB .. r+2KB (width 3KB+1, the upper
// part leaves room for the fill-in of partial pivoting).  A system with a
// coupling pair outside the band is flagged and solved on the host with the
// dense path instead.

// One thread per band entry; grid.y = system.
__global__ void k_eps_band_entries(MarkerDev m, GridScalars g, const int* __restrict__ sys_first,
                                   const int* __restrict__ sys_n, int KB, double* __restrict__ band,
                                   const long long* __restrict__ band_off) {
    const int sys = blockIdx.y;
    const int first = sys_first[sys], n = sys_n[sys];
    const int W = 3 * KB + 1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * W) return;
    const int r = (int)(t / W), q = (int)(t - (long long)r * W);
    const int c = r - KB + q;
    double a = 0.0;
    if (q <= 2 * KB && c >= 0 && c < n) {
        const int mi = first + r, mj = first + c;
        const double Dx = g.Dx;
        const double xj = m.pos[2 * mj] / Dx, yj = m.pos[2 * mj + 1] / Dx;
        const int cnt = m.supp_cnt[mi];
        for (int s = 0; s < cnt; s++) {
            const double dx = fabs(xj - m.supp_ij[(mi * kSupp + s) * 2]);
            const double dy = fabs(yj - m.supp_ij[(mi * kSupp + s) * 2 + 1]);
            a += m.supp_delta[mi * kSupp + s] * (dirac64(dx) * dirac64(dy));
        }
        a = a * m.ds[mj];
    }
    band[band_off[sys] + t] = a;
}

// Pairs outside the band that could couple (support of i within 1.5 of
// x_i, delta of j within 1.5 of the support: |x_i - x_j| < 3 spacings).
// One thread per pair r < c - KB (the test is symmetric); grid.y = system.
__global__ void k_eps_far_check(MarkerDev m, GridScalars g, const int* __restrict__ sys_first,
                                const int* __restrict__ sys_n, int KB, int* __restrict__ outside) {
    const int sys = blockIdx.y;
    const int first = sys_first[sys], n = sys_n[sys];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * n) return;
    const int r = (int)(t / n), c = (int)(t - (long long)r * n);
    if (c <= r + KB) return;
    const double lim = 3.01 * g.Dx;
    const int mi = first + r, mj = first + c;
    if (fabs(m.pos[2 * mi] - m.pos[2 * mj]) < lim && fabs(m.pos[2 * mi + 1] - m.pos[2 * mj + 1]) < lim)
        outside[sys] = 1;
}

// Canonical elimination (partial pivoting, first maximum) of one banded
// system per wavefront, right-hand side of ones, writing epsilon.
//
// With lower and upper bandwidth KB, every non-zero that step k of the
// dense elimination reads or writes lies in rows k..k+KB and columns
// k..k+2KB (row swaps widen the upper band by KB).  Lane c of the wavefront
// keeps column k+c of that window in registers (w[r] = a(k+r, k+c)); each
// step broadcasts column k, picks the pivot, swaps, forms the multipliers,
// updates the other columns (each a_ij - l_i a_kj exactly once from the same
// operands as the dense algorithm), emits row k of U, and slides the window
// one row and one column.  Back substitution then runs on one lane in
// increasing column order.  The skipped dense operations all involve an
// exact zero, so the values equal the dense elimination's.  The band is
// staged in LDS (rows of U replace rows of the band as they are emitted)
// when it fits in lds_rows rows.
template <int KB>
__global__ __launch_bounds__(64) void k_eps_band_solve(const int* __restrict__ sys_first, const int* __restrict__ sys_n,
                                                       const int* __restrict__ outside, double* __restrict__ band,
                                                       const long long* __restrict__ band_off,
                                                       double* __restrict__ ywork, double* __restrict__ eps,
                                                       int lds_rows) {
    constexpr int R = KB + 1, C = 2 * KB + 1, W = 3 * KB + 1;
    extern __shared__ double sm[];
    const int sys = blockIdx.x;
    if (outside[sys]) return;
    const int first = sys_first[sys], n = sys_n[sys];
    const bool in_lds = n <= lds_rows;
    double* G = band + band_off[sys];
    double* B = in_lds ? sm : G;
    double* y = in_lds ? sm + (long long)n * W : ywork + first;
    const int c = threadIdx.x;
    if (in_lds) {
        for (int t = c; t < n * W; t += 64) B[t] = G[t];
        __syncthreads();
    }
    auto A0 = [&](int r, int col) -> double {
        if (c >= C || r >= n || col < 0 || col >= n) return 0.0;
        const int q = col - r + KB;
        return (q >= 0 && q < W) ? B[r * W + q] : 0.0;
    };
    double w[R], yw[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        w[r] = A0(r, c);
        yw[r] = r < n ? 1.0 : 0.0;
    }
    for (int k = 0; k < n; k++) {
        const double incoming = A0(k + R, k + 1 + c);  // row entering the window next
        const int rmax = min(KB, n - 1 - k);
        double col[R];
#pragma unroll
        for (int r = 0; r < R; r++) col[r] = __shfl(w[r], 0);
        int p = 0;
        double best = fabs(col[0]);
#pragma unroll
        for (int r = 1; r < R; r++)
            if (r <= rmax) {
                const double a = fabs(col[r]);
                if (a > best) {
                    best = a;
                    p = r;
                }
            }
        if (best != 0.0) {
            if (p != 0) {
                const double tw = w[0], tc = col[0], ty = yw[0];
#pragma unroll
                for (int r = 1; r < R; r++)
                    if (r == p) {
                        w[0] = w[r];
                        w[r] = tw;
                        col[0] = col[r];
                        col[r] = tc;
                        yw[0] = yw[r];
                        yw[r] = ty;
                    }
            }
            const double piv = col[0];
#pragma unroll
            for (int r = 1; r < R; r++)
                if (r <= rmax) {
                    const double l = col[r] / piv;
                    if (c == 0) w[r] = l;
                    else w[r] = w[r] - l * w[0];
                    yw[r] = yw[r] - l * yw[0];
                }
        }
        // row k of U (in place of row k of the band) and the final y_k
        if (c < C && k + c < n) B[k * W + (c + KB)] = w[0];
        if (c == 0) y[k] = yw[0];
        // slide: drop row k and column k, take in row k+R
#pragma unroll
        for (int r = 0; r < R - 1; r++) {
            w[r] = __shfl_down(w[r + 1], 1);
            yw[r] = yw[r + 1];
        }
        w[R - 1] = incoming;
        yw[R - 1] = k + R < n ? 1.0 : 0.0;
    }
    __syncthreads();
    if (c == 0) {
        double xw[2 * KB];
#pragma unroll
        for (int q = 0; q < 2 * KB; q++) xw[q] = 0.0;
        for (int i = n - 1; i >= 0; i--) {
            const double* Ui = B + i * W + KB;
            double s = y[i];
#pragma unroll
            for (int q = 0; q < 2 * KB; q++)
                if (i + 1 + q < n) s = s - Ui[1 + q] * xw[q];
            const double x = s / Ui[0];
            eps[first + i] = x;
#pragma unroll
            for (int q = 2 * KB - 1; q > 0; q--) xw[q] = xw[q - 1];
            xw[0] = x;
        }
    }
}

// ObjectsClass::ibmKernelInterp body (IBMNode.cpp:26-58): rho and rho*u at
// the supports (GridClass::macroscopic values: force_xy only) interpolated
// with the support deltas, then F = 2 (Dt/Dx rho_I U - (rho u)_I).
template <typename T>
__global__ void k_interpolate(const T* __restrict__ f, MarkerDev m, GridScalars g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m.n) return;
    double iR = 0.0, mx = 0.0, my = 0.0;
    const int cnt = m.supp_cnt[k];
    for (int s = 0; s < cnt; s++) {
        const long long id = m.supp_site[k * kSupp + s];
        double fs[kNV], rho, ux, uy;
        load_site(f, g.S, id, fs);
        moments64(fs, g.Fx, g.Fy, rho, ux, uy);
        const double d = m.supp_delta[k * kSupp + s];
        iR += rho * d * 1.0 * 1.0;
        mx += rho * ux * d * 1.0 * 1.0;
        my += rho * uy * d * 1.0 * 1.0;
    }
    const double vs = g.velScale;
    m.force[2 * k] = 2.0 * (vs * iR * m.vel[2 * k] - mx);
    m.force[2 * k + 1] = 2.0 * (vs * iR * m.vel[2 * k + 1] - my);
}

// ---------------------------------------------------------------- spread
// The reference adds marker contributions to each site in marker order
// (ORDERED, IBMNode.cpp:67-83).  Here every (marker, support) pair takes a
// slot at its site, each touched site gets a compact index, the pairs are
// scattered into per-site lists, and one thread per site sorts its list by
// marker and sums it from zero: the same additions in the same order.

__global__ void k_spread_clear(double* __restrict__ fibm, long long S, const int* __restrict__ uniq,
                               const int* __restrict__ uniq_n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *uniq_n) return;
    const long long id = uniq[k];
    fibm[id] = 0.0;
    fibm[S + id] = 0.0;
}

__global__ void k_spread_reset(int* uniq_n) { *uniq_n = 0; }

__global__ void k_spread_count(MarkerDev m, SpreadScratch sc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
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

__global__ void k_spread_scatter(MarkerDev m, SpreadScratch sc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
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

__global__ void k_spread_reduce(MarkerDev m, SpreadScratch sc, double* __restrict__ fibm, long long S) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
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

void launch_eps_band(const MarkerDev& m, const GridScalars& g, const int* sys_first, const int* sys_n,
                     const int* host_n, int nsys, int KB, double* band, const long long* band_off, int* outside,
                     double* ywork, hipStream_t s) {
    if (nsys <= 0 || KB != 8) return;
    int max_n = 0;
    for (int k = 0; k < nsys; k++) max_n = max_n > host_n[k] ? max_n : host_n[k];
    const int W = 3 * KB + 1;
    const long long ne = (long long)max_n * W, np = (long long)max_n * max_n;
    k_eps_band_entries<<<dim3((unsigned)((ne + 255) / 256), nsys), 256, 0, s>>>(m, g, sys_first, sys_n, KB, band,
                                                                                 band_off);
    k_eps_far_check<<<dim3((unsigned)((np + 255) / 256), nsys), 256, 0, s>>>(m, g, sys_first, sys_n, KB, outside);
    // band, then right-hand side, in LDS for systems of up to lds_rows rows
    const int lds_rows = (int)((60 * 1024) / (sizeof(double) * (W + 1)));
    const int rows = max_n < lds_rows ? max_n : lds_rows;
    k_eps_band_solve<8><<<nsys, 64, sizeof(double) * (size_t)rows * (W + 1), s>>>(sys_first, sys_n, outside, band,
                                                                                   band_off, ywork, m.eps, rows);
}

void launch_find_support(const MarkerDev& m, const GridScalars& g, int flexible_only, hipStream_t s) {
    if (m.n <= 0) return;
    k_find_support<<<(m.n + 127) / 128, 128, 0, s>>>(m, g, flexible_only);
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
    k_interpolate<T><<<(m.n + 127) / 128, 128, 0, s>>>(f, m, g);
}
template void launch_interpolate<double>(const double*, const MarkerDev&, const GridScalars&, hipStream_t);
template void launch_interpolate<float>(const float*, const MarkerDev&, const GridScalars&, hipStream_t);

void launch_spread(const MarkerDev& m, const GridScalars& g, double* fibm, SpreadScratch& sc, bool dense_clear,
                   hipStream_t s) {
    const int np = m.n * kSupp;
    const int cap = sc.cap;
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

}  // namespace life
